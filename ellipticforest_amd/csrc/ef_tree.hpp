// ef_tree.hpp -- descriptors and launchers of the tree kernels (merges of the build stage,
// upward and downward sweeps), shared by the host planner (ef_api.hip) and ef_tree.hip.
//
// Every kernel here is written for gfx950 with these invariants, which the launchers check
// on the host before each launch (launch_check / check_range in ef_host.hpp):
//  * no scratch (private arrays are indexed only by compile-time constants), static LDS
//    below 64 KiB, no dynamic LDS;
//  * every address a launch can form lies inside the device allocation its descriptor
//    names, over the whole extent the kernel indexes (hipMemGetAddressRange);
//  * operands read as 16-byte vectors are 16-byte aligned with an even leading dimension.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ef_device.hpp"

namespace efk {

// ---------------------------------------------------------------------------------------
// Interface and boundary tables of a merge (HPSAlgorithm.hpp:752-993).
// Children c = 0..3 are alpha (SW), beta (SE), gamma (NW), omega (NE).  Interfaces
// r = 0..3 are [alpha|gamma, beta|omega, alpha|beta, gamma|omega]; each has a first child
// (its DtN rows enter X and S_RHS with -) and a second (+).  Parent boundary blocks
// q = 0..7 in WESN order are (child, side) = aW, gW, bE, oE, aS, bS, gN, oN.
// ---------------------------------------------------------------------------------------
__host__ __device__ inline int if_c1(int r) { return r == 0 ? 0 : r == 1 ? 1 : r == 2 ? 0 : 2; }
__host__ __device__ inline int if_c2(int r) { return r == 0 ? 2 : r == 1 ? 3 : r == 2 ? 1 : 3; }
__host__ __device__ inline int if_s1(int r) { return r < 2 ? (int)N : (int)E; }
__host__ __device__ inline int if_s2(int r) { return r < 2 ? (int)S : (int)W; }
__host__ __device__ inline int q_child(int q) { return q == 0 ? 0 : q == 1 ? 2 : q == 2 ? 1 : q == 3 ? 3 : q - 4; }
__host__ __device__ inline int q_side(int q) { return q >> 1; }
// side of child c on interface r, or -1
__host__ __device__ inline int side_on_if(int c, int r) {
    if (if_c1(r) == c) return if_s1(r);
    if (if_c2(r) == c) return if_s2(r);
    return -1;
}
// the two interfaces child c lies on (CH_IF), slot 0 and 1
__host__ __device__ inline int child_if(int c, int slot) {
    // alpha: {0, 2}, beta: {1, 2}, gamma: {0, 3}, omega: {1, 3}
    return slot == 0 ? (c & 1) : 2 + (c >> 1);
}
// children's g (applyS_, HPSAlgorithm.hpp:1194-1230): alpha = [aW, ab, aS, ag], beta = [ab, bE,
// bS, bo], gamma = [gW, go, ag, gN], omega = [go, oE, bo, oN].  A child's side lies either on
// tau's boundary (block q of tau's g: half q & 1 of tau's side q >> 1) or on an interface
// (block r of the interface potential).
__host__ __device__ inline int child_src_q(int c, int s) {
    // side s of child c on tau's boundary: the WESN block q of (c, s), or -1
    const bool west = (c & 1) == 0, south = (c >> 1) == 0;
    if (s == W && west) return c == 0 ? 0 : 1;
    if (s == E && !west) return c == 1 ? 2 : 3;
    if (s == S && south) return c == 0 ? 4 : 5;
    if (s == N && !south) return c == 2 ? 6 : 7;
    return -1;
}
__host__ __device__ inline int child_src_r(int c, int s) {
    for (int r = 0; r < 4; r++)
        if (side_on_if(c, r) == s) return r;
    return 0;
}

// ---------------------------------------------------------------------------------------
// Batched two-segment GEMM on v_mfma_f64_16x16x4_f64, row-major operands:
//   D = alpha0 A0 B0 + alpha1 A1 B1 + beta C        (segment 1 absent when K1 == 0,
//                                                    C absent when null)
// A0 is M x K0 (lda0), B0 is K0 x N (ldb0), ...  C may alias D (read, then written, by the
// same lane).  One wave computes a (16 RT) x (16 TN) block; the blocks of all problems of a
// launch form one flat list (problem p owns waves start[p] .. start[p+1]-1).
// Split-K (kc > 0): the concatenated k range of the two segments is cut into nks chunks of kc (a
// multiple of 16); a launch with such a problem runs on tree_gemm_split_kernel, one workgroup per
// 16 x 16 tile and one wave per chunk, the chunks' partials summed in chunk order through LDS
// (fixed order, so a column's result does not depend on the other columns of the batch).
// ---------------------------------------------------------------------------------------
struct TreeGemm {
    const double* A0;
    const double* B0;
    const double* A1;
    const double* B1;
    const double* C;
    double* D;
    int lda0, ldb0, K0, lda1, ldb1, K1;  // element (i, j) of B_s at B_s[i * ldb_s + j * bcs]
    int ldc, ldd, M, N;                  // element (i, j) of C at C[i * ldc + j * ccs], of D likewise
    int bcs, ccs, dcs, kc;               // column strides (0 means 1); split-K chunk length (0: none)
    double alpha0, alpha1, beta;
    int nks, pad;                        // split-K chunk count (set by plan_gemm)
    // Block-gather mode (dn > 0, segment 0 only, ldb0 = 1): row kk of B0 is B0[d2[r] + i] -
    // B0[d1[r] + i] with r = kk / dn, i = kk % dn (dn a multiple of 4): the four interface
    // blocks of mergeW_'s hDiff read straight from the children's h, each block the difference
    // of two children.  dsub = 0: row kk is B0[d2[r] + i] alone (d1 unused): the four sides of
    // a node's g read where they lie.
    int dn, dsub;
    long long d2[4], d1[4];
};
static_assert(sizeof(TreeGemm) == 208, "TreeGemm layout");

// F32: operands rounded to FP32, v_mfma_f32_16x16x4_f32 with FP32 accumulation, FP64 output
// (+ beta C in FP64); the mixed-precision build's S and T products.
template <int RT, int TN, bool F32>
__global__ void tree_gemm_kernel(const TreeGemm* probs, const int* start, const int* wprob, int nwaves);
__global__ void tree_gemm_split_kernel(const TreeGemm* probs, const int* start, const int* gprob, int ngroups);

// ---------------------------------------------------------------------------------------
// In-place blocked Gauss-Jordan inversion with partial pivoting of a batch of B equally
// sized m x m matrices (m <= GJ_MAXM), GJ_NB columns per step, ping-pong buffers Wa -> Wb.
// Step k (columns k .. k+nb):
//   panel        LU with partial pivoting of each chunk of up to GJ_CH panel rows (one
//                workgroup per chunk, rows in registers): nb candidate pivot rows per chunk;
//                with several chunks a second pass over all candidates picks the pivots.  The
//                pass that picks the pivots also forms P^-1 of the pivot block, the row
//                interchanges (ipiv, LAPACK convention) and the position -> row map src
//   mult         multipliers M (m x nb): P^-1 - I on the pivot positions, -A(i, panel) P^-1
//                elsewhere
//   update       Wb(i, :) = Wa(src(i), :) + M(i, :) Wa(pivot rows, :) outside the panel,
//                Wb(i, panel) = M(i, :) + [i is pivot position]
// With one chunk (m - k <= GJ_CH) the pivots are exactly those of partial pivoting; with
// more, the candidates of each chunk compete in panel_final (tournament pivoting).
// After the last step: perm computes the column order undoing the interchanges (last to
// first) and gather writes the inverse X^-1(:, j) = W(:, idx(j)).
// ---------------------------------------------------------------------------------------
constexpr int GJ_NB = 32;
constexpr int GJ_CH = 512;    // panel rows per chunk: two lanes of a 1024-thread workgroup per row
constexpr int GJ_MAXCH = 8;
constexpr int GJ_MAXM = GJ_CH * GJ_MAXCH;

struct GJArgs {
    const double* Wa;   // B x m x m, current
    double* Wb;         // B x m x m, next
    int* cand;          // B x GJ_MAXCH x GJ_NB candidate positions of each chunk (-1: none)
    int* cfin;          // B x GJ_NB pivot positions of the step, in pivot order (-1: none)
    int* src;           // B x m: position -> row of Wa; after perm: column map of the inverse
    int* ipiv;          // B x m: interchange of position c with ipiv[c] (c <= ipiv[c])
    int* prow;          // B x GJ_NB: pivot rows (positions in Wa) of this step
    int* info;          // B: first zero pivot (column + 1) or 0, kept across steps
    double* Pinv;       // B x GJ_NB x GJ_NB
    double* Mm;         // B x m x GJ_NB
    double* Xi;         // gather: B x m x m output (stride m*m)
    int m, B, k, nch;
};

template <int STAGE, int WAVES, int RPL, bool FIN>
__global__ void gj_panel_kernel(GJArgs a);
__global__ void gj_mult_kernel(GJArgs a);
__global__ void gj_update_kernel(GJArgs a);
__global__ void gj_perm_kernel(GJArgs a);
__global__ void gj_gather_kernel(GJArgs a);

// ---------------------------------------------------------------------------------------
// Merge assembly (createMatrixBlocks_ + mergeX_ + the block-diagonal part of mergeT_,
// HPSAlgorithm.hpp:799-968, with the WESN order of reorderOperators_, :979-993): from the
// four children's DtN maps (each 4n x 4n, coarsened), X (4n x 4n) into the inversion work
// buffer and T (8n x 8n, WESN rows and columns) = diag blocks T_c(tau side, tau side).
// ---------------------------------------------------------------------------------------
struct AsmDesc {
    const double* Tc[4];
    double* X;
    double* T;
    int n, pad;
};
static_assert(sizeof(AsmDesc) == 56, "AsmDesc layout");
__global__ void tree_assemble_kernel(const AsmDesc* d);

// coarsen_ (HPSAlgorithm.hpp:707-741): dst (4 m/2 square) = blkdiag4(L21) src (4m square) blkdiag4(L12)
__global__ void tree_coarsen_kernel(const CoarsenDesc* d);

// ---------------------------------------------------------------------------------------
// Sweeps over the vector arena V: vrows rows by k right-hand sides, right-hand side major
// (element (row, rhs) at rhs * vrows + row).
// ---------------------------------------------------------------------------------------
// upwards4to1 gather (coarsenUpwards_ + hDiff of mergeW_ + boundary part of mergeH_):
//   hd(r) = h_c2(side) - h_c1(side) on the four interfaces, h_tau(q) = h_child(q)(side q)
__global__ void tree_up_gather_kernel(const UpDesc* d, double* V, int k, long long vrows);
// uncoarsen_ step: g <- blkdiag4(L12(nfine)) g
__global__ void tree_uncoarsen_kernel(const VecStepDesc* d, double* V, int k, long long vrows);
// applyS_ scatter: children's g from tau's g (WESN) and the interface potential
__global__ void tree_side_copy_kernel(const SideCopy* d, double* V, int k, long long vrows);
__global__ void tree_side_interp_kernel(const SideCopy* d, double* V, int k, long long vrows);
// root Dirichlet data g = r / a (solveStage with b = 0)
__global__ void tree_root_dirichlet_kernel(const double* r, const double* a, double* g, int m, int k, long long vrows);
// root Robin system A = diag(a) + diag(b) T, rhs = r - b h
__global__ void tree_root_robin_kernel(const double* T, const double* a, const double* b, const double* r,
                                       const double* h, double* A, double* rhs, int m, int k, long long vrows);

// ---------------------------------------------------------------------------------------
// Host side (ef_tree.hip)
// ---------------------------------------------------------------------------------------
// A GEMM launch: problems and the flat wave list, on the device.
struct GemmLaunch {
    TreeGemm* probs = nullptr;
    int* start = nullptr;
    int* wprob = nullptr;         // problem of each wave (of each workgroup for split-K launches)
    int nprob = 0, nwaves = 0, rt = 2, tn = 2;
    int kw = 0;                   // > 1: split-K launch, waves per 16 x 16 tile
    bool f32 = false;  // FP32 matrix cores (mixed precision)
};
// Checks every problem (extents inside their allocations, alignment), sets the split-K chunks of
// a split launch, stages problems and wave offsets in `store` (device memory the caller keeps
// alive; valid on the device after store.flush) and returns the launch.  tn = 1 (16 x 16 blocks) suits N <= 16, tn = 2 (32 x 32) the rest.
struct DevStore;
// rt = 0 picks the row tile from tn (16 rows for tn = 1 or problems of <= 16 rows, else 32);
// rt = 1 or 2 forces it (2 only with tn = 2).  The tile shape never changes a result bit: every
// output element sums its k range in the same order.
GemmLaunch plan_gemm(const std::vector<TreeGemm>& probs, int tn, DevStore& store, hipStream_t s, bool f32 = false,
                     int rt = 0);
void run_gemm(const GemmLaunch& g, hipStream_t s);

// Inversion of `B` matrices of order m held in wa (B x m x m, destroyed); X^-1 into xi
// (B x m x m).  Work buffers from `ws`, info (B ints) accumulates first zero pivots.
struct GJWork {
    double* wb = nullptr;
    int* cand = nullptr;
    int* cfin = nullptr;
    int* src = nullptr;
    int* ipiv = nullptr;
    int* prow = nullptr;
    double* Pinv = nullptr;
    double* Mm = nullptr;
};
size_t gj_work_doubles(int m, int B);  // doubles of wb + Pinv + Mm
size_t gj_work_ints(int m, int B);     // ints of cand + src + ipiv + prow
void run_gj_inverse(double* wa, double* xi, int m, int B, const GJWork& w, int* info, hipStream_t s);

}  // namespace efk
