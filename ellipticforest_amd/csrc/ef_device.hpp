// ef_device.hpp -- descriptors shared by the host planner and the gfx950 kernels.
//
// Layout conventions (all row-major, FP64 unless stated):
//  * a boundary vector / DtN block order is WESN, each side ordered by increasing
//    coordinate (src/Patches/FiniteVolume/FiniteVolumeSolver.cpp:296-353);
//  * a leaf cell vector is indexed j + i*N, x-major with y fastest
//    (FiniteVolumeSolver.cpp:17-19);
//  * an internal node tau with child interface size n stores
//      S    4n x 8n   interior potential from WESN boundary data (columns already in
//                     WESN order, i.e. after reorderOperators_, HPSAlgorithm.hpp:979)
//      Xi   4n x 4n   inverse of the interface Schur complement X (mergeX_, :870)
//      H    4 blocks of 2n x 2n, child c's tau-side rows x its two interface columns
//      T    8n x 8n   DtN of tau in WESN order
//  * multi-RHS vectors (h, g, w, u_int) live in per-batch arenas of R rows and k columns,
//    right-hand side major: element (row, rhs) is at rhs * R + row, so one right-hand side's
//    node vector is contiguous;
//  * leaf f / u for k right-hand sides: [rhs][leaf][cell];
//  * a node's boundary data g is read side by side: side s (WESN) of a node of side length m
//    is m consecutive rows at an arena offset of its own, which is either the node's g (when
//    the sweep writes it there) or the segment of an ancestor's interface potential or of the
//    root's g that the side lies on.
#pragma once
#include <cstddef>
#include <cstdint>

namespace efk {

// Pointers read from descriptors are generic to the compiler; loads through them become FLAT
// instructions, which also count against the LDS counter, so every wait on a shuffle or LDS
// read would wait for outstanding HBM loads as well.  Kernels cast such pointers to the
// global address space (device compilation only; the host pass never dereferences them).
#if defined(__HIP_DEVICE_COMPILE__)
#define EF_GLOBAL __attribute__((address_space(1)))
#else
#define EF_GLOBAL
#endif
template <class T>
__host__ __device__ inline const EF_GLOBAL T* gmem(const T* p) { return (const EF_GLOBAL T*)p; }
template <class T>
__host__ __device__ inline EF_GLOBAL T* gmem(T* p) { return (EF_GLOBAL T*)p; }

enum Side { W = 0, E = 1, S = 2, N = 3 };

template <class T>
__host__ __device__ inline T pick4(const T (&a)[4], int c) {  // a[c] without indexing a private array
    return c == 0 ? a[0] : c == 1 ? a[1] : c == 2 ? a[2] : a[3];
}

// coarsen_ step (HPSAlgorithm.hpp:707-741): T (4m x 4m) -> L21 T L12 (2m x 2m) per side block.
struct CoarsenDesc {
    const double* src;  // 4m x 4m
    double* dst;        // 2m x 2m ... i.e. (4 * m/2)^2
    int m;              // fine side length
};

// Per-leaf view used by the leaf sweep kernels.
struct LeafDesc {
    int cls;        // unused by the block-tridiagonal kernels (-1)
    int node;       // node id
    long long h_off; // row offset of h in the vector arena (4N rows)
    long long gs[4]; // row offsets of the four sides of g (N rows each)
    const double* coef;  // variable coefficients: 6 N*N arrays (alpha, bW, bE, bS, bN, lambda) or null
    double* Sinv;        // variable coefficients: N blocks of N x N, inverses of the block-LU pivots
    double* stencil;     // variable coefficients: 5 N*N arrays cW, cE, cS, cN, cC
    double* T;           // leaf DtN (4N x 4N)
    double* SinvP;       // symmetric matrices: the inverse pivot blocks packed (upper triangles), or null
    double dx, dy;
    long long pstride;   // entries from one packed block of the leaf to its next
};

// Upward family step (upwards4to1, HPSAlgorithm.hpp:529-551) for one merge.
struct UpDesc {
    int n;
    int tag[4];             // coarsenings of each child's h
    int clen[4];            // child vector length before coarsening (4 * size)
    long long ch_off[4];    // child h offsets (rows)
    long long hd_off;       // hDiff work (4n rows)
    long long w_off;        // w (4n rows)
    long long h_off;        // tau h (8n rows)
    const double* Xi;       // 4n x 4n (unused by the gather)
    const double* H;        // unused by the gather
};

// Boundary data of a child written out by the downward sweep (split1to4, HPSAlgorithm.hpp:
// 562-577): the four sides, len rows each, copied from their sources (segments of the parent's
// boundary data or interface potential) to dst + s * len.
struct SideCopy {
    long long src[4];
    long long dst;
    int len, pad;
};

// Batched patch solves by fast diagonalisation (constant coefficients per patch).
enum PatchOp { PATCH_OP_SOLVE = 0, PATCH_OP_MAP_D2N = 1, PATCH_OP_DTN = 2 };

struct PatchCoef {
    double cx, cy, lam;  // stencil: c_W = c_E = cx, c_S = c_N = cy, centre -2cx - 2cy + lam
    double dx, dy;
};

struct PatchFDArgs {
    int N, npatches, nrhs, op;
    const PatchCoef* coef;
    const int* cls;         // patch -> coefficient entry, or null (one entry per patch)
    const double* QA;       // basis in the MFMA A layout, lane-swizzled: [NB*NB*4][64]
    const double* QB;       // basis in the MFMA B layout, lane-swizzled
    const double* mu;       // eigenvalues of K, padded to 16*NB entries
    const double* dinv;     // per coefficient entry, 1/D on the padded (16 NB)^2 modes (0 where padded), or null
    const double* f;        // loads [nrhs][npatches][N*N], or null (zero)
    const double* g;        // boundary data, or null (zero); see g_off
    double* out;            // u [nrhs][npatches][N*N] | h (see h_off) | T [npatches][4N][4N]
    // Vector addressing.  Null offsets: g and h are [nrhs][npatches][4N].  Otherwise the
    // multi-RHS arena layout: element t of h of patch p, rhs c at c * vk + h_off[p] + t, and
    // element i of side s of g at c * vk + g_off[4 p + s] + i.
    const long long* g_off;
    const long long* h_off;
    int vk;
};

// Batched GEMM with arbitrary strides (transposes are free): for every batch entry b,
// C_b = A_b B_b, A_b(i, k) = A[b*sa + i*ra + k*ca], B_b(k, j) = B[b*sb + k*rb + j*cb],
// C_b(i, j) = C[b*sc + i*rc + j*cc].  M, N multiples of 16, K a multiple of 4.
struct GemmStridedArgs {
    const double* A; long long sa, ra, ca;
    const double* B; long long sb, rb, cb;
    double* C; long long sc, rc, cc;
    int M, N, K, batch;
};

// Element-wise steps of the fast diagonalisation of large patches (N > 32), on buffers of
// `items` N x N cell arrays.  Item t is (rhs, patch) = (t / npatches, t % npatches) for the
// solves, or DtN column `col0 + t` of patch `patch0` for DtN chunks.
struct PatchLargeArgs {
    int N, npatches, op, items;
    int patch0, col0;          // DtN chunks
    const PatchCoef* coef;
    const int* cls;
    const double* mu;          // N eigenvalues
    const double* f;           // [nrhs][npatches][N*N] or null
    const double* g;           // [nrhs][npatches][4N] or null
    double* work;              // items x N*N
    double* out;               // u (solve) | h | T
    long long item0;           // first item of this chunk
};

// One interpolation step on vectors in an arena (uncoarsen_: L12 of fine size nfine per side).
struct VecStepDesc {
    long long src_off, dst_off;
    int nfine;
};


// The planner fills these on the host and the kernels read them on the device: both
// compiler passes must agree on every size and offset (x86-64 and amdgcn are both LP64).
static_assert(sizeof(CoarsenDesc) == 24, "CoarsenDesc");
static_assert(sizeof(LeafDesc) == 112, "LeafDesc");
static_assert(offsetof(LeafDesc, h_off) == 8, "LeafDesc.h_off");
static_assert(offsetof(LeafDesc, coef) == 48, "LeafDesc.coef");
static_assert(sizeof(UpDesc) == 112, "UpDesc");
static_assert(offsetof(UpDesc, ch_off) == 40, "UpDesc.ch_off");
static_assert(offsetof(UpDesc, Xi) == 96, "UpDesc.Xi");
static_assert(sizeof(SideCopy) == 48, "SideCopy");
static_assert(sizeof(VecStepDesc) == 24, "VecStepDesc");
static_assert(sizeof(PatchCoef) == 40, "PatchCoef");
static_assert(sizeof(PatchFDArgs) == 112, "PatchFDArgs");
static_assert(offsetof(PatchFDArgs, coef) == 16, "PatchFDArgs.coef");
static_assert(offsetof(PatchFDArgs, vk) == 104, "PatchFDArgs.vk");

}  // namespace efk
