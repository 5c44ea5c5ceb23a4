// ef_kernels.hip -- gfx950 kernels of the HPS factor-and-solve path.
//
// Every kernel is batched over all patches of one tree level (or all leaves) so a
// level is a handful of launches, whatever the number of patches.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include "ef_device.hpp"
#include "ef_kernels.hpp"

namespace efk {

typedef double d4 __attribute__((ext_vector_type(4)));

// boundary cell of WESN position b (0..4N): (i, j)
__device__ inline void bcell(int b, int Nn, int& i, int& j) {
    const int s = b / Nn, t = b % Nn;
    if (s == W) { i = 0; j = t; }
    else if (s == E) { i = Nn - 1; j = t; }
    else if (s == S) { i = t; j = 0; }
    else { i = t; j = Nn - 1; }
}

__global__ void fill_kernel(double* p, double v, long long n) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) p[e] = v;
}

__global__ void copy_kernel(const double* src, double* dst, long long n) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) dst[e] = src[e];
}

}  // namespace efk

namespace efk {

// ---------------------------------------------------------------------------
// Variable-coefficient leaves: batched block-tridiagonal LU of the 5-point matrix.
// With cell index j + i*N, block row i (an x-line) couples to block i-1 through
// diag(cW(i,:)) and to block i+1 through diag(cE(i,:)); the diagonal block D_i is
// tridiagonal in j with cS/cN off the diagonal and the centre coefficient (minus the
// neighbour coefficient on a physical side, ghost = 2g - u, FiniteVolumeSolver.cpp:100-184).
// Block LU: S_0 = D_0, S_i = D_i - diag(cW_i) S_{i-1}^-1 diag(cE_{i-1}); the kernels keep
// Sinv_i = S_i^-1.  Solve: z_0 = Sinv_0 b_0, z_i = Sinv_i (b_i - cW_i .* z_{i-1}),
// x_{N-1} = z_{N-1}, x_i = z_i - Sinv_i (cE_i .* x_{i+1}).
// ---------------------------------------------------------------------------
constexpr int BT_MAXN = 32;

// stencil of one cell (FiniteVolumeSolver.cpp:64-85)
__device__ inline void bt_stencil(const double* coef, int NN, int I, double dx2, double dy2, double& cW, double& cE,
                                  double& cS, double& cN, double& cC) {
    const double a = coef[I], bW = coef[NN + I], bE = coef[2 * NN + I], bS = coef[3 * NN + I], bN = coef[4 * NN + I];
    const double lam = coef[5 * NN + I];
    cW = a * bW / dx2; cE = a * bE / dx2; cS = a * bS / dy2; cN = a * bN / dy2;
    cC = -a * ((bE + bW) / dx2 + (bN + bS) / dy2) + lam;
}

// One workgroup per leaf.  Gauss-Jordan with partial pivoting of [S | I] in LDS per block.
__global__ void __launch_bounds__(256) bt_factor_kernel(const LeafDesc* leaves, int Nn) {
    __shared__ double aug[BT_MAXN][2 * BT_MAXN + 1];
    __shared__ double prev[BT_MAXN][BT_MAXN + 1];
    __shared__ double cw[BT_MAXN], ce_prev[BT_MAXN];
    __shared__ int piv_row;
    const LeafDesc L = leaves[blockIdx.x];
    const int tid = threadIdx.x;
    const int NN = Nn * Nn;
    const double dx2 = L.dx * L.dx, dy2 = L.dy * L.dy;
    // stencil of every cell
    for (int I = tid; I < NN; I += 256) {
        double cW, cE, cS, cN, cC;
        bt_stencil(L.coef, NN, I, dx2, dy2, cW, cE, cS, cN, cC);
        L.stencil[I] = cW; L.stencil[NN + I] = cE; L.stencil[2 * NN + I] = cS; L.stencil[3 * NN + I] = cN;
        L.stencil[4 * NN + I] = cC;
    }
    __syncthreads();
    for (int i = 0; i < Nn; i++) {
        // S_i = D_i - diag(cW_i) Sinv_{i-1} diag(cE_{i-1}), augmented with I
        for (int j = tid; j < Nn; j += 256) {
            cw[j] = L.stencil[i * Nn + j];
            ce_prev[j] = (i > 0) ? L.stencil[NN + (i - 1) * Nn + j] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < Nn * 2 * Nn; e += 256) {
            const int r = e / (2 * Nn), c = e % (2 * Nn);
            double v = 0.0;
            if (c < Nn) {
                const int I = r + i * Nn;
                if (c == r) {
                    v = L.stencil[4 * NN + I];
                    if (i == 0) v -= L.stencil[I];
                    if (i == Nn - 1) v -= L.stencil[NN + I];
                    if (r == 0) v -= L.stencil[2 * NN + I];
                    if (r == Nn - 1) v -= L.stencil[3 * NN + I];
                } else if (c == r - 1) v = L.stencil[2 * NN + I];
                else if (c == r + 1) v = L.stencil[3 * NN + I];
                if (i > 0) v -= cw[r] * prev[r][c] * ce_prev[c];
            } else {
                v = (c - Nn == r) ? 1.0 : 0.0;
            }
            aug[r][c] = v;
        }
        __syncthreads();
        for (int k = 0; k < Nn; k++) {
            if (tid == 0) {
                int p = k;
                double best = fabs(aug[k][k]);
                for (int r = k + 1; r < Nn; r++) {
                    const double v = fabs(aug[r][k]);
                    if (v > best) { best = v; p = r; }
                }
                piv_row = p;
            }
            __syncthreads();
            const int p = piv_row;
            if (p != k)
                for (int c = tid; c < 2 * Nn; c += 256) {
                    const double t = aug[k][c];
                    aug[k][c] = aug[p][c];
                    aug[p][c] = t;
                }
            __syncthreads();
            const double pk = aug[k][k];
            const double inv = (pk != 0.0) ? 1.0 / pk : 0.0;
            // eliminate column k from every other row, then normalise row k
            for (int e = tid; e < Nn * 2 * Nn; e += 256) {
                const int r = e / (2 * Nn), c = e % (2 * Nn);
                if (r != k && c != k) aug[r][c] -= aug[r][k] * inv * aug[k][c];
            }
            __syncthreads();
            for (int e = tid; e < Nn * 2 * Nn; e += 256) {
                const int r = e / (2 * Nn), c = e % (2 * Nn);
                if (r != k && c == k) aug[r][c] = 0.0;
            }
            for (int c = tid; c < 2 * Nn; c += 256) aug[k][c] *= inv;
            __syncthreads();
        }
        double* Si = L.Sinv + (long long)i * NN;
        for (int e = tid; e < NN; e += 256) {
            const int r = e / Nn, c = e % Nn;
            const double v = aug[r][Nn + c];
            Si[e] = v;
            prev[r][c] = v;
        }
        __syncthreads();
    }
}

// x of lane `src` in every lane (src uniform)
__device__ inline double wave_read_lane(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}

// The same factorisation with one wave per leaf, N = NT in {8, 16, 32}: lane c holds column c
// of [S_i | I] (2 NT columns, rows in registers).  Partial pivoting by pivot order instead of
// row exchanges: step k takes the largest |S(r, k)| among the rows not yet used, divides that
// row by it and removes column k from every other row, so no register is ever indexed by a
// run-time value; the row that pivots at step k holds row k of the inverse at the end, and the
// inverse goes to LDS in row order (for the next block's S) and from there to Sinv.  The stencil
// of each block row is computed one cell per lane (as bt_stencil) into LDS and written out.
// PIVOT = false is for diagonally dominant matrices (the host checks every row), whose Schur
// complements S_i stay diagonally dominant, so Gauss-Jordan needs no pivot search: step k is
// unrolled, pivots on row k itself and leaves the inverse in natural row order.
template <int NT, bool PIVOT>
__global__ void __launch_bounds__(256) bt_factor_wave_kernel(const LeafDesc* leaves, int nleaves) {
    constexpr int NN = NT * NT;
    __shared__ double inv_all[4][NT][NT + 1];
    __shared__ double row_all[4][6][NT];  // cW, cE, cS, cN, cC of cells (i, r); cE of (i - 1, r)
    __shared__ int pinv_all[4][NT];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int leaf = blockIdx.x * 4 + wave;
    if (leaf >= nleaves) return;  // uniform per wave; no barrier below
    double (*inv)[NT + 1] = inv_all[wave];
    double (*srow)[NT] = row_all[wave];
    int* pinv = pinv_all[wave];
    const LeafDesc L = leaves[leaf];
    const double dx2 = L.dx * L.dx, dy2 = L.dy * L.dy;
    const auto* cf = gmem((const double*)L.coef);
    auto* stw = gmem(L.stencil);
    // stencil of cell I (FiniteVolumeSolver.cpp:64-85), as bt_stencil
    auto stencil = [&](int I, double& cW, double& cE, double& cS, double& cN, double& cC) {
        const double a = cf[I], bW = cf[NN + I], bE = cf[2 * NN + I], bS = cf[3 * NN + I], bN = cf[4 * NN + I];
        const double lam = cf[5 * NN + I];
        cW = a * bW / dx2; cE = a * bE / dx2; cS = a * bS / dy2; cN = a * bN / dy2;
        cC = -a * ((bE + bW) / dx2 + (bN + bS) / dy2) + lam;
    };
    const int c = (lane < NT) ? lane : lane - NT;  // column of S (lane < NT) or of I
    double col[NT];
    for (int i = 0; i < NT; i++) {
        // the stencil of block row i, one cell per lane (written out for the solves too), and cE
        // of block row i - 1 for the coupling: into LDS, read below by every lane
        if (lane < NT) {
            double cW, cE, cS, cN, cC;
            const int I = i * NT + lane;
            stencil(I, cW, cE, cS, cN, cC);
            srow[0][lane] = cW; srow[1][lane] = cE; srow[2][lane] = cS; srow[3][lane] = cN; srow[4][lane] = cC;
            stw[I] = cW; stw[NN + I] = cE; stw[2 * NN + I] = cS; stw[3 * NN + I] = cN; stw[4 * NN + I] = cC;
        } else if (lane < 2 * NT && i > 0) {
            double cW, cE, cS, cN, cC;
            stencil((i - 1) * NT + c, cW, cE, cS, cN, cC);
            srow[5][c] = cE;
        }
        __builtin_amdgcn_wave_barrier();
        const double ce_prev = (i > 0) ? srow[5][c] : 0.0;
#pragma unroll
        for (int r = 0; r < NT; r++) {
            const double cW = srow[0][r], cE = srow[1][r], cS = srow[2][r], cN = srow[3][r], cC = srow[4][r];
            double v;
            if (lane < NT) {
                v = 0.0;
                if (c == r) {
                    v = cC;
                    if (i == 0) v -= cW;
                    if (i == NT - 1) v -= cE;
                    if (r == 0) v -= cS;
                    if (r == NT - 1) v -= cN;
                } else if (c == r - 1) {
                    v = cS;
                } else if (c == r + 1) {
                    v = cN;
                }
                if (i > 0) v -= cW * inv[r][c] * ce_prev;
            } else {
                v = (lane < 2 * NT && c == r) ? 1.0 : 0.0;
            }
            col[r] = v;
        }
        if (!PIVOT) {
#pragma unroll
            for (int k = 0; k < NT; k++) {
                const double pk = wave_read_lane(col[k], k);
                const double ip = (pk != 0.0) ? 1.0 / pk : 0.0;
                const double rowk = col[k] * ip;
#pragma unroll
                for (int r = 0; r < NT; r++) {
                    if (r == k) continue;
                    const double l = wave_read_lane(col[r], k);
                    col[r] = fma(-l, rowk, col[r]);
                }
                col[k] = rowk;
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < NT; r++)
                if (lane >= NT && lane < 2 * NT) inv[r][c] = col[r];
        }
        unsigned used = 0;
        for (int k = 0; PIVOT && k < NT; k++) {
            // pivot: the largest unused |S(r, k)|, found by lane k on its own column
            double best = -1.0;
            int bi = 0;
#pragma unroll
            for (int r = 0; r < NT; r++) {
                const double a = fabs(col[r]);
                const bool take = !((used >> r) & 1u) && a > best;
                best = take ? a : best;
                bi = take ? r : bi;
            }
            const int pr = __builtin_amdgcn_readlane(bi, k);
            used |= 1u << pr;
            double rowp = 0.0;
#pragma unroll
            for (int r = 0; r < NT; r++) rowp = (r == pr) ? col[r] : rowp;
            const double pk = wave_read_lane(rowp, k);
            const double ip = (pk != 0.0) ? 1.0 / pk : 0.0;
            rowp *= ip;
#pragma unroll
            for (int r = 0; r < NT; r++) {
                const double l = wave_read_lane(col[r], k);
                col[r] = (r == pr) ? rowp : fma(-l, rowp, col[r]);
            }
            if (lane == 0) pinv[pr] = k;
        }
        // the inverse in row order: row pinv[r] of Sinv_i is row r of the right half
        if (PIVOT) {
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < NT; r++)
                if (lane >= NT && lane < 2 * NT) inv[pinv[r]][c] = col[r];
        }
        __builtin_amdgcn_wave_barrier();
        auto* Si = gmem(L.Sinv) + (long long)i * NN;
        for (int e = lane; e < NN; e += 64) Si[e] = inv[e / NT][e % NT];
        __builtin_amdgcn_wave_barrier();
    }
}

template __global__ void bt_factor_wave_kernel<8, true>(const LeafDesc*, int);
template __global__ void bt_factor_wave_kernel<16, true>(const LeafDesc*, int);
template __global__ void bt_factor_wave_kernel<32, true>(const LeafDesc*, int);
template __global__ void bt_factor_wave_kernel<8, false>(const LeafDesc*, int);
template __global__ void bt_factor_wave_kernel<16, false>(const LeafDesc*, int);
template __global__ void bt_factor_wave_kernel<32, false>(const LeafDesc*, int);

void launch_bt_factor(const LeafDesc* leaves, int Nn, int nleaves, hipStream_t s, bool dominant) {
    const unsigned blocks = (unsigned)((nleaves + 3) / 4);
#define EF_BT_FACTOR(NT_)                                                                                 \
    do {                                                                                                   \
        if (dominant) hipLaunchKernelGGL((bt_factor_wave_kernel<NT_, false>), dim3(blocks), dim3(256), 0, s, leaves, nleaves); \
        else hipLaunchKernelGGL((bt_factor_wave_kernel<NT_, true>), dim3(blocks), dim3(256), 0, s, leaves, nleaves); \
    } while (0)
    if (Nn == 8) EF_BT_FACTOR(8);
    else if (Nn == 16) EF_BT_FACTOR(16);
    else if (Nn == 32) EF_BT_FACTOR(32);
#undef EF_BT_FACTOR
    else
        hipLaunchKernelGGL(bt_factor_kernel, dim3(nleaves), dim3(256), 0, s, leaves, Nn);
}

// Block substitution of one right-hand side b (N x N, in LDS) in place -> solution.
__device__ inline void bt_substitute(const LeafDesc& L, int Nn, double* b, double* tmp) {
    const int tid = threadIdx.x;
    const int NN = Nn * Nn;
    // forward: z_i = Sinv_i (b_i - cW_i .* z_{i-1})
    for (int i = 0; i < Nn; i++) {
        for (int j = tid; j < Nn; j += 256) {
            double v = b[i * Nn + j];
            if (i > 0) v -= L.stencil[i * Nn + j] * b[(i - 1) * Nn + j];
            tmp[j] = v;
        }
        __syncthreads();
        const double* Si = L.Sinv + (long long)i * NN;
        for (int r = tid; r < Nn; r += 256) {
            double s = 0.0;
            for (int c = 0; c < Nn; c++) s += Si[r * Nn + c] * tmp[c];
            b[i * Nn + r] = s;
        }
        __syncthreads();
    }
    // backward: x_i = z_i - Sinv_i (cE_i .* x_{i+1})
    for (int i = Nn - 2; i >= 0; i--) {
        for (int j = tid; j < Nn; j += 256) tmp[j] = L.stencil[NN + i * Nn + j] * b[(i + 1) * Nn + j];
        __syncthreads();
        const double* Si = L.Sinv + (long long)i * NN;
        for (int r = tid; r < Nn; r += 256) {
            double s = 0.0;
            for (int c = 0; c < Nn; c++) s += Si[r * Nn + c] * tmp[c];
            b[i * Nn + r] -= s;
        }
        __syncthreads();
    }
}

// Leaf solve for every (leaf, rhs) with variable coefficients.  Modes: 0 particular
// Neumann data (g = 0, h out), 1 solve with g (u out), 2 mapD2N with g (h out).
__global__ void __launch_bounds__(256) bt_solve_kernel(const LeafDesc* leaves, int Nn, int nleaves, int k, int mode,
                                                       const double* f, int f_zero, const double* garena,
                                                       double* harena, double* u, long long vcs) {
    __shared__ double b[BT_MAXN * BT_MAXN];
    __shared__ double tmp[BT_MAXN];
    const int leaf = blockIdx.x, rhs = blockIdx.y;
    const LeafDesc L = leaves[leaf];
    const int tid = threadIdx.x;
    const int NN = Nn * Nn;
    const double* fl = f_zero ? nullptr : f + ((long long)rhs * nleaves + leaf) * NN;
    // side sd of g: Nn rows at the leaf's own offset for that side
    const double* g = (mode != 0) ? garena + (long long)rhs * vcs : nullptr;
    const double* gW = g ? g + L.gs[W] : nullptr;
    const double* gE = g ? g + L.gs[E] : nullptr;
    const double* gS = g ? g + L.gs[S] : nullptr;
    const double* gN = g ? g + L.gs[N] : nullptr;
    for (int e = tid; e < NN; e += 256) {
        const int i = e / Nn, j = e % Nn;
        double v = fl ? fl[e] : 0.0;
        if (g && (i == 0 || j == 0 || i == Nn - 1 || j == Nn - 1)) {
            double add = 0.0;
            if (i == 0) add += L.stencil[e] * gW[j];
            if (i == Nn - 1) add += L.stencil[NN + e] * gE[j];
            if (j == 0) add += L.stencil[2 * NN + e] * gS[i];
            if (j == Nn - 1) add += L.stencil[3 * NN + e] * gN[i];
            v += -2.0 * add;
        }
        b[e] = v;
    }
    __syncthreads();
    bt_substitute(L, Nn, b, tmp);
    if (mode == 1) {
        double* ul = u + ((long long)rhs * nleaves + leaf) * NN;
        for (int e = tid; e < NN; e += 256) ul[e] = b[e];
    } else {
        double* h = harena + (long long)rhs * vcs + L.h_off;
        for (int t = tid; t < 4 * Nn; t += 256) {
            int i, j;
            bcell(t, Nn, i, j);
            const int sd = t / Nn;
            const double dd = (sd < 2) ? L.dx : L.dy;
            const double sg = (sd == W || sd == S) ? 2.0 / dd : -2.0 / dd;
            h[t] = sg * (b[i * Nn + j] - (g ? g[pick4(L.gs, sd) + (t - sd * Nn)] : 0.0));
        }
    }
}

// The same substitution with one wave per (leaf, rhs), for N = NT in {8, 16, 32}.
// Lane (j, h): j = lane % NT is the cell row index within a block row (an x-line), h =
// lane / NT one of G = 64 / NT groups that each own C = NT / G columns of the inverse pivot
// blocks.  The block rows of the solution sit in the wave's own slice of LDS (lane j holds
// element j); a block product Sinv_i t is C FMAs per lane on columns fetched from lane c by
// shuffles, then a butterfly over the G groups.  Sinv_{i+1} is loaded while step i computes
// (two register buffers, the block-row loop unrolled by two).
template <int NT>
__device__ inline double bt_wave_matvec(const double (&Srow)[NT * NT / 64], double t, int h) {
    constexpr int G = 64 / NT, C = NT / G;
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < C; m++) acc = fma(Srow[m], __shfl(t, h * C + m), acc);
#pragma unroll
    for (int o = NT; o < 64; o <<= 1) acc += __shfl_xor(acc, o);
    return acc;
}

template <int NT>
__device__ inline void bt_wave_load(const double* Si, int j, int h, double (&Srow)[NT * NT / 64]) {
    constexpr int G = 64 / NT, C = NT / G;
    const auto* row = gmem(Si) + j * NT + h * C;
#pragma unroll
    for (int m = 0; m < C; m++) Srow[m] = row[m];
}

template <int NT>
__global__ void __launch_bounds__(256) bt_solve_wave_kernel(const LeafDesc* leaves, int nleaves, int k, int mode,
                                                            const double* f, int f_zero, const double* garena,
                                                            double* harena, double* u, long long vcs) {
    constexpr int NN = NT * NT, C = NT * NT / 64;
    __shared__ double zs_all[4][NN];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const long long item = (long long)blockIdx.x * 4 + wave;
    if (item >= (long long)nleaves * k) return;  // uniform per wave; no barrier below
    double* zs = zs_all[wave];
    const int rhs = (int)(item / nleaves), leaf = (int)(item % nleaves);
    const int j = lane % NT, h = lane / NT;
    const LeafDesc L = leaves[leaf];
    const auto* fl = f_zero ? nullptr : gmem(f) + ((long long)rhs * nleaves + leaf) * NN;
    // side sd of g: NT rows at the leaf's own offset for that side
    const auto* g = (mode != 0) ? gmem(garena) + (long long)rhs * vcs : nullptr;
    const auto* cWv = gmem(L.stencil);
    const auto* cEv = cWv + NN;
    const auto* cSv = cWv + 2 * NN;
    const auto* cNv = cWv + 3 * NN;
    auto rhs_row = [&](int i) {
        const int I = i * NT + j;
        double b = fl ? fl[I] : 0.0;
        if (g) {
            double add = 0.0;
            if (i == 0) add += cWv[I] * g[L.gs[W] + j];
            if (i == NT - 1) add += cEv[I] * g[L.gs[E] + j];
            if (j == 0) add += cSv[I] * g[L.gs[S] + i];
            if (j == NT - 1) add += cNv[I] * g[L.gs[N] + i];
            b += -2.0 * add;
        }
        return b;
    };
    double Sa[C], Sb[C];
    // forward: z_i = Sinv_i (b_i - cW_i .* z_{i-1}); even i in Sa, odd i in Sb.  The next block
    // row's inverse, right-hand side and west coefficients are loaded while this one computes.
    bt_wave_load<NT>(L.Sinv, j, h, Sa);
    double zprev = 0.0;
    double bn = rhs_row(0), cwn = 0.0;
    for (int i = 0; i < NT; i += 2) {
        double b = bn, cw = cwn;
        bt_wave_load<NT>(L.Sinv + (long long)(i + 1) * NN, j, h, Sb);
        bn = rhs_row(i + 1);
        cwn = cWv[(i + 1) * NT + j];
        zprev = bt_wave_matvec<NT>(Sa, (i > 0) ? b - cw * zprev : b, h);
        if (h == 0) zs[i * NT + j] = zprev;
        b = bn;
        cw = cwn;
        if (i + 2 < NT) {
            bt_wave_load<NT>(L.Sinv + (long long)(i + 2) * NN, j, h, Sa);
            bn = rhs_row(i + 2);
            cwn = cWv[(i + 2) * NT + j];
        }
        zprev = bt_wave_matvec<NT>(Sb, b - cw * zprev, h);
        if (h == 0) zs[(i + 1) * NT + j] = zprev;
    }
    // backward: x_i = z_i - Sinv_i (cE_i .* x_{i+1}) for i = N-2 .. 0, x_{N-1} = z_{N-1} = zprev;
    // pairs (i, i-1) for i = N-2, N-4, .., 2 (Sinv_i in Sa, Sinv_{i-1} in Sb), then i = 0 (Sa).
    // z_i is read before the product, whose shuffles order every lane's read before the write.
    double xnext = zprev;
    bt_wave_load<NT>(L.Sinv + (long long)(NT - 2) * NN, j, h, Sa);
    double cen = cEv[(NT - 2) * NT + j];
    for (int i = NT - 2; i >= 2; i -= 2) {
        double ce = cen;
        bt_wave_load<NT>(L.Sinv + (long long)(i - 1) * NN, j, h, Sb);
        cen = cEv[(i - 1) * NT + j];
        double zi = zs[i * NT + j];
        double x = zi - bt_wave_matvec<NT>(Sa, ce * xnext, h);
        if (h == 0) zs[i * NT + j] = x;
        xnext = x;
        ce = cen;
        bt_wave_load<NT>(L.Sinv + (long long)(i - 2) * NN, j, h, Sa);
        cen = cEv[(i - 2) * NT + j];
        zi = zs[(i - 1) * NT + j];
        x = zi - bt_wave_matvec<NT>(Sb, ce * xnext, h);
        if (h == 0) zs[(i - 1) * NT + j] = x;
        xnext = x;
    }
    {
        const double zi = zs[j];
        const double x = zi - bt_wave_matvec<NT>(Sa, cen * xnext, h);
        if (h == 0) zs[j] = x;
    }
    // the wave's LDS writes are in order before its reads below; on the CPU emulation this is
    // where the lanes meet
    __builtin_amdgcn_wave_barrier();
    if (mode == 1) {
        auto* ul = gmem(u) + ((long long)rhs * nleaves + leaf) * NN;
        for (int e = lane; e < NN; e += 64) ul[e] = zs[e];
        return;
    }
    // Neumann data of the edge cells: h_W = (2/dx)(u_W - g_W), h_E = -(2/dx)(u_E - g_E),
    // h_S = (2/dy)(u_S - g_S), h_N = -(2/dy)(u_N - g_N)
    auto* hv = gmem(harena) + (long long)rhs * vcs + L.h_off;
    const double sx = 2.0 / L.dx, sy = 2.0 / L.dy;
    auto gv = [&](int t) {
        const int sd = t / NT;
        return g ? g[pick4(L.gs, sd) + (t - sd * NT)] : 0.0;
    };
    for (int t = lane; t < 4 * NT; t += 64) {
        const int sd = t / NT, m = t % NT;
        double v;
        if (sd == W) v = sx * (zs[m] - gv(t));
        else if (sd == E) v = -sx * (zs[(NT - 1) * NT + m] - gv(t));
        else if (sd == S) v = sy * (zs[m * NT] - gv(t));
        else v = -sy * (zs[m * NT + NT - 1] - gv(t));
        hv[t] = v;
    }
}

template __global__ void bt_solve_wave_kernel<8>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                 double*, double*, long long);
template __global__ void bt_solve_wave_kernel<16>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                  double*, double*, long long);
template __global__ void bt_solve_wave_kernel<32>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                  double*, double*, long long);

// Leaf DtN with variable coefficients on the matrix cores, N = NT in {16, 32}: the same
// substitution as bt_dtn_kernel for all 4N boundary columns at once, as block products
// Z_i = Sinv_i (R_i - cW_i .* Z_{i-1}) and X_i = Z_i - Sinv_i (cE_i .* X_{i+1}) on
// v_mfma_f64_16x16x4.  Wave w of the workgroup (one leaf per workgroup) owns the N columns of
// side w, for all rows, so the waves never meet.  An accumulator tile (slot v of lane (q, r)
// = row 4v + q, column r) is the B operand of the next product as it stands (k-step ks takes
// rows 4 ks + q: tile ks / 4, slot ks % 4): the column scaling by cW or cE and the boundary
// forcing R are applied in registers.  Z_i goes to the leaf's slice of `work` ([i][row][4N
// columns]; each lane reads back only what it wrote) for the backward sweep, and each X_i's
// edge entries go straight into T as the sweep produces them.
template <int NT>
__global__ void __launch_bounds__(256) bt_dtn_mfma_kernel(const LeafDesc* leaves, int first, double* work) {
    constexpr int NN = NT * NT, NB = 4 * NT, RT = NT / 16, CT = NT / 16, KS = NT / 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    const LeafDesc L = leaves[first + blockIdx.x];
    auto* Zw = gmem(work) + (long long)blockIdx.x * NT * NN * 4;
    const auto* st = gmem((const double*)L.stencil);
    auto* Tl = gmem(L.T);
    const int c0 = wave * NT;  // columns of side `wave`
    // forcing of this lane's columns c = c0 + 16 t + r (position m on side `wave`): the boundary
    // cell (fi, fk) of the column and -2 c_side there (FiniteVolumeSolver.cpp:100-184)
    int fi[CT], fk[CT];
    double fv[CT];
#pragma unroll
    for (int t = 0; t < CT; t++) {
        const int m = 16 * t + r;
        fi[t] = (wave == W) ? 0 : (wave == E) ? NT - 1 : m;
        fk[t] = (wave == S) ? 0 : (wave == N) ? NT - 1 : m;
        fv[t] = -2.0 * st[wave * NN + fi[t] * NT + fk[t]];
    }
    const double sx = 2.0 / L.dx, sy = 2.0 / L.dy;
    // T entries of block row i: W (i = 0) and E (i = N-1) rows, S (k = 0) and N (k = N-1) entries
    auto emit = [&](int i, const d4 (&X)[RT][CT]) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = rt * 16 + 4 * v + q, c = c0 + 16 * t + r;
                    const double x = X[rt][t][v];
                    auto put = [&](int a, double sg) { Tl[(long long)a * NB + c] = sg * (x - (a == c ? 1.0 : 0.0)); };
                    if (i == 0) put(W * NT + k, sx);
                    if (i == NT - 1) put(E * NT + k, -sx);
                    if (k == 0) put(S * NT + i, sy);
                    if (k == NT - 1) put(N * NT + i, -sy);
                }
    };
    auto zaddr = [&](int i, int rt, int t, int v) {
        return ((long long)i * NT + rt * 16 + 4 * v + q) * NB + c0 + 16 * t + r;
    };
    d4 Zp[RT][CT], Zc[RT][CT];
    // Everything block i reads from memory (Sinv_i in the A layout, the cW or cE column scaling,
    // and in the backward sweep Z_i) is loaded while block i -+ 1 is computed, and nothing else
    // is loaded meanwhile: the compiler's waits count loads in issue order, so the block's own
    // products never wait for memory.
    double a_all[KS][RT], a_nxt[KS][RT], c_all[KS], c_nxt[KS];
    auto load_a = [&](int i, int side, double (&A)[KS][RT], double (&C)[KS]) {
        const auto* Si = gmem((const double*)L.Sinv) + (long long)i * NN;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int rt = 0; rt < RT; rt++) A[ks][rt] = Si[(rt * 16 + r) * NT + 4 * ks + q];
            C[ks] = st[side * NN + i * NT + 4 * ks + q];  // cW (side 0) or cE (side 1) of block row i
        }
    };
    load_a(0, 0, a_all, c_all);
    // forward
    for (int i = 0; i < NT; i++) {
        // next: block i + 1 with its cW, or for the backward sweep's first step block N - 2 with
        // its cE
        if (i + 1 < NT) load_a(i + 1, 0, a_nxt, c_nxt);
        else load_a(NT - 2, 1, a_nxt, c_nxt);
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++) Zc[rt][t] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + q;
            const double cw = (i > 0) ? c_all[ks] : 0.0;
            const double (&a)[RT] = a_all[ks];
#pragma unroll
            for (int t = 0; t < CT; t++) {
                double b = (i > 0) ? -cw * Zp[ks / 4][t][ks % 4] : 0.0;
                if (fi[t] == i && fk[t] == k) b += fv[t];
#pragma unroll
                for (int rt = 0; rt < RT; rt++) Zc[rt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b, Zc[rt][t], 0, 0, 0);
            }
        }
        if (i < NT - 1) {
#pragma unroll
            for (int rt = 0; rt < RT; rt++)
#pragma unroll
                for (int t = 0; t < CT; t++)
#pragma unroll
                    for (int v = 0; v < 4; v++) Zw[zaddr(i, rt, t, v)] = Zc[rt][t][v];
        }
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++) Zp[rt][t] = Zc[rt][t];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int rt = 0; rt < RT; rt++) a_all[ks][rt] = a_nxt[ks][rt];
            c_all[ks] = c_nxt[ks];
        }
    }
    // backward: X_{N-1} = Z_{N-1}.  Z_i (the accumulator's start) is read one block ahead too;
    // Z_{N-1} never went to memory, the forward sweep stored Z_0 .. Z_{N-2}.
    d4 zn[RT][CT];
    auto load_z = [&](int i) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++)
#pragma unroll
                for (int v = 0; v < 4; v++) zn[rt][t][v] = Zw[zaddr(i, rt, t, v)];
    };
    load_z(NT - 2);
    emit(NT - 1, Zp);
    for (int i = NT - 2; i >= 0; i--) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++) Zc[rt][t] = zn[rt][t];
        const int in = i > 0 ? i - 1 : 0;
        load_a(in, 1, a_nxt, c_nxt);
        load_z(in);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int k = 4 * ks + q;
            const double ce = c_all[ks];
            const double (&a)[RT] = a_all[ks];
#pragma unroll
            for (int t = 0; t < CT; t++) {
                const double b = -ce * Zp[ks / 4][t][ks % 4];
#pragma unroll
                for (int rt = 0; rt < RT; rt++) Zc[rt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b, Zc[rt][t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int rt = 0; rt < RT; rt++) a_all[ks][rt] = a_nxt[ks][rt];
            c_all[ks] = c_nxt[ks];
        }
        emit(i, Zc);
#pragma unroll
        for (int rt = 0; rt < RT; rt++)
#pragma unroll
            for (int t = 0; t < CT; t++) Zp[rt][t] = Zc[rt][t];
    }
}

template __global__ void bt_dtn_mfma_kernel<16>(const LeafDesc*, int, double*);
template __global__ void bt_dtn_mfma_kernel<32>(const LeafDesc*, int, double*);

void launch_bt_dtn(const LeafDesc* leaves, int Nn, int nleaves, double* work, long long chunk, hipStream_t s) {
    for (int first = 0; first < nleaves; first += (int)chunk) {
        const unsigned cnt = (unsigned)std::min<long long>(chunk, nleaves - first);
        if (Nn == 16)
            hipLaunchKernelGGL(bt_dtn_mfma_kernel<16>, dim3(cnt), dim3(256), 0, s, leaves, first, work);
        else if (Nn == 32)
            hipLaunchKernelGGL(bt_dtn_mfma_kernel<32>, dim3(cnt), dim3(256), 0, s, leaves, first, work);
        else
            hipLaunchKernelGGL(bt_dtn_kernel, dim3(cnt), dim3(256), 0, s, leaves, Nn, first, work);
    }
}

// ---------------------------------------------------------------------------
// Symmetric matrices (alpha beta the same seen from both cells of every interior face, which
// holds for constant alpha): the pivot blocks S_i and their inverses are symmetric, and the
// solve reads only the upper triangles, NT (NT + 1) / 2 entries per block instead of NT^2.
// Memory layout of one block: NT / 2 pairs of rows, pair p = row p from the diagonal on
// (NT - p entries), then row NT - 1 - p from the diagonal on (p + 1 entries): NT + 1 entries
// per pair.  In LDS the pairs sit NT + 2 entries apart, so that the rows a lane reads along
// lie in different banks.
// ---------------------------------------------------------------------------
template <int NT>
struct BtPack {
    static constexpr int NP = NT * (NT + 1) / 2;  // entries per block
    static constexpr int PAIR = NT + 1;           // entries per pair of rows in memory
    static constexpr int LSIZE = NT / 2 * (NT + 2);  // LDS entries per block
    static constexpr int R = (NP + 63) / 64;      // entries per lane when a wave loads a block
    // LDS index of S[a][b], a <= b
    static __device__ inline int lds(int a, int b) {
        return a < NT / 2 ? (NT + 1) * a + b : (NT + 1) * (NT - 1) + NT - (NT + 2) * a + b;
    }
};

// SinvP of every leaf from its Sinv: entry (a, b) of the upper triangle = (S[a][b] + S[b][a]) / 2
// (the computed inverse of a symmetric block is symmetric up to rounding).  Block i of a leaf
// starts i * pstride entries after its block 0: with the blocks of all leaves interleaved
// (block i of leaf l at (i * nleaves + l) * NP), the waves of a launch, which step through
// their leaves' blocks at about the same pace, read neighbouring memory.
__global__ void bt_pack_kernel(const LeafDesc* leaves, int nleaves, int Nn) {
    const int NP = Nn * (Nn + 1) / 2, PAIR = Nn + 1;
    const long long per = (long long)Nn * NP;
    const long long total = (long long)nleaves * per;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int leaf = (int)(e / per);
        const long long r = e % per;
        const int i = (int)(r / NP), q = (int)(r % NP);
        const int p = q / PAIR, o = q % PAIR;
        int a, b;
        if (o < Nn - p) {
            a = p;
            b = p + o;
        } else {
            a = Nn - 1 - p;
            b = a + (o - (Nn - p));
        }
        const double* S = leaves[leaf].Sinv + (long long)i * Nn * Nn;
        leaves[leaf].SinvP[(long long)i * leaves[leaf].pstride + q] = 0.5 * (S[a * Nn + b] + S[b * Nn + a]);
    }
}

// x[l] + x[l ^ 32] in every lane: the two halves of the wave exchanged by v_permlane32_swap
__device__ inline double wave_half_sum(double x) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto sl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto sh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double((int)sh[0], (int)sl[0]) + __hiloint2double((int)sh[1], (int)sl[1]);
}

// One wave per (leaf, rhs) as bt_solve_wave_kernel, on the packed blocks.  The 2 NT - 1 block
// steps (forward i = 0 .. NT-1, then backward i = NT-2 .. 0) form one sequence; while step s
// computes on block s in one of the wave's two LDS slots, block s + 1 waits in the other and
// block s + 2 is on its way into registers, so the backward sweep's first blocks arrive during
// the last forward steps.  A block product: group 0 publishes the vector in LDS, lane (j, h)
// gathers S[j][c] for its C columns c from the packed upper triangle, and the G groups are
// summed.  The rows z_i of the forward sweep go to zbuf ([rhs][leaf][N*N]; the solution itself
// when mode == 1), every lane writing and later reading back element j, and the backward sweep
// overwrites them with x_i; LDS holds two blocks and the vector (8.9 KB per wave at N = 32), so
// four waves per SIMD fit.
template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
bt_solve_sym_kernel(const LeafDesc* leaves, int nleaves, int k, int mode, const double* f, int f_zero,
                    const double* garena, double* harena, double* zbuf, long long vcs) {
    using PK = BtPack<NT>;
    constexpr int NN = NT * NT, G = 64 / NT, C = NT / G, R = PK::R, STEPS = 2 * NT - 1;
    __shared__ double sp_all[4][2][PK::LSIZE];
    __shared__ __attribute__((aligned(16))) double tv_all[4][NT];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const long long item = (long long)blockIdx.x * 4 + wave;
    if (item >= (long long)nleaves * k) return;  // uniform per wave; no barrier below
    double* tv = tv_all[wave];
    const int rhs = (int)(item / nleaves), leaf = (int)(item % nleaves);
    const int j = lane % NT, h = lane / NT;
    const LeafDesc L = leaves[leaf];
    const auto* fl = f_zero ? nullptr : gmem(f) + ((long long)rhs * nleaves + leaf) * NN;
    // side sd of g: NT rows at the leaf's own offset for that side
    const auto* g = (mode != 0) ? gmem(garena) + (long long)rhs * vcs : nullptr;
    auto* zl = gmem(zbuf) + ((long long)rhs * nleaves + leaf) * NN;
    const auto* cWv = gmem(L.stencil);
    const auto* cEv = cWv + NN;
    const auto* cSv = cWv + 2 * NN;
    const auto* cNv = cWv + 3 * NN;
    const auto* SP = gmem((const double*)L.SinvP);
    auto rhs_row = [&](int i) {
        const int I = i * NT + j;
        double b = fl ? fl[I] : 0.0;
        if (g) {
            double add = 0.0;
            if (i == 0) add += cWv[I] * g[L.gs[W] + j];
            if (i == NT - 1) add += cEv[I] * g[L.gs[E] + j];
            if (j == 0) add += cSv[I] * g[L.gs[S] + i];
            if (j == NT - 1) add += cNv[I] * g[L.gs[N] + i];
            b += -2.0 * add;
        }
        return b;
    };
    auto blk = [](int s) { return s < NT ? s : 2 * NT - 2 - s; };
    auto load = [&](int s, double (&Rr)[R]) {
        const auto* src = SP + (long long)blk(s) * L.pstride;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = lane + 64 * r;
            Rr[r] = (e < PK::NP) ? src[e] : 0.0;
        }
    };
    auto stage = [&](const double (&Rr)[R], int slot) {
        double* sp = sp_all[wave][slot];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = lane + 64 * r;
            if (e < PK::NP) sp[e + e / PK::PAIR] = Rr[r];
        }
    };
    // gather indices of this lane's columns c = h C + m: S[j][c] = sp[row + c] for c >= j (row j of
    // the upper triangle), sp[col0 + m col1 + j] for c < j (row c)
    const int cfirst = h * C;
    const int row = PK::lds(j, 0);
    const int col0 = (cfirst < NT / 2) ? (NT + 1) * cfirst : (NT + 1) * (NT - 1) + NT - (NT + 2) * cfirst;
    const int col1 = (cfirst < NT / 2) ? (NT + 1) : -(NT + 2);
    auto matvec = [&](const double* sp, double t) {
        if (h == 0) tv[j] = t;
        __builtin_amdgcn_wave_barrier();
        double acc0 = 0.0, acc1 = 0.0;
        if constexpr (C == 1) {
            acc0 = sp[(cfirst >= j) ? row + cfirst : col0 + j] * tv[cfirst];
        } else {
            // two columns per iteration, unrolled by two: fewer gathers in flight at a time keep
            // the register count within four waves per SIMD
#pragma unroll 2
            for (int m = 0; m < C; m += 2) {
                const int c = cfirst + m;
                const int i0 = (c >= j) ? row + c : col0 + m * col1 + j;
                const int i1 = (c + 1 >= j) ? row + c + 1 : col0 + (m + 1) * col1 + j;
                acc0 += sp[i0] * tv[c];
                acc1 += sp[i1] * tv[c + 1];
            }
        }
        double acc = acc0 + acc1;
#pragma unroll
        for (int o = NT; o < 32; o <<= 1) acc += __shfl_xor(acc, o);
        return wave_half_sum(acc);
    };
    double Rr[R];
    load(0, Rr);
    stage(Rr, 0);
    load(1, Rr);
    stage(Rr, 1);
    load(2, Rr);
    double zprev = 0.0, xnext = 0.0;
    double bn = rhs_row(0), cwn = 0.0, cen = 0.0;
    // step s on the block in LDS: forward z_i = Sinv_i (b_i - cW_i .* z_{i-1}) for s < NT, backward
    // x_i = z_i - Sinv_i (cE_i .* x_{i+1}) after (x_{N-1} = z_{N-1})
    auto step = [&](int s) {
        const double* sp = sp_all[wave][s & 1];
        if (s < NT) {
            const int i = s;
            const double b = bn, cw = cwn;
            if (i + 1 < NT) {
                bn = rhs_row(i + 1);
                cwn = cWv[(i + 1) * NT + j];
            } else {
                cen = cEv[(NT - 2) * NT + j];
            }
            zprev = matvec(sp, (i > 0) ? b - cw * zprev : b);
            zl[i * NT + j] = zprev;
            if (i == NT - 1) xnext = zprev;
        } else {
            const int i = 2 * NT - 2 - s;
            const double ce = cen;
            if (i > 0) cen = cEv[(i - 1) * NT + j];
            const double zi = zl[i * NT + j];
            const double x = zi - matvec(sp, ce * xnext);
            zl[i * NT + j] = x;
            xnext = x;
        }
    };
    for (int s = 0; s < STEPS; s++) {
        step(s);
        // slot s & 1 is free once step s has read it (the wave's LDS accesses run in order)
        if (s + 2 < STEPS) {
            stage(Rr, s & 1);
            if (s + 3 < STEPS) load(s + 3, Rr);
        }
    }
    if (mode == 1) return;  // the solution is in zbuf
    // Neumann data of the edge cells: h_W = (2/dx)(u_W - g_W), h_E = -(2/dx)(u_E - g_E),
    // h_S = (2/dy)(u_S - g_S), h_N = -(2/dy)(u_N - g_N), from elements other lanes wrote
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    auto* hv = gmem(harena) + (long long)rhs * vcs + L.h_off;
    const double sx = 2.0 / L.dx, sy = 2.0 / L.dy;
    auto gv = [&](int t) {
        const int sd = t / NT;
        return g ? g[pick4(L.gs, sd) + (t - sd * NT)] : 0.0;
    };
    for (int t = lane; t < 4 * NT; t += 64) {
        const int sd = t / NT, mm = t % NT;
        double v;
        if (sd == W) v = sx * (zl[mm] - gv(t));
        else if (sd == E) v = -sx * (zl[(NT - 1) * NT + mm] - gv(t));
        else if (sd == S) v = sy * (zl[mm * NT] - gv(t));
        else v = -sy * (zl[mm * NT + NT - 1] - gv(t));
        hv[t] = v;
    }
}

template __global__ void bt_solve_sym_kernel<8>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                double*, double*, long long);
template __global__ void bt_solve_sym_kernel<16>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                 double*, double*, long long);
template __global__ void bt_solve_sym_kernel<32>(const LeafDesc*, int, int, int, const double*, int, const double*,
                                                 double*, double*, long long);

void launch_bt_solve(const LeafDesc* leaves, int Nn, int nleaves, int k, int mode, const double* f, int f_zero,
                     const double* garena, double* harena, double* u, hipStream_t s, bool packed, double* zwork,
                     long long vcs) {
    const unsigned blocks = (unsigned)(((long long)nleaves * k + 3) / 4);
    // the full blocks (Sinv) are always kept: sides without a packed kernel use them
    double* zbuf = (mode == 1) ? u : zwork;
    if (packed && Nn == 8)
        hipLaunchKernelGGL(bt_solve_sym_kernel<8>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f, f_zero,
                           garena, harena, zbuf, vcs);
    else if (packed && Nn == 16)
        hipLaunchKernelGGL(bt_solve_sym_kernel<16>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f,
                           f_zero, garena, harena, zbuf, vcs);
    else if (packed && Nn == 32)
        hipLaunchKernelGGL(bt_solve_sym_kernel<32>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f,
                           f_zero, garena, harena, zbuf, vcs);
    else if (Nn == 8)
        hipLaunchKernelGGL(bt_solve_wave_kernel<8>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f, f_zero,
                           garena, harena, u, vcs);
    else if (Nn == 16)
        hipLaunchKernelGGL(bt_solve_wave_kernel<16>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f,
                           f_zero, garena, harena, u, vcs);
    else if (Nn == 32)
        hipLaunchKernelGGL(bt_solve_wave_kernel<32>, dim3(blocks), dim3(256), 0, s, leaves, nleaves, k, mode, f,
                           f_zero, garena, harena, u, vcs);
    else
        hipLaunchKernelGGL(bt_solve_kernel, dim3(nleaves, k), dim3(256), 0, s, leaves, Nn, nleaves, k, mode, f, f_zero,
                           garena, harena, u, vcs);
}

// Leaf DtN (buildD2N, FiniteVolumeSolver.cpp:355-455) with variable coefficients: column b is
// driven by -2 c_b at the boundary cell of b; Z (N blocks of N x 4N per leaf) is global work.
__global__ void __launch_bounds__(256) bt_dtn_kernel(const LeafDesc* leaves, int Nn, int first, double* work) {
    const LeafDesc L = leaves[first + blockIdx.x];
    const int tid = threadIdx.x;
    const int NN = Nn * Nn, NB = 4 * Nn;
    double* Z = work + (long long)blockIdx.x * Nn * Nn * NB;  // block i: Z + i*N*4N, row r, col b
    // right-hand side block rows, then forward substitution, all 4N columns at once
    for (int i = 0; i < Nn; i++) {
        double* Zi = Z + (long long)i * Nn * NB;
        const double* Si = L.Sinv + (long long)i * NN;
        for (int e = tid; e < Nn * NB; e += 256) {
            const int r = e / NB, bcol = e % NB;
            double acc = 0.0;
            for (int c = 0; c < Nn; c++) {
                int ib, jb;
                bcell(bcol, Nn, ib, jb);
                double v = 0.0;
                if (ib == i && jb == c) {
                    const int sb = bcol / Nn;
                    v = -2.0 * L.stencil[sb * NN + ib * Nn + jb];
                }
                if (i > 0) v -= L.stencil[i * Nn + c] * Z[(long long)(i - 1) * Nn * NB + c * NB + bcol];
                acc += Si[r * Nn + c] * v;
            }
            Zi[r * NB + bcol] = acc;
        }
        __syncthreads();
    }
    for (int i = Nn - 2; i >= 0; i--) {
        double* Zi = Z + (long long)i * Nn * NB;
        const double* Si = L.Sinv + (long long)i * NN;
        // x_i = z_i - Sinv_i (cE_i .* x_{i+1}); each thread owns one (r, b) of block i
        for (int e = tid; e < Nn * NB; e += 256) {
            const int r = e / NB, bcol = e % NB;
            double acc = 0.0;
            for (int c = 0; c < Nn; c++)
                acc += Si[r * Nn + c] * L.stencil[NN + i * Nn + c] * Z[(long long)(i + 1) * Nn * NB + c * NB + bcol];
            Zi[r * NB + bcol] -= acc;
        }
        __syncthreads();
    }
    // T[a][b] = sgn_a (2/d_a) (u_cell(a) - delta_ab)
    for (int e = tid; e < NB * NB; e += 256) {
        const int a = e / NB, bcol = e % NB;
        int ia, ja;
        bcell(a, Nn, ia, ja);
        const double ua = Z[(long long)ia * Nn * NB + ja * NB + bcol];
        const int sa = a / Nn;
        const double da = (sa < 2) ? L.dx : L.dy;
        const double sg = (sa == W || sa == S) ? 2.0 / da : -2.0 / da;
        L.T[e] = sg * (ua - (a == bcol ? 1.0 : 0.0));
    }
}

}  // namespace efk
