// ef_tree.hip -- gfx950 kernels of the tree levels of the HPS path: the 4-to-1 merges of
// the build stage (HPSAlgorithm::merge4to1 and its steps, src/HPSAlgorithm.hpp:497-518,
// 707-1009) and the family steps of the upward and downward sweeps (upwards4to1 :529-551,
// split1to4 :562-577), batched over whole tree levels.
//
// The reference forms X, S_RHS and H element by element, calls dgesv for S = X^-1 S_RHS and
// dgemm for T = blkdiag + H S (Matrix.hpp:827-953), one family at a time.  Here:
//  * X is inverted once per merge (blocked Gauss-Jordan with partial pivoting, every O(m^3)
//    step a batched MFMA product), so the upward sweep's mergeW_ is a product with X^-1
//    instead of a re-factorisation per solve;
//  * S = X^-1 S_RHS and T = blkdiag + H S read S_RHS and H straight out of the children's
//    DtN maps (they are blocks of them, with signs), skipping their zero blocks;
//  * every product runs on one register-resident batched GEMM kernel whose waves read their
//    operands in the v_mfma_f64_16x16x4_f64 layouts (no LDS, no scratch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "ef_host.hpp"
#include "ef_kernels.hpp"
#include "ef_tree.hpp"

namespace efk {

typedef double td4 __attribute__((ext_vector_type(4)));
typedef double td2 __attribute__((ext_vector_type(2)));

__device__ inline td4 mfma_f64(double a, double b, td4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <class T>
__device__ inline T pick4(const T (&a)[4], int c) {
    return c == 0 ? a[0] : c == 1 ? a[1] : c == 2 ? a[2] : a[3];
}

// ---------------------------------------------------------------------------------------
// Batched GEMM
// ---------------------------------------------------------------------------------------
// alpha * A B over k in [0, K) into this wave's 32 x (16 TN) block.  a0p / a1p: the lane's two
// A rows (r and r + 16, clamped into the matrix); bp[j]: B column (clamped) of column tile j.
// Within each block of 16 k values lane group q takes k = kb + 4q + s at MFMA step s: A is
// read as two 16-byte vectors per lane and row (a 128-byte row segment per lane group), B as
// one row segment of 16 consecutive columns per lane group and step.
template <int TN>
__device__ inline void gemm_segment(const double* __restrict__ a0p, const double* __restrict__ a1p,
                                    const double* const (&bp)[TN], long long ldb, int K, double alpha,
                                    td4 (&acc)[2][TN]) {
    const int q = (int)(threadIdx.x & 63) >> 4;
    int kb = 0;
    for (; kb + 16 <= K; kb += 16) {
        const int k = kb + 4 * q;
        const td2 x0 = *(const td2*)(a0p + k), x1 = *(const td2*)(a0p + k + 2);
        const td2 y0 = *(const td2*)(a1p + k), y1 = *(const td2*)(a1p + k + 2);
        const double a0[4] = {alpha * x0[0], alpha * x0[1], alpha * x1[0], alpha * x1[1]};
        const double a1[4] = {alpha * y0[0], alpha * y0[1], alpha * y1[0], alpha * y1[1]};
        double b[TN][4];
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int s = 0; s < 4; s++) b[j][s] = bp[j][(long long)(k + s) * ldb];
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                acc[0][j] = mfma_f64(a0[s], b[j][s], acc[0][j]);
                acc[1][j] = mfma_f64(a1[s], b[j][s], acc[1][j]);
            }
    }
    if (kb < K) {  // partial last block: k >= K contributes zero (addresses clamped to K - 1)
        const int k = kb + 4 * q;
        double a0[4], a1[4], b[TN][4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const bool ok = k + s < K;
            const int kk = ok ? k + s : K - 1;
            const double x = a0p[kk], y = a1p[kk];
            a0[s] = ok ? alpha * x : 0.0;
            a1[s] = ok ? alpha * y : 0.0;
#pragma unroll
            for (int j = 0; j < TN; j++) b[j][s] = bp[j][(long long)kk * ldb];
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                acc[0][j] = mfma_f64(a0[s], b[j][s], acc[0][j]);
                acc[1][j] = mfma_f64(a1[s], b[j][s], acc[1][j]);
            }
    }
}

template <int TN>
__global__ void __launch_bounds__(256) tree_gemm_kernel(const TreeGemm* __restrict__ probs,
                                                         const int* __restrict__ start, int nprob) {
    const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (w >= start[nprob]) return;  // uniform per wave
    int lo = 0, hi = nprob - 1;     // last problem whose first wave is <= w
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const TreeGemm p = probs[lo];
    const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    constexpr int CW = 16 * TN;
    const int tn = (p.N + CW - 1) / CW;
    const int t = w - start[lo];
    const int row0 = (t / tn) * 32, col0 = (t % tn) * CW;
    const int ra = min(row0 + r, p.M - 1), rb = min(row0 + 16 + r, p.M - 1);
    int cj[TN];
#pragma unroll
    for (int j = 0; j < TN; j++) cj[j] = min(col0 + 16 * j + r, p.N - 1);
    td4 acc[2][TN];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = td4{0.0, 0.0, 0.0, 0.0};
    if (p.K0 > 0) {
        const double* bp[TN];
#pragma unroll
        for (int j = 0; j < TN; j++) bp[j] = p.B0 + cj[j];
        gemm_segment<TN>(p.A0 + (long long)ra * p.lda0, p.A0 + (long long)rb * p.lda0, bp, p.ldb0, p.K0, p.alpha0, acc);
    }
    if (p.K1 > 0) {
        const double* bp[TN];
#pragma unroll
        for (int j = 0; j < TN; j++) bp[j] = p.B1 + cj[j];
        gemm_segment<TN>(p.A1 + (long long)ra * p.lda1, p.A1 + (long long)rb * p.lda1, bp, p.ldb1, p.K1, p.alpha1, acc);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = row0 + 16 * i + 4 * v + q, col = col0 + 16 * j + r;
                if (row < p.M && col < p.N) {
                    double o = acc[i][j][v];
                    if (p.C) o += p.beta * p.C[(long long)row * p.ldc + col];
                    p.D[(long long)row * p.ldd + col] = o;
                }
            }
}

template __global__ void tree_gemm_kernel<1>(const TreeGemm*, const int*, int);
template __global__ void tree_gemm_kernel<2>(const TreeGemm*, const int*, int);

// ---------------------------------------------------------------------------------------
// Gauss-Jordan inversion
// ---------------------------------------------------------------------------------------
// Tournament LU of the panel (columns k .. k+nb) with partial pivoting.  One workgroup of 1024
// threads handles up to GJ_CH rows; each row is held by two adjacent lanes, 16 panel columns
// each, in registers.  Stage 0: the rows of chunk `blockIdx.x` (positions k + ch GJ_CH + i);
// stage 1: the candidates of every chunk (a.cand).  The nb rows pivoted on, in order, are the
// chunk's candidates, or with one chunk / in stage 1 the pivots of the step (a.cfin).  Ties go
// to the lower position, as idamax takes the first maximum.  Register indices are compile-time
// constants (one instantiation of panel_column per column), so the rows never leave registers.
constexpr int GJ_HALF = GJ_NB / 2;
struct PanelRows {
    double v[GJ_HALF];  // columns 16 h .. 16 h + 15 of the row
    int pos;            // position of the row in the matrix
    int h;              // which half of the row
    bool act;           // row present and not yet pivoted on
};
constexpr int GJ_LEAF_THREADS = 2 * GJ_CH;
constexpr int GJ_LEAF_WAVES = GJ_LEAF_THREADS / 64;
struct PanelShared {
    double red_v[GJ_LEAF_WAVES];
    int red_p[GJ_LEAF_WAVES];
    double prow[GJ_NB];
    int piv;
};
constexpr int GJ_NONE = 0x7fffffff;

template <int J>
__device__ __attribute__((always_inline)) inline void panel_column(PanelRows& R, PanelShared& sh, int nbb, int* out) {
    constexpr int HJ = J / GJ_HALF, JC = J % GJ_HALF;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (J >= nbb) {  // uniform
        if (tid == 0) out[J] = -1;
        return;
    }
    const bool own = R.h == HJ;  // this lane holds column J of its row
    double bv = -1.0;
    int bp = GJ_NONE;
    if (R.act && own) { bv = fabs(R.v[JC]); bp = R.pos; }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int op = __shfl_xor(bp, off);
        if (ov > bv || (ov == bv && op < bp)) { bv = ov; bp = op; }
    }
    if (lane == 0) { sh.red_v[wave] = bv; sh.red_p[wave] = bp; }
    __syncthreads();
    if (tid == 0) {
        double best = sh.red_v[0];
        int bpos = sh.red_p[0];
        for (int w2 = 1; w2 < GJ_LEAF_WAVES; w2++)
            if (sh.red_v[w2] > best || (sh.red_v[w2] == best && sh.red_p[w2] < bpos)) { best = sh.red_v[w2]; bpos = sh.red_p[w2]; }
        sh.piv = bpos;
        out[J] = (bpos == GJ_NONE) ? -1 : bpos;
    }
    __syncthreads();
    const int pp = sh.piv;
    if (R.act && R.pos == pp) {  // both lanes of the pivot row publish their halves
#pragma unroll
        for (int c = 0; c < GJ_HALF; c++) sh.prow[GJ_HALF * R.h + c] = R.v[c];
        R.act = false;
    }
    __syncthreads();
    // multiplier of the row: computed by the lane holding column J, passed to its partner
    double lown = 0.0;
    if (pp != GJ_NONE && R.act && own) {
        const double piv = sh.prow[J];
        lown = R.v[JC] * ((piv != 0.0) ? 1.0 / piv : 0.0);
    }
    const double lpart = __shfl_xor(lown, 1);
    const double l = own ? lown : lpart;
    if (pp != GJ_NONE && R.act) {
#pragma unroll
        for (int c = 0; c < GJ_HALF; c++) {
            const int col = GJ_HALF * R.h + c;
            if (col > J) R.v[c] -= l * sh.prow[col];
        }
    }
    __syncthreads();  // prow, red_* and piv are rewritten by the next column
}

template <int... J>
__device__ __attribute__((always_inline)) inline void panel_columns(std::integer_sequence<int, J...>, PanelRows& R,
                                                                    PanelShared& sh, int nbb, int* out) {
    (panel_column<J>(R, sh, nbb, out), ...);
}

template <int STAGE>
__global__ void __launch_bounds__(GJ_LEAF_THREADS) gj_panel_leaf_kernel(GJArgs a) {
    __shared__ PanelShared sh;
    const int b = blockIdx.y, ch = blockIdx.x;
    const int m = a.m, k = a.k, nbb = min(GJ_NB, m - k);
    const int tid = threadIdx.x;
    const double* W = a.Wa + (long long)b * m * m;
    PanelRows R;
    R.h = tid & 1;
    const int i = tid >> 1;  // row of the chunk
    int* out;
    if (STAGE == 0) {
        R.pos = k + ch * GJ_CH + i;
        R.act = R.pos < m;
        out = (a.nch == 1) ? a.cfin + (long long)b * GJ_NB : a.cand + ((long long)b * GJ_MAXCH + ch) * GJ_NB;
    } else {
        const int p = (i < a.nch * GJ_NB) ? a.cand[(long long)b * GJ_MAXCH * GJ_NB + i] : -1;
        R.act = p >= k && p < m;
        R.pos = R.act ? p : GJ_NONE;
        out = a.cfin + (long long)b * GJ_NB;
    }
    const double* rowp = W + (long long)(R.act ? R.pos : k) * m + k + GJ_HALF * R.h;
#pragma unroll
    for (int c = 0; c < GJ_HALF; c++) {
        R.v[c] = 0.0;
        if (R.act && GJ_HALF * R.h + c < nbb) R.v[c] = rowp[c];  // columns k + 16h + c < m of a row of the matrix
    }
    panel_columns(std::make_integer_sequence<int, GJ_NB>{}, R, sh, nbb, out);
}

template __global__ void gj_panel_leaf_kernel<0>(GJArgs);
template __global__ void gj_panel_leaf_kernel<1>(GJArgs);

// LU (in the order chosen) of the pivot block, P^-1, interchanges and the position -> row map
// of the step.  One workgroup per matrix.
__global__ void __launch_bounds__(256) gj_panel_final_kernel(GJArgs a) {
    __shared__ double pb[GJ_NB][GJ_NB + 1];
    __shared__ int piv[GJ_NB];
    __shared__ double Li[GJ_NB][GJ_NB + 1], Ui[GJ_NB][GJ_NB + 1];
    __shared__ int top_row[GJ_NB], opos[GJ_NB], orow[GJ_NB];
    __shared__ int nout_s;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int m = a.m, k = a.k, nbb = min(GJ_NB, m - k);
    const double* W = a.Wa + (long long)b * m * m;
    if (tid < GJ_NB) {
        const int p = (tid < nbb) ? a.cfin[(long long)b * GJ_NB + tid] : -1;
        piv[tid] = (p >= k && p < m) ? p : -1;
    }
    __syncthreads();
    for (int e = tid; e < GJ_NB * GJ_NB; e += 256) {
        const int i = e / GJ_NB, c = e % GJ_NB;
        double x = 0.0;
        if (i < nbb && c < nbb && piv[i] >= 0) x = W[(long long)piv[i] * m + k + c];
        pb[i][c] = x;
    }
    __syncthreads();
    // LU of the pivot rows in pivot order (the order partial pivoting chose them in)
    for (int j = 0; j < nbb; j++) {
        const double pj = pb[j][j];
        const double inv = (pj != 0.0) ? 1.0 / pj : 0.0;
        if (tid > j && tid < nbb) pb[tid][j] *= inv;
        __syncthreads();
        for (int e = tid; e < GJ_NB * GJ_NB; e += 256) {
            const int i = e / GJ_NB, c = e % GJ_NB;
            if (i > j && i < nbb && c > j && c < nbb) pb[i][c] -= pb[i][j] * pb[j][c];
        }
        __syncthreads();
    }
    // inverses of L11 (unit lower) and U11 (upper, with the diagonal), column c per thread
    if (tid < GJ_NB) {
        const int c = tid;
        for (int i = 0; i < GJ_NB; i++) { Li[i][c] = 0.0; Ui[i][c] = 0.0; }
        if (c < nbb) {
            Li[c][c] = 1.0;
            for (int i = c + 1; i < nbb; i++) {
                double s = 0.0;
                for (int t = c; t < i; t++) s += pb[i][t] * Li[t][c];
                Li[i][c] = -s;
            }
            const double ucc = pb[c][c];
            Ui[c][c] = (ucc != 0.0) ? 1.0 / ucc : 0.0;
            for (int i = c - 1; i >= 0; i--) {
                double s = 0.0;
                for (int t = i + 1; t <= c; t++) s += pb[i][t] * Ui[t][c];
                const double uii = pb[i][i];
                Ui[i][c] = (uii != 0.0) ? -s / uii : 0.0;
            }
        }
    }
    __syncthreads();
    double* Pinv = a.Pinv + (long long)b * GJ_NB * GJ_NB;
    for (int e = tid; e < GJ_NB * GJ_NB; e += 256) {
        const int i = e / GJ_NB, c = e % GJ_NB;
        double s = 0.0;
        if (i < nbb && c < nbb)
            for (int t = max(i, c); t < nbb; t++) s += Ui[i][t] * Li[t][c];
        Pinv[e] = s;
    }
    // interchanges (LAPACK order: position k+j with the chosen row's current position)
    if (tid == 0) {
        int* prow = a.prow + (long long)b * GJ_NB;
        int* ipiv = a.ipiv + (long long)b * m;
        int nout = 0, zero = 0;
        for (int j = 0; j < GJ_NB; j++) top_row[j] = k + j;
        for (int j = 0; j < GJ_NB; j++) {
            const int target = (j < nbb) ? piv[j] : -1;
            prow[j] = (target >= 0) ? target : min(k + j, m - 1);
            if (j >= nbb) continue;
            if (zero == 0 && (target < 0 || pb[j][j] == 0.0)) zero = j + 1;
            const int apos = k + j;
            if (target < 0) { ipiv[apos] = apos; continue; }
            int t = -1;
            for (int jj = j; jj < nbb && t < 0; jj++)
                if (top_row[jj] == target) t = k + jj;
            for (int o = 0; o < nout && t < 0; o++)
                if (orow[o] == target) t = opos[o];
            if (t < 0) t = target;
            const int row_a = top_row[j];
            if (t != apos) {
                top_row[j] = target;
                if (t < k + nbb) top_row[t - k] = row_a;
                else {
                    int o = 0;
                    while (o < nout && opos[o] != t) o++;
                    if (o == nout && nout < GJ_NB) { opos[nout] = t; nout++; }
                    if (o < GJ_NB) orow[o] = row_a;
                }
            }
            ipiv[apos] = t;
        }
        nout_s = nout;
        if (zero && a.info[b] == 0) a.info[b] = k + zero;
    }
    __syncthreads();
    int* src = a.src + (long long)b * m;
    const int nout = nout_s;
    for (int i = tid; i < m; i += 256) {
        int s = i;
        if (i >= k && i < k + nbb) s = top_row[i - k];
        else
            for (int o = 0; o < nout; o++)
                if (opos[o] == i) s = orow[o];
        src[i] = s;
    }
}

// Multipliers of the step: M(i, :) = P^-1(j, :) - e_j on pivot position k + j, -A(src(i), panel) P^-1 elsewhere.
__global__ void __launch_bounds__(256) gj_mult_kernel(GJArgs a) {
    __shared__ double Pi[GJ_NB][GJ_NB + 1];
    const int b = blockIdx.y, tid = threadIdx.x;
    const int m = a.m, k = a.k, nbb = min(GJ_NB, m - k);
    const double* Pg = a.Pinv + (long long)b * GJ_NB * GJ_NB;
    for (int e = tid; e < GJ_NB * GJ_NB; e += 256) Pi[e / GJ_NB][e % GJ_NB] = Pg[e];
    __syncthreads();
    const long long e = (long long)blockIdx.x * 256 + tid;
    if (e >= (long long)m * GJ_NB) return;
    const int i = (int)(e / GJ_NB), c = (int)(e % GJ_NB);
    const int j = i - k;
    double val = 0.0;
    if (c < nbb) {
        if (j >= 0 && j < nbb) {
            val = Pi[j][c] - (j == c ? 1.0 : 0.0);
        } else {
            const int s = a.src[(long long)b * m + i];
            const double* row = a.Wa + (long long)b * m * m + (long long)s * m + k;
            double acc = 0.0;
            for (int t = 0; t < nbb; t++) acc += row[t] * Pi[t][c];
            val = -acc;
        }
    }
    a.Mm[(long long)b * m * GJ_NB + e] = val;
}

// Wb = Wa(src rows) + M Wa(pivot rows) outside the panel; the panel column block is M + I on
// the pivot positions.  One wave per 32 x 32 block of the matrix.
__global__ void __launch_bounds__(256) gj_update_kernel(GJArgs a) {
    const int b = blockIdx.y;
    const int m = a.m, k = a.k, nbb = min(GJ_NB, m - k);
    const int tpr = (m + 31) / 32;
    const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (w >= tpr * tpr) return;
    const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    const int row0 = (w / tpr) * 32, col0 = (w % tpr) * 32;
    const double* Wa = a.Wa + (long long)b * m * m;
    double* Wb = a.Wb + (long long)b * m * m;
    const double* Mm = a.Mm + (long long)b * m * GJ_NB;
    const int* src = a.src + (long long)b * m;
    const int* prow = a.prow + (long long)b * GJ_NB;
    if (col0 == k) {  // k is a multiple of 32: this block column is the panel
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = row0 + 16 * i + 4 * v + q, col = col0 + 16 * j + r;
                    if (row < m && col < m)
                        Wb[(long long)row * m + col] = Mm[(long long)row * GJ_NB + (col - k)] + (row == col ? 1.0 : 0.0);
                }
        return;
    }
    const int ra = min(row0 + r, m - 1), rb = min(row0 + 16 + r, m - 1);
    const int c0 = min(col0 + r, m - 1), c1 = min(col0 + 16 + r, m - 1);
    td4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = td4{0.0, 0.0, 0.0, 0.0};
    for (int kb = 0; kb < nbb; kb += 16) {
        const int k0 = kb + 4 * q;
        double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const bool ok = k0 + s < nbb;
            const int kk = ok ? k0 + s : nbb - 1;
            const double x = Mm[(long long)ra * GJ_NB + kk], y = Mm[(long long)rb * GJ_NB + kk];
            a0[s] = ok ? x : 0.0;
            a1[s] = ok ? y : 0.0;
            const double* brow = Wa + (long long)prow[kk] * m;
            b0[s] = brow[c0];
            b1[s] = brow[c1];
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
            acc[0][0] = mfma_f64(a0[s], b0[s], acc[0][0]);
            acc[0][1] = mfma_f64(a0[s], b1[s], acc[0][1]);
            acc[1][0] = mfma_f64(a1[s], b0[s], acc[1][0]);
            acc[1][1] = mfma_f64(a1[s], b1[s], acc[1][1]);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = row0 + 16 * i + 4 * v + q, col = col0 + 16 * j + r;
                if (row < m && col < m)
                    Wb[(long long)row * m + col] = acc[i][j][v] + Wa[(long long)src[row] * m + col];
            }
}

// Column order of the inverse: the interchanges undone from the last to the first.
__global__ void __launch_bounds__(256) gj_perm_kernel(GJArgs a) {
    __shared__ int idx[GJ_MAXM];
    const int b = blockIdx.x, tid = threadIdx.x, m = a.m;
    for (int i = tid; i < m; i += 256) idx[i] = i;
    __syncthreads();
    if (tid == 0) {
        const int* ipiv = a.ipiv + (long long)b * m;
        for (int c = m - 1; c >= 0; c--) {
            const int p = min(max(ipiv[c], 0), m - 1);
            if (p != c) {
                const int t = idx[c];
                idx[c] = idx[p];
                idx[p] = t;
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < m; i += 256) a.src[(long long)b * m + i] = idx[i];
}

// X^-1(i, j) = W(i, idx(j)), written to the operator arena.
__global__ void __launch_bounds__(256) gj_gather_kernel(GJArgs a) {
    const int b = blockIdx.y, m = a.m;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)m * m) return;
    const int i = (int)(e / m), j = (int)(e % m);
    const int c = min(max(a.src[(long long)b * m + j], 0), m - 1);
    a.Xi[(long long)b * m * m + e] = a.Wa[(long long)b * m * m + (long long)i * m + c];
}

// ---------------------------------------------------------------------------------------
// Merge assembly
// ---------------------------------------------------------------------------------------
// block (sa, sb) element (i, j) of a child DtN map (4n x 4n, WESN blocks of n)
__device__ inline double tblk(const double* T, int n, int sa, int sb, int i, int j) {
    return T[(long long)(sa * n + i) * (4 * n) + sb * n + j];
}

__global__ void __launch_bounds__(256) tree_assemble_kernel(const AsmDesc* __restrict__ descs) {
    const AsmDesc d = descs[blockIdx.y];
    const int n = d.n;
    const long long n2 = (long long)n * n;
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= 80 * n2) return;
    if (e < 16 * n2) {  // X (mergeX_, HPSAlgorithm.hpp:870-895), interface order [ag, bo, ab, go]
        const int row = (int)(e / (4 * n)), col = (int)(e % (4 * n));
        const int ri = row / n, i = row % n, ci = col / n, j = col % n;
        const int c1 = if_c1(ri), s1 = if_s1(ri), c2 = if_c2(ri), s2 = if_s2(ri);
        double v;
        if (ri == ci) {
            v = tblk(pick4(d.Tc, c1), n, s1, s1, i, j) - tblk(pick4(d.Tc, c2), n, s2, s2, i, j);
        } else {
            const int sc1 = side_on_if(c1, ci), sc2 = side_on_if(c2, ci);
            v = 0.0;
            if (sc1 >= 0) v = tblk(pick4(d.Tc, c1), n, s1, sc1, i, j);
            else if (sc2 >= 0) v = -tblk(pick4(d.Tc, c2), n, s2, sc2, i, j);
        }
        d.X[e] = v;
        return;
    }
    e -= 16 * n2;
    // T block diagonal (mergeT_ :940-968 before H S is added), WESN rows and columns
    const int row = (int)(e / (8 * n)), col = (int)(e % (8 * n));
    const int q = row / n, i = row % n, q2 = col / n, j = col % n;
    double v = 0.0;
    if (q_child(q) == q_child(q2)) v = tblk(pick4(d.Tc, q_child(q)), n, q_side(q), q_side(q2), i, j);
    d.T[e] = v;
}

// L12 of fine size nf (SpecialMatrices.hpp:136-173), column b: rows 0 (b <= 2), 2b-1 and 2b
// (b >= 1), 2b+1 and 2b+2 (b <= nc-2), nf-1 (b >= nc-3).
__global__ void __launch_bounds__(256) tree_coarsen_kernel(const CoarsenDesc* __restrict__ descs) {
    const CoarsenDesc d = descs[blockIdx.y];
    const int m = d.m, mc = m / 2;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)(4 * mc) * (4 * mc)) return;
    const int R = (int)(e / (4 * mc)), C = (int)(e % (4 * mc));
    const int sa = R / mc, ar = R % mc, sb = C / mc, bc = C % mc;
    const double* r0 = d.src + (long long)(sa * m + 2 * ar) * (4 * m) + sb * m;
    const double* r1 = r0 + 4 * m;
    auto tv = [&](int t) { return 0.5 * r0[t] + 0.5 * r1[t]; };  // L21 row ar against column t
    double s = 0.0;
    if (bc <= 2) s += (bc == 0 ? 1.40625 : bc == 1 ? -0.5625 : 0.15625) * tv(0);
    if (bc >= 1) s += 0.25 * tv(2 * bc - 1) + 0.75 * tv(2 * bc);
    if (bc <= mc - 2) s += 0.75 * tv(2 * bc + 1) + 0.25 * tv(2 * bc + 2);
    if (bc >= mc - 3) s += (bc == mc - 1 ? 1.40625 : bc == mc - 2 ? -0.5625 : 0.15625) * tv(m - 1);
    d.dst[e] = s;
}

// ---------------------------------------------------------------------------------------
// Sweeps
// ---------------------------------------------------------------------------------------
// entry i of side `side` of a child's h (side length clen / 4), coarsened `tag` times by L21
// (pairwise averages, nested as coarsenUpwards_ applies them one step at a time)
__device__ inline double coarse_h(const double* V, long long off, int clen, int tag, int side, int i, int k, int rhs) {
    const int sl = clen / 4;
    const long long base = off + (long long)side * sl;
    if (tag == 0) return V[(base + i) * k + rhs];
    const int w = 1 << tag;
    auto x = [&](int t) { return V[(base + (long long)i * w + t) * k + rhs]; };
    if (tag == 1) return 0.5 * x(0) + 0.5 * x(1);
    double s = 0.0;  // tag >= 2: blocks of four in the nested order, then their mean
    for (int t = 0; t < w; t += 4) {
        const double p = 0.5 * x(t) + 0.5 * x(t + 1), u = 0.5 * x(t + 2) + 0.5 * x(t + 3);
        s += 0.5 * p + 0.5 * u;
    }
    return s * (4.0 / (double)w);
}

__global__ void __launch_bounds__(256) tree_up_gather_kernel(const UpDesc* __restrict__ descs, double* V, int k) {
    const UpDesc d = descs[blockIdx.y];
    const int n = d.n;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= 12LL * n * k) return;
    const int rhs = (int)(e % k);
    const int row = (int)(e / k);
    if (row < 4 * n) {  // hDiff of mergeW_ (HPSAlgorithm.hpp:1059-1087)
        const int r = row / n, i = row % n;
        const int c1 = if_c1(r), c2 = if_c2(r);
        const double h2 = coarse_h(V, pick4(d.ch_off, c2), pick4(d.clen, c2), pick4(d.tag, c2), if_s2(r), i, k, rhs);
        const double h1 = coarse_h(V, pick4(d.ch_off, c1), pick4(d.clen, c1), pick4(d.tag, c1), if_s1(r), i, k, rhs);
        V[(d.hd_off + row) * k + rhs] = h2 - h1;
    } else {  // boundary part of mergeH_ (:1098-1117), WESN order (:1128-1150)
        const int rr = row - 4 * n, q = rr / n, i = rr % n, c = q_child(q);
        V[(d.h_off + rr) * k + rhs] = coarse_h(V, pick4(d.ch_off, c), pick4(d.clen, c), pick4(d.tag, c), q_side(q), i, k, rhs);
    }
}

__global__ void __launch_bounds__(256) tree_uncoarsen_kernel(const VecStepDesc* __restrict__ descs, double* V, int k) {
    const VecStepDesc d = descs[blockIdx.y];
    const int nf = d.nfine, nc = nf / 2;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= 4LL * nf * k) return;
    const int rhs = (int)(e % k), row = (int)(e / k);
    const int side = row / nf, i = row % nf;
    const double* src = V + (d.src_off + (long long)side * nc) * k + rhs;
    auto g = [&](int j) { return src[(long long)j * k]; };
    double s;
    if (i == 0) s = 1.40625 * g(0) + (-0.5625) * g(1) + 0.15625 * g(2);
    else if (i == nf - 1) s = 0.15625 * g(nc - 3) + (-0.5625) * g(nc - 2) + 1.40625 * g(nc - 1);
    else if (i & 1) { const int j = (i - 1) / 2; s = 0.75 * g(j) + 0.25 * g(j + 1); }
    else { const int j = (i - 2) / 2; s = 0.25 * g(j) + 0.75 * g(j + 1); }
    V[(d.dst_off + row) * k + rhs] = s;
}

// children's g (applyS_, HPSAlgorithm.hpp:1194-1230): alpha = [aW, ab, aS, ag], beta = [ab, bE,
// bS, bo], gamma = [gW, go, ag, gN], omega = [go, oE, bo, oN].  A child's side lies either on
// tau's boundary (block q of tau's g) or on an interface (block r of the interface potential).
__device__ inline int child_src_q(int c, int s) {
    // side s of child c on tau's boundary: the WESN block q of (c, s), or -1
    const bool west = (c & 1) == 0, south = (c >> 1) == 0;
    if (s == W && west) return c == 0 ? 0 : 1;
    if (s == E && !west) return c == 1 ? 2 : 3;
    if (s == S && south) return c == 0 ? 4 : 5;
    if (s == N && !south) return c == 2 ? 6 : 7;
    return -1;
}
__device__ inline int child_src_r(int c, int s) {
    for (int r = 0; r < 4; r++)
        if (side_on_if(c, r) == s) return r;
    return 0;
}

__global__ void __launch_bounds__(256) tree_down_scatter_kernel(const DownDesc* __restrict__ descs, double* V, int k) {
    const DownDesc d = descs[blockIdx.y];
    const int n = d.n;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= 16LL * n * k) return;
    const int rhs = (int)(e % k), row = (int)(e / k);
    const int c = row / (4 * n), rr = row % (4 * n);
    const int s = rr / n, i = rr % n;
    const int q = child_src_q(c, s);
    double v;
    if (q >= 0) v = V[(d.g_off + (long long)q * n + i) * k + rhs];
    else v = V[(d.ui_off + (long long)child_src_r(c, s) * n + i) * k + rhs];
    V[(pick4(d.cg_off, c) + rr) * k + rhs] = v;
}

// g = r / a: the LU solve of diag(a) in solveStage (HPSAlgorithm.hpp:402-420); r is [rhs][m],
// g is in arena layout [row][rhs].
__global__ void __launch_bounds__(256) tree_root_dirichlet_kernel(const double* r, const double* a, double* g, int m, int k) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)m * k) return;
    const int i = (int)(e / k), c = (int)(e % k);
    g[e] = r[(long long)c * m + i] / a[i];
}

// A = diag(a) + diag(b) T and rhs = r - b h (HPSAlgorithm.hpp:402-420), rhs in arena layout.
__global__ void __launch_bounds__(256) tree_root_robin_kernel(const double* T, const double* a, const double* b, const double* r,
                                                              const double* h, double* A, double* rhs, int m, int k) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nA = (long long)m * m;
    if (e < nA) {
        const int i = (int)(e / m), j = (int)(e % m);
        A[e] = b[i] * T[e] + (i == j ? a[i] : 0.0);
    } else if (e < nA + (long long)m * k) {
        const long long t = e - nA;
        const int i = (int)(t / k), c = (int)(t % k);
        rhs[t] = r[(long long)c * m + i] - b[i] * (h ? h[t] : 0.0);
    }
}

// ---------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------
GemmLaunch plan_gemm(const std::vector<TreeGemm>& probs, int tn, DevStore& store, hipStream_t s) {
    GemmLaunch g;
    g.tn = tn;
    if (probs.empty()) return g;
    std::vector<int> start(probs.size() + 1, 0);
    const int cw = 16 * tn;
    long long total = 0;
    for (size_t i = 0; i < probs.size(); i++) {
        const TreeGemm& p = probs[i];
        start[i] = (int)total;
        if (p.M <= 0 || p.N <= 0) continue;
        // extents: the kernel clamps rows to M-1, columns to N-1 and k to K-1
        check_matrix(p.D, p.M, p.N, p.ldd, "GEMM D");
        if (p.C) check_matrix(p.C, p.M, p.N, p.ldc, "GEMM C");
        const double* As[2] = {p.A0, p.A1};
        const double* Bs[2] = {p.B0, p.B1};
        const int lda[2] = {p.lda0, p.lda1}, ldb[2] = {p.ldb0, p.ldb1}, K[2] = {p.K0, p.K1};
        for (int sg = 0; sg < 2; sg++) {
            if (K[sg] <= 0) continue;
            check_matrix(As[sg], p.M, K[sg], lda[sg], "GEMM A");
            check_matrix(Bs[sg], K[sg], p.N, ldb[sg], "GEMM B");
            if (((uintptr_t)As[sg] & 15) != 0 || (lda[sg] & 1) != 0)
                throw EfError(EF_ERR_INTERNAL, "GEMM A operand not 16-byte aligned with an even leading dimension");
        }
        total += (long long)((p.M + 31) / 32) * ((p.N + cw - 1) / cw);
        if (total > 0x7fffffffLL) throw EfError(EF_ERR_INTERNAL, "GEMM launch too large");
    }
    start[probs.size()] = (int)total;
    g.nprob = (int)probs.size();
    g.nwaves = (int)total;
    g.probs = store.put(probs, s);
    g.start = store.put(start, s);
    return g;
}

void run_gemm(const GemmLaunch& g, hipStream_t s) {
    if (g.nwaves == 0) return;
    const dim3 grid(blocks_for(g.nwaves, 4)), block(256);
    if (g.tn == 1) EF_LAUNCH("tree_gemm_kernel<1>", tree_gemm_kernel<1>, grid, block, s, g.probs, g.start, g.nprob);
    else EF_LAUNCH("tree_gemm_kernel<2>", tree_gemm_kernel<2>, grid, block, s, g.probs, g.start, g.nprob);
}

size_t gj_work_doubles(int m, int B) {
    return (size_t)B * ((size_t)m * m + (size_t)GJ_NB * GJ_NB + (size_t)m * GJ_NB);
}
size_t gj_work_ints(int m, int B) {
    return (size_t)B * ((size_t)GJ_MAXCH * GJ_NB + GJ_NB + 2 * (size_t)m + GJ_NB);
}

void run_gj_inverse(double* wa, double* xi, int m, int B, const GJWork& w, int* info, hipStream_t s) {
    if (B <= 0) return;
    if (m < 1 || m > GJ_MAXM)
        throw EfError(EF_ERR_ARG, "interface system of order " + std::to_string(m) + " exceeds the inversion limit " +
                                      std::to_string(GJ_MAXM));
    const size_t mm = (size_t)m * m * B;
    check_range(wa, mm * 8, "GJ matrices");
    check_range(w.wb, mm * 8, "GJ second buffer");
    check_range(xi, mm * 8, "GJ inverse output");
    check_range(w.cand, (size_t)B * GJ_MAXCH * GJ_NB * 4, "GJ candidates");
    check_range(w.cfin, (size_t)B * GJ_NB * 4, "GJ pivots");
    check_range(w.src, (size_t)B * m * 4, "GJ row map");
    check_range(w.ipiv, (size_t)B * m * 4, "GJ interchanges");
    check_range(w.prow, (size_t)B * GJ_NB * 4, "GJ pivot rows");
    check_range(w.Pinv, (size_t)B * GJ_NB * GJ_NB * 8, "GJ pivot-block inverse");
    check_range(w.Mm, (size_t)B * m * GJ_NB * 8, "GJ multipliers");
    check_range(info, (size_t)B * 4, "GJ info");
    GJArgs a{};
    a.cand = w.cand; a.cfin = w.cfin; a.src = w.src; a.ipiv = w.ipiv; a.prow = w.prow; a.info = info;
    a.Pinv = w.Pinv; a.Mm = w.Mm; a.Xi = xi;
    a.m = m; a.B = B;
    double* cur = wa;
    double* nxt = w.wb;
    const int tpr = (m + 31) / 32;
    for (int k = 0; k < m; k += GJ_NB) {
        a.Wa = cur; a.Wb = nxt; a.k = k;
        a.nch = (m - k + GJ_CH - 1) / GJ_CH;
        EF_LAUNCH("gj_panel_leaf_kernel<0>", gj_panel_leaf_kernel<0>, dim3(a.nch, B), dim3(GJ_LEAF_THREADS), s, a);
        if (a.nch > 1)
            EF_LAUNCH("gj_panel_leaf_kernel<1>", gj_panel_leaf_kernel<1>, dim3(1, B), dim3(GJ_LEAF_THREADS), s, a);
        EF_LAUNCH("gj_panel_final_kernel", gj_panel_final_kernel, dim3(B), dim3(256), s, a);
        EF_LAUNCH("gj_mult_kernel", gj_mult_kernel, dim3(blocks_for((long long)m * GJ_NB), B), dim3(256), s, a);
        EF_LAUNCH("gj_update_kernel", gj_update_kernel, dim3(blocks_for((long long)tpr * tpr, 4), B), dim3(256), s, a);
        std::swap(cur, nxt);
    }
    a.Wa = cur; a.Wb = nxt; a.k = 0;
    EF_LAUNCH("gj_perm_kernel", gj_perm_kernel, dim3(B), dim3(256), s, a);
    EF_LAUNCH("gj_gather_kernel", gj_gather_kernel, dim3(blocks_for((long long)m * m), B), dim3(256), s, a);
}

}  // namespace efk
