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
#include <type_traits>
#include <utility>
#include <vector>

#include "ef_host.hpp"
#include "ef_kernels.hpp"
#include "ef_tree.hpp"

namespace efk {

typedef double td4 __attribute__((ext_vector_type(4)));
typedef double td2 __attribute__((ext_vector_type(2)));
// loads through the global address space: the descriptors' pointers are generic, and a generic
// (FLAT) load also counts against the LDS counter, so the compiler waits for every load in
// flight before each use -- which defeats the GEMM's software pipeline
__device__ inline td2 gld2(const double* p) { return *(const EF_GLOBAL td2*)gmem(p); }
__device__ inline double gld(const double* p) { return *gmem(p); }

typedef float tf4 __attribute__((ext_vector_type(4)));

__device__ inline td4 mfma_f64(double a, double b, td4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// v_mfma_f32_16x16x4_f32: the same operand layout (lane l: A[l%16][l/16], B[l/16][l%16]);
// the accumulator holds rows 4 (l/16) + v of column l%16 (the f64 instruction: rows 4 v + l/16)
__device__ inline tf4 mfma_f32(double a, double b, tf4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, c, 0, 0, 0);
}
template <bool F32>
using GemmAcc = std::conditional_t<F32, tf4, td4>;
template <bool F32>
__device__ inline GemmAcc<F32> mfma_acc(double a, double b, GemmAcc<F32> c) {
    if constexpr (F32) return mfma_f32(a, b, c);
    else return mfma_f64(a, b, c);
}

// lane `src`'s value of x in every lane (src uniform)
__device__ inline double readlane_f64(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}



// ---------------------------------------------------------------------------------------
// Batched GEMM
// ---------------------------------------------------------------------------------------
// alpha * A B over k in [0, K) into this wave's 32 x (16 TN) block.  a0p / a1p: the lane's two
// A rows (r and r + 16, clamped into the matrix); bp[j]: B column (clamped) of column tile j.
// Within each block of 16 k values lane group q takes k = kb + 4q + s at MFMA step s: A is
// read as two 16-byte vectors per lane and row (a 128-byte row segment per lane group), B as
// one row segment of 16 consecutive columns per lane group and step.
// block-gather mode of segment 0: row offsets of each block's segment (hDiff: the two children's
// segments, subtracted; g: the side's segment alone)
// (MODE 1 subtracts the d1 segment, MODE 2 reads the d2 segment alone; MODE 0: B contiguous
// along k, MODE 3: B strided along k)
struct DiffRows {
    int dn, kbase;  // block size; absolute row of the segment's first k
    long long d2[4], d1[4];
};

template <int RT, int TN, bool F32, int MODE>
__device__ inline void gemm_segment(const double* __restrict__ a0p, const double* __restrict__ a1p,
                                    const double* const (&bp)[TN], long long ldb, int K, double alpha,
                                    GemmAcc<F32> (&acc)[2][TN], const DiffRows& dr) {
    constexpr bool DIFF = MODE == 1 || MODE == 2;
    const int q = (int)(threadIdx.x & 63) >> 4;
    // hDiff rows: absolute row ka lies in interface block r = ka / dn; both children's offsets
    auto drow = [&](int ka, long long& o2, long long& o1) {
        const int r = ka / dr.dn, i = ka - r * dr.dn;
        o2 = pick4(dr.d2, r) + i;
        o1 = pick4(dr.d1, r) + i;
    };
    // the operands of block kb: A rows (two 16-byte vectors per row) and B columns
    // MODE 1 keeps both children's rows and subtracts them in the products: a subtraction at
    // load time would wait for the loads as soon as they are issued
    struct Blk {
        td2 x0, x1, y0, y1;
        double b[TN][4];
        double c[MODE == 1 ? TN : 1][4];
    };
    auto load = [&](int kb, Blk& o) {
        const int k = kb + 4 * q;
        o.x0 = gld2(a0p + k);
        o.x1 = gld2(a0p + k + 2);
        if (RT == 2) {
            o.y0 = gld2(a1p + k);
            o.y1 = gld2(a1p + k + 2);
        }
        if (DIFF) {  // four consecutive rows lie in one block (dn % 4 == 0)
            long long o2, o1;
            drow(dr.kbase + k, o2, o1);
            if (MODE == 1) {
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const td2 u0 = gld2(bp[j] + o2), u1 = gld2(bp[j] + o2 + 2);
                    const td2 v0 = gld2(bp[j] + o1), v1 = gld2(bp[j] + o1 + 2);
                    o.b[j][0] = u0[0]; o.b[j][1] = u0[1]; o.b[j][2] = u1[0]; o.b[j][3] = u1[1];
                    o.c[j][0] = v0[0]; o.c[j][1] = v0[1]; o.c[j][2] = v1[0]; o.c[j][3] = v1[1];
                }
            } else {
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    const td2 u0 = gld2(bp[j] + o2), u1 = gld2(bp[j] + o2 + 2);
                    o.b[j][0] = u0[0]; o.b[j][1] = u0[1]; o.b[j][2] = u1[0]; o.b[j][3] = u1[1];
                }
            }
        } else if (MODE == 0) {  // B contiguous along k: two 16-byte vectors per column
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const td2 u0 = gld2(bp[j] + k), u1 = gld2(bp[j] + k + 2);
                o.b[j][0] = u0[0]; o.b[j][1] = u0[1]; o.b[j][2] = u1[0]; o.b[j][3] = u1[1];
            }
        } else {
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int s = 0; s < 4; s++) o.b[j][s] = gld(bp[j] + (long long)(k + s) * ldb);
        }
    };
    auto mma = [&](const Blk& o) {
        const double a0[4] = {alpha * o.x0[0], alpha * o.x0[1], alpha * o.x1[0], alpha * o.x1[1]};
        double a1[4] = {0.0, 0.0, 0.0, 0.0};
        if (RT == 2) {
            a1[0] = alpha * o.y0[0]; a1[1] = alpha * o.y0[1]; a1[2] = alpha * o.y1[0]; a1[3] = alpha * o.y1[1];
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const double bv = MODE == 1 ? o.b[j][s] - o.c[MODE == 1 ? j : 0][s] : o.b[j][s];
                acc[0][j] = mfma_acc<F32>(a0[s], bv, acc[0][j]);
                if (RT == 2) acc[1][j] = mfma_acc<F32>(a1[s], bv, acc[1][j]);
            }
    };
    const int nfull = K / 16;
    int kb = 0;
    if (nfull > 0) {
        // software pipeline of depth PD: block b lives in buf[b % PD]; while block b's MFMAs run,
        // blocks b + 1 .. b + PD - 1 are in flight, and block b + PD is issued into the slot b
        // frees -- no buffer is copied to another (a copy waits for its loads).  Measured at 64
        // right-hand sides, adaptive Poisson sweep GEMMs: PD = 2 0.871 ms, 3 0.969 ms, 4 0.996 ms
        // (more registers, fewer waves per SIMD); a cur/next pair with a copy 0.948 ms.
        // The steady-state loop issues its loads unconditionally, so the loop head sees one
        // order of outstanding loads and waits for the older block only (a conditional load
        // there makes the wait counts cover both blocks).
        constexpr int PD = 2;
        Blk buf[PD];
        if (nfull >= PD) {
#pragma unroll
            for (int j = 0; j < PD; j++) load(16 * j, buf[j]);
            int i = 0;
            for (; i + 2 * PD <= nfull; i += PD) {
#pragma unroll
                for (int j = 0; j < PD; j++) {
                    mma(buf[j]);
                    load(16 * (i + j + PD), buf[j]);
                }
            }
            // PD <= nfull - i < 2 PD: one round with the last loads, then the blocks left
#pragma unroll
            for (int j = 0; j < PD; j++) {
                mma(buf[j]);
                if (i + j + PD < nfull) load(16 * (i + j + PD), buf[j]);
            }
            i += PD;
#pragma unroll
            for (int j = 0; j < PD - 1; j++)
                if (i + j < nfull) mma(buf[j]);
        } else {
            load(0, buf[0]);
            mma(buf[0]);
        }
        kb = 16 * nfull;
    }
    if (kb < K) {  // partial last block: k >= K contributes zero (addresses clamped to K - 1)
        const int k = kb + 4 * q;
        double a0[4], a1[4], b[TN][4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const bool ok = k + s < K;
            const int kk = ok ? k + s : K - 1;
            const double x = gld(a0p + kk);
            a0[s] = ok ? alpha * x : 0.0;
            if (RT == 2) {
                const double y = gld(a1p + kk);
                a1[s] = ok ? alpha * y : 0.0;
            }
            if (DIFF) {
                long long o2, o1;
                drow(dr.kbase + kk, o2, o1);
#pragma unroll
                for (int j = 0; j < TN; j++) b[j][s] = MODE == 1 ? gld(bp[j] + o2) - gld(bp[j] + o1) : gld(bp[j] + o2);
            } else {
#pragma unroll
                for (int j = 0; j < TN; j++) b[j][s] = gld(bp[j] + (long long)kk * ldb);
            }
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                acc[0][j] = mfma_acc<F32>(a0[s], b[j][s], acc[0][j]);
                if (RT == 2) acc[1][j] = mfma_acc<F32>(a1[s], b[j][s], acc[1][j]);
            }
    }
}

template <int RT, int TN, bool F32>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
tree_gemm_kernel(const TreeGemm* __restrict__ probs,
                                                         const int* __restrict__ start,
                                                         const int* __restrict__ wprob, int nwaves) {
    const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (w >= nwaves) return;  // uniform per wave
    const int lo = __builtin_amdgcn_readfirstlane(wprob[w]);  // the wave's problem (a table, not a search)
    const TreeGemm p = probs[lo];
    const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    constexpr int RW = 16 * RT, CW = 16 * TN;
    const int tm = (p.M + RW - 1) / RW, tn = (p.N + CW - 1) / CW;
    const int t = w - start[lo];
    const int row0 = (t / tn) * RW, col0 = (t % tn) * CW;
    int rows[RT];
#pragma unroll
    for (int i = 0; i < RT; i++) rows[i] = min(row0 + 16 * i + r, p.M - 1);
    int cj[TN];
#pragma unroll
    for (int j = 0; j < TN; j++) cj[j] = min(col0 + 16 * j + r, p.N - 1);
    GemmAcc<F32> acc[2][TN];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = GemmAcc<F32>{0, 0, 0, 0};
    // this wave sums both segments over their whole k range (split-K launches run on
    // tree_gemm_split_kernel)
    const int kb0 = 0, ke0 = p.K0, kb1 = 0, ke1 = p.K1;
    // C (when the product adds to it) loaded before the products, so its latency overlaps
    // them; rows and columns clamped as the stores are guarded
    double cv[RT][TN][4];
    const bool has_c = p.C != nullptr;
    if (has_c) {
#pragma unroll
        for (int i = 0; i < RT; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = min(row0 + 16 * i + (F32 ? 4 * q + v : 4 * v + q), p.M - 1);
                    const int col = min(col0 + 16 * j + r, p.N - 1);
                    cv[i][j][v] = gld(p.C + (long long)row * p.ldc + (long long)col * p.ccs);
                }
    }
    DiffRows dr;
    dr.dn = p.dn;
    dr.kbase = kb0;
#pragma unroll
    for (int r4 = 0; r4 < 4; r4++) { dr.d2[r4] = p.d2[r4]; dr.d1[r4] = p.d1[r4]; }
    if (ke0 > kb0) {
        const double* bp[TN];
        if (p.dn > 0) {  // hDiff mode: column bases, rows resolved per interface block
#pragma unroll
            for (int j = 0; j < TN; j++) bp[j] = p.B0 + (long long)cj[j] * p.bcs;
            const double* a0 = p.A0 + (long long)rows[0] * p.lda0 + kb0;
            const double* a1 = p.A0 + (long long)rows[RT - 1] * p.lda0 + kb0;
            if (p.dsub) gemm_segment<RT, TN, F32, 1>(a0, a1, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
            else gemm_segment<RT, TN, F32, 2>(a0, a1, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
        } else {
#pragma unroll
            for (int j = 0; j < TN; j++) bp[j] = p.B0 + (long long)kb0 * p.ldb0 + (long long)cj[j] * p.bcs;
            const double* a0 = p.A0 + (long long)rows[0] * p.lda0 + kb0;
            const double* a1 = p.A0 + (long long)rows[RT - 1] * p.lda0 + kb0;
            if (p.ldb0 == 1) gemm_segment<RT, TN, F32, 0>(a0, a1, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
            else gemm_segment<RT, TN, F32, 3>(a0, a1, bp, p.ldb0, ke0 - kb0, p.alpha0, acc, dr);
        }
    }
    if (ke1 > kb1) {
        const double* bp[TN];
#pragma unroll
        for (int j = 0; j < TN; j++) bp[j] = p.B1 + (long long)kb1 * p.ldb1 + (long long)cj[j] * p.bcs;
        const double* a0 = p.A1 + (long long)rows[0] * p.lda1 + kb1;
        const double* a1 = p.A1 + (long long)rows[RT - 1] * p.lda1 + kb1;
        if (p.ldb1 == 1) gemm_segment<RT, TN, F32, 0>(a0, a1, bp, 1, ke1 - kb1, p.alpha1, acc, dr);
        else gemm_segment<RT, TN, F32, 3>(a0, a1, bp, p.ldb1, ke1 - kb1, p.alpha1, acc, dr);
    }
    // with RT = 1 both accumulator rows hold the same 16 rows; only the first is stored
#pragma unroll
    for (int i = 0; i < RT; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = row0 + 16 * i + (F32 ? 4 * q + v : 4 * v + q), col = col0 + 16 * j + r;
                if (row < p.M && col < p.N) {
                    double o = (double)acc[i][j][v];
                    if (has_c) o += p.beta * cv[i][j][v];
                    *gmem(p.D + (long long)row * p.ldd + (long long)col * p.dcs) = o;
                }
            }
}

template __global__ void tree_gemm_kernel<1, 1, false>(const TreeGemm*, const int*, const int*, int);
template __global__ void tree_gemm_kernel<1, 2, false>(const TreeGemm*, const int*, const int*, int);
template __global__ void tree_gemm_kernel<2, 2, false>(const TreeGemm*, const int*, const int*, int);
template __global__ void tree_gemm_kernel<1, 1, true>(const TreeGemm*, const int*, const int*, int);
template __global__ void tree_gemm_kernel<1, 2, true>(const TreeGemm*, const int*, const int*, int);
template __global__ void tree_gemm_kernel<2, 2, true>(const TreeGemm*, const int*, const int*, int);

// Split-K inside the workgroup: workgroup g computes the 16 x 16 output tile g - start[p] of
// problem p = gprob[g]; its wave w sums chunk w, k in [w kc, (w + 1) kc), of the concatenated k
// range of the two segments (a chunk may span both), and wave 0 adds the chunks' partials in
// chunk order, then beta C.  The chunks depend on the problem only, never on N, so a column's
// result does not depend on the batch; nothing goes through HBM between the chunks and no second
// launch sums them.  blockDim = 64 x (the launch's largest chunk count, <= 16).
__global__ void __launch_bounds__(1024) tree_gemm_split_kernel(const TreeGemm* __restrict__ probs,
                                                                const int* __restrict__ start,
                                                                const int* __restrict__ gprob, int ngroups) {
    __shared__ double part[16][4][64];  // [chunk][accumulator slot][lane], 32 KiB
    const int g = (int)blockIdx.x;
    if (g >= ngroups) return;  // uniform per workgroup
    const int lo = __builtin_amdgcn_readfirstlane(gprob[g]);
    const TreeGemm p = probs[lo];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    const int tn = (p.N + 15) / 16;
    const int t = g - start[lo];
    const int row0 = (t / tn) * 16, col0 = (t % tn) * 16;
    const int rowc = min(row0 + r, p.M - 1), cj = min(col0 + r, p.N - 1);
    GemmAcc<false> acc[2][1];
    acc[0][0] = GemmAcc<false>{0, 0, 0, 0};
    acc[1][0] = GemmAcc<false>{0, 0, 0, 0};
    const int K = p.K0 + p.K1, kbeg = wave * p.kc, kend = min(kbeg + p.kc, K);
    // wave 0 loads C before its products (it alone writes D)
    double cv[4] = {0.0, 0.0, 0.0, 0.0};
    if (wave == 0 && p.C) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = min(row0 + 4 * v + q, p.M - 1);
            cv[v] = gld(p.C + (long long)row * p.ldc + (long long)cj * p.ccs);
        }
    }
    if (kbeg < K) {  // uniform per wave
        const int kb0 = min(kbeg, p.K0), ke0 = min(kend, p.K0);
        const int kb1 = max(kbeg, p.K0) - p.K0, ke1 = max(kend, p.K0) - p.K0;
        DiffRows dr;
        dr.dn = p.dn;
        dr.kbase = kb0;
#pragma unroll
        for (int r4 = 0; r4 < 4; r4++) { dr.d2[r4] = p.d2[r4]; dr.d1[r4] = p.d1[r4]; }
        if (ke0 > kb0) {
            const double* bp[1];
            const double* a0 = p.A0 + (long long)rowc * p.lda0 + kb0;
            if (p.dn > 0) {
                bp[0] = p.B0 + (long long)cj * p.bcs;
                if (p.dsub) gemm_segment<1, 1, false, 1>(a0, a0, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
                else gemm_segment<1, 1, false, 2>(a0, a0, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
            } else {
                bp[0] = p.B0 + (long long)kb0 * p.ldb0 + (long long)cj * p.bcs;
                if (p.ldb0 == 1) gemm_segment<1, 1, false, 0>(a0, a0, bp, 1, ke0 - kb0, p.alpha0, acc, dr);
                else gemm_segment<1, 1, false, 3>(a0, a0, bp, p.ldb0, ke0 - kb0, p.alpha0, acc, dr);
            }
        }
        if (ke1 > kb1) {
            const double* bp[1] = {p.B1 + (long long)kb1 * p.ldb1 + (long long)cj * p.bcs};
            const double* a1 = p.A1 + (long long)rowc * p.lda1 + kb1;
            if (p.ldb1 == 1) gemm_segment<1, 1, false, 0>(a1, a1, bp, 1, ke1 - kb1, p.alpha1, acc, dr);
            else gemm_segment<1, 1, false, 3>(a1, a1, bp, p.ldb1, ke1 - kb1, p.alpha1, acc, dr);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int v = 0; v < 4; v++) part[wave][v][lane] = acc[0][0][v];
    }
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int row = row0 + 4 * v + q, col = col0 + r;
        if (row < p.M && col < p.N) {
            double o = 0.0;
            o += acc[0][0][v];
            for (int c = 1; c < p.nks; c++) o += part[c][v][lane];
            if (p.C) o += p.beta * cv[v];
            *gmem(p.D + (long long)row * p.ldd + (long long)col * p.dcs) = o;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Gauss-Jordan inversion
// ---------------------------------------------------------------------------------------
// Tournament LU of the panel (columns k .. k+nb) with partial pivoting, fused with the step's
// pivot bookkeeping.  One workgroup of 64 WAVES threads holds up to 32 WAVES panel rows, each row
// in two adjacent lanes (16 columns each) in registers.  STAGE 0: the rows of chunk blockIdx.x
// (positions k + ch * rows + i); STAGE 1: the candidates of every chunk (a.cand).  Ties go to
// the lower position, as idamax takes the first maximum.  FIN = false: the nb rows pivoted on,
// in order, are the chunk's candidates.  FIN = true (one chunk, or stage 1): they are the pivots
// of the step, and the same workgroup forms P^-1 of the pivot block from the rows it published
// (the multiplier of each eliminated entry is kept in place, so the pivot row of column J holds
// row J of [L11 \ U11]), the LAPACK-order interchanges (ipiv) and the position -> row map (src).
// Tournament LU of the panel (columns k .. k+nb) with partial pivoting, fused with the step's
// pivot bookkeeping.  One workgroup of 64 WAVES threads holds up to 64 WAVES RPL panel rows,
// RPL whole rows (32 columns) per lane in registers, row q NT + tid of the chunk in slot q.
// STAGE 0: the rows of chunk blockIdx.x (positions k + ch * rows + i); STAGE 1: the candidates
// of every chunk (a.cand).  Ties go to the lower position, as idamax takes the first maximum.
// FIN = false: the nb rows pivoted on, in order, are the chunk's candidates.  FIN = true (one
// chunk, or stage 1): they are the pivots of the step, and the same workgroup forms P^-1 of
// the pivot block, the LAPACK-order interchanges (ipiv) and the position -> row map (src).
// Per column (all 32 unrolled, so register indices are compile-time constants): each lane's
// best row, one DPP argmax per wave, the wave's best row published to LDS with its value, one
// barrier, then every thread combines the waves' candidates and reads the pivot row from the
// winner's slot (slots alternate by column parity, so column J + 1 never overwrites what a
// slower wave still reads for column J).
constexpr int GJ_NONE = 0x7fffffff;
template <int RPL>
struct PanelRows {
    double v[RPL][GJ_NB];  // the rows (multipliers left of the pivot column)
    int pos[RPL];          // position of each row at the start of the step (its name)
    int cur[RPL];          // its position now, after the step's interchanges so far
    bool act[RPL];         // row present and not yet pivoted on
};
// Pivot order key: current position, then name (both < GJ_MAXM = 2^12).
__device__ inline int piv_key(int cur, int pos) { return (cur << 12) | pos; }
template <int WAVES, bool FIN>
struct PanelShared {
    union {
        double cand[2][WAVES][GJ_NB];   // each wave's best row, by column parity
        double A[2][GJ_NB][GJ_NB];      // FIN: in-place Gauss-Jordan on P, ping-pong
    } u;
    double red_v[2][WAVES];
    int red_p[2][WAVES];
    int piv[GJ_NB];  // pivot position chosen for each column (-1: none)
    int top_row[GJ_NB], opos[GJ_NB], orow[GJ_NB];
    int nout;
};

// Wave reductions into lane 63 through DPP lane shifts: row shifts 1, 2, 4, 8 leave each
// 16-lane row's result in its lane 15; row_bcast:15 folds rows 0 -> 1 and 2 -> 3, row_bcast:31
// folds lane 31 into rows 2 and 3.  Lanes without a source lane (or in rows outside the row
// mask) see the identity.
template <int CTRL, int ROWMASK>
__device__ inline double max_dpp(double v) {  // identity -1 (every candidate is >= 0)
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(-1.0), __double2hiint(v), CTRL, ROWMASK, 0xf, false);
    return fmax(v, __hiloint2double(hi, lo));
}
template <int CTRL, int ROWMASK>
__device__ inline int min_dpp(int p) {  // identity GJ_NONE
    return min(p, __builtin_amdgcn_update_dpp(GJ_NONE, p, CTRL, ROWMASK, 0xf, false));
}
__device__ inline double wave_max(double v) {
    v = max_dpp<0x111, 0xf>(v);
    v = max_dpp<0x112, 0xf>(v);
    v = max_dpp<0x114, 0xf>(v);
    v = max_dpp<0x118, 0xf>(v);
    v = max_dpp<0x142, 0xa>(v);
    v = max_dpp<0x143, 0xc>(v);
    return v;
}
__device__ inline int wave_min(int p) {
    p = min_dpp<0x111, 0xf>(p);
    p = min_dpp<0x112, 0xf>(p);
    p = min_dpp<0x114, 0xf>(p);
    p = min_dpp<0x118, 0xf>(p);
    p = min_dpp<0x142, 0xa>(p);
    p = min_dpp<0x143, 0xc>(p);
    return p;
}
__device__ inline double readlane_dyn(double x, int src) { return readlane_f64(x, src); }

// Column J: the pivot is the largest |entry| among the active rows, the lowest current position
// among equal ones (idamax takes the first maximum of the interchanged column): a wave max of
// the values, then a wave min of the (position, name) keys of the lanes that hold that maximum.
// After the choice the row at position k + J takes the pivot row's position (the interchange).
// One wave reads the pivot row straight from its owner lane (readlane); several waves publish
// their candidate rows in LDS.
template <int J, int WAVES, int RPL, int STAGE, bool FIN>
__device__ __attribute__((always_inline)) inline void panel_column(PanelRows<RPL>& R, PanelShared<WAVES, FIN>& sh,
                                                                   int base, int top0) {
    constexpr int BUF = J & 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double lv = -1.0;  // the lane's best row
    int lk = GJ_NONE;
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const double x = R.act[q] ? fabs(R.v[q][J]) : -1.0;
        const int key = piv_key(R.cur[q], R.pos[q]);
        const bool take = x > lv || (x == lv && R.act[q] && key < lk);
        lk = take ? key : lk;
        lv = take ? x : lv;
    }
    const double wmax = readlane_f64(wave_max(lv), 63);
    const int wkey = __builtin_amdgcn_readlane(wave_min((lv == wmax) ? lk : GJ_NONE), 63);
    const int wbp = (wkey == GJ_NONE) ? GJ_NONE : (wkey & 4095);
    int pkey = wkey;
    double pr[GJ_NB];  // pivot row, columns J..
    int pp = wbp;
    if (WAVES == 1) {
        if (tid == 0) sh.piv[J] = (pp == GJ_NONE) ? -1 : pp;
        if (pp == GJ_NONE) return;  // uniform: no row left in the panel
        // STAGE 0: row named base + q * 64 + lane; STAGE 1 (several waves) never gets here
        const int off = pp - base, ol = off & 63, oq = off >> 6;
#pragma unroll
        for (int c = J; c < GJ_NB; c++) {
            double x = R.v[0][c];
#pragma unroll
            for (int q = 1; q < RPL; q++) x = (oq == q) ? R.v[q][c] : x;
            pr[c] = readlane_dyn(x, ol);
        }
    } else {
#pragma unroll
        for (int q = 0; q < RPL; q++)
            if (R.act[q] && R.pos[q] == wbp) {  // the wave's best row: columns J.. are what the pivot row needs
#pragma unroll
                for (int c = J; c < GJ_NB; c++) sh.u.cand[BUF][wave][c] = R.v[q][c];
            }
        if (lane == 63) { sh.red_v[BUF][wave] = wmax; sh.red_p[BUF][wave] = wkey; }
        __syncthreads();
        double best = sh.red_v[BUF][0];
        int ww = 0;
        pkey = sh.red_p[BUF][0];
#pragma unroll
        for (int w2 = 1; w2 < WAVES; w2++) {
            const double v2 = sh.red_v[BUF][w2];
            const int p2 = sh.red_p[BUF][w2];
            const bool take = v2 > best || (v2 == best && p2 < pkey);
            best = take ? v2 : best;
            pkey = take ? p2 : pkey;
            ww = take ? w2 : ww;
        }
        pp = (pkey == GJ_NONE) ? GJ_NONE : (pkey & 4095);
        if (tid == 0) sh.piv[J] = (pp == GJ_NONE) ? -1 : pp;
        if (pp == GJ_NONE) return;  // uniform: no row left in the panel
        const double* prow = sh.u.cand[BUF][ww];
#pragma unroll
        for (int c = J; c < GJ_NB; c++) pr[c] = prow[c];
    }
    const int pcur = pkey >> 12, top = top0 + J;
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        if (R.act[q] && R.pos[q] == pp) R.act[q] = false;
        else if (R.act[q] && R.cur[q] == top) R.cur[q] = pcur;  // the interchange
    }
    const double piv = pr[J];
    const double rp = (piv != 0.0) ? 1.0 / piv : 0.0;
    // branch-free: rows that are not active (pivoted on, absent) take l = 0 and keep their values
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const double l = R.act[q] ? R.v[q][J] * rp : 0.0;  // multiplier, kept in place
        R.v[q][J] = R.act[q] ? l : R.v[q][J];
#pragma unroll
        for (int c = J + 1; c < GJ_NB; c++) R.v[q][c] -= l * pr[c];
    }
}

template <int WAVES, int RPL, int STAGE, bool FIN, int... J>
__device__ __attribute__((always_inline)) inline void panel_columns(std::integer_sequence<int, J...>, PanelRows<RPL>& R,
                                                                    PanelShared<WAVES, FIN>& sh, int nbb, int base,
                                                                    int top0) {
    // columns past nbb (a last step narrower than GJ_NB) are skipped; nbb is uniform
    ((J < nbb ? panel_column<J, WAVES, RPL, STAGE, FIN>(R, sh, base, top0) : void()), ...);
}

template <int STAGE, int WAVES, int RPL, bool FIN>
__global__ void __launch_bounds__(64 * WAVES) gj_panel_kernel(GJArgs a) {
    __shared__ PanelShared<WAVES, FIN> sh;
    constexpr int NT = 64 * WAVES, ROWS = NT * RPL;
    const int b = blockIdx.y, ch = blockIdx.x;
    const int m = a.m, k = a.k, nbb = min(GJ_NB, m - k);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* W = a.Wa + (long long)b * m * m;
    PanelRows<RPL> R;
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const int i = q * NT + tid;  // row of the chunk
        if (STAGE == 0) {
            R.pos[q] = k + ch * ROWS + i;
            R.act[q] = R.pos[q] < m;
        } else {
            const int p = (i < a.nch * GJ_NB) ? a.cand[(long long)b * GJ_MAXCH * GJ_NB + i] : -1;
            R.act[q] = p >= k && p < m;
            R.pos[q] = R.act[q] ? p : GJ_NONE;
        }
        R.cur[q] = R.pos[q];
        const double* rowp = W + (long long)(R.act[q] ? R.pos[q] : k) * m + k;
#pragma unroll
        for (int c = 0; c < GJ_NB; c++)  // columns k + c < m of a row of the matrix
            R.v[q][c] = (R.act[q] && c < nbb) ? rowp[c] : 0.0;
    }
    if (tid < GJ_NB) sh.piv[tid] = -1;
    static_assert(WAVES > 1 || STAGE == 0, "one-wave panels hold chunk rows in position order");
    // position k + J is the top of column J (a chunk's own first rows in the tournament stage)
    panel_columns<WAVES, RPL, STAGE, FIN>(std::make_integer_sequence<int, GJ_NB>{}, R, sh, nbb, k + ch * ROWS,
                                          STAGE == 0 ? k + ch * ROWS : k);
    __syncthreads();
    if (!FIN) {
        if (tid < GJ_NB) a.cand[((long long)b * GJ_MAXCH + ch) * GJ_NB + tid] = sh.piv[tid];
        return;
    }
    // P^-1: in-place Gauss-Jordan on P in LDS by the whole workgroup, P = the pivot rows (in
    // pivot order) as they were at the start of the step.  Step j leaves column j holding column j
    // of the inverse: rp = 1/p_jj on the diagonal, -a_rj rp elsewhere (what [P | I] elimination
    // puts in the identity's column j), and row j scaled by rp.  The order is the one partial
    // pivoting chose, so the multipliers below the diagonal are bounded by 1 and no further
    // pivoting is needed.  Beyond nbb the block is the identity.
    for (int e = tid; e < GJ_NB * GJ_NB; e += NT) {
        const int r = e / GJ_NB, c = e % GJ_NB;
        double x = (r == c) ? 1.0 : 0.0;
        const int pr = sh.piv[r];
        if (r < nbb && c < nbb && pr >= 0) x = W[(long long)pr * m + k + c];
        sh.u.A[0][r][c] = x;
    }
    __syncthreads();
    int zero = 0;
    for (int j = 0; j < nbb; j++) {
        const int cur = j & 1, nxt = cur ^ 1;
        const double pv = sh.u.A[cur][j][j];
        if (zero == 0 && (pv == 0.0 || sh.piv[j] < 0)) zero = j + 1;
        const double rp = (pv != 0.0) ? 1.0 / pv : 0.0;
        // NT is a multiple of GJ_NB: each thread keeps one column c and rows r0 + i NT / GJ_NB;
        // every read of the step is issued before its writes
        constexpr int RR = GJ_NB * GJ_NB / NT;
        const int c = tid % GJ_NB, r0 = tid / GJ_NB;
        const double rowj = (c == j) ? rp : sh.u.A[cur][j][c] * rp;
        double arj[RR], arc[RR];
#pragma unroll
        for (int i = 0; i < RR; i++) {
            const int r = r0 + i * (NT / GJ_NB);
            arj[i] = sh.u.A[cur][r][j];
            arc[i] = sh.u.A[cur][r][c];
        }
#pragma unroll
        for (int i = 0; i < RR; i++) {
            const int r = r0 + i * (NT / GJ_NB);
            double out;
            if (r == j) out = rowj;
            else if (c == j) out = -arj[i] * rp;
            else out = arc[i] - arj[i] * rowj;
            sh.u.A[nxt][r][c] = out;
        }
        __syncthreads();
    }
    {
        const int fin = nbb & 1;
        double* Pinv = a.Pinv + (long long)b * GJ_NB * GJ_NB;
        for (int e = tid; e < GJ_NB * GJ_NB; e += NT) {
            const int r = e / GJ_NB, c = e % GJ_NB;
            Pinv[e] = (r < nbb && c < nbb) ? sh.u.A[fin][r][c] : 0.0;
        }
    }
    if (wave == 0) {
        // interchanges (LAPACK order: position k+j with the chosen row's current position)
        int top = k + lane;      // row now at position k + lane (lanes < nbb)
        int op = -1, orw = -1;   // entry `lane` of the positions beyond the panel that changed row
        int nout = 0, mine = k + lane;
        for (int j = 0; j < nbb; j++) {
            const int target = sh.piv[j];
            const int apos = k + j;
            int t = apos;
            if (target >= 0) {
                const unsigned long long m1 = __ballot(lane >= j && lane < nbb && top == target);
                if (m1) {
                    t = k + __ffsll(m1) - 1;
                } else {
                    const unsigned long long m2 = __ballot(lane < nout && orw == target);
                    t = m2 ? __shfl(op, __ffsll(m2) - 1) : target;
                }
                const int row_a = __shfl(top, j);
                if (t != apos) {
                    if (lane == j) top = target;
                    if (t < k + nbb) {
                        if (lane == t - k) top = row_a;
                    } else {
                        const unsigned long long m3 = __ballot(lane < nout && op == t);
                        const int o = m3 ? __ffsll(m3) - 1 : nout;
                        if (lane == o) { op = t; orw = row_a; }
                        if (!m3) nout++;
                    }
                }
            }
            if (lane == j) mine = t;
        }
        if (lane < GJ_NB) {
            sh.top_row[lane] = top;
            sh.opos[lane] = op;
            sh.orow[lane] = orw;
            const int target = (lane < nbb) ? sh.piv[lane] : -1;
            a.prow[(long long)b * GJ_NB + lane] = (target >= 0) ? target : min(k + lane, m - 1);
            if (lane < nbb) a.ipiv[(long long)b * m + k + lane] = mine;
        }
        if (lane == 0) {
            sh.nout = nout;
            if (zero && a.info[b] == 0) a.info[b] = k + zero;
        }
    }
    __syncthreads();
    int* src = a.src + (long long)b * m;
    const int nout = sh.nout;
    for (int r = tid; r < m; r += NT) {
        int s = r;
        if (r >= k && r < k + nbb) s = sh.top_row[r - k];
        else
            for (int o = 0; o < nout; o++)
                if (sh.opos[o] == r) s = sh.orow[o];
        src[r] = s;
    }
}

template __global__ void gj_panel_kernel<0, 1, 1, true>(GJArgs);
template __global__ void gj_panel_kernel<0, 1, 2, true>(GJArgs);
template __global__ void gj_panel_kernel<0, 2, 2, true>(GJArgs);
template __global__ void gj_panel_kernel<0, 4, 2, true>(GJArgs);
template __global__ void gj_panel_kernel<0, 4, 2, false>(GJArgs);
template __global__ void gj_panel_kernel<1, 4, 1, true>(GJArgs);

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
    // the block's current values, loaded before the products so their latency overlaps them
    double cv[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = min(row0 + 16 * i + 4 * v + q, m - 1);
            const double* crow = Wa + (long long)src[row] * m;
            cv[i][0][v] = crow[c0];
            cv[i][1][v] = crow[c1];
        }
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
                if (row < m && col < m) Wb[(long long)row * m + col] = acc[i][j][v] + cv[i][j][v];
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
// (pairwise averages, nested as coarsenUpwards_ applies them one step at a time); Vr is the
// right-hand side's column of the arena
__device__ inline double coarse_h(const double* Vr, long long off, int clen, int tag, int side, int i) {
    const int sl = clen / 4;
    const long long base = off + (long long)side * sl;
    if (tag == 0) return Vr[base + i];
    const int w = 1 << tag;
    auto x = [&](int t) { return Vr[base + (long long)i * w + t]; };
    if (tag == 1) return 0.5 * x(0) + 0.5 * x(1);
    double s = 0.0;  // tag >= 2: blocks of four in the nested order, then their mean
    for (int t = 0; t < w; t += 4) {
        const double p = 0.5 * x(t) + 0.5 * x(t + 1), u = 0.5 * x(t + 2) + 0.5 * x(t + 3);
        s += 0.5 * p + 0.5 * u;
    }
    return s * (4.0 / (double)w);
}

// The vector kernels below (up gather, uncoarsen, side copy, side interpolation) take one
// descriptor per blockIdx.y and stride over its (rhs, row) elements with gridDim.x workgroups:
// the host sizes gridDim.x to the launch's total work, not to its longest descriptor, so short
// descriptors do not leave thousands of empty workgroups to dispatch.
__global__ void __launch_bounds__(256) tree_up_gather_kernel(const UpDesc* __restrict__ descs, double* V, int k,
                                                              long long vrows) {
    const UpDesc d = descs[blockIdx.y];
    const int n = d.n;
    const long long total = 12LL * n * k;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int rhs = (int)(e / (12 * n)), row = (int)(e % (12 * n));
        double* Vr = V + (long long)rhs * vrows;
        if (row < 4 * n) {  // hDiff of mergeW_ (HPSAlgorithm.hpp:1059-1087)
            const int r = row / n, i = row % n;
            const int c1 = if_c1(r), c2 = if_c2(r);
            const double h2 = coarse_h(Vr, pick4(d.ch_off, c2), pick4(d.clen, c2), pick4(d.tag, c2), if_s2(r), i);
            const double h1 = coarse_h(Vr, pick4(d.ch_off, c1), pick4(d.clen, c1), pick4(d.tag, c1), if_s1(r), i);
            Vr[d.hd_off + row] = h2 - h1;
        } else {  // boundary part of mergeH_ (:1098-1117), WESN order (:1128-1150)
            const int rr = row - 4 * n, q = rr / n, i = rr % n, c = q_child(q);
            Vr[d.h_off + rr] = coarse_h(Vr, pick4(d.ch_off, c), pick4(d.clen, c), pick4(d.tag, c), q_side(q), i);
        }
    }
}

// value i of a side prolonged from nc coarse values src to nf = 2 nc (uncoarsen_'s L12:
// quadratic end weights, 3/4-1/4 in between)
__device__ inline double prolong(const double* src, int nc, int nf, int i) {
    if (i == 0) return 1.40625 * src[0] + (-0.5625) * src[1] + 0.15625 * src[2];
    if (i == nf - 1) return 0.15625 * src[nc - 3] + (-0.5625) * src[nc - 2] + 1.40625 * src[nc - 1];
    if (i & 1) {
        const int j = (i - 1) / 2;
        return 0.75 * src[j] + 0.25 * src[j + 1];
    }
    const int j = (i - 2) / 2;
    return 0.25 * src[j] + 0.75 * src[j + 1];
}

__global__ void __launch_bounds__(256) tree_uncoarsen_kernel(const VecStepDesc* __restrict__ descs, double* V, int k,
                                                              long long vrows) {
    const VecStepDesc d = descs[blockIdx.y];
    const int nf = d.nfine, nc = nf / 2;
    const long long total = 4LL * nf * k;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int rhs = (int)(e / (4 * nf)), row = (int)(e % (4 * nf));
        const int side = row / nf, i = row % nf;
        double* Vr = V + (long long)rhs * vrows;
        Vr[d.dst_off + row] = prolong(Vr + d.src_off + (long long)side * nc, nc, nf, i);
    }
}

// A child's boundary data written out, where the sweeps cannot read sides in place (sizes that
// do not allow the 4-row operand blocks): side s of descriptor blockIdx.y from src[s], each
// thread striding over (rhs, row) elements.
__global__ void __launch_bounds__(256) tree_side_copy_kernel(const SideCopy* __restrict__ descs, double* V, int k,
                                                              long long vrows) {
    const SideCopy d = descs[blockIdx.y];
    const int len = d.len;
    const long long total = 4LL * len * k;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int rhs = (int)(e / (4 * len)), row = (int)(e % (4 * len));
        const int s = row / len, i = row - s * len;
        double* Vr = V + (long long)rhs * vrows;
        Vr[d.dst + row] = Vr[pick4(d.src, s) + i];
    }
}

// The first uncoarsen_ step of a coarsened child straight from its sides' sources: the coarse
// side s (len values at src[s]) prolonged to 2 len values at dst + s 2 len, with the weights of
// tree_uncoarsen_kernel (HPSAlgorithm.hpp:1165-1190).
__global__ void __launch_bounds__(256) tree_side_interp_kernel(const SideCopy* __restrict__ descs, double* V, int k,
                                                                long long vrows) {
    const SideCopy d = descs[blockIdx.y];
    const int nc = d.len, nf = 2 * nc;
    const long long total = 4LL * nf * k;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int rhs = (int)(e / (4 * nf)), row = (int)(e % (4 * nf));
        const int side = row / nf, i = row % nf;
        double* Vr = V + (long long)rhs * vrows;
        Vr[d.dst + row] = prolong(Vr + pick4(d.src, side), nc, nf, i);
    }
}

// g = r / a: the LU solve of diag(a) in solveStage (HPSAlgorithm.hpp:402-420); r is [rhs][m],
// g is column rhs of the arena (g + rhs * vrows).
__global__ void __launch_bounds__(256) tree_root_dirichlet_kernel(const double* r, const double* a, double* g, int m, int k,
                                                                   long long vrows) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)m * k) return;
    const int c = (int)(e / m), i = (int)(e % m);
    g[(long long)c * vrows + i] = r[e] / a[i];
}

// A = diag(a) + diag(b) T and rhs = r - b h (HPSAlgorithm.hpp:402-420); rhs is [rhs][m], h a
// column of the arena (h + rhs * vrows) or null.
__global__ void __launch_bounds__(256) tree_root_robin_kernel(const double* T, const double* a, const double* b, const double* r,
                                                              const double* h, double* A, double* rhs, int m, int k,
                                                              long long vrows) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nA = (long long)m * m;
    if (e < nA) {
        const int i = (int)(e / m), j = (int)(e % m);
        A[e] = b[i] * T[e] + (i == j ? a[i] : 0.0);
    } else if (e < nA + (long long)m * k) {
        const long long t = e - nA;
        const int c = (int)(t / m), i = (int)(t % m);
        rhs[t] = r[t] - b[i] * (h ? h[(long long)c * vrows + i] : 0.0);
    }
}

// ---------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------
GemmLaunch plan_gemm(const std::vector<TreeGemm>& probs_in, int tn, DevStore& store, hipStream_t s, bool f32, int rt) {
    GemmLaunch g;
    g.tn = tn;
    g.f32 = f32;
    g.rt = (tn == 1) ? 1 : 2;
    if (probs_in.empty()) return g;
    int maxM = 0;
    for (const TreeGemm& p : probs_in) maxM = std::max(maxM, p.M);
    if (maxM <= 16) g.rt = 1;  // 16-row tiles: a 32-row tile would compute a clamped copy of the same rows
    if (rt == 1 || (rt == 2 && tn == 2)) g.rt = rt;  // an explicit choice (the tile autotuner)
    std::vector<TreeGemm> probs = probs_in;
    // Split-K: when a problem asks for k chunks (kc > 0), the whole launch runs on
    // tree_gemm_split_kernel with one wave per chunk of a 16 x 16 tile; problems that asked for
    // none are cut into chunks of the same count (rounded to 16).  The chunks depend on the
    // launch's problems only, never on N.
    int kw = 1;
    for (TreeGemm& p : probs)
        if (p.M > 0 && p.N > 0 && p.kc > 0) {
            if (p.kc % 16 != 0) throw EfError(EF_ERR_INTERNAL, "split-K chunk must be a multiple of 16");
            const int K = p.K0 + p.K1;
            // at most 16 chunks (waves per workgroup): longer ranges get longer chunks
            if ((K + p.kc - 1) / p.kc > 16) p.kc = ((K + 15) / 16 + 15) / 16 * 16;
            kw = std::max(kw, (K + p.kc - 1) / p.kc);
        }
    if (kw > 16) throw EfError(EF_ERR_INTERNAL, "split-K launch with more than 16 chunks per tile");
    if (kw > 1 && f32) throw EfError(EF_ERR_INTERNAL, "split-K launch on the FP32 matrix cores");
    if (kw > 1) {
        g.kw = kw;
        g.rt = g.tn = 1;
    }
    std::vector<int> start(probs.size() + 1, 0);
    const int rw = 16 * g.rt, cw = 16 * g.tn;
    long long total = 0;
    for (size_t i = 0; i < probs.size(); i++) {
        TreeGemm& p = probs[i];
        start[i] = (int)total;
        p.nks = 1;
        if (p.bcs == 0) p.bcs = 1;
        if (p.ccs == 0) p.ccs = 1;
        if (p.dcs == 0) p.dcs = 1;
        if (p.M <= 0 || p.N <= 0) { p.kc = 0; continue; }
        // extents: the kernel clamps rows to M-1, columns to N-1 and k to K-1
        check_strided(p.D, p.M, p.N, p.ldd, p.dcs, "GEMM D");
        if (p.C) check_strided(p.C, p.M, p.N, p.ldc, p.ccs, "GEMM C");
        const double* As[2] = {p.A0, p.A1};
        const double* Bs[2] = {p.B0, p.B1};
        const int lda[2] = {p.lda0, p.lda1}, ldb[2] = {p.ldb0, p.ldb1}, K[2] = {p.K0, p.K1};
        if (p.dn > 0) {  // block-gather mode: every block's segments inside B0's allocation
            if (p.dn % 4 != 0 || p.ldb0 != 1 || p.K0 != 4 * p.dn || p.K1 > 0 || (p.dsub != 0 && p.dsub != 1))
                throw EfError(EF_ERR_INTERNAL, "block-gather GEMM needs four blocks of a multiple of 4 rows, contiguous");
            for (int r4 = 0; r4 < 4; r4++)
                for (int h2 = 0; h2 < 1 + p.dsub; h2++) {
                    const long long o = h2 ? p.d1[r4] : p.d2[r4];
                    if (o < 0 || (o & 1)) throw EfError(EF_ERR_INTERNAL, "block-gather GEMM row offset negative or odd");
                    check_strided(p.B0 + o, p.dn, p.N, 1, p.bcs, "GEMM gathered block");
                }
        }
        for (int sg = 0; sg < 2; sg++) {
            if (K[sg] <= 0) continue;
            check_matrix(As[sg], p.M, K[sg], lda[sg], "GEMM A");
            if (sg == 0 && p.dn > 0) {
                if (((uintptr_t)As[0] & 15) != 0 || (lda[0] & 1) != 0 || ((uintptr_t)p.B0 & 15) != 0 || (p.bcs & 1) != 0)
                    throw EfError(EF_ERR_INTERNAL, "block-gather GEMM operands not 16-byte aligned");
                continue;
            }
            check_strided(Bs[sg], K[sg], p.N, ldb[sg], p.bcs, "GEMM B");
            if (((uintptr_t)As[sg] & 15) != 0 || (lda[sg] & 1) != 0)
                throw EfError(EF_ERR_INTERNAL, "GEMM A operand not 16-byte aligned with an even leading dimension");
            if (ldb[sg] == 1 && (((uintptr_t)Bs[sg] & 15) != 0 || (p.bcs & 1) != 0))
                throw EfError(EF_ERR_INTERNAL, "GEMM B operand contiguous along k but not 16-byte aligned");
        }
        if (g.kw > 1) {
            const int Kt = p.K0 + p.K1;
            if (p.kc == 0) p.kc = std::max(16, ((Kt + g.kw - 1) / g.kw + 15) / 16 * 16);
            p.nks = (Kt + p.kc - 1) / p.kc;
            if (p.nks > g.kw) throw EfError(EF_ERR_INTERNAL, "split-K chunks exceed the launch's waves per tile");
            // a chunk that starts in segment 1 starts at an even k there (16-byte operand reads)
            if (p.K1 > 0 && (p.K0 & 1)) throw EfError(EF_ERR_INTERNAL, "split-K over two segments needs an even K0");
        } else {
            p.kc = 0;
        }
        total += (long long)((p.M + rw - 1) / rw) * ((p.N + cw - 1) / cw);
        if (total > 0x7fffffffLL) throw EfError(EF_ERR_INTERNAL, "GEMM launch too large");
    }
    start[probs.size()] = (int)total;
    g.nprob = (int)probs.size();
    g.nwaves = (int)total;  // waves (tile kernel) or workgroups (split kernel)
    std::vector<int> wprob((size_t)total);
    for (size_t i = 0; i < probs.size(); i++)
        for (int w = start[i]; w < start[i + 1]; w++) wprob[w] = (int)i;
    g.probs = store.stage(probs);
    g.start = store.stage(start);
    g.wprob = store.stage(wprob);
    return g;
}

void run_gemm(const GemmLaunch& g, hipStream_t s) {
    if (g.nwaves == 0) return;
    if (g.kw > 1) {
        EF_LAUNCH("tree_gemm_split_kernel", tree_gemm_split_kernel, dim3(g.nwaves), dim3(64 * g.kw), s, g.probs,
                  g.start, g.wprob, g.nwaves);
        return;
    }
    const dim3 grid(blocks_for(g.nwaves, 4)), block(256);
#define EF_GEMM_LAUNCH(RT_, TN_, F32_, NAME_)                                                          \
    EF_LAUNCH(NAME_, (tree_gemm_kernel<RT_, TN_, F32_>), grid, block, s, g.probs, g.start, g.wprob, g.nwaves)
    if (g.f32) {
        if (g.tn == 1) EF_GEMM_LAUNCH(1, 1, true, "tree_gemm_kernel<1,1,f32>");
        else if (g.rt == 1) EF_GEMM_LAUNCH(1, 2, true, "tree_gemm_kernel<1,2,f32>");
        else EF_GEMM_LAUNCH(2, 2, true, "tree_gemm_kernel<2,2,f32>");
    } else {
        if (g.tn == 1) EF_GEMM_LAUNCH(1, 1, false, "tree_gemm_kernel<1,1>");
        else if (g.rt == 1) EF_GEMM_LAUNCH(1, 2, false, "tree_gemm_kernel<1,2>");
        else EF_GEMM_LAUNCH(2, 2, false, "tree_gemm_kernel<2,2>");
    }
#undef EF_GEMM_LAUNCH
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
        const int rows = m - k;
        // rows per workgroup: 64 WAVES RPL (chunks of GJ_CH = 512 rows; up to 256 candidates)
        if (rows <= 64)
            EF_LAUNCH("gj_panel_kernel<0,1,1,fin>", (gj_panel_kernel<0, 1, 1, true>), dim3(1, B), dim3(64), s, a);
        else if (rows <= 128)
            EF_LAUNCH("gj_panel_kernel<0,1,2,fin>", (gj_panel_kernel<0, 1, 2, true>), dim3(1, B), dim3(64), s, a);
        else if (rows <= 256)
            EF_LAUNCH("gj_panel_kernel<0,2,2,fin>", (gj_panel_kernel<0, 2, 2, true>), dim3(1, B), dim3(128), s, a);
        else if (a.nch == 1)
            EF_LAUNCH("gj_panel_kernel<0,4,2,fin>", (gj_panel_kernel<0, 4, 2, true>), dim3(1, B), dim3(256), s, a);
        else {
            EF_LAUNCH("gj_panel_kernel<0,4,2>", (gj_panel_kernel<0, 4, 2, false>), dim3(a.nch, B), dim3(256), s, a);
            EF_LAUNCH("gj_panel_kernel<1,4,1,fin>", (gj_panel_kernel<1, 4, 1, true>), dim3(1, B), dim3(256), s, a);
        }
        EF_LAUNCH("gj_mult_kernel", gj_mult_kernel, dim3(blocks_for((long long)m * GJ_NB), B), dim3(256), s, a);
        EF_LAUNCH("gj_update_kernel", gj_update_kernel, dim3(blocks_for((long long)tpr * tpr, 4), B), dim3(256), s, a);
        std::swap(cur, nxt);
    }
    a.Wa = cur; a.Wb = nxt; a.k = 0;
    EF_LAUNCH("gj_perm_kernel", gj_perm_kernel, dim3(B), dim3(256), s, a);
    EF_LAUNCH("gj_gather_kernel", gj_gather_kernel, dim3(blocks_for((long long)m * m), B), dim3(256), s, a);
}

}  // namespace efk
