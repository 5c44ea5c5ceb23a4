// ef_patch_large.hip -- fast diagonalisation of large patches (N > 32) as batched GEMMs.
//
// Same mathematics as patch_fd_kernel (U = Q ((Q^T B Q) ./ D) Q^T, FiniteVolumeSolver.cpp
// :25-294 with constant coefficients), but the N x N arrays no longer fit one wave's
// registers: each step is a launch over a chunk of items held in HBM work buffers padded
// to Np = 16 ceil(N / 16) (zero padding; padded modes are zeroed at the scaling step).
//   fold   B = f - 2 c g on the edge cells (or a unit boundary column for DtN chunks)
//   GEMM   W2 = Q^T W1,  W1 = W2 Q,  scale W1 ./ D,  W2 = Q W1,  W1 = W2 Q^T
//   out    u, or the Neumann data h / DtN column from the edge cells of u
// The GEMM is register-resident like the small-patch kernel: one wave per 32 x 32 block of
// output (2 x 2 tiles of 16 x 16), operands read straight from memory in the
// v_mfma_f64_16x16x4_f64 layouts (A[r][k0 + q], B[k0 + q][r] for lane 16q + r), no LDS and no
// scratch.
#include <hip/hip_runtime.h>
#include "ef_device.hpp"
#include "ef_kernels.hpp"

namespace efk {

typedef double ld4 __attribute__((ext_vector_type(4)));

// AK: A is contiguous along k (ca == 1); BK: B is contiguous along k (rb == 1).
template <bool AK, bool BK>
__global__ void __launch_bounds__(256) gemm_strided_f64_kernel(GemmStridedArgs a) {
    // one wave per 32 x 32 block of C: 2 x 2 tiles, so each A and B value loaded feeds two
    // MFMAs.  On the last block row / column a missing second tile points at the first (loads
    // in range, the duplicate is not stored).
    //
    // Within each block of 16 k values, lane group q takes k = kb + 4q + s at step s (the
    // reduction order is free as long as A and B agree): an operand contiguous along k is
    // then read as 4 consecutive doubles per lane (two 16-byte loads, the four lane groups
    // covering a row's 128 bytes), and one contiguous along the tile's rows or columns as 16
    // consecutive doubles per lane group and step -- every cache line fetched is used whole.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    const int tn = (a.N + 31) / 32, tm = (a.M + 31) / 32;
    const long long w = (long long)blockIdx.x * 4 + wave;
    if (w >= (long long)tm * tn) return;  // uniform per wave
    const long long b = blockIdx.y;
    const int row0 = (int)(w / tn) * 32, col0 = (int)(w % tn) * 32;
    const bool r1 = row0 + 16 < a.M, c1 = col0 + 16 < a.N;
    const double* A0 = a.A + b * a.sa + (long long)(row0 + r) * a.ra;
    const double* A1 = r1 ? A0 + 16LL * a.ra : A0;
    const double* B0 = a.B + b * a.sb + (long long)(col0 + r) * a.cb;
    const double* B1 = c1 ? B0 + 16LL * a.cb : B0;
    ld4 c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    // four consecutive k of operand X from k-index k (contiguous: one 32-byte aligned vector)
    auto quad = [](const double* X, long long stride, int k, bool contiguous, double (&v)[4]) {
        if (contiguous) {
            const ld4 x = *(const ld4*)(X + k);
            v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) v[s] = X[(long long)(k + s) * stride];
        }
    };
    // K is a multiple of 16 (Np)
#pragma unroll 2
    for (int kb = 0; kb < a.K; kb += 16) {
        const int k = kb + 4 * q;
        double a0[4], a1[4], b0[4], b1[4];
        quad(A0, a.ca, k, AK, a0);
        quad(A1, a.ca, k, AK, a1);
        quad(B0, a.rb, k, BK, b0);
        quad(B1, a.rb, k, BK, b1);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], b0[s], c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], b1[s], c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], b0[s], c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], b1[s], c11, 0, 0, 0);
        }
    }
    double* C = a.C + b * a.sc;
    auto put = [&](const ld4& acc, int rb, int cb) {
#pragma unroll
        for (int v = 0; v < 4; v++) C[(long long)(rb + 4 * v + q) * a.rc + (long long)(cb + r) * a.cc] = acc[v];
    };
    put(c00, row0, col0);
    if (c1) put(c01, row0, col0 + 16);
    if (r1) put(c10, row0 + 16, col0);
    if (r1 && c1) put(c11, row0 + 16, col0 + 16);
}

template __global__ void gemm_strided_f64_kernel<false, false>(GemmStridedArgs);
template __global__ void gemm_strided_f64_kernel<true, false>(GemmStridedArgs);
template __global__ void gemm_strided_f64_kernel<false, true>(GemmStridedArgs);
template __global__ void gemm_strided_f64_kernel<true, true>(GemmStridedArgs);

// item t of the chunk -> (patch, rhs, DtN column or -1)
__device__ inline void large_item(const PatchLargeArgs& a, long long t, int& patch, int& rhs, int& col) {
    const long long it = a.item0 + t;
    if (a.op == PATCH_OP_DTN) {
        patch = a.patch0;
        rhs = 0;
        col = a.col0 + (int)t;
    } else {
        rhs = (int)(it / a.npatches);
        patch = (int)(it % a.npatches);
        col = -1;
    }
}

__device__ inline double large_g(const PatchLargeArgs& a, const double* g, int col, int t) {
    if (col >= 0) return (t == col) ? 1.0 : 0.0;
    return g ? g[t] : 0.0;
}

// work[t][i][j] (Np x Np) = f - 2 c g at the edge cells, zero in the padding
__global__ void patch_large_fold_kernel(PatchLargeArgs a) {
    const int Nn = a.N, Np = (Nn + 15) / 16 * 16;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)a.items * Np * Np) return;
    const long long t = e / ((long long)Np * Np);
    const int i = (int)((e / Np) % Np), j = (int)(e % Np);
    double v = 0.0;
    if (i < Nn && j < Nn) {
        int patch, rhs, col;
        large_item(a, t, patch, rhs, col);
        const long long slot = (long long)rhs * a.npatches + patch;
        const PatchCoef cf = a.coef[a.cls ? a.cls[patch] : patch];
        if (a.f && col < 0) v = a.f[slot * Nn * Nn + (long long)i * Nn + j];
        const double* g = (a.g && col < 0) ? a.g + slot * 4 * Nn : nullptr;
        double add = 0.0;
        if (i == 0) add += cf.cx * large_g(a, g, col, W * Nn + j);
        if (i == Nn - 1) add += cf.cx * large_g(a, g, col, E * Nn + j);
        if (j == 0) add += cf.cy * large_g(a, g, col, S * Nn + i);
        if (j == Nn - 1) add += cf.cy * large_g(a, g, col, N * Nn + i);
        v += -2.0 * add;
    }
    a.work[e] = v;
}

// work[t][k][l] /= cx mu_k + cy mu_l + lam, zero on padded modes
__global__ void patch_large_scale_kernel(PatchLargeArgs a) {
    const int Nn = a.N, Np = (Nn + 15) / 16 * 16;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)a.items * Np * Np) return;
    const long long t = e / ((long long)Np * Np);
    const int k = (int)((e / Np) % Np), l = (int)(e % Np);
    if (k >= Nn || l >= Nn) {
        a.work[e] = 0.0;
        return;
    }
    int patch, rhs, col;
    large_item(a, t, patch, rhs, col);
    const PatchCoef cf = a.coef[a.cls ? a.cls[patch] : patch];
    a.work[e] = a.work[e] / (cf.cx * a.mu[k] + cf.cy * a.mu[l] + cf.lam);
}

// u out, or h / DtN column from the edge cells: one thread per cell (solve) or per
// boundary position (Neumann data)
__global__ void patch_large_out_kernel(PatchLargeArgs a) {
    const int Nn = a.N, Np = (Nn + 15) / 16 * 16;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (a.op == PATCH_OP_SOLVE) {
        if (e >= (long long)a.items * Nn * Nn) return;
        const long long t = e / ((long long)Nn * Nn);
        const int i = (int)((e / Nn) % Nn), j = (int)(e % Nn);
        int patch, rhs, col;
        large_item(a, t, patch, rhs, col);
        const long long slot = (long long)rhs * a.npatches + patch;
        a.out[slot * Nn * Nn + (long long)i * Nn + j] = a.work[t * Np * Np + (long long)i * Np + j];
        return;
    }
    if (e >= (long long)a.items * 4 * Nn) return;
    const long long t = e / (4 * Nn);
    const int b = (int)(e % (4 * Nn));
    int patch, rhs, col;
    large_item(a, t, patch, rhs, col);
    const long long slot = (long long)rhs * a.npatches + patch;
    const PatchCoef cf = a.coef[a.cls ? a.cls[patch] : patch];
    const double* g = (a.g && col < 0) ? a.g + slot * 4 * Nn : nullptr;
    const int sd = b / Nn, m = b % Nn;
    int i, j;
    if (sd == W) { i = 0; j = m; }
    else if (sd == E) { i = Nn - 1; j = m; }
    else if (sd == S) { i = m; j = 0; }
    else { i = m; j = Nn - 1; }
    const double uu = a.work[t * Np * Np + (long long)i * Np + j];
    const double dd = (sd < 2) ? cf.dx : cf.dy;
    const double sg = (sd == W || sd == S) ? 2.0 / dd : -2.0 / dd;
    const double h = sg * (uu - large_g(a, g, col, b));
    if (a.op == PATCH_OP_DTN) a.out[(long long)patch * 16 * Nn * Nn + (long long)b * 4 * Nn + col] = h;
    else a.out[slot * 4 * Nn + b] = h;
}

}  // namespace efk
