// ef_patch_kernels.hip -- batched leaf-patch solves by fast diagonalisation on the f64 MFMA.
//
// A constant-coefficient patch (FISHPACK semantics, or the five-point stencil with
// constant alpha, beta, lambda; FiniteVolumeSolver.cpp:25-294) has, with x-major cell
// values U[i][j] and ghost value 2g - u on every side,
//     A U = cx K U + cy U K + lam U,   K = tridiag(1, -2, 1) with -3 at both ends,
// and K = Q diag(mu) Q^T with Q[i][k] = nrm_k sin(pi (k+1)(i+1/2)/N),
// mu_k = -4 sin^2(pi (k+1)/(2N)).  Hence
//     U = Q ((Q^T B Q) ./ D) Q^T,   D[k][l] = cx mu_k + cy mu_l + lam,
// where B is the load with the boundary data folded in (-2 c g at the adjacent cell).
//
// One wave per (patch, right-hand side) at a time, four independent waves per workgroup
// looping over the items with a grid stride (the basis tables stay in registers), no LDS.
// Every product is done on v_mfma_f64_16x16x4_f64 in registers.  With q = lane/16 and
// r = lane%16, the gfx950 operand layouts for a 16x16 block are
//     A operand, step s: A[r][4s + q]        B operand, step s: B[4s + q][r]
//     accumulator slot v:  C[4v + q][r]
// so an accumulator block IS the B operand of the next product, and the accumulator of
// M^T is the A operand of M.  The chain is arranged so no value ever changes lanes:
//     Z1^T = B^T Q          (A = B^T loaded straight from memory, B = Q)
//     Z2   = Z1 Q           (A = Z1 from the accumulator of Z1^T)
//     Z3   = Z2 ./ D        (in the accumulator)
//     Z4^T = Z3^T Q^T       (A = Z3^T from the accumulator of Z3)
//     U    = Z4 Q^T         (A = Z4 from the accumulator of Z4^T)
// Patches with N <= 16 use one padded 16x16 block, N <= 32 a 2x2 block grid (NB).
#include <hip/hip_runtime.h>
#include "ef_device.hpp"
#include "ef_kernels.hpp"

namespace efk {

typedef double pd4 __attribute__((ext_vector_type(4)));

// C = A B with A given per (bi, bk, s) lane values and B per (bk, bj, s).
template <int NB, class FA, class FB>
__device__ inline void patch_mm(FA a, FB b, pd4 (&C)[NB][NB]) {
#pragma unroll
    for (int bi = 0; bi < NB; bi++)
#pragma unroll
        for (int bj = 0; bj < NB; bj++) {
            pd4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int bk = 0; bk < NB; bk++)
#pragma unroll
                for (int s = 0; s < 4; s++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a(bi, bk, s), b(bk, bj, s), acc, 0, 0, 0);
            C[bi][bj] = acc;
        }
}

template <int NB>
__global__ void __launch_bounds__(256) patch_fd_kernel(PatchFDArgs p) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    const int Nn = p.N;
    const long long total = (p.op == PATCH_OP_DTN) ? (long long)p.npatches * 4 * Nn : (long long)p.npatches * p.nrhs;

    // Q in both operand layouts, loaded once per wave: QA[x][y][s] = Q[16x + r][16y + 4s + q],
    // QB[x][y][s] = Q[16x + 4s + q][16y + r]
    double QA[NB][NB][4], QB[NB][NB][4];
#pragma unroll
    for (int x = 0; x < NB; x++)
#pragma unroll
        for (int y = 0; y < NB; y++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                QA[x][y][s] = p.QA[(((x * NB) + y) * 4 + s) * 64 + lane];
                QB[x][y][s] = p.QB[(((x * NB) + y) * 4 + s) * 64 + lane];
            }

    // grid-stride loop over (patch, rhs) items; the trip count is uniform per wave
    for (long long item = (long long)blockIdx.x * 4 + wave; item < total; item += (long long)gridDim.x * 4) {
        int patch, rhs, col = -1;
        if (p.op == PATCH_OP_DTN) {
            patch = (int)(item / (4 * Nn));
            col = (int)(item % (4 * Nn));
            rhs = 0;
        } else {
            rhs = (int)(item / p.npatches);
            patch = (int)(item % p.npatches);
        }
        const PatchCoef cf = p.coef[p.cls ? p.cls[patch] : patch];
        const long long slot = (long long)rhs * p.npatches + patch;
        const double* f = p.f ? p.f + slot * Nn * Nn : nullptr;
        const double* g = nullptr;
        long long gs = 1;
        if (p.g && col < 0) {
            if (p.g_off) { g = p.g + p.g_off[patch] * p.vk + rhs; gs = p.vk; }
            else g = p.g + slot * 4 * Nn;
        }
        // boundary value at WESN position t (side*N + i)
        auto gval = [&](int t) -> double {
            if (col >= 0) return (t == col) ? 1.0 : 0.0;
            return g ? g[t * gs] : 0.0;
        };
        const bool use_g = (col >= 0) || (g != nullptr);

        // B^T in the A layout: element (bi, bk, s) is B[i][j] with i = 16bk + 4s + q, j = 16bi + r
        double BT[NB][NB][4];
#pragma unroll
        for (int bi = 0; bi < NB; bi++)
#pragma unroll
            for (int bk = 0; bk < NB; bk++)
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int i = 16 * bk + 4 * s + q, j = 16 * bi + r;
                    double v = 0.0;
                    if (i < Nn && j < Nn) {
                        if (f) v = f[i * Nn + j];
                        if (use_g && (i == 0 || j == 0 || i == Nn - 1 || j == Nn - 1)) {
                            double add = 0.0;
                            if (i == 0) add += cf.cx * gval(W * Nn + j);
                            if (i == Nn - 1) add += cf.cx * gval(E * Nn + j);
                            if (j == 0) add += cf.cy * gval(S * Nn + i);
                            if (j == Nn - 1) add += cf.cy * gval(N * Nn + i);
                            v += -2.0 * add;
                        }
                    }
                    BT[bi][bk][s] = v;
                }

        pd4 Z[NB][NB], Y[NB][NB];
        // Z1^T = B^T Q
        patch_mm<NB>([&](int bi, int bk, int s) { return BT[bi][bk][s]; },
                     [&](int bk, int bj, int s) { return QB[bk][bj][s]; }, Z);
        // Z2 = Z1 Q, A operand of Z1 = accumulator of Z1^T with the block indices swapped
        patch_mm<NB>([&](int bi, int bk, int s) { return Z[bk][bi][s]; },
                     [&](int bk, int bj, int s) { return QB[bk][bj][s]; }, Y);
        // Z3 = Z2 ./ D; padded modes (k or l >= N) are zero
#pragma unroll
        for (int bi = 0; bi < NB; bi++)
#pragma unroll
            for (int bj = 0; bj < NB; bj++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int kk = 16 * bi + 4 * v + q, ll = 16 * bj + r;
                    double z = 0.0;
                    if (kk < Nn && ll < Nn) z = Y[bi][bj][v] / (cf.cx * p.mu[kk] + cf.cy * p.mu[ll] + cf.lam);
                    Y[bi][bj][v] = z;
                }
        // Z4^T = Z3^T Q^T: A operand of Z3^T = accumulator of Z3 swapped; B operand of Q^T = QA swapped
        patch_mm<NB>([&](int bi, int bk, int s) { return Y[bk][bi][s]; },
                     [&](int bk, int bj, int s) { return QA[bj][bk][s]; }, Z);
        // U = Z4 Q^T
        patch_mm<NB>([&](int bi, int bk, int s) { return Z[bk][bi][s]; },
                     [&](int bk, int bj, int s) { return QA[bj][bk][s]; }, Y);

        if (p.op == PATCH_OP_SOLVE) {
            double* u = p.out + slot * Nn * Nn;
#pragma unroll
            for (int bi = 0; bi < NB; bi++)
#pragma unroll
                for (int bj = 0; bj < NB; bj++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int i = 16 * bi + 4 * v + q, j = 16 * bj + r;
                        if (i < Nn && j < Nn) u[i * Nn + j] = Y[bi][bj][v];
                    }
            continue;
        }
        // Neumann data (mapD2N, FiniteVolumeSolver.cpp:296-353):
        //   h_W = (2/dx)(u_W - g_W), h_E = -(2/dx)(u_E - g_E), h_S = (2/dy)(u_S - g_S), h_N = -(2/dy)(u_N - g_N)
        // written to out[t * ostride] of this (patch, rhs) or of DtN column col.
        double* h;
        long long ostride;
        if (col >= 0) {
            h = p.out + (long long)patch * 16 * Nn * Nn + col;
            ostride = 4 * Nn;
        } else if (p.h_off) {
            h = p.out + p.h_off[patch] * p.vk + rhs;
            ostride = p.vk;
        } else {
            h = p.out + slot * 4 * Nn;
            ostride = 1;
        }
        const double sx = 2.0 / cf.dx, sy = 2.0 / cf.dy;
#pragma unroll
        for (int bi = 0; bi < NB; bi++)
#pragma unroll
            for (int bj = 0; bj < NB; bj++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = 16 * bi + 4 * v + q, j = 16 * bj + r;
                    if (i >= Nn || j >= Nn) continue;
                    const double uu = Y[bi][bj][v];
                    if (i == 0) h[(long long)(W * Nn + j) * ostride] = sx * (uu - gval(W * Nn + j));
                    if (i == Nn - 1) h[(long long)(E * Nn + j) * ostride] = -sx * (uu - gval(E * Nn + j));
                    if (j == 0) h[(long long)(S * Nn + i) * ostride] = sy * (uu - gval(S * Nn + i));
                    if (j == Nn - 1) h[(long long)(N * Nn + i) * ostride] = -sy * (uu - gval(N * Nn + i));
                }
    }
}

template __global__ void patch_fd_kernel<1>(PatchFDArgs p);
template __global__ void patch_fd_kernel<2>(PatchFDArgs p);

}  // namespace efk
