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
// Work: one wave per (patch, right-hand side) item, four independent waves per
// workgroup, a grid-stride loop over the items (the basis tables stay in registers); for
// N <= 16 the next item's load and boundary data are fetched while the current one is
// computed, for N > 16 a second wave per SIMD hides the loads instead.
// No LDS.  Every product is done on v_mfma_f64_16x16x4_f64 in registers.  With
// q = lane/16 and r = lane%16, the gfx950 operand layouts for a 16x16 block are
//     A operand, step s: A[r][4s + q]        B operand, step s: B[4s + q][r]
//     accumulator slot v:  C[4v + q][r]
// so an accumulator block IS the B operand of the next product, and the accumulator of
// M^T is the A operand of M.  The chain is arranged so no value ever changes lanes:
//     Z1^T = B^T Q          (A = B^T loaded straight from memory, B = Q)
//     Z2   = Z1 Q           (A = Z1 from the accumulator of Z1^T)
//     Z3   = Z2 ./ D        (in the accumulator)
//     Z4^T = Z3^T Q^T       (A = Z3^T from the accumulator of Z3)
//     U    = Z4 Q^T         (A = Z4 from the accumulator of Z4^T)
// Patches with N <= 16 use one padded 16x16 block, N <= 32 a 2x2 block grid (NB);
// FULL marks N = 16 NB (no padding, no bounds tests).  The boundary vector of an item
// (4N values) is read with one coalesced load per 64 entries and handed to the lanes
// that need it by cross-lane shuffles.
#include <hip/hip_runtime.h>
#include <type_traits>
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

// Inputs of one item held in registers: the load in the A layout of B^T and the
// boundary vector spread over the lanes (entry t in lane t % 64 of register t / 64).
template <int NB>
struct PatchItemIn {
    double F[NB][NB][4];
    double G[NB];
    double T[4];  // N = 16 Neumann data / DtN: 1/D of this lane's modes from the entry's table
    int ce;       // the item's coefficient entry (uniform)
};

// Read-only index arrays of PatchFDArgs, passed to the kernel as separate __restrict__
// arguments: only then may the compiler read them with scalar loads inside the loop that
// stores the results (a vector load there would have to be waited for like the data loads).
struct PatchIdx {
    const int* cls;
    const PatchCoef* coef;
    const long long* g_off;
    const long long* h_off;
};

// HF, HG (unpadded variants): the launch has f, resp. g; without them nothing is loaded for them
// and the values are zeros.
template <int NB, bool FULL, int OP, bool HF = true, bool HG = true, int NV = 0>
__device__ inline void patch_fetch(const PatchFDArgs& p, const PatchIdx& ix, int patch, int rhs, int lane,
                                   PatchItemIn<NB>& in) {
    // FULL: N = 16 NB; NV: N fixed at compile time for a padded variant (0: p.N)
    const int Nn = FULL ? 16 * NB : (NV ? NV : p.N), q = lane >> 4, r = lane & 15;
    const long long slot = (long long)rhs * p.npatches + patch;
    in.ce = ix.cls ? ix.cls[patch] : patch;
    if (FULL) {
        // Every load is issued, with no branch around it and nothing that waits for it here:
        // a load inside a branch leaves the compiler unsure how many loads are in flight
        // after the branch, and it then waits for all of them -- the prefetched item
        // included.  Whether f and g exist is known at compile time (HF, HG).  The loads are
        // global (not flat) loads: a flat load in flight makes the compiler wait for every
        // load before any other wait.
#pragma unroll
        for (int bi = 0; bi < NB; bi++)
#pragma unroll
            for (int bk = 0; bk < NB; bk++)
#pragma unroll
                for (int s = 0; s < 4; s++)
                    in.F[bi][bk][s] = HF ? gmem(p.f + slot * Nn * Nn)[(16 * bk + 4 * s + q) * Nn + 16 * bi + r] : 0.0;
        if (NB == 1 && OP != PATCH_OP_SOLVE) {
            // the entry's 1/D table (the basis table when there is none: the kernel then
            // computes 1/D and ignores these values).  Measured: it saves the Neumann-data
            // variant 5 % and costs the solve 7 % (more registers, fewer waves), which computes.
            const double* tb = p.dinv ? p.dinv + (long long)in.ce * 256 : p.QB;
#pragma unroll
            for (int v = 0; v < 4; v++) in.T[v] = gmem(tb)[(4 * v + q) * 16 + r];
        }
        if (OP == PATCH_OP_DTN) return;
        if (!HG) {
#pragma unroll
            for (int m = 0; m < NB; m++) in.G[m] = 0.0;
            return;
        }
        // 4N = 64 NB entries, side t / N at offset so[side] from gb (the arena's per-side row
        // offsets of the patch, or consecutive sides); only the offsets are chosen in a branch
        const double* gb;
        long long so[4];
        if (ix.g_off) {
            gb = p.g + (long long)rhs * p.vk;
#pragma unroll
            for (int sd = 0; sd < 4; sd++) so[sd] = ix.g_off[4LL * patch + sd];
        } else {
            gb = p.g + slot * 4 * Nn;
#pragma unroll
            for (int sd = 0; sd < 4; sd++) so[sd] = (long long)sd * Nn;
        }
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int t = 64 * m + lane, sd = t / Nn;
            in.G[m] = gmem(gb)[pick4(so, sd) + (t - sd * Nn)];
        }
        return;
    }
    const double* f = p.f ? p.f + slot * Nn * Nn : nullptr;
#pragma unroll
    for (int bi = 0; bi < NB; bi++)
#pragma unroll
        for (int bk = 0; bk < NB; bk++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int i = 16 * bk + 4 * s + q, j = 16 * bi + r;
                double v = 0.0;
                if (f) {
                    if (FULL) {
                        v = f[i * Nn + j];
                    } else {
                        // clamped address, then select: no divergent branch around the load
                        const double x = f[min(i, Nn - 1) * Nn + min(j, Nn - 1)];
                        v = (i < Nn && j < Nn) ? x : 0.0;
                    }
                }
                in.F[bi][bk][s] = v;
            }
    if (OP == PATCH_OP_DTN) return;
    if (p.g) {
        const double* gb;
        long long so[4];
        if (ix.g_off) {
            gb = p.g + (long long)rhs * p.vk;
#pragma unroll
            for (int sd = 0; sd < 4; sd++) so[sd] = ix.g_off[4LL * patch + sd];
        } else {
            gb = p.g + slot * 4 * Nn;
#pragma unroll
            for (int sd = 0; sd < 4; sd++) so[sd] = (long long)sd * Nn;
        }
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int t = 64 * m + lane, sd = min(t / Nn, 3);
            const double x = gb[pick4(so, sd) + min(t - sd * Nn, Nn - 1)];
            in.G[m] = t < 4 * Nn ? x : 0.0;
        }
    } else {
#pragma unroll
        for (int m = 0; m < NB; m++) in.G[m] = 0.0;
    }
}

// Boundary value at WESN position t of the current item (t may differ per lane).
template <int NB, int OP>
__device__ inline double patch_g(const PatchItemIn<NB>& in, int t, int col) {
    if (OP == PATCH_OP_DTN) return (t == col) ? 1.0 : 0.0;
    if (NB == 1) return __shfl(in.G[0], t & 63);
    const double a = __shfl(in.G[0], t & 63);
    const double b = __shfl(in.G[NB - 1], t & 63);
    return (t < 64) ? a : b;
}

// Waves per SIMD (minimum asked of the compiler): five for the N = 16 solve and four for the
// N = 16 Neumann data, which overlaps one wave's loads with another's MFMA chain (both compile
// to under 81 registers, six waves; seven would spill).  NB = 2 fetches each
// item when its turn comes instead of one item ahead (a prefetched item costs 64 registers),
// which fits two waves per SIMD in 256 registers: a second wave hides the loads better than
// the prefetch did (N = 32, 4,096 patches x 4 right-hand sides: 186 -> 145 us).  The NB = 2
// DtN would spill at two waves and keeps one.
template <int NB, bool FULL, int OP>
constexpr int patch_fd_waves() {
    if (NB == 1) return !FULL ? 1 : (OP == PATCH_OP_SOLVE ? 5 : (OP == PATCH_OP_MAP_D2N ? 4 : 1));
    return OP == PATCH_OP_DTN ? 1 : 2;
}

template <int NB, bool FULL, int OP, int NV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(patch_fd_waves<NB, FULL, OP>(), 8)))
patch_fd_kernel(PatchFDArgs p, const int* __restrict__ cls, const PatchCoef* __restrict__ coef,
                const long long* __restrict__ g_off, const long long* __restrict__ h_off) {
    const PatchIdx ix{cls, coef, g_off, h_off};
    // the wave index is uniform: keep it (and every item index and base address derived
    // from it) in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q = lane >> 4, r = lane & 15;
    const int Nn = FULL ? 16 * NB : (NV ? NV : p.N);  // compile-time for FULL (N = 16 NB) and NV
    // items: (rhs, patch) rhs-major, or (patch, DtN column) patch-major
    const int inner = (OP == PATCH_OP_DTN) ? 4 * Nn : p.npatches;
    const int outer = (OP == PATCH_OP_DTN) ? p.npatches : p.nrhs;
    const long long first = (long long)blockIdx.x * 4 + wave;
    if (first >= (long long)inner * outer) return;  // uniform per wave
    const int stride = (int)gridDim.x * 4;
    const int step_o = stride / inner, step_i = stride % inner;

    // basis tables, loaded once per wave: QA[x][y][s] = Q[16x + r][16y + 4s + q],
    // QB[x][y][s] = Q[16x + 4s + q][16y + r]; eigenvalues of this lane's modes
    double QA[NB][NB][4], QB[NB][NB][4], muk[NB][4], mul[NB];
#pragma unroll
    for (int x = 0; x < NB; x++) {
#pragma unroll
        for (int y = 0; y < NB; y++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                QA[x][y][s] = p.QA[(((x * NB) + y) * 4 + s) * 64 + lane];
                QB[x][y][s] = p.QB[(((x * NB) + y) * 4 + s) * 64 + lane];
            }
#pragma unroll
        for (int v = 0; v < 4; v++) muk[x][v] = p.mu[16 * x + 4 * v + q];
        mul[x] = p.mu[16 * x + r];
    }

    // The item loop, compiled for each combination of f and g present (unpadded variants; the
    // padded ones test f and g as they load)
    auto run = [&](auto HFc, auto HGc) {
        constexpr bool HF = decltype(HFc)::value, HG = decltype(HGc)::value;
        int o = (int)(first / inner), in_i = (int)(first % inner);
        constexpr bool PREFETCH = NB == 1;
        PatchItemIn<NB> cur{}, nxt{};
        if (PREFETCH) {
            const int patch = (OP == PATCH_OP_DTN) ? o : in_i, rhs = (OP == PATCH_OP_DTN) ? 0 : o;
            patch_fetch<NB, FULL, OP, HF, HG, NV>(p, ix, patch, rhs, lane, cur);
        }
        // One item: (o, in_i) from cur, with (o2, i2) fetched into nxt meanwhile (NB = 1).  The
        // loop below alternates the two buffers, so no item is ever copied between registers (a
        // copy would wait for the prefetch to arrive at the end of the item it overlaps).
        auto item = [&](const int o, const int in_i, const int o2, const int i2, PatchItemIn<NB>& cur,
                        PatchItemIn<NB>& nxt) {
            const int patch = (OP == PATCH_OP_DTN) ? o : in_i;
            const int rhs = (OP == PATCH_OP_DTN) ? 0 : o;
            const int col = (OP == PATCH_OP_DTN) ? in_i : -1;
            if (!PREFETCH) patch_fetch<NB, FULL, OP, HF, HG, NV>(p, ix, patch, rhs, lane, cur);
            if (PREFETCH) {
                // past the last item the current one is fetched again: the fetch is never skipped
                const bool more = o2 < outer;
                const int oo = more ? o2 : o, ii = more ? i2 : in_i;
                const int patch2 = (OP == PATCH_OP_DTN) ? oo : ii, rhs2 = (OP == PATCH_OP_DTN) ? 0 : oo;
                patch_fetch<NB, FULL, OP, HF, HG, NV>(p, ix, patch2, rhs2, lane, nxt);
                // keep these loads ahead of the current item's work: issued later, they would
                // shorten the time they have to arrive
                __builtin_amdgcn_sched_barrier(0);
            }
            const PatchCoef cf = ix.coef[cur.ce];
            const double* dinv_row = p.dinv ? p.dinv + (long long)cur.ce * (16 * NB) * (16 * NB) : nullptr;
            const bool has_g = (OP == PATCH_OP_DTN) || p.g != nullptr;

            // fold the boundary data into the load: -2 (c_W g_W + c_E g_E) on the x-edge
            // cells, -2 (c_S g_S + c_N g_N) on the y-edge cells
            double BT[NB][NB][4];
            // N = 16: element s of lane (q, r) is cell (i, j) = (4s + q, r), both in the load (A
            // layout of B^T) and in the solution (accumulator layout of U).  x-edges: (s, q) = (0, 0)
            // takes g_W[r] (entry r), (3, 3) takes g_E[r] (entry 16 + r); y-edges: r = 0 takes
            // g_S[4s + q] (entry 32 + 4s + q), r = 15 takes g_N[4s + q] (entry 48 + 4s + q).  The
            // shuffled values serve the folding here and the Neumann data at the end.
            double xg = 0.0, yg[4] = {0.0, 0.0, 0.0, 0.0};
            if (NB == 1 && FULL && OP != PATCH_OP_DTN) {
                const double ax = -2.0 * cf.cx, ay = (r == 0 || r == 15) ? -2.0 * cf.cy : 0.0;
                const int ybase = (r == 0) ? 32 + q : 48 + q;
                xg = __shfl(cur.G[0], q == 0 ? r : 16 + r);  // G is zero without g
    #pragma unroll
                for (int s = 0; s < 4; s++) {
                    yg[s] = __shfl(cur.G[0], ybase + 4 * s);
                    double v = fma(ay, yg[s], cur.F[0][0][s]);
                    if ((s == 0 && q == 0) || (s == 3 && q == 3)) v = fma(ax, xg, v);
                    BT[0][0][s] = v;
                }
            } else if (NB == 2 && FULL && OP != PATCH_OP_DTN) {
                // N = 32: element s of block (bi, bk) of lane (q, r) is cell (i, j) = (16 bk + 4s + q,
                // 16 bi + r).  G[0] holds g_W (entries 0-31) and g_E (32-63), G[1] g_S and g_N.
                // x-edges: (bk, s, q) = (0, 0, 0) takes g_W[j], (1, 3, 3) takes g_E[j];
                // y-edges: (bi, r) = (0, 0) takes g_S[i], (1, 15) takes g_N[i].
                const double ax = -2.0 * cf.cx, ay = -2.0 * cf.cy;
                const int xsel = (q == 0) ? 0 : 32, ysel = (r == 0) ? 0 : 32;
    #pragma unroll
                for (int bi = 0; bi < NB; bi++) {
                    const double xg2 = __shfl(cur.G[0], xsel + 16 * bi + r);
                    const bool yedge = (bi == 0) ? (r == 0) : (r == 15);
    #pragma unroll
                    for (int bk = 0; bk < NB; bk++)
    #pragma unroll
                        for (int s = 0; s < 4; s++) {
                            const double yg2 = __shfl(cur.G[NB - 1], ysel + 16 * bk + 4 * s + q);
                            const double v = fma(yedge ? ay : 0.0, yg2, cur.F[bi][bk][s]);
                            const bool xedge = (bk == 0 && s == 0 && q == 0) || (bk == 1 && s == 3 && q == 3);
                            BT[bi][bk][s] = fma(xedge ? ax : 0.0, xg2, v);
                        }
                }
            } else {
    #pragma unroll
            for (int bi = 0; bi < NB; bi++) {
                const int j = 16 * bi + r;
                const double gw = has_g ? patch_g<NB, OP>(cur, W * Nn + min(j, Nn - 1), col) : 0.0;
                const double ge = has_g ? patch_g<NB, OP>(cur, E * Nn + min(j, Nn - 1), col) : 0.0;
    #pragma unroll
                for (int bk = 0; bk < NB; bk++)
    #pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const int i = 16 * bk + 4 * s + q;
                        double v = cur.F[bi][bk][s];
                        if (has_g) {
                            const double gsv = patch_g<NB, OP>(cur, S * Nn + min(i, Nn - 1), col);
                            const double gnv = patch_g<NB, OP>(cur, N * Nn + min(i, Nn - 1), col);
                            const bool inb = FULL || (i < Nn && j < Nn);
                            const double ax = (i == 0 ? gw : 0.0) + (i == Nn - 1 ? ge : 0.0);
                            const double ay = (j == 0 ? gsv : 0.0) + (j == Nn - 1 ? gnv : 0.0);
                            v += inb ? -2.0 * (cf.cx * ax + cf.cy * ay) : 0.0;
                        }
                        BT[bi][bk][s] = v;
                    }
            }
            }

            pd4 Z[NB][NB], Y[NB][NB];
            // Z1^T = B^T Q
            patch_mm<NB>([&](int bi, int bk, int s) { return BT[bi][bk][s]; },
                         [&](int bk, int bj, int s) { return QB[bk][bj][s]; }, Z);
            // Z2 = Z1 Q, A operand of Z1 = accumulator of Z1^T with the block indices swapped
            patch_mm<NB>([&](int bi, int bk, int s) { return Z[bk][bi][s]; },
                         [&](int bk, int bj, int s) { return QB[bk][bj][s]; }, Y);
            // Z3 = Z2 ./ D; padded modes (k or l >= N) are zero
            if (NB == 1 && FULL) {
                if (OP != PATCH_OP_SOLVE && p.dinv) {
                    // the table came with the item (a uniform branch, no memory access either side)
    #pragma unroll
                    for (int v = 0; v < 4; v++) Y[0][0][v] *= cur.T[v];
                } else {
    #pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const double d = cf.cx * muk[0][v] + cf.cy * mul[0] + cf.lam;
                        double rc = __builtin_amdgcn_rcp(d);
                        rc = fma(rc, fma(-d, rc, 1.0), rc);
                        rc = fma(rc, fma(-d, rc, 1.0), rc);
                        Y[0][0][v] *= rc;
                    }
                }
            } else
    #pragma unroll
            for (int bi = 0; bi < NB; bi++)
    #pragma unroll
                for (int bj = 0; bj < NB; bj++)
    #pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int kk = 16 * bi + 4 * v + q, ll = 16 * bj + r;
                        if (p.dinv) {
                            Y[bi][bj][v] *= dinv_row[kk * (16 * NB) + ll];
                        } else {
                            const double d = cf.cx * muk[bi][v] + cf.cy * mul[bj] + cf.lam;
                            // 1/d: hardware estimate and two Newton steps (relative error ~ 1 ulp)
                            double rc = __builtin_amdgcn_rcp(d);
                            rc = fma(rc, fma(-d, rc, 1.0), rc);
                            rc = fma(rc, fma(-d, rc, 1.0), rc);
                            const double z = Y[bi][bj][v] * rc;
                            Y[bi][bj][v] = (FULL || (kk < Nn && ll < Nn)) ? z : 0.0;
                        }
                    }
            // Z4^T = Z3^T Q^T: A operand of Z3^T = accumulator of Z3 swapped; B operand of Q^T = QA swapped
            patch_mm<NB>([&](int bi, int bk, int s) { return Y[bk][bi][s]; },
                         [&](int bk, int bj, int s) { return QA[bj][bk][s]; }, Z);
            // U = Z4 Q^T
            patch_mm<NB>([&](int bi, int bk, int s) { return Z[bk][bi][s]; },
                         [&](int bk, int bj, int s) { return QA[bj][bk][s]; }, Y);

            if (OP == PATCH_OP_SOLVE) {
                double* u = p.out + ((long long)rhs * p.npatches + patch) * Nn * Nn;
    #pragma unroll
                for (int bi = 0; bi < NB; bi++)
    #pragma unroll
                    for (int bj = 0; bj < NB; bj++)
    #pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int i = 16 * bi + 4 * v + q, j = 16 * bj + r;
                            // streaming stores: u is not read again by this launch
                            if (FULL || (i < Nn && j < Nn)) __builtin_nontemporal_store(Y[bi][bj][v], gmem(u) + i * Nn + j);
                        }
            } else {
                // Neumann data (mapD2N, FiniteVolumeSolver.cpp:296-353):
                //   h_W = (2/dx)(u_W - g_W), h_E = -(2/dx)(u_E - g_E),
                //   h_S = (2/dy)(u_S - g_S), h_N = -(2/dy)(u_N - g_N)
                // into h of this (patch, rhs), or column col of the patch's DtN.
                double* h;
                long long ostride;
                if (OP == PATCH_OP_DTN) {
                    h = p.out + (long long)patch * 16 * Nn * Nn + col;
                    ostride = 4 * Nn;
                } else if (ix.h_off) {
                    h = p.out + (long long)rhs * p.vk + ix.h_off[patch];
                    ostride = 1;
                } else {
                    h = p.out + ((long long)rhs * p.npatches + patch) * 4 * Nn;
                    ostride = 1;
                }
                const double sx = 2.0 / cf.dx, sy = 2.0 / cf.dy;
                if (NB == 1 && FULL && OP == PATCH_OP_MAP_D2N) {
                    // Lane t makes entry t of h (4N = 64 entries), so the item ends in one store
                    // issued by every lane.  u of lane (q, r), slot v is cell (4v + q, r); entry t
                    // needs: W (t < 16) cell (0, t): lane t, slot 0; E cell (15, t - 16): lane
                    // 48 + t - 16, slot 3; S cell (i, 0), i = t - 32: lane 16 (i % 4), slot i / 4;
                    // N cell (i, 15), i = t - 48: lane 16 (i % 4) + 15, slot i / 4.  Its boundary
                    // value g[t] is already in lane t.
                    const int side = lane >> 4, e = lane & 15;
                    const int src = side == W ? e : side == E ? 48 + e : 16 * (e & 3) + (side == N ? 15 : 0);
                    const int slot = side == W ? 0 : side == E ? 3 : e >> 2;
                    const double u0 = __shfl(Y[0][0][0], src), u1 = __shfl(Y[0][0][1], src);
                    const double u2 = __shfl(Y[0][0][2], src), u3 = __shfl(Y[0][0][3], src);
                    const double uu = slot == 0 ? u0 : slot == 1 ? u1 : slot == 2 ? u2 : u3;
                    const double sc = side == W ? sx : side == E ? -sx : side == S ? sy : -sy;
                    h[(long long)lane * ostride] = sc * (uu - cur.G[0]);
                    return;
                }
                if (NB == 2 && FULL && OP == PATCH_OP_MAP_D2N) {
                    // Lane t makes entries t (W, E) and 64 + t (S, N) of h (4N = 128 entries): two
                    // stores issued by every lane.  u of lane (q, r), block (bi, bj), slot v is cell
                    // (16 bi + 4v + q, 16 bj + r).  W[j] = cell (0, j): lane j % 16, block (0, j / 16),
                    // slot 0; E[j] = cell (31, j): lane 48 + j % 16, block (1, j / 16), slot 3;
                    // S[i] = cell (i, 0): lane 16 (i % 4), block (i / 16, 0), slot (i / 4) % 4;
                    // N[i] = cell (i, 31): lane 16 (i % 4) + 15, block (i / 16, 1), slot (i / 4) % 4.
                    {
                        const bool east = lane >= 32;
                        const int j = lane & 31, bj = j >> 4;
                        const int src = (east ? 48 : 0) + (j & 15);
                        // W takes block (0, bj) slot 0, E block (1, bj) slot 3
                        const double w0 = __shfl(Y[0][0][0], src), w1 = __shfl(Y[0][1][0], src);
                        const double e0 = __shfl(Y[1][0][3], src), e1 = __shfl(Y[1][1][3], src);
                        const double uu = east ? (bj ? e1 : e0) : (bj ? w1 : w0);
                        h[(long long)lane * ostride] = (east ? -sx : sx) * (uu - cur.G[0]);
                    }
                    {
                        const bool north = lane >= 32;
                        const int i = lane & 31, bi = i >> 4, v = (i >> 2) & 3;
                        const int src = 16 * (i & 3) + (north ? 15 : 0);
                        double uu = 0.0;
    #pragma unroll
                        for (int x = 0; x < 2; x++)
    #pragma unroll
                            for (int y = 0; y < 2; y++)
    #pragma unroll
                                for (int w = 0; w < 4; w++) {
                                    const double t = __shfl(Y[x][y][w], src);
                                    uu = (bi == x && (north ? 1 : 0) == y && v == w) ? t : uu;
                                }
                        h[(long long)(64 + lane) * ostride] = (north ? -sy : sy) * (uu - cur.G[1]);
                    }
                    return;
                }
    #pragma unroll
                for (int bi = 0; bi < NB; bi++)
    #pragma unroll
                    for (int bj = 0; bj < NB; bj++)
    #pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int i = 16 * bi + 4 * v + q, j = 16 * bj + r;
                            const int ic = min(i, Nn - 1), jc = min(j, Nn - 1);
                            const double uu = Y[bi][bj][v];
                            const double gw = has_g ? patch_g<NB, OP>(cur, W * Nn + jc, col) : 0.0;
                            const double ge = has_g ? patch_g<NB, OP>(cur, E * Nn + jc, col) : 0.0;
                            const double gsv = has_g ? patch_g<NB, OP>(cur, S * Nn + ic, col) : 0.0;
                            const double gnv = has_g ? patch_g<NB, OP>(cur, N * Nn + ic, col) : 0.0;
                            if (i >= Nn || j >= Nn) continue;
                            if (i == 0) h[(long long)(W * Nn + j) * ostride] = sx * (uu - gw);
                            if (i == Nn - 1) h[(long long)(E * Nn + j) * ostride] = -sx * (uu - ge);
                            if (j == 0) h[(long long)(S * Nn + i) * ostride] = sy * (uu - gsv);
                            if (j == Nn - 1) h[(long long)(N * Nn + i) * ostride] = -sy * (uu - gnv);
                        }
            }
        };
        auto advance = [&](int oo, int ii, int& o2, int& i2) {
            o2 = oo + step_o;
            i2 = ii + step_i;
            if (i2 >= inner) { i2 -= inner; o2++; }
        };
        while (o < outer) {
            int o2, i2;
            advance(o, in_i, o2, i2);
            item(o, in_i, o2, i2, cur, nxt);
            o = o2;
            in_i = i2;
            if (!PREFETCH || o >= outer) continue;
            advance(o, in_i, o2, i2);
            item(o, in_i, o2, i2, nxt, cur);
            o = o2;
            in_i = i2;
        }
    };
    using T1 = std::true_type;
    using T0 = std::false_type;
    const bool hf = p.f != nullptr, hg = p.g != nullptr;
    if (!FULL || (hf && hg)) run(T1{}, T1{});
    else if (hf) run(T1{}, T0{});
    else if (hg) run(T0{}, T1{});
    else run(T0{}, T0{});
}

#define EF_PATCH_INSTANCES(NB, FULL, NV)                                                                      \
    template __global__ void patch_fd_kernel<NB, FULL, PATCH_OP_SOLVE, NV>(PatchFDArgs, const int*, const PatchCoef*, \
                                                                      const long long*, const long long*);   \
    template __global__ void patch_fd_kernel<NB, FULL, PATCH_OP_MAP_D2N, NV>(PatchFDArgs, const int*,              \
                                                                        const PatchCoef*, const long long*, \
                                                                        const long long*);                  \
    template __global__ void patch_fd_kernel<NB, FULL, PATCH_OP_DTN, NV>(PatchFDArgs, const int*, const PatchCoef*, \
                                                                    const long long*, const long long*);
EF_PATCH_INSTANCES(1, true, 0)
EF_PATCH_INSTANCES(1, false, 0)
EF_PATCH_INSTANCES(1, false, 4)
EF_PATCH_INSTANCES(1, false, 8)
EF_PATCH_INSTANCES(1, false, 12)
EF_PATCH_INSTANCES(2, true, 0)
EF_PATCH_INSTANCES(2, false, 0)
EF_PATCH_INSTANCES(2, false, 24)

}  // namespace efk
