// ef_kernels.hpp -- kernel declarations and launch constants (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include "ef_device.hpp"

namespace efk {

__global__ void bt_factor_kernel(const LeafDesc* leaves, int Nn);
template <int NT, bool PIVOT>
__global__ void bt_factor_wave_kernel(const LeafDesc* leaves, int nleaves);
// Stencil and block-LU factors (Sinv) of every leaf with variable coefficients: one wave per
// leaf for N in {8, 16, 32}, the workgroup-per-leaf kernel otherwise.  dominant: every leaf's
// five-point matrix is diagonally dominant (cells_dominant), so the wave kernel may skip the
// pivot search.
void launch_bt_factor(const LeafDesc* leaves, int Nn, int nleaves, hipStream_t s, bool dominant = false);
__global__ void bt_solve_kernel(const LeafDesc* leaves, int Nn, int nleaves, int k, int mode, const double* f,
                                int f_zero, const double* garena, double* harena, double* u, long long vcs);
__global__ void bt_dtn_kernel(const LeafDesc* leaves, int Nn, int first, double* work);
template <int NT>
__global__ void bt_dtn_mfma_kernel(const LeafDesc* leaves, int first, double* work);
// DtN of every leaf with variable coefficients (through each leaf's T) in launches of `chunk`
// leaves; `work` holds chunk * N * N * 4N doubles.  N in {16, 32} on the matrix cores.
void launch_bt_dtn(const LeafDesc* leaves, int Nn, int nleaves, double* work, long long chunk, hipStream_t s);
template <int NT>
__global__ void bt_solve_wave_kernel(const LeafDesc* leaves, int nleaves, int k, int mode, const double* f,
                                     int f_zero, const double* garena, double* harena, double* u, long long vcs);
template <int NT>
__global__ void bt_solve_sym_kernel(const LeafDesc* leaves, int nleaves, int k, int mode, const double* f,
                                    int f_zero, const double* garena, double* harena, double* zbuf, long long vcs);
__global__ void bt_pack_kernel(const LeafDesc* leaves, int nleaves, int Nn);
// Leaf solves with variable coefficients: one wave per (leaf, rhs) for N in {8, 16, 32},
// the workgroup-per-leaf kernel otherwise.  Modes as bt_solve_kernel.  With `packed`, every
// leaf's SinvP holds its packed symmetric inverse pivot blocks (bt_pack_kernel), which the
// kernel for N in {8, 16, 32} reads instead of Sinv; in modes 0 and 2 it needs zwork,
// nleaves * k * N * N doubles of work space.  Boundary data and Neumann data of leaf L and
// right-hand side c: garena / harena + c * vcs + L.g_off / L.h_off, 4N consecutive doubles.
void launch_bt_solve(const LeafDesc* leaves, int Nn, int nleaves, int k, int mode, const double* f, int f_zero,
                     const double* garena, double* harena, double* u, hipStream_t s, bool packed = false,
                     double* zwork = nullptr, long long vcs = 0);
// Packed storage of the symmetric inverse pivot blocks: entries per block.
inline long long bt_packed_entries(int Nn) { return (long long)Nn * (Nn + 1) / 2; }
constexpr int BT_MAX_N = 32;  // largest leaf side of the block-tridiagonal path (LDS tiles)
template <int NB, bool FULL, int OP, int NV = 0>
__global__ void patch_fd_kernel(PatchFDArgs p, const int* __restrict__ cls, const PatchCoef* __restrict__ coef,
                                const long long* __restrict__ g_off, const long long* __restrict__ h_off);
template <bool AK, bool BK>
__global__ void gemm_strided_f64_kernel(GemmStridedArgs a);
__global__ void patch_large_fold_kernel(PatchLargeArgs a);
__global__ void patch_large_scale_kernel(PatchLargeArgs a);
__global__ void patch_large_out_kernel(PatchLargeArgs a);
__global__ void fill_kernel(double* p, double v, long long n);
__global__ void copy_kernel(const double* src, double* dst, long long n);

}  // namespace efk
