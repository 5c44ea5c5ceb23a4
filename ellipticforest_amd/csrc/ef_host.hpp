// ef_host.hpp -- host-side helpers shared by the C ABI sources of libefhps.so: the
// error type that every entry point turns into an EF_ERR_* code, HIP status checks and
// an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "ef_device.hpp"
#include "ef_hps.h"

namespace efk {
void set_last_error(const std::string& msg);  // thread-local message of ef_last_error()
// Host <-> device copies through a process-wide pinned staging buffer (hipHostMalloc),
// ordered on stream s and complete on return, so no caller memory is ever page-locked.
void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);
void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s);
// Fast-diagonalisation tables of an N x N patch for patch_fd_kernel (NB = 1 for N <= 16,
// 2 for N <= 32): the basis in both MFMA operand layouts and the eigenvalues of K.
void patch_fd_tables(int N, int NB, std::vector<double>& QA, std::vector<double>& QB, std::vector<double>& mu);
// 1/D of every coefficient entry on the padded (16 NB)^2 modes, D[k][l] = cx mu_k + cy mu_l + lam
// (zero on padded modes), entry after entry.
void patch_fd_dinv(int N, int NB, const std::vector<double>& mu, const std::vector<PatchCoef>& coef,
                   std::vector<double>& out);
// Launch patch_fd_kernel for a.op over at most max_blocks workgroups on stream s.
void launch_patch_fd(const PatchFDArgs& a, int NB, long long max_blocks, hipStream_t s);
// Workgroup cap of patch launches (EF_PATCH_MAX_BLOCKS or the default).
long long patch_default_blocks();
// True when the five-point matrix of a patch with per-cell coefficients c ([6][N*N]) is
// symmetric to 1e-14 relative (then its inverse pivot blocks can be kept packed).
bool cells_symmetric(const double* c, int N);
// Every row of the patch's five-point matrix diagonally dominant (block LU without pivoting).
bool cells_dominant(const double* c, int N, double dx, double dy);
}

// A context is one GPU.
struct ef_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
};

namespace {

struct EfError : std::runtime_error {
    int code;
    EfError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define EF_HIP(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) throw EfError(EF_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define EF_LAUNCH_CHECK_NAMED(name)                                                                  \
    do {                                                                                             \
        hipError_t e_ = hipGetLastError();                                                           \
        if (e_ != hipSuccess) throw EfError(EF_ERR_DEVICE, std::string("launch of ") + name + ": " + hipGetErrorString(e_)); \
    } while (0)
#define EF_LAUNCH_CHECK() EF_LAUNCH_CHECK_NAMED(__func__ + std::string(":") + std::to_string(__LINE__))

inline unsigned blocks_for(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// Owning device buffer.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count) EF_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    // Host -> device copy ordered on the stream the kernels that read it run on, and
    // complete before the host vector can go away.
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(v.size());
        if (v.empty()) return;
        efk::copy_h2d(p, v.data(), v.size() * sizeof(T), s);
    }
};

inline int fail(const EfError& e) {
    efk::set_last_error(e.what());
    return e.code;
}

// Every C entry point body sits between these: no exception crosses the ABI.
#define EF_API_BEGIN try {
#define EF_API_END                                              \
    }                                                           \
    catch (const EfError& e) { return fail(e); }                \
    catch (const std::bad_alloc&) {                             \
        efk::set_last_error("host out of memory");              \
        return EF_ERR_INTERNAL;                                 \
    }                                                           \
    catch (const std::exception& e) {                           \
        efk::set_last_error(e.what());                          \
        return EF_ERR_INTERNAL;                                 \
    }                                                           \
    return EF_OK;

}  // namespace
