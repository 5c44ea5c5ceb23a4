// ef_host.hpp -- host-side helpers shared by the C ABI sources of libefhps.so: the error type
// that every entry point turns into an EF_ERR_* code, HIP status checks, an owning device
// buffer, and the launch checks every kernel launch of the tree path goes through.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
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
// Device address of the no-allocation scratch array on `device` (current), taken once per device.
void noalloc_setup(int device);

struct EfError : std::runtime_error {
    int code;
    EfError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// EF_SYNC_CHECK=1: every launch of the tree path prints its name before it starts and is
// waited for and checked right after, so a failing kernel is named by the last line printed.
inline bool sync_check() {
    static const bool on = [] {
        const char* e = std::getenv("EF_SYNC_CHECK");
        return e && std::strcmp(e, "0") != 0;
    }();
    return on;
}
}  // namespace efk

#define EF_HIP(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw efk::EfError(EF_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define EF_LAUNCH_CHECK_NAMED(name)                                                                  \
    do {                                                                                             \
        hipError_t e_ = hipGetLastError();                                                           \
        if (e_ != hipSuccess)                                                                        \
            throw efk::EfError(EF_ERR_DEVICE, std::string("launch of ") + name + ": " + hipGetErrorString(e_)); \
    } while (0)
#define EF_LAUNCH_CHECK() EF_LAUNCH_CHECK_NAMED(__func__ + std::string(":") + std::to_string(__LINE__))

namespace efk {

inline unsigned blocks_for(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// gridDim.x of a descriptor-per-row vector launch (cnt descriptors of at most n elements each,
// the kernel striding over its elements): about 2048 workgroups in all, never more than n needs.
// One workgroup per 256 elements of the longest descriptor had dispatched mostly empty
// workgroups on levels with hundreds of short descriptors.
inline unsigned vec_blocks(long long n, int cnt) {
    const long long cap = cnt > 0 ? (2048 + cnt - 1) / cnt : 1;
    const long long b = blocks_for(n);
    return (unsigned)(b < cap ? b : (cap < 1 ? 1 : cap));
}

// Grid and block of a launch against the limits of a gfx950 dispatch; the name is printed
// first in EF_SYNC_CHECK mode.
inline void pre_launch(const char* name, dim3 grid, dim3 block) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0 || grid.y > 65535 || grid.z > 65535 || block.x == 0 ||
        block.x * block.y * block.z > 1024 || (unsigned long long)grid.x * block.x > 0xffffffffULL)
        throw EfError(EF_ERR_INTERNAL, std::string("launch of ") + name + ": grid/block outside the dispatch limits");
    if (sync_check()) {
        std::fprintf(stderr, "[ef-launch] %s grid=(%u,%u,%u) block=%u\n", name, grid.x, grid.y, grid.z, block.x);
        std::fflush(stderr);
    }
}
inline void post_launch(const char* name, hipStream_t s) {
    EF_LAUNCH_CHECK_NAMED(name);
    if (sync_check()) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess)
            throw EfError(EF_ERR_DEVICE, std::string("kernel ") + name + " failed: " + hipGetErrorString(e));
    }
}

// Launch with the checks above (no dynamic LDS in the tree path).
#define EF_LAUNCH(name, kern, grid, block, stream, ...)                     \
    do {                                                                    \
        const dim3 g_ = (grid), b_ = (block);                              \
        efk::pre_launch(name, g_, b_);                                      \
        hipLaunchKernelGGL(kern, g_, b_, 0, stream, __VA_ARGS__);           \
        efk::post_launch(name, stream);                                     \
    } while (0)

// [p, p + bytes) must lie inside one device allocation (the runtime's record of it).
inline void check_range(const void* p, size_t bytes, const char* what) {
    if (bytes == 0) return;
    if (p == nullptr) throw EfError(EF_ERR_INTERNAL, std::string("null device pointer: ") + what);
    void* base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
        (void)hipGetLastError();
        throw EfError(EF_ERR_INTERNAL, std::string("pointer outside every device allocation: ") + what);
    }
    const uintptr_t b = (uintptr_t)base, q = (uintptr_t)p;
    if (q < b || q + bytes > b + size)
        throw EfError(EF_ERR_INTERNAL, std::string("device range past the end of its allocation: ") + what);
}
// A row-major matrix of `rows` rows, `cols` used columns, leading dimension ld (elements of T).
template <class T>
inline void check_matrix(const T* p, long long rows, long long cols, long long ld, const char* what) {
    if (rows <= 0 || cols <= 0) return;
    if (ld < cols) throw EfError(EF_ERR_INTERNAL, std::string("leading dimension below the column count: ") + what);
    check_range(p, (size_t)(((rows - 1) * ld + cols) * (long long)sizeof(T)), what);
}

// Element (i, j) of a rows x cols array at p[i * rs + j * cs] (elements of T).
template <class T>
inline void check_strided(const T* p, long long rows, long long cols, long long rs, long long cs, const char* what) {
    if (rows <= 0 || cols <= 0) return;
    if (rs < 0 || cs < 0) throw EfError(EF_ERR_INTERNAL, std::string("negative stride: ") + what);
    check_range(p, (size_t)(((rows - 1) * rs + (cols - 1) * cs + 1) * (long long)sizeof(T)), what);
}

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
    // complete before the host vector can go away.  The allocation is kept when it is
    // large enough (n is the capacity).
    void upload(const std::vector<T>& v, hipStream_t s) {
        ensure(v.size());
        if (v.empty()) return;
        copy_h2d(p, v.data(), v.size() * sizeof(T), s);
    }
};

// Descriptor tables of one plan, packed into a few device blocks.  stage() copies a table
// into the host mirror of the current block and returns its device address at once;
// flush() uploads everything staged since the last flush (one copy per block).  Nothing a
// kernel reads from here is valid on the device before flush().
struct DevStore {
    struct Block {
        DevBuf<char> dev;
        std::vector<char> host;
        size_t used = 0, flushed = 0;
    };
    static constexpr size_t BLOCK = size_t(1) << 20, ALIGN = 256;
    std::vector<std::unique_ptr<Block>> blocks;
    std::vector<std::unique_ptr<DevBuf<char>>> bufs;  // device-only work (alloc_doubles)
    template <class T>
    T* stage(const std::vector<T>& v) {
        if (v.empty()) return nullptr;
        const size_t bytes = v.size() * sizeof(T);
        Block* b = blocks.empty() ? nullptr : blocks.back().get();
        size_t off = b ? (b->used + ALIGN - 1) / ALIGN * ALIGN : 0;
        if (!b || off + bytes > b->host.size()) {
            auto nb = std::make_unique<Block>();
            const size_t cap = std::max(BLOCK, (bytes + ALIGN - 1) / ALIGN * ALIGN);
            nb->dev.alloc(cap);
            nb->host.resize(cap);
            b = nb.get();
            blocks.push_back(std::move(nb));
            off = 0;
        }
        std::memcpy(b->host.data() + off, v.data(), bytes);
        b->used = off + bytes;
        return (T*)(b->dev.p + off);
    }
    bool pending() const {
        for (auto& b : blocks)
            if (b->used > b->flushed) return true;
        return false;
    }
    void flush(hipStream_t s) {
        for (auto& b : blocks)
            if (b->used > b->flushed) {
                copy_h2d(b->dev.p + b->flushed, b->host.data() + b->flushed, b->used - b->flushed, s);
                b->flushed = b->used;
            }
    }
    double* alloc_doubles(size_t n) {
        auto b = std::make_unique<DevBuf<char>>();
        b->alloc(n * sizeof(double));
        double* out = (double*)b->p;
        bufs.push_back(std::move(b));
        return out;
    }
    void clear() {
        blocks.clear();
        bufs.clear();
    }
};

// Static resources of a kernel against the invariants of the tree path: no scratch, static
// LDS below 64 KiB.
inline void check_kernel_resources(const void* fn, const char* name) {
    hipFuncAttributes a;
    EF_HIP(hipFuncGetAttributes(&a, fn));
    if (a.localSizeBytes != 0)
        throw EfError(EF_ERR_INTERNAL, std::string("kernel ") + name + " uses scratch memory (" +
                                           std::to_string(a.localSizeBytes) + " B per lane)");
    if (a.sharedSizeBytes > 64 * 1024)
        throw EfError(EF_ERR_INTERNAL, std::string("kernel ") + name + " uses more than 64 KiB of LDS");
}

inline int fail(const EfError& e) {
    set_last_error(e.what());
    return e.code;
}

}  // namespace efk

// Every C entry point body sits between these: no exception crosses the ABI.
#define EF_API_BEGIN try {
#define EF_API_END                                              \
    }                                                           \
    catch (const efk::EfError& e) { return efk::fail(e); }      \
    catch (const std::bad_alloc&) {                             \
        efk::set_last_error("host out of memory");              \
        return EF_ERR_INTERNAL;                                 \
    }                                                           \
    catch (const std::exception& e) {                           \
        efk::set_last_error(e.what());                          \
        return EF_ERR_INTERNAL;                                 \
    }                                                           \
    return EF_OK;

// A context is one GPU.
struct ef_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
};
