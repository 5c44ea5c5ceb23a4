// hip_cpu_shim.h -- CPU emulation of the HIP subset used by ellipticforest_amd/csrc,
// for checking kernel indexing on the host (test infrastructure only).
//
// * every device allocation is an mmap'd block whose end touches a PROT_NONE page,
//   so any read or write past the end of a buffer faults at once on the CPU;
// * a workgroup is 256 std::threads with a block barrier (__syncthreads) and a
//   64-lane barrier per wave; the f64 16x16x4 MFMA is emulated with the lane layout
//   of gfx950 (A[l%16][l/16], B[l/16][l%16], D slot v of lane l = D[4v + l/16][l%16]);
// * blocks run one after another; dynamic shared memory is a guarded buffer too.
#pragma once
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __constant__
#define __launch_bounds__(x)
#define __align__(n) __attribute__((aligned(n)))
using std::max;
using std::min;

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x, y, z; };
extern thread_local emu_idx threadIdx;
extern emu_idx blockIdx;
extern dim3 blockDim, gridDim;
void __syncthreads();
void emu_wave_sync();
void* emu_dyn_smem();

typedef int hipError_t;
typedef void* hipStream_t;
enum { hipSuccess = 0 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize };
struct hipFuncAttributes { size_t sharedSizeBytes = 0; size_t localSizeBytes = 0; int numRegs = 0; int maxThreadsPerBlock = 1024; };
inline const char* hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipMalloc(void** p, size_t bytes);
template <class T> hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
hipError_t hipFree(void* p);
// allocation containing ptr (base and size in bytes), as the HIP runtime reports it
hipError_t hipMemGetAddressRange(void** base, size_t* size, void* ptr);
// EF_EMU_PTRCHECK=1: report every aligned 8-byte word of a kernel argument or of a host-to-device
// copy that points into mapped memory outside every device allocation.  Such a word is a host
// address handed to the device, which works here (one address space) and faults a GPU.
bool emu_ptrcheck_on();
void emu_ptrcheck(const void* p, size_t n, const char* what, const void* kernel, int index);
// Stream capture: kernels launched while a capture is open are recorded, not run; copies,
// fills and synchronisations are errors then (the product must capture kernels only).
bool emu_capturing();
enum { hipErrorStreamCaptureUnsupported = 900 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
struct EmuGraph;
typedef EmuGraph* hipGraph_t;
typedef EmuGraph* hipGraphExec_t;
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) {
    if (emu_capturing()) return hipErrorStreamCaptureUnsupported;
    if (n) std::memcpy(d, s, n);
    if (n && k == hipMemcpyHostToDevice && emu_ptrcheck_on()) emu_ptrcheck(d, n, "host-to-device copy", nullptr, -1);
    return 0;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
inline hipError_t hipMemset(void* d, int v, size_t n) {
    if (emu_capturing()) return hipErrorStreamCaptureUnsupported;
    std::memset(d, v, n);
    return 0;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return 0; }
inline hipError_t hipSetDevice(int) { return 0; }
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = (hipStream_t)0x1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return emu_capturing() ? hipErrorStreamCaptureUnsupported : 0; }
#define hipHostMallocDefault 0u
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { *p = std::malloc(n ? n : 1); return *p ? 0 : 2; }
inline hipError_t hipHostFree(void* p) { std::free(p); return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*) { *a = hipFuncAttributes{}; return 0; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return 0; }
// device symbols are host objects here: their address is the symbol's own
#define HIP_SYMBOL(X) ((const void*)&(X))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return 0; }

// events: host wall clock (launches run synchronously in the emulation)
typedef double* hipEvent_t;
inline double emu_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new double(0.0); return 0; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return 0; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { *e = emu_now_ms(); return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
enum { hipErrorNotReady = 600, hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = new double(0.0); return 0; }
inline hipError_t hipEventQuery(hipEvent_t) { return 0; }  // emulated work is complete on return
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(*b - *a); return 0; }

typedef double emu_d4 __attribute__((ext_vector_type(4)));
emu_d4 emu_mfma_f64(double a, double b, emu_d4 c, int, int, int);
// cross-lane read within a wave (ds_bpermute semantics: lane reads src % 64 of its wave)
double __shfl(double v, int src, int width = 64);
inline double __shfl_xor(double v, int mask, int width = 64) { return __shfl(v, (int)(threadIdx.x % 64) ^ mask, width); }
#define __builtin_amdgcn_mfma_f64_16x16x4f64 emu_mfma_f64
typedef float emu_f4 __attribute__((ext_vector_type(4)));
emu_f4 emu_mfma_f32(float a, float b, emu_f4 c, int, int, int);
#define __builtin_amdgcn_mfma_f32_16x16x4f32 emu_mfma_f32
inline int __double2loint(double x) { unsigned long long u; std::memcpy(&u, &x, 8); return (int)(unsigned)u; }
inline int __double2hiint(double x) { unsigned long long u; std::memcpy(&u, &x, 8); return (int)(unsigned)(u >> 32); }
inline double __hiloint2double(int hi, int lo) {
    const unsigned long long u = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}
// wave vote: bit l set when lane l's predicate holds
unsigned long long __ballot(int pred);
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
// v_readlane_b32: lane src's value in every lane (src uniform); 32-bit values are exact in a double
inline int __builtin_amdgcn_readlane(int x, int src) { return (int)__shfl((double)x, src); }
// v_permlane32_swap: lanes 32..63 of the first operand trade places with lanes 0..31 of the
// second; element 0 of the result is the first operand afterwards, element 1 the second
struct emu_u2 {
    unsigned v[2];
    unsigned operator[](int i) const { return v[i]; }
};
inline emu_u2 __builtin_amdgcn_permlane32_swap(unsigned old, unsigned src, bool, bool) {
    const int l = (int)(threadIdx.x % 64);
    const unsigned from_src = (unsigned)__shfl((double)src, l ^ 32);  // exact: 32-bit values
    const unsigned from_old = (unsigned)__shfl((double)old, l ^ 32);
    return {{l >= 32 ? from_src : old, l < 32 ? from_old : src}};
}
// v_mov_b32 with a DPP source (the controls the product uses): quad_perm, row_shl/row_shr 1..15,
// row_bcast:15 (lane 15 of the previous row) and row_bcast:31 (lane 31, rows 2 and 3).  Lanes
// in rows or banks masked off keep `old`; lanes without a source lane get 0 with bound_ctrl, else `old`.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    const int l = (int)(threadIdx.x % 64), i = l & 15, r = l >> 4;
    int s = -1;
    if (ctrl >= 0 && ctrl <= 0xff) s = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
    else if (ctrl >= 0x101 && ctrl <= 0x10f) { if (i + (ctrl - 0x100) < 16) s = l + (ctrl - 0x100); }
    else if (ctrl >= 0x111 && ctrl <= 0x11f) { if (i - (ctrl - 0x110) >= 0) s = l - (ctrl - 0x110); }
    else if (ctrl == 0x142) { if (r >= 1) s = 16 * r - 1; }
    else if (ctrl == 0x143) { if (r >= 2) s = 31; }
    else { fprintf(stderr, "emulator: DPP control 0x%x not modelled\n", ctrl); abort(); }
    const int v = (int)(long long)__shfl((double)src, s < 0 ? l : s);  // every lane takes part
    if (!((row_mask >> r) & 1) || !((bank_mask >> (i >> 2)) & 1)) return old;
    if (s < 0) return bound_ctrl ? 0 : old;
    return v;
}
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
// a scheduling barrier on the GPU (lanes of a wave run in lockstep); here the lanes meet
#define __builtin_amdgcn_wave_barrier() emu_wave_sync()
// memory fences: the emulated lanes share the host's coherent memory; ordering between lanes
// comes from the wave and block barriers
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
// instruction-scheduling fence of the GPU compiler: nothing to do on the CPU
#define __builtin_amdgcn_sched_barrier(mask) ((void)(mask))
#ifndef __ATOMIC_RELEASE
#define __ATOMIC_RELEASE 3
#define __ATOMIC_ACQUIRE 2
#endif
// only applied to values that are uniform across the wave by construction
#define __builtin_amdgcn_readfirstlane(x) (x)

void emu_run_grid(dim3 g, dim3 b, size_t shm, const std::function<void()>& body);
template <class F, class... A>
void emu_launch(F* f, dim3 g, dim3 b, size_t shm, A... args) {
    if (emu_ptrcheck_on()) {
        int i = 0;
        (emu_ptrcheck(&args, sizeof(args), "kernel argument", (const void*)f, i++), ...);
    }
    emu_run_grid(g, b, shm, [=]() { f(args...); });
}
#define hipLaunchKernelGGL(k, g, b, shm, s, ...) emu_launch(k, dim3(g), dim3(b), shm, ##__VA_ARGS__)
