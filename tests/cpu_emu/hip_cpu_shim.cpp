// hip_cpu_shim.cpp -- runtime of the CPU emulation (see hip_cpu_shim.h).
#include "hip_cpu_shim.h"

#include <dlfcn.h>

#include <cstdint>
#include <map>
#include <mutex>

thread_local emu_idx threadIdx;
emu_idx blockIdx;
dim3 blockDim, gridDim;

static const size_t PAGE = 4096;
static std::map<void*, std::pair<void*, size_t>> g_maps;
static std::map<void*, size_t> g_sizes;  // user bytes of each allocation
static std::mutex g_mu;

// Each buffer sits inside its own reservation with 1 GiB of PROT_NONE on both sides, so
// a stray access anywhere near it faults.  EF_EMU_GUARD=front aligns the buffer to the
// front guard (catches underruns at once); the default aligns its end to the back guard.
static const size_t GUARD = 1ULL << 30;
static void* guarded_alloc(size_t bytes, void** base_out, size_t* len_out) {
    const size_t body = ((bytes + PAGE - 1) / PAGE) * PAGE;
    const size_t len = body + 2 * GUARD;
    char* base = (char*)mmap(nullptr, len, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (base == MAP_FAILED) return nullptr;
    char* rw = base + GUARD;
    if (mprotect(rw, body, PROT_READ | PROT_WRITE) != 0) return nullptr;
    const char* mode = getenv("EF_EMU_GUARD");
    const bool front = mode && std::strcmp(mode, "front") == 0;
    *base_out = base;
    *len_out = len;
    // Fill of new memory: all-ones by default (a NaN pattern, so reads of unwritten memory show
    // up in results; reads as -1 where an int is read); EF_EMU_FILL=zero or =rand instead, since a
    // GPU's fresh memory holds arbitrary values and an unwritten index read there is wild.
    const char* fill = getenv("EF_EMU_FILL");
    if (fill && std::strcmp(fill, "zero") == 0) {
        std::memset(rw, 0, body);
    } else if (fill && std::strcmp(fill, "rand") == 0) {
        static uint64_t x = 0x9e3779b97f4a7c15ULL;
        for (size_t i = 0; i < body; i += 8) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            std::memcpy(rw + i, &x, 8);
        }
    } else {
        std::memset(rw, 0xff, body);
    }
    if (front) return rw;
    return rw + body - ((bytes + 7) / 8) * 8;  // end of the buffer touches the back guard
}

hipError_t hipMalloc(void** p, size_t bytes) {
    void* base;
    size_t len;
    void* q = guarded_alloc(bytes ? bytes : 8, &base, &len);
    if (!q) return 2;
    std::lock_guard<std::mutex> lk(g_mu);
    g_maps[q] = {base, len};
    g_sizes[q] = bytes ? bytes : 8;
    *p = q;
    return 0;
}

hipError_t hipMemGetAddressRange(void** base, size_t* size, void* ptr) {
    std::lock_guard<std::mutex> lk(g_mu);
    const uintptr_t w = (uintptr_t)ptr;
    for (auto& m : g_sizes) {
        const uintptr_t b = (uintptr_t)m.first;
        if (w >= b && w < b + m.second) {
            *base = m.first;
            *size = m.second;
            return 0;
        }
    }
    return 1;
}
bool emu_ptrcheck_on() {
    static const bool on = getenv("EF_EMU_PTRCHECK") && std::strcmp(getenv("EF_EMU_PTRCHECK"), "0") != 0;
    return on;
}

static bool in_device_allocation(uintptr_t w) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& m : g_maps) {
        const uintptr_t b = (uintptr_t)m.second.first;
        if (w >= b && w < b + m.second.second) return true;
    }
    return false;
}

void emu_ptrcheck(const void* p, size_t n, const char* what, const void* kernel, int index) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t off = 0; off + 8 <= n; off += 8) {
        uintptr_t w;
        std::memcpy(&w, c + off, 8);
        if (w < 0x10000 || w > 0x7fffffffffffULL || (w & 7)) continue;
        if (in_device_allocation(w)) continue;
        void* page = (void*)(w & ~(uintptr_t)(PAGE - 1));
        if (msync(page, PAGE, MS_ASYNC) != 0) continue;  // not mapped: an ordinary integer
        Dl_info info{};
        const char* name = (kernel && dladdr(kernel, &info) && info.dli_sname) ? info.dli_sname : "-";
        fprintf(stderr, "EMU-PTRCHECK host address 0x%llx in %s %d at byte %zu of %zu (kernel %s)\n",
                (unsigned long long)w, what, index, off, n, name);
        if (std::strcmp(getenv("EF_EMU_PTRCHECK"), "abort") == 0) abort();
    }
}

hipError_t hipFree(void* p) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_maps.find(p);
    if (it == g_maps.end()) { fprintf(stderr, "hipFree of unknown pointer\n"); abort(); }
    munmap(it->second.first, it->second.second);
    g_maps.erase(it);
    g_sizes.erase(p);
    return 0;
}

// ----- workgroup emulation -------------------------------------------------------
static const int NT = 1024;  // largest workgroup; a launch uses blockDim.x of them
static int g_nt = 256;
static std::barrier<>* g_block_bar = nullptr;
static std::barrier<>* g_wave_bar[NT / 64];
// Operand slots of the wave-collective operations, double-buffered by call parity: a lane
// writes slot parity p, passes the wave barrier and reads slot p; the next collective uses
// slot 1 - p, so one barrier per call suffices (no lane can write slot p again before every
// lane has passed the next call's barrier, i.e. finished reading slot p).
static double g_mfma_a[2][NT], g_mfma_b[2][NT];
static thread_local int g_par = 0;
static void* g_dyn = nullptr;
static void* g_dyn_base = nullptr;
static size_t g_dyn_len = 0;

void __syncthreads() { g_block_bar->arrive_and_wait(); }
void emu_wave_sync() { g_wave_bar[threadIdx.x / 64]->arrive_and_wait(); }
void* emu_dyn_smem() { return g_dyn; }

emu_d4 emu_mfma_f64(double a, double b, emu_d4 c, int, int, int) {
    const int t = threadIdx.x, w = t / 64 * 64, l = t % 64;
    const int p = g_par;
    g_par ^= 1;
    g_mfma_a[p][t] = a;
    g_mfma_b[p][t] = b;
    emu_wave_sync();
    emu_d4 out = c;
    for (int v = 0; v < 4; v++) {
        const int row = 4 * v + l / 16, col = l % 16;
        double s = 0.0;
        for (int k = 0; k < 4; k++) s += g_mfma_a[p][w + row + 16 * k] * g_mfma_b[p][w + col + 16 * k];
        out[v] += s;
    }
    return out;
}

emu_f4 emu_mfma_f32(float a, float b, emu_f4 c, int, int, int) {
    const int t = threadIdx.x, w = t / 64 * 64, l = t % 64;
    const int p = g_par;
    g_par ^= 1;
    g_mfma_a[p][t] = a;
    g_mfma_b[p][t] = b;
    emu_wave_sync();
    emu_f4 out = c;
    for (int v = 0; v < 4; v++) {
        const int row = 4 * (l / 16) + v, col = l % 16;  // the f32 accumulator layout
        float s = 0.0f;
        for (int k = 0; k < 4; k++) s += (float)g_mfma_a[p][w + row + 16 * k] * (float)g_mfma_b[p][w + col + 16 * k];
        out[v] += s;
    }
    return out;
}

unsigned long long __ballot(int pred) {
    const int t = threadIdx.x, w = t / 64 * 64;
    const int p = g_par;
    g_par ^= 1;
    g_mfma_a[p][t] = pred != 0 ? 1.0 : 0.0;
    emu_wave_sync();
    unsigned long long m = 0;
    for (int l = 0; l < 64; l++)
        if (g_mfma_a[p][w + l] != 0.0) m |= 1ULL << l;
    return m;
}

double __shfl(double v, int src, int) {
    const int t = threadIdx.x, w = t / 64 * 64;
    const int p = g_par;
    g_par ^= 1;
    g_mfma_a[p][t] = v;
    emu_wave_sync();
    return g_mfma_a[p][w + (src & 63)];
}

// Persistent pools of workers, one pool per block size (one worker per lane of a block).  Per
// block the main thread publishes the body, releases the pool's workers through its start
// barrier and waits on its end barrier.  The block and wave barriers are rebuilt per block
// because a finishing thread drops out of them, as an exited wave stops counting at s_barrier.
static const std::function<void()>* g_body = nullptr;
struct EmuPool {
    std::barrier<>* start;
    std::barrier<>* end;
};
static std::map<int, EmuPool> g_pools;

static EmuPool& pool_for(int nt) {
    auto it = g_pools.find(nt);
    if (it != g_pools.end()) return it->second;
    EmuPool P{new std::barrier<>(nt + 1), new std::barrier<>(nt + 1)};
    g_pools[nt] = P;
    for (int t = 0; t < nt; t++)
        std::thread([t, P]() {
            threadIdx = {(unsigned)t, 0, 0};
            for (;;) {
                P.start->arrive_and_wait();
                g_par = 0;  // every block starts its wave collectives on slot 0
                (*g_body)();
                g_wave_bar[t / 64]->arrive_and_drop();
                g_block_bar->arrive_and_drop();
                P.end->arrive_and_wait();
            }
        }).detach();
    return g_pools[nt];
}

struct EmuGraph {
    std::vector<std::function<void()>> nodes;
};
static thread_local EmuGraph* g_capture = nullptr;
bool emu_capturing() { return g_capture != nullptr; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (g_capture) return hipErrorStreamCaptureUnsupported;
    g_capture = new EmuGraph();
    return 0;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    *g = g_capture;
    g_capture = nullptr;
    return *g ? 0 : hipErrorStreamCaptureUnsupported;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
    *e = new EmuGraph(*g);
    return 0;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    if (g_capture) return hipErrorStreamCaptureUnsupported;
    for (auto& n : e->nodes) n();
    return 0;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return 0; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return 0; }

void emu_run_grid(dim3 g, dim3 b, size_t shm, const std::function<void()>& body) {
    if (g_capture) {  // recorded with its arguments, run at hipGraphLaunch
        std::function<void()> f = body;
        g_capture->nodes.push_back([=]() { emu_run_grid(g, b, shm, f); });
        return;
    }
    if (b.x % 64 != 0 || b.x > (unsigned)NT || b.y != 1 || b.z != 1) {
        fprintf(stderr, "emulator supports 1-D blocks of 64..1024 threads (multiples of 64) only\n");
        abort();
    }
    g_nt = (int)b.x;
    EmuPool& pool = pool_for(g_nt);
    gridDim = g;
    blockDim = b;
    if (shm) g_dyn = guarded_alloc(shm, &g_dyn_base, &g_dyn_len);
    g_body = &body;
    // Block order: in order by default; EF_EMU_ORDER=reverse or =shuffle runs the blocks of a
    // grid in another order, so results that depend on the order (a race between blocks of one
    // launch, which a GPU runs concurrently) change.
    std::vector<emu_idx> order;
    for (unsigned z = 0; z < g.z; z++)
        for (unsigned y = 0; y < g.y; y++)
            for (unsigned x = 0; x < g.x; x++) order.push_back({x, y, z});
    const char* ord = getenv("EF_EMU_ORDER");
    if (ord && std::strcmp(ord, "reverse") == 0) std::reverse(order.begin(), order.end());
    if (ord && std::strcmp(ord, "shuffle") == 0) {
        static uint64_t x = 0x2545f4914f6cdd1dULL;
        for (size_t i = order.size(); i > 1; i--) {
            x ^= x << 13; x ^= x >> 7; x ^= x << 17;
            std::swap(order[i - 1], order[x % i]);
        }
    }
    for (const emu_idx& bi : order) {
        blockIdx = bi;
        std::barrier<> bb(g_nt);
        g_block_bar = &bb;
        for (int i = 0; i < g_nt / 64; i++) g_wave_bar[i] = new std::barrier<>(64);
        if (shm) {
            // fresh dynamic shared memory per block, filled as EF_EMU_FILL says (see guarded_alloc)
            const char* fill = getenv("EF_EMU_FILL");
            if (fill && std::strcmp(fill, "zero") == 0) {
                std::memset(g_dyn, 0, shm);
            } else if (fill && std::strcmp(fill, "rand") == 0) {
                static uint64_t x = 0x853c49e6748fea9bULL;
                for (size_t i = 0; i + 8 <= shm; i += 8) {
                    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
                    std::memcpy((char*)g_dyn + i, &x, 8);
                }
            } else {
                std::memset(g_dyn, 0xff, shm);
            }
        }
        pool.start->arrive_and_wait();
        pool.end->arrive_and_wait();
        for (int i = 0; i < g_nt / 64; i++) delete g_wave_bar[i];
    }
    if (shm) { munmap(g_dyn_base, g_dyn_len); g_dyn = nullptr; }
}
