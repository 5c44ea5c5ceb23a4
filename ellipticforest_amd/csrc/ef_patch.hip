// ef_patch.hip -- C ABI of batched leaf-patch operations: FiniteVolumeSolver::solve,
// mapD2N, particularNeumannData and buildD2N (src/Patches/FiniteVolume/
// FiniteVolumeSolver.cpp:25-462) over many independent patches at once.
//
// Constant coefficients within every patch (FISHPACK semantics, or the five-point
// stencil with constant alpha, beta, lambda) run on patch_fd_kernel (fast
// diagonalisation on the f64 MFMA, one wave per patch and right-hand side).  Patches
// whose per-cell coefficients vary run on the block-tridiagonal kernels (factorised
// once at creation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ef_hps.h"
#include "ef_kernels.hpp"
#include "ef_host.hpp"

using namespace efk;

struct ef_patches_s {
    ef_ctx ctx = nullptr;
    int N = 0, P = 0, NB = 1;
    long long max_blocks = 4096;
    bool variable = false;
    bool dominant = false;         // variable, every patch diagonally dominant: no pivoting
    bool packed = false;           // variable and symmetric: packed inverse pivot blocks (N 8, 16, 32)
    bool large = false;            // N > 32: batched-GEMM fast diagonalisation
    int Np = 0;                    // padded side of the large path
    hipStream_t stream = nullptr, own_stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ev_recorded = false;
    std::vector<PatchCoef> coef_h;
    // fast-diagonalisation tables: one coefficient entry per distinct (cx, cy, lam, dx, dy)
    DevBuf<PatchCoef> coef;
    DevBuf<int> cls;               // patch -> entry
    DevBuf<double> QA, QB, mu, dinv;
    // block-tridiagonal path
    std::vector<LeafDesc> ld_h;
    DevBuf<LeafDesc> leafdesc;
    DevBuf<double> vcoef, vstencil, vSinv, vSinvP, vwork, vz;  // vz: forward rows of the packed solve
    // large path: padded basis (Np x Np), work buffers
    DevBuf<double> Qp, W1, W2;
    // staging of the host-buffer entry points
    DevBuf<double> gbuf, fbuf, obuf;

    ~ef_patches_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace efk {

// True when the five-point matrix of the patch (FiniteVolumeSolver.cpp:64-85, cell I = j + i N
// with i along x) is symmetric to 1e-14 relative: across every interior face the coupling
// alpha beta / h^2 is the same seen from both cells.  The physical-side terms change the
// diagonal only.
bool cells_symmetric(const double* c, int N) {
    const long long NN = (long long)N * N;
    const double *a = c, *bW = c + NN, *bE = c + 2 * NN, *bS = c + 3 * NN, *bN = c + 4 * NN;
    auto close = [](double x, double y) { return std::fabs(x - y) <= 1e-14 * std::max(std::fabs(x), std::fabs(y)); };
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const long long I = (long long)i * N + j;
            if (i > 0 && !close(a[I] * bW[I], a[I - N] * bE[I - N])) return false;
            if (j > 0 && !close(a[I] * bS[I], a[I - 1] * bN[I - 1])) return false;
        }
    return true;
}


// True when every row of the five-point matrix of the patch (the stencil of
// FiniteVolumeSolver.cpp:64-85, the ghost terms folded into the diagonal on the physical
// sides) is diagonally dominant, |diagonal| >= sum of |off-diagonals| (to 1e-12 relative):
// alpha, beta > 0 and lambda <= 0 give it.  Its block-LU then needs no pivoting.
bool cells_dominant(const double* c, int N, double dx, double dy) {
    const long long NN = (long long)N * N;
    const double *a = c, *bW = c + NN, *bE = c + 2 * NN, *bS = c + 3 * NN, *bN = c + 4 * NN, *lam = c + 5 * NN;
    const double dx2 = dx * dx, dy2 = dy * dy;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const long long I = (long long)i * N + j;
            const double cW = a[I] * bW[I] / dx2, cE = a[I] * bE[I] / dx2;
            const double cS = a[I] * bS[I] / dy2, cN = a[I] * bN[I] / dy2;
            double d = -a[I] * ((bE[I] + bW[I]) / dx2 + (bN[I] + bS[I]) / dy2) + lam[I];
            double off = 0.0;
            if (i == 0) d -= cW; else off += std::fabs(cW);
            if (i == N - 1) d -= cE; else off += std::fabs(cE);
            if (j == 0) d -= cS; else off += std::fabs(cS);
            if (j == N - 1) d -= cN; else off += std::fabs(cN);
            if (!(std::fabs(d) >= off * (1.0 - 1e-12))) return false;
        }
    return true;
}

// Workgroups of a patch launch: several rounds of every CU's resident waves (256 CUs x
// 4 SIMDs x 5 waves / 4 waves per workgroup = 1,280 at a time); each wave
// then loops over its items.  EF_PATCH_MAX_BLOCKS, read when a batch is created,
// overrides it (tuning sweeps, tests).
long long patch_default_blocks() {
    const char* e = std::getenv("EF_PATCH_MAX_BLOCKS");
    const long long x = e ? std::atoll(e) : 0;
    return x > 0 ? x : 4096LL;
}

// Basis of K = tridiag(1, -2, 1) with -3 ends (ghost 2g - u on both sides), padded to
// 16*NB, and the MFMA lane tables: QA[x][y][s] = Q[16x + r][16y + 4s + q],
// QB[x][y][s] = Q[16x + 4s + q][16y + r], lane = 16q + r.
void patch_fd_tables(int N, int NB, std::vector<double>& QA, std::vector<double>& QB, std::vector<double>& mu) {
    if (NB < 1 || NB > 2 || N < 1 || N > 16 * NB) throw EfError(EF_ERR_INTERNAL, "patch side outside the padded basis");
    const int Np = 16 * NB;
    std::vector<double> Q((size_t)Np * Np, 0.0);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < N; k++) {
            const double nrm = (k == N - 1) ? std::sqrt(1.0 / N) : std::sqrt(2.0 / N);
            Q[(size_t)i * Np + k] = nrm * std::sin(M_PI * (double)(k + 1) * ((double)i + 0.5) / (double)N);
        }
    mu.assign(Np, 0.0);
    for (int k = 0; k < N; k++) {
        const double s = std::sin(M_PI * (double)(k + 1) / (2.0 * N));
        mu[k] = -4.0 * s * s;
    }
    QA.assign((size_t)NB * NB * 4 * 64, 0.0);
    QB.assign((size_t)NB * NB * 4 * 64, 0.0);
    for (int x = 0; x < NB; x++)
        for (int y = 0; y < NB; y++)
            for (int s = 0; s < 4; s++)
                for (int lane = 0; lane < 64; lane++) {
                    const int q = lane / 16, r = lane % 16;
                    const size_t at = ((size_t)(x * NB + y) * 4 + s) * 64 + lane;
                    QA[at] = Q[(size_t)(16 * x + r) * Np + 16 * y + 4 * s + q];
                    QB[at] = Q[(size_t)(16 * x + 4 * s + q) * Np + 16 * y + r];
                }
}

void patch_fd_dinv(int N, int NB, const std::vector<double>& mu, const std::vector<PatchCoef>& coef,
                   std::vector<double>& out) {
    if (NB < 1 || NB > 2 || N < 1 || N > 16 * NB || mu.size() < (size_t)N)
        throw EfError(EF_ERR_INTERNAL, "patch side outside the padded basis");
    const int Np = 16 * NB;
    out.assign(coef.size() * (size_t)Np * Np, 0.0);
    for (size_t c = 0; c < coef.size(); c++)
        for (int k = 0; k < N; k++)
            for (int l = 0; l < N; l++)
                out[(c * Np + k) * Np + l] = 1.0 / (coef[c].cx * mu[k] + coef[c].cy * mu[l] + coef[c].lam);
}

void launch_patch_fd(const PatchFDArgs& a, int NB, long long max_blocks, hipStream_t s) {
    if (NB < 1 || NB > 2 || a.N < 1 || a.N > 16 * NB) throw EfError(EF_ERR_INTERNAL, "patch side outside the padded basis");
    const long long waves =
        (a.op == PATCH_OP_DTN) ? (long long)a.npatches * 4 * a.N : (long long)a.npatches * a.nrhs;
    if (waves == 0) return;
    const unsigned blocks = (unsigned)std::min<long long>((waves + 3) / 4, max_blocks);
    const bool full = (a.N == 16 * NB);
    const dim3 grid(blocks), block(256);
    const int op = a.op;
#define EF_PATCH_LAUNCH(NB_, FULL_, NV_)                                                                         \
    do {                                                                                                          \
        if (op == PATCH_OP_SOLVE)                                                                                 \
            hipLaunchKernelGGL((patch_fd_kernel<NB_, FULL_, PATCH_OP_SOLVE, NV_>), grid, block, 0, s, a, a.cls, a.coef, \
                               a.g_off, a.h_off);                                                                 \
        else if (op == PATCH_OP_MAP_D2N)                                                                          \
            hipLaunchKernelGGL((patch_fd_kernel<NB_, FULL_, PATCH_OP_MAP_D2N, NV_>), grid, block, 0, s, a, a.cls,   \
                               a.coef, a.g_off, a.h_off);                                                         \
        else                                                                                                      \
            hipLaunchKernelGGL((patch_fd_kernel<NB_, FULL_, PATCH_OP_DTN, NV_>), grid, block, 0, s, a, a.cls, a.coef,   \
                               a.g_off, a.h_off);                                                                 \
    } while (0)
    // N = 4, 8, 12 and 24 have their own padded variants with the side known at compile time
    if (NB == 1) {
        if (full) EF_PATCH_LAUNCH(1, true, 0);
        else if (a.N == 4) EF_PATCH_LAUNCH(1, false, 4);
        else if (a.N == 8) EF_PATCH_LAUNCH(1, false, 8);
        else if (a.N == 12) EF_PATCH_LAUNCH(1, false, 12);
        else EF_PATCH_LAUNCH(1, false, 0);
    } else {
        if (full) EF_PATCH_LAUNCH(2, true, 0);
        else if (a.N == 24) EF_PATCH_LAUNCH(2, false, 24);
        else EF_PATCH_LAUNCH(2, false, 0);
    }
#undef EF_PATCH_LAUNCH
    EF_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------
// Work space of ef_patches_apply_device_noalloc: a file-scope device array, one copy per GPU,
// held until the process ends.  In host code its name is only a host-side stand-in; the
// device address comes from hipGetSymbolAddress, taken once per device at context creation
// (efk::noalloc_setup) with that device current, and only that address is used.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 22;  // 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch;  // device -> address of its copy of `scratch`

double* scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    return it == g_scratch.end() ? nullptr : it->second;
}

// Doubles of work space an apply of this batch needs beyond its inputs and outputs.
size_t apply_scratch_doubles(const ef_patches_s* p, int op) {
    if (p->variable) return (p->packed && op == PATCH_OP_MAP_D2N) ? (size_t)p->P * p->N * p->N : 0;
    if (p->large) return 2 * (size_t)p->Np * p->Np;  // one item in each of the two GEMM buffers
    return 0;
}

// Device address of the no-allocation entry point's scratch array on the current device,
// taken once per device (the library's context creation calls this with the device current).
void noalloc_setup(int device) {
    if (scratch_address(device)) return;
    double* ptr = nullptr;
    EF_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw EfError(EF_ERR_DEVICE, "no device address for the no-allocation scratch array");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    g_scratch[device] = ptr;
}

}  // namespace efk

namespace {

void launch_fd(ef_patches_s* p, int op, const double* g, const double* f, double* out, int nrhs) {
    PatchFDArgs a{};
    a.N = p->N;
    a.npatches = p->P;
    a.nrhs = nrhs;
    a.op = op;
    a.coef = p->coef.p;
    a.cls = p->cls.p;
    a.QA = p->QA.p;
    a.QB = p->QB.p;
    a.mu = p->mu.p;
    a.dinv = p->dinv.p;
    a.f = f;
    a.g = g;
    a.out = out;
    a.g_off = nullptr;
    a.h_off = nullptr;
    a.vk = 1;
    efk::launch_patch_fd(a, p->NB, p->max_blocks, p->stream);
}

// Large patches: chunks of items through fold, four GEMMs, scaling and the output step.
void launch_gemm_strided(const GemmStridedArgs& g, hipStream_t s) {
    const long long blocks32 = (long long)((g.M + 31) / 32) * ((g.N + 31) / 32);  // one wave each
    const dim3 grid((unsigned)((blocks32 + 3) / 4), (unsigned)g.batch), block(256);
    // an operand is read as 32-byte vectors along k when it is contiguous along k and every
    // row / column it starts from is 32-byte aligned
    auto aligned = [](const double* x, long long s0, long long s1) {
        return ((uintptr_t)x % 32 == 0) && s0 % 4 == 0 && s1 % 4 == 0;
    };
    const bool ak = g.ca == 1 && aligned(g.A, g.sa, g.ra);
    const bool bk = g.rb == 1 && aligned(g.B, g.sb, g.cb);
    if (ak && bk) hipLaunchKernelGGL((gemm_strided_f64_kernel<true, true>), grid, block, 0, s, g);
    else if (ak) hipLaunchKernelGGL((gemm_strided_f64_kernel<true, false>), grid, block, 0, s, g);
    else if (bk) hipLaunchKernelGGL((gemm_strided_f64_kernel<false, true>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((gemm_strided_f64_kernel<false, false>), grid, block, 0, s, g);
    EF_LAUNCH_CHECK();
}

// W1, W2: work buffers of at least wcap doubles each (null: the batch's own, grown as needed).
// Items are independent, so the chunking does not change any result.
void launch_large(ef_patches_s* p, int op, const double* g, const double* f, double* out, int nrhs, int patch0 = 0,
                  double* W1 = nullptr, double* W2 = nullptr, long long wcap = 0) {
    const int N = p->N, Np = p->Np;
    const long long per = (long long)Np * Np;
    const long long total = (op == PATCH_OP_DTN) ? 4LL * N : (long long)p->P * nrhs;
    long long chunk = std::max<long long>(1, std::min<long long>({65535LL, total, (16LL << 20) / per}));
    if (W1) {
        chunk = std::min(chunk, wcap / per);
        if (chunk < 1) throw EfError(EF_ERR_ARG, "work buffers too small for one large patch");
    } else {
        p->W1.ensure((size_t)(chunk * per));
        p->W2.ensure((size_t)(chunk * per));
        W1 = p->W1.p;
        W2 = p->W2.p;
    }
    for (long long first = 0; first < total; first += chunk) {
        const int items = (int)std::min(chunk, total - first);
        PatchLargeArgs a{};
        a.N = N;
        a.npatches = p->P;
        a.op = op;
        a.items = items;
        a.patch0 = patch0;
        a.col0 = (int)first;
        a.coef = p->coef.p;
        a.cls = p->cls.p;
        a.mu = p->mu.p;
        a.f = f;
        a.g = g;
        a.work = W1;
        a.out = out;
        a.item0 = (op == PATCH_OP_DTN) ? 0 : first;
        hipLaunchKernelGGL(patch_large_fold_kernel, dim3(blocks_for(items * per)), dim3(256), 0, p->stream, a);
        EF_LAUNCH_CHECK();
        GemmStridedArgs m{};
        m.M = m.N = m.K = Np;
        m.batch = items;
        // W2 = Q^T W1
        m.A = p->Qp.p; m.sa = 0; m.ra = 1; m.ca = Np;
        m.B = W1; m.sb = per; m.rb = Np; m.cb = 1;
        m.C = W2; m.sc = per; m.rc = Np; m.cc = 1;
        launch_gemm_strided(m, p->stream);
        // W1 = W2 Q
        m.A = W2; m.sa = per; m.ra = Np; m.ca = 1;
        m.B = p->Qp.p; m.sb = 0; m.rb = Np; m.cb = 1;
        m.C = W1;
        launch_gemm_strided(m, p->stream);
        hipLaunchKernelGGL(patch_large_scale_kernel, dim3(blocks_for(items * per)), dim3(256), 0, p->stream, a);
        EF_LAUNCH_CHECK();
        // W2 = Q W1
        m.A = p->Qp.p; m.sa = 0; m.ra = Np; m.ca = 1;
        m.B = W1; m.sb = per; m.rb = Np; m.cb = 1;
        m.C = W2;
        launch_gemm_strided(m, p->stream);
        // W1 = W2 Q^T
        m.A = W2; m.sa = per; m.ra = Np; m.ca = 1;
        m.B = p->Qp.p; m.sb = 0; m.rb = 1; m.cb = Np;
        m.C = W1;
        launch_gemm_strided(m, p->stream);
        const long long nout = (op == PATCH_OP_SOLVE) ? (long long)items * N * N : (long long)items * 4 * N;
        hipLaunchKernelGGL(patch_large_out_kernel, dim3(blocks_for(nout)), dim3(256), 0, p->stream, a);
        EF_LAUNCH_CHECK();
    }
}

// True when every per-cell array of the patch is constant and the four face betas agree.
bool patch_constant(const double* c, long long NN) {
    const double a0 = c[0], b0 = c[NN], l0 = c[5 * NN];
    for (long long e = 0; e < NN; e++)
        if (c[e] != a0 || c[NN + e] != b0 || c[2 * NN + e] != b0 || c[3 * NN + e] != b0 || c[4 * NN + e] != b0 ||
            c[5 * NN + e] != l0)
            return false;
    return true;
}

void record_begin(ef_patches_s* p) {
    if (!p->ev[0]) {
        EF_HIP(hipEventCreate(&p->ev[0]));
        EF_HIP(hipEventCreate(&p->ev[1]));
    }
    EF_HIP(hipEventRecord(p->ev[0], p->stream));
}
void record_end(ef_patches_s* p) {
    EF_HIP(hipEventRecord(p->ev[1], p->stream));
    p->ev_recorded = true;
}

// Block-tridiagonal path, one right-hand side per launch ([patch][...] layouts with k = 1).
// vz: work space of the packed Neumann maps, P N^2 doubles (null: the batch's own, grown as needed).
void launch_bt(ef_patches_s* p, int op, const double* g, const double* f, double* out, int nrhs, double* vz = nullptr) {
    const long long NN = (long long)p->N * p->N, B4 = 4LL * p->N;
    if (p->packed && op == PATCH_OP_MAP_D2N && !vz) {
        p->vz.ensure((size_t)p->P * NN);
        vz = p->vz.p;
    }
    if (op == PATCH_OP_MAP_D2N && g == nullptr) {
        for (int c = 0; c < nrhs; c++) {
            launch_bt_solve(p->leafdesc.p, p->N, p->P, 1, 0, f ? f + c * p->P * NN : nullptr, f ? 0 : 1,
                            (const double*)nullptr, out + c * p->P * B4, (double*)nullptr, p->stream, p->packed,
                            vz);
            EF_LAUNCH_CHECK();
        }
        return;
    }
    if (g == nullptr) throw EfError(EF_ERR_ARG, "boundary data g is required");
    const int mode = (op == PATCH_OP_SOLVE) ? 1 : 2;
    for (int c = 0; c < nrhs; c++) {
        launch_bt_solve(p->leafdesc.p, p->N, p->P, 1, mode, f ? f + c * p->P * NN : nullptr, f ? 0 : 1,
                        g + c * p->P * B4, mode == 2 ? out + c * p->P * B4 : (double*)nullptr,
                        mode == 1 ? out + c * p->P * NN : (double*)nullptr, p->stream, p->packed, vz);
        EF_LAUNCH_CHECK();
    }
}

void apply(ef_patches_s* p, int op, const double* g, const double* f, double* out, int nrhs) {
    if (op != PATCH_OP_SOLVE && op != PATCH_OP_MAP_D2N) throw EfError(EF_ERR_ARG, "op must be EF_PATCH_SOLVE or EF_PATCH_MAP_D2N");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    if (!out) throw EfError(EF_ERR_ARG, "null output");
    if (p->variable) launch_bt(p, op, g, f, out, nrhs);
    else if (p->large) launch_large(p, op, g, f, out, nrhs);
    else launch_fd(p, op, g, f, out, nrhs);
}

}  // namespace

extern "C" {

int ef_patches_create(ef_ctx ctx, const ef_patch_batch* b, const ef_coeffs* c, ef_patches* out) {
    EF_API_BEGIN
    if (!ctx || !b || !c || !out) throw EfError(EF_ERR_ARG, "null argument");
    if (b->npatches < 1) throw EfError(EF_ERR_ARG, "npatches must be >= 1");
    if (b->N < 4 || b->N > 1024) throw EfError(EF_ERR_ARG, "patch side N must be in [4, 1024]");
    if (!b->bounds) throw EfError(EF_ERR_ARG, "null patch bounds");
    if (c->solver_type != EF_SOLVER_FISHPACK && c->solver_type != EF_SOLVER_FIVE_POINT)
        throw EfError(EF_ERR_ARG, "unknown solver type");
    EF_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ef_patches_s> p(new ef_patches_s());
    p->ctx = ctx;
    p->N = b->N;
    p->P = b->npatches;
    p->NB = (b->N <= 16) ? 1 : 2;
    p->large = b->N > 32;
    p->Np = (b->N + 15) / 16 * 16;
    p->max_blocks = efk::patch_default_blocks();
    EF_HIP(hipStreamCreate(&p->own_stream));
    p->stream = p->own_stream;
    const int N = p->N;
    const long long NN = (long long)N * N;
    p->coef_h.resize(p->P);
    for (int i = 0; i < p->P; i++) {
        const double* bd = b->bounds + 4LL * i;
        const double dx = (bd[1] - bd[0]) / N, dy = (bd[3] - bd[2]) / N;
        if (!(dx > 0.0) || !(dy > 0.0)) throw EfError(EF_ERR_ARG, "patch " + std::to_string(i) + " has empty bounds");
        PatchCoef& k = p->coef_h[i];
        k.dx = dx;
        k.dy = dy;
        if (c->solver_type == EF_SOLVER_FISHPACK) {
            k.cx = 1.0 / (dx * dx); k.cy = 1.0 / (dy * dy); k.lam = c->lambda0;
        } else if (c->cell == nullptr) {
            k.cx = c->alpha0 * c->beta0 / (dx * dx); k.cy = c->alpha0 * c->beta0 / (dy * dy); k.lam = c->lambda0;
        } else {
            const double* cc = c->cell + (long long)i * 6 * NN;
            if (!patch_constant(cc, NN)) p->variable = true;
            k.cx = cc[0] * cc[NN] / (dx * dx); k.cy = cc[0] * cc[NN] / (dy * dy); k.lam = cc[5 * NN];
        }
    }
    if (p->variable && p->N > BT_MAX_N)
        throw EfError(EF_ERR_ARG, "coefficients that vary within a patch need N <= " + std::to_string(BT_MAX_N));
    if (p->variable) {
        const size_t P = p->P;
        p->vcoef.alloc(P * 6 * NN);
        efk::copy_h2d(p->vcoef.p, c->cell, sizeof(double) * P * 6 * NN, p->stream);
        p->vstencil.alloc(P * 5 * NN);
        p->vSinv.alloc(P * N * NN);
        p->packed = (N == 8 || N == 16 || N == 32);
        for (size_t i = 0; i < P && p->packed; i++) p->packed = efk::cells_symmetric(c->cell + i * 6 * NN, N);
        const long long NP = bt_packed_entries(N);
        if (p->packed) p->vSinvP.alloc(P * N * NP);
        p->ld_h.assign(P, LeafDesc{});
        for (size_t i = 0; i < P; i++) {
            LeafDesc& L = p->ld_h[i];
            L.cls = -1;
            L.node = (int)i;
            L.h_off = (long long)i * 4 * N;
            for (int sd = 0; sd < 4; sd++) L.gs[sd] = (long long)i * 4 * N + (long long)sd * N;
            L.coef = p->vcoef.p + i * 6 * NN;
            L.stencil = p->vstencil.p + i * 5 * NN;
            L.Sinv = p->vSinv.p + i * N * NN;
            L.T = nullptr;
            L.SinvP = p->packed ? p->vSinvP.p + i * NP : nullptr;  // block-major: block b at b * P * NP
            L.pstride = (long long)P * NP;
            L.dx = p->coef_h[i].dx;
            L.dy = p->coef_h[i].dy;
        }
        p->leafdesc.upload(p->ld_h, p->stream);
        bool dominant = true;
        for (size_t i = 0; i < P && dominant; i++)
            dominant = efk::cells_dominant(c->cell + i * 6 * NN, N, p->coef_h[i].dx, p->coef_h[i].dy);
        p->dominant = dominant;
        launch_bt_factor(p->leafdesc.p, N, p->P, p->stream, dominant);
        EF_LAUNCH_CHECK();
        if (p->packed) {
            hipLaunchKernelGGL(bt_pack_kernel, dim3((unsigned)std::min<long long>(blocks_for(P * N * NP), 65535)),
                               dim3(256), 0, p->stream, p->leafdesc.p, p->P, N);
            EF_LAUNCH_CHECK();
        }
    } else {
        std::vector<double> QA, QB, mu;
        if (p->large) {
            // padded basis, row-major Np x Np, and the N eigenvalues
            const int Np = p->Np;
            std::vector<double> Q((size_t)Np * Np, 0.0);
            mu.assign(N, 0.0);
            for (int i = 0; i < N; i++)
                for (int k = 0; k < N; k++) {
                    const double nrm = (k == N - 1) ? std::sqrt(1.0 / N) : std::sqrt(2.0 / N);
                    Q[(size_t)i * Np + k] = nrm * std::sin(M_PI * (double)(k + 1) * ((double)i + 0.5) / (double)N);
                }
            for (int k = 0; k < N; k++) {
                const double sk = std::sin(M_PI * (double)(k + 1) / (2.0 * N));
                mu[k] = -4.0 * sk * sk;
            }
            p->Qp.upload(Q, p->stream);
        } else {
            efk::patch_fd_tables(N, p->NB, QA, QB, mu);
            p->QA.upload(QA, p->stream);
            p->QB.upload(QB, p->stream);
        }
        p->mu.upload(mu, p->stream);
        // distinct coefficient entries (a leaf set has a handful: one per level and shape)
        std::map<std::vector<uint64_t>, int> seen;
        auto bits = [](double v) { uint64_t u; std::memcpy(&u, &v, 8); return u; };
        std::vector<PatchCoef> entries;
        std::vector<int> cls(p->P);
        for (int i = 0; i < p->P; i++) {
            const PatchCoef& k = p->coef_h[i];
            std::vector<uint64_t> key = {bits(k.cx), bits(k.cy), bits(k.lam), bits(k.dx), bits(k.dy)};
            auto it = seen.find(key);
            if (it == seen.end()) {
                cls[i] = (int)entries.size();
                seen[key] = cls[i];
                entries.push_back(k);
            } else cls[i] = it->second;
        }
        p->coef.upload(entries, p->stream);
        p->cls.upload(cls, p->stream);
        // 1/D tables unless they would outweigh the loads themselves (then computed in the kernel)
        const size_t Np = 16 * (size_t)p->NB;
        if (!p->large && entries.size() * Np * Np <= (size_t)p->P * NN / 4) {
            std::vector<double> dinv;
            efk::patch_fd_dinv(N, p->NB, mu, entries, dinv);
            p->dinv.upload(dinv, p->stream);
        }
    }
    EF_HIP(hipStreamSynchronize(p->stream));
    *out = p.release();
    EF_API_END
}

int ef_patches_destroy(ef_patches p) {
    EF_API_BEGIN
    if (!p) return EF_OK;
    (void)hipSetDevice(p->ctx->device);
    delete p;
    EF_API_END
}

int ef_patches_set_stream(ef_patches p, void* stream) {
    EF_API_BEGIN
    if (!p) throw EfError(EF_ERR_ARG, "null handle");
    EF_HIP(hipStreamSynchronize(p->stream));
    p->stream = stream ? (hipStream_t)stream : p->own_stream;
    EF_API_END
}

int ef_patches_info(ef_patches p, int* info) {
    EF_API_BEGIN
    if (!p || !info) throw EfError(EF_ERR_ARG, "null argument");
    info[0] = p->P;
    info[1] = p->N;
    info[2] = p->variable ? (p->packed ? 2 : 1) : 0;
    EF_API_END
}

int ef_patches_apply_device(ef_patches p, int op, const double* g_dev, const double* f_dev, double* out_dev, int nrhs) {
    EF_API_BEGIN
    if (!p) throw EfError(EF_ERR_ARG, "null handle");
    EF_HIP(hipSetDevice(p->ctx->device));
    record_begin(p);
    apply(p, op, g_dev, f_dev, out_dev, nrhs);
    record_end(p);
    EF_API_END
}

int ef_patches_apply_device_noalloc(ef_patches p, int op, const double* g_dev, const double* f_dev, double* out_dev,
                                    int nrhs) {
    EF_API_BEGIN
    if (!p) throw EfError(EF_ERR_ARG, "null handle");
    if (op != PATCH_OP_SOLVE && op != PATCH_OP_MAP_D2N) throw EfError(EF_ERR_ARG, "op must be EF_PATCH_SOLVE or EF_PATCH_MAP_D2N");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    if (!out_dev) throw EfError(EF_ERR_ARG, "null output");
    EF_HIP(hipSetDevice(p->ctx->device));
    double* work = scratch_address(p->ctx->device);
    if (!work) throw EfError(EF_ERR_STATE, "no scratch address for this device (the context did not set it up)");
    const size_t need = apply_scratch_doubles(p, op);
    if (need > NOALLOC_SCRATCH_DOUBLES)
        throw EfError(EF_ERR_ARG, "this batch needs " + std::to_string(need) + " doubles of work space; the "
                                  "no-allocation entry point has " + std::to_string(NOALLOC_SCRATCH_DOUBLES));
    if (p->variable) launch_bt(p, op, g_dev, f_dev, out_dev, nrhs, work);
    else if (p->large) {
        const long long half = (long long)(NOALLOC_SCRATCH_DOUBLES / 2);
        launch_large(p, op, g_dev, f_dev, out_dev, nrhs, 0, work, work + half, half);
    } else launch_fd(p, op, g_dev, f_dev, out_dev, nrhs);
    EF_API_END
}

int ef_patches_apply(ef_patches p, int op, const double* g, const double* f, double* out, int nrhs) {
    EF_API_BEGIN
    if (!p || !out) throw EfError(EF_ERR_ARG, "null argument");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    EF_HIP(hipSetDevice(p->ctx->device));
    const size_t NN = (size_t)p->N * p->N, B4 = 4 * (size_t)p->N, P = p->P;
    const size_t gcnt = (size_t)nrhs * P * B4, fcnt = (size_t)nrhs * P * NN;
    const size_t ocnt = (op == PATCH_OP_SOLVE) ? fcnt : gcnt;
    if (g) {
        p->gbuf.ensure(gcnt);
        efk::copy_h2d(p->gbuf.p, g, gcnt * sizeof(double), p->stream);
    }
    if (f) {
        p->fbuf.ensure(fcnt);
        efk::copy_h2d(p->fbuf.p, f, fcnt * sizeof(double), p->stream);
    }
    p->obuf.ensure(ocnt);
    record_begin(p);
    apply(p, op, g ? p->gbuf.p : nullptr, f ? p->fbuf.p : nullptr, p->obuf.p, nrhs);
    record_end(p);
    efk::copy_d2h(out, p->obuf.p, ocnt * sizeof(double), p->stream);
    EF_HIP(hipStreamSynchronize(p->stream));
    EF_API_END
}

int ef_patches_dtn(ef_patches p, double* T) {
    EF_API_BEGIN
    if (!p || !T) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(p->ctx->device));
    const size_t per = 16 * (size_t)p->N * p->N;
    p->obuf.ensure(per * p->P);
    record_begin(p);
    if (p->variable) {
        // bt_dtn_kernel writes each leaf's T through its descriptor
        for (int i = 0; i < p->P; i++) p->ld_h[i].T = p->obuf.p + (size_t)i * per;
        p->leafdesc.upload(p->ld_h, p->stream);
        // work space of the forward sweep: N * N * 4N doubles per leaf, up to 2 GiB per launch
        // (2,048 leaves at N = 32, enough workgroups to fill the GPU)
        const long long zper = (long long)p->N * p->N * 4 * p->N;
        const long long chunk = std::max<long long>(1, std::min<long long>(p->P, (2048LL << 20) / (zper * 8)));
        p->vwork.ensure((size_t)(chunk * zper));
        launch_bt_dtn(p->leafdesc.p, p->N, p->P, p->vwork.p, chunk, p->stream);
        EF_LAUNCH_CHECK();
    } else if (p->large) {
        for (int i = 0; i < p->P; i++) launch_large(p, PATCH_OP_DTN, nullptr, nullptr, p->obuf.p, 1, i);
    } else {
        launch_fd(p, PATCH_OP_DTN, nullptr, nullptr, p->obuf.p, 1);
    }
    record_end(p);
    efk::copy_d2h(T, p->obuf.p, per * p->P * sizeof(double), p->stream);
    EF_HIP(hipStreamSynchronize(p->stream));
    EF_API_END
}

int ef_patches_sync(ef_patches p) {
    EF_API_BEGIN
    if (!p) throw EfError(EF_ERR_ARG, "null handle");
    EF_HIP(hipStreamSynchronize(p->stream));
    EF_API_END
}

int ef_patches_stream(ef_patches p, void** stream) {
    EF_API_BEGIN
    if (!p || !stream) throw EfError(EF_ERR_ARG, "null argument");
    *stream = (void*)p->stream;
    EF_API_END
}

int ef_patches_last_ms(ef_patches p, double* ms) {
    EF_API_BEGIN
    if (!p || !ms) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipStreamSynchronize(p->stream));
    *ms = -1.0;
    if (p->ev_recorded) {
        float t = 0.0f;
        EF_HIP(hipEventElapsedTime(&t, p->ev[0], p->ev[1]));
        *ms = t;
    }
    EF_API_END
}

}  // extern "C"

struct ef_timer_s {
    ef_ctx ctx = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool recorded[2] = {false, false};
    ~ef_timer_s() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

extern "C" {

int ef_timer_create(ef_ctx ctx, ef_timer* out) {
    EF_API_BEGIN
    if (!ctx || !out) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ef_timer_s> t(new ef_timer_s());
    t->ctx = ctx;
    EF_HIP(hipEventCreate(&t->ev[0]));
    EF_HIP(hipEventCreate(&t->ev[1]));
    *out = t.release();
    EF_API_END
}

int ef_timer_destroy(ef_timer t) {
    EF_API_BEGIN
    delete t;
    EF_API_END
}

int ef_timer_record(ef_timer t, void* stream, int which) {
    EF_API_BEGIN
    if (!t) throw EfError(EF_ERR_ARG, "null timer");
    if (which != 0 && which != 1) throw EfError(EF_ERR_ARG, "which must be 0 (start) or 1 (stop)");
    EF_HIP(hipSetDevice(t->ctx->device));
    EF_HIP(hipEventRecord(t->ev[which], (hipStream_t)stream));
    t->recorded[which] = true;
    EF_API_END
}

int ef_timer_elapsed_ms(ef_timer t, double* ms) {
    EF_API_BEGIN
    if (!t || !ms) throw EfError(EF_ERR_ARG, "null argument");
    if (!t->recorded[0] || !t->recorded[1]) throw EfError(EF_ERR_STATE, "timer events not recorded");
    EF_HIP(hipEventSynchronize(t->ev[1]));
    float v = 0.0f;
    EF_HIP(hipEventElapsedTime(&v, t->ev[0], t->ev[1]));
    *ms = v;
    EF_API_END
}

}  // extern "C"
