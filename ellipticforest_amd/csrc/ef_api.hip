// ef_api.hip -- host side of libefhps.so: tree planner, device arenas, stage drivers
// and the C ABI declared in include/ef_hps.h.
//
// The reference visits families one at a time through p4est callbacks
// (Quadtree::merge/split, src/Quadtree.hpp:322-360) and keeps every operator in a
// per-node std::vector.  Here the tree is planned once into level-ordered groups and
// device arenas; each stage is a fixed sequence of batched launches per tree level, and the
// build reads the device back once (the zero-pivot flags of every merge) at its end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "ef_hps.h"
#include "ef_host.hpp"
#include "ef_kernels.hpp"
#include "ef_tree.hpp"

using namespace efk;

namespace {

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Tree plan
// ---------------------------------------------------------------------------
struct NodeP {
    int level = 0, ix = 0, iy = 0, parent = -1;
    int child[4] = {-1, -1, -1, -1};
    bool leaf = false;
    int leaf_index = -1;
    double xl = 0, xu = 0, yl = 0, yu = 0;
    int s = 0;    // patch size after its own merge (grid nx): N for leaves
    int tag = 0;  // n_coarsens set by the parent's merge
    int n = 0;    // internal: child interface size
    int group = -1, slot = -1;  // internal: merge group and position in it
    // operator arena offsets (elements)
    long long T_off = -1;               // internal: own DtN (8n x 8n)
    std::vector<long long> Tc_off;      // coarsened DtN after each coarsening step
    long long S_off = -1, Xi_off = -1;
    // vector arena row offsets
    long long h_off = -1, g_off = -1, gc_off = -1, gt_off = -1, hd_off = -1, w_off = -1, ui_off = -1;
    long long gsrc[4] = {-1, -1, -1, -1};  // where side s of g (s rows) is read in the downward sweep
};

struct Group {  // merges at one depth with one interface size, contiguous in the arenas
    int depth = 0, n = 0;
    std::vector<int> nodes;
    long long Xi0 = 0;  // Xi arena offset of the first merge (stride 16 n^2)
    int info0 = 0;      // first zero-pivot flag of the group
};

struct Plan {
    int N = 0;
    double xl = 0, xu = 0, yl = 0, yu = 0;
    std::vector<NodeP> nodes;
    std::vector<int> leaves;                   // node ids, z-order
    std::vector<std::vector<int>> internal_at; // internal node ids per depth
    std::vector<Group> groups;                 // build order: deepest depth first
    int maxdepth = 0, ninternal = 0;
    long long T_elems = 0, S_elems = 0, Xi_elems = 0;
    long long vec_rows = 0;
    long long work_X = 0;                      // max over groups of B m^2
    long long gj_d = 0, gj_i = 0;              // GJ work (doubles, ints) of the largest group
    int max_n = 0;
    bool side_reads = false;                   // sides of uncoarsened nodes read in place (gsrc)
};

int build_node(Plan& P, int parent, int level, int ix, int iy, double xl, double xu, double yl, double yu,
               const ef_mesh_desc& d, int& cursor) {
    if (level > 30) throw EfError(EF_ERR_ARG, "leaf list is not a pre-order quadtree leaf list (depth > 30)");
    NodeP nd;
    nd.level = level; nd.ix = ix; nd.iy = iy; nd.parent = parent;
    nd.xl = xl; nd.xu = xu; nd.yl = yl; nd.yu = yu;
    P.nodes.push_back(nd);
    const int id = (int)P.nodes.size() - 1;
    if (cursor >= d.nleaves) throw EfError(EF_ERR_ARG, "leaf list ended early");
    if (d.level[cursor] == level) {
        if (d.ix[cursor] != ix || d.iy[cursor] != iy)
            throw EfError(EF_ERR_ARG, "leaf list is not a pre-order quadtree leaf list");
        P.nodes[id].leaf = true;
        P.nodes[id].leaf_index = (int)P.leaves.size();
        P.leaves.push_back(id);
        cursor++;
        return id;
    }
    if (d.level[cursor] < level) throw EfError(EF_ERR_ARG, "leaf list is not a pre-order quadtree leaf list");
    // child quadrants exactly as FiniteVolumeNodeFactory::createChildNode (:15-67)
    const double xm = (xl + xu) / 2.0, ym = (yl + yu) / 2.0;
    const double cb[4][4] = {{xl, xm, yl, ym}, {xm, xu, yl, ym}, {xl, xm, ym, yu}, {xm, xu, ym, yu}};
    for (int c = 0; c < 4; c++) {
        const int cid = build_node(P, id, level + 1, 2 * ix + (c & 1), 2 * iy + (c >> 1), cb[c][0], cb[c][1], cb[c][2],
                                   cb[c][3], d, cursor);
        P.nodes[id].child[c] = cid;
    }
    return id;
}

int ilog2(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) l++;
    return l;
}

// Sizes and coarsening tags (tagPatchesForCoarsening_, HPSAlgorithm.hpp:676-696;
// mergePatch_ :1004-1009 makes the parent 2n cells wide), merge groups, arena offsets.
// Where the downward sweep reads each node's boundary data g, side by side.  The root's g and
// the g of a coarsened node (interpolated by the uncoarsen steps) are written to the node's own
// rows.  Any other node's side s is a segment of its parent's side or interface potential
// (split1to4 copies exactly that segment, HPSAlgorithm.hpp:562-577), so it is read there and
// never copied -- when every block of the sweeps' gathered GEMM operands is a multiple of 4
// rows at even offsets (the MFMA operand reads); otherwise every node's g is written out.
void plan_sides(Plan& P) {
    bool ok = true;
    for (const NodeP& nd : P.nodes)
        if (!nd.leaf && (nd.s % 4 != 0 || nd.n % 2 != 0)) ok = false;
    P.side_reads = ok;
    auto own = [](NodeP& nd) {
        for (int sd = 0; sd < 4; sd++) nd.gsrc[sd] = nd.g_off + (long long)sd * nd.s;
    };
    own(P.nodes[0]);
    for (int dd = 0; dd <= P.maxdepth; dd++)
        for (int id : P.internal_at[dd]) {
            const NodeP& nd = P.nodes[id];
            for (int c = 0; c < 4; c++) {
                NodeP& ch = P.nodes[nd.child[c]];
                if (!ok || ch.tag > 0) {
                    own(ch);
                    continue;
                }
                for (int sd = 0; sd < 4; sd++) {
                    const int q = child_src_q(c, sd);
                    ch.gsrc[sd] = q >= 0 ? nd.gsrc[q >> 1] + (long long)(q & 1) * nd.n
                                         : nd.ui_off + (long long)child_src_r(c, sd) * nd.n;
                }
            }
        }
}

void plan_tree(Plan& P, const ef_mesh_desc& d) {
    if (d.N < 4 || (d.N & 1)) throw EfError(EF_ERR_ARG, "N must be even and >= 4");
    if (d.N > 32)
        throw EfError(EF_ERR_ARG, "the mesh path takes leaf patches of up to 32 x 32 cells (single grids of any size "
                                  "run on ef_patches_*)");
    if (d.nleaves < 1 || !d.level || !d.ix || !d.iy) throw EfError(EF_ERR_ARG, "empty leaf list");
    if (!(d.xu > d.xl) || !(d.yu > d.yl)) throw EfError(EF_ERR_ARG, "empty domain");
    P.N = d.N; P.xl = d.xl; P.xu = d.xu; P.yl = d.yl; P.yu = d.yu;
    int cursor = 0;
    build_node(P, -1, 0, 0, 0, d.xl, d.xu, d.yl, d.yu, d, cursor);
    if (cursor != d.nleaves) throw EfError(EF_ERR_ARG, "leaf list has entries beyond a complete quadtree");
    P.maxdepth = 0;
    for (int id = (int)P.nodes.size() - 1; id >= 0; id--) {  // reverse pre-order: children first
        NodeP& nd = P.nodes[id];
        P.maxdepth = std::max(P.maxdepth, nd.level);
        if (nd.leaf) { nd.s = P.N; continue; }
        int smin = 1 << 30;
        for (int c = 0; c < 4; c++) smin = std::min(smin, P.nodes[nd.child[c]].s);
        nd.n = smin;
        nd.s = 2 * smin;
        for (int c = 0; c < 4; c++) {
            NodeP& ch = P.nodes[nd.child[c]];
            ch.tag = ilog2(ch.s) - ilog2(smin);
        }
        P.max_n = std::max(P.max_n, nd.n);
    }
    if (4LL * P.max_n > GJ_MAXM) throw EfError(EF_ERR_ARG, "interface systems larger than the inversion limit");
    P.internal_at.assign(P.maxdepth + 1, {});
    for (int id = 0; id < (int)P.nodes.size(); id++)
        if (!P.nodes[id].leaf) P.internal_at[P.nodes[id].level].push_back(id);
    // groups, deepest first; operator arenas in group order
    long long T = 0, S = 0, Xi = 0;
    int info = 0;
    for (int dd = P.maxdepth; dd >= 0; dd--) {
        std::map<int, std::vector<int>> byn;
        for (int id : P.internal_at[dd]) byn[P.nodes[id].n].push_back(id);
        for (auto& kv : byn) {
            Group g;
            g.depth = dd; g.n = kv.first; g.nodes = kv.second;
            g.Xi0 = Xi;
            g.info0 = info;
            const long long n = g.n, B = (long long)g.nodes.size();
            for (size_t i = 0; i < g.nodes.size(); i++) {
                NodeP& nd = P.nodes[g.nodes[i]];
                nd.group = (int)P.groups.size();
                nd.slot = (int)i;
                nd.T_off = T; T += 64 * n * n;
                nd.S_off = S; S += 32 * n * n;
                nd.Xi_off = Xi; Xi += 16 * n * n;
            }
            info += (int)B;
            P.work_X = std::max(P.work_X, B * 16 * n * n);
            P.gj_d = std::max(P.gj_d, (long long)gj_work_doubles((int)(4 * n), (int)B));
            P.gj_i = std::max(P.gj_i, (long long)gj_work_ints((int)(4 * n), (int)B));
            P.groups.push_back(g);
        }
    }
    P.ninternal = info;
    for (auto& nd : P.nodes) {  // coarsened copies of tagged nodes
        nd.Tc_off.clear();
        int m = nd.s;
        for (int t = 0; t < nd.tag; t++) {
            m /= 2;
            nd.Tc_off.push_back(T);
            T += 16LL * m * m;
        }
    }
    // vector arena rows
    long long rows = 0;
    for (auto& nd : P.nodes) {
        nd.h_off = rows; rows += 4LL * nd.s;
        nd.g_off = rows; rows += 4LL * nd.s;
        if (nd.tag > 0) { nd.gc_off = rows; rows += 4LL * (nd.s >> nd.tag); }
        else nd.gc_off = nd.g_off;
        if (nd.tag >= 2) { nd.gt_off = rows; rows += 4LL * nd.s; }
        if (!nd.leaf) {
            nd.hd_off = rows; rows += 4LL * nd.n;
            nd.w_off = rows; rows += 4LL * nd.n;
            nd.ui_off = rows; rows += 4LL * nd.n;
        }
    }
    P.T_elems = T; P.S_elems = S; P.Xi_elems = Xi; P.vec_rows = rows;
    plan_sides(P);
}

// ---------------------------------------------------------------------------
// Context and mesh
// ---------------------------------------------------------------------------
struct GroupPlan {  // device descriptors of one merge group
    std::vector<const CoarsenDesc*> csteps;
    std::vector<int> csteps_cnt, csteps_maxm;
    const AsmDesc* asmd = nullptr;
    GemmLaunch gS, gT;
};

struct LevelSweep {  // device descriptors of one depth of the sweeps, k right-hand sides
    int cnt = 0, nmax = 0;
    int tile[3] = {0, 0, 0};  // row tile x 10 + column tile (16-element units) of the w, h and u products
    int up_cnt = 0, up_nmax = 0;  // merges whose hDiff and boundary h are gathered (a coarsened child)
    const UpDesc* up = nullptr;
    GemmLaunch gw, gh;
    const SideCopy* copies = nullptr;  // children's g written out (without side reads)
    int copy_cnt = 0;
    const SideCopy* interps = nullptr;  // coarsened children's first uncoarsen step, from the sources
    int interp_cnt = 0;
    GemmLaunch gu;
    std::vector<const VecStepDesc*> unc;
    std::vector<int> unc_cnt, unc_nf;
};

struct SweepPlan {
    std::vector<LevelSweep> lv;
    DevStore store;
};

}  // namespace

constexpr int EF_MAX_SLOTS = 8;

// Serving workload loads (one request = 64 doubles): f = sum of up to 8 Gaussians
// amp exp(-|p - c|^2 / (2 w^2)) at leaf cell centres (x_i = xl + dx/2 + i dx, cell j + i N),
// r = sum of 8 modes a cos(kx x + ky y + phi) at the root boundary points.  Each Gaussian
// separates, exp(-dx^2 / (2 w^2)) exp(-dy^2 / (2 w^2)) (equal to the exponential of the sum up to
// rounding), and leaves of one column of the mesh share their x centres (of one row, their y
// centres): serve_factor_kernel evaluates the one-dimensional factors once per distinct centre
// and request (fac[rhs][q][v], v over the nxv distinct x centres, then the nyv y centres), and
// serve_cells_kernel forms each cell from 8 products of them.
__global__ void __launch_bounds__(256) serve_factor_kernel(const double* __restrict__ xy, int nxv, int nyv,
                                                            const double* __restrict__ params,
                                                            double* __restrict__ fac) {
    const int rhs = blockIdx.y, nv = nxv + nyv;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= 8LL * nv) return;
    const int q = (int)(e / nv), v = (int)(e % nv), axis = v >= nxv ? 1 : 0;
    const double* P = params + 64LL * rhs;
    const double amp = P[4 * q], c = P[4 * q + 1 + axis], w = P[4 * q + 3], d = xy[v] - c;
    fac[(long long)rhs * 8 * nv + e] = (amp != 0.0) ? exp(-(d * d) / (2.0 * w * w)) : 0.0;
}

// f[rhs][leaf][cell]: blockIdx.y = rhs, workgroups stride over the leaves; a leaf's 16 N factors
// (its x and y centre columns, offsets col[2 leaf], col[2 leaf + 1]) are staged in LDS and each
// cell sums 8 products of them.  The next leaf's factors are loaded while the current leaf's
// cells are formed (two LDS buffers, one barrier per leaf).
__global__ void __launch_bounds__(256) serve_cells_kernel(const int* __restrict__ col, int nl, int N, int nxv,
                                                           int nyv, const double* __restrict__ params,
                                                           const double* __restrict__ fac, double* __restrict__ f) {
    __shared__ double sf[2][2][8][32];  // [buffer][axis][Gaussian][cell index along the axis], N <= 32
    const int rhs = blockIdx.y, NN = N * N, nv = nxv + nyv, tid = threadIdx.x;
    const double* P = params + 64LL * rhs;
    const double* F = fac + (long long)rhs * 8 * nv;
    double pre[2];  // 16 N <= 512 factors: at most two per thread
    auto fetch = [&](int lf) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int t = tid + 256 * u;
            if (t < 16 * N) {
                const int axis = t / (8 * N), rem = t % (8 * N), q = rem / N, i = rem % N;
                const int v = axis ? nxv + col[2 * lf + 1] + i : col[2 * lf] + i;
                pre[u] = F[(long long)q * nv + v];
            }
        }
    };
    auto stash = [&](int b) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int t = tid + 256 * u;
            if (t < 16 * N) {
                const int axis = t / (8 * N), rem = t % (8 * N), q = rem / N, i = rem % N;
                sf[b][axis][q][i] = pre[u];
            }
        }
    };
    int leaf = blockIdx.x, buf = 0;
    if (leaf >= nl) return;  // uniform per workgroup
    fetch(leaf);
    stash(0);
    __syncthreads();
    for (; leaf < nl; leaf += gridDim.x) {
        const int next = leaf + (int)gridDim.x;
        if (next < nl) fetch(next);  // in flight while this leaf's cells are formed
        double* fo = f + ((long long)rhs * nl + leaf) * NN;
        for (int cell = tid; cell < NN; cell += 256) {
            const int i = cell / N, j = cell % N;
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) acc += P[4 * q] * (sf[buf][0][q][i] * sf[buf][1][q][j]);
            fo[cell] = acc;
        }
        if (next < nl) stash(buf ^ 1);  // the other buffer was last read before the previous barrier
        __syncthreads();
        buf ^= 1;
    }
}

__global__ void __launch_bounds__(256) serve_bdata_kernel(const double* __restrict__ params, int nrhs,
                                                           const double* __restrict__ pts, int m,
                                                           double* __restrict__ r) {
    const long long total = (long long)nrhs * m;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long long)gridDim.x * 256) {
        const long long rhs = t / m;
        const int p = (int)(t % m);
        const double x = pts[p], y = pts[m + p];
        const double* P = params + 64 * rhs + 32;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) acc += P[4 * q] * cos(P[4 * q + 1] * x + P[4 * q + 2] * y + P[4 * q + 3]);
        r[t] = acc;
    }
}

// FP64 residual of the uniform-mesh five-point system, res = f - A u, for the refinement of a
// mixed-precision solve.  On a uniform mesh the HPS interface condition (equal and opposite
// normal derivatives, h_L + h_R = 0) gives g = (u_L + u_R) / 2 at every interface, so the
// leaf stencils with ghost 2g - u (FiniteVolumeSolver.cpp:100-184) are the global 5-point
// operator: a cell's neighbour across a leaf side is the neighbouring leaf's edge cell, and on
// the physical boundary the ghost is 2 g_D - u with g_D = r / a at the root boundary point.
// Stencil: per cell (variable coefficients, cW cE cS cN cC) or per leaf class.
struct ResidualArgs {
    const double* u;
    const double* f;  // null: zero load
    double* res;
    int k, nl, N, M;
    const int* nbr;  // nl x 4 neighbour leaf (W, E, S, N) or -1
    const int* pos;  // nl x 2 (ix, iy) on the uniform level
    const double* r;
    const double* a;
    const double* stencil;  // variable: 5 N^2 per leaf, or null
    const PatchCoef* coef;  // constant: per class
    const int* cls;
};
__global__ void __launch_bounds__(256) mixed_residual_kernel(ResidualArgs p) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long NN = (long long)p.N * p.N, per = (long long)p.nl * NN;
    if (e >= (long long)p.k * per) return;
    const long long rhs = e / per, rem = e % per;
    const int leaf = (int)(rem / NN), I = (int)(rem % NN), N = p.N, i = I / N, j = I % N;
    double cW, cE, cS, cN, cC;
    if (p.stencil) {
        const double* st = p.stencil + (long long)leaf * 5 * NN;
        cW = st[I]; cE = st[NN + I]; cS = st[2 * NN + I]; cN = st[3 * NN + I]; cC = st[4 * NN + I];
    } else {
        const PatchCoef c = p.coef[p.cls[leaf]];
        cW = cE = c.cx; cS = cN = c.cy; cC = -2.0 * c.cx - 2.0 * c.cy + c.lam;
    }
    const double* U = p.u + rhs * per;
    const double uc = U[(long long)leaf * NN + I];
    const int ix = p.pos[2 * leaf], iy = p.pos[2 * leaf + 1], M = p.M;
    const double* R = p.r + rhs * 4LL * M;
    auto ghost = [&](int side, int t) { return 2.0 * (R[side * M + t] / p.a[side * M + t]) - uc; };
    double vW, vE, vS, vN;
    if (i > 0) vW = U[(long long)leaf * NN + I - N];
    else { const int nb = p.nbr[4 * leaf]; vW = nb >= 0 ? U[(long long)nb * NN + (N - 1) * N + j] : ghost(0, iy * N + j); }
    if (i < N - 1) vE = U[(long long)leaf * NN + I + N];
    else { const int nb = p.nbr[4 * leaf + 1]; vE = nb >= 0 ? U[(long long)nb * NN + j] : ghost(1, iy * N + j); }
    if (j > 0) vS = U[(long long)leaf * NN + I - 1];
    else { const int nb = p.nbr[4 * leaf + 2]; vS = nb >= 0 ? U[(long long)nb * NN + i * N + N - 1] : ghost(2, ix * N + i); }
    if (j < N - 1) vN = U[(long long)leaf * NN + I + 1];
    else { const int nb = p.nbr[4 * leaf + 3]; vN = nb >= 0 ? U[(long long)nb * NN + i * N] : ghost(3, ix * N + i); }
    const double Au = cC * uc + cW * vW + cE * vE + cS * vS + cN * vN;
    p.res[e] = (p.f ? p.f[e] : 0.0) - Au;
}

__global__ void __launch_bounds__(256) axpy_kernel(double* y, const double* x, long long n) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) y[e] += x[e];
}

struct ef_mesh_s {
    ef_ctx ctx = nullptr;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    Plan P;
    bool cache_operators = false, homogeneous_rhs = false;
    bool built = false;
    int solver_type = EF_SOLVER_FISHPACK;
    int precision = EF_PRECISION_FP64;
    // leaf classes
    std::vector<PatchCoef> classes_h;  // one per distinct (stencil, dx, dy)
    std::vector<int> leaf_class;       // per leaf
    int tleaf_class0 = -1;
    DevBuf<PatchCoef> classes;
    DevBuf<int> leaf_cls;                  // per leaf, on the device
    DevBuf<long long> leaf_goff, leaf_hoff;  // vector-arena rows of each leaf's g and h
    DevBuf<LeafDesc> leafdesc;             // block-tridiagonal path
    DevBuf<double> QA, QB, mu, dinv, leafT;  // fast-diagonalisation tables, leaf DtN
    // variable-coefficient leaves (block-tridiagonal path)
    bool variable = false;
    bool dominant = false;                 // variable, every leaf diagonally dominant: no pivoting
    bool packed = false;                   // variable and symmetric: packed inverse pivot blocks
    DevBuf<double> vcoef, vstencil, vSinv, vSinvP, vwork, vz;
    double* coef_pin = nullptr;  // page-locked staging of the per-cell coefficients (doubles: coef_pin_n)
    size_t coef_pin_n = 0;
    // operator arenas
    DevBuf<double> Tar, Sar, Xiar;
    // build work: X of a group, Gauss-Jordan work, zero-pivot flags of every merge
    DevBuf<double> wX, gjd;
    DevBuf<int> gji, info;
    std::vector<GroupPlan> gplans;
    DevStore bstore;
    std::vector<const void*> plan_sig;  // arena and leaf-DtN addresses gplans was made for
    // vectors
    int kcap = 0;
    int k_up = 0;  // k of the last upwards call
    DevBuf<double> vec;
    DevBuf<double> fbuf, ubuf, rbuf, abuf, bbuf;  // host-call staging
    const double* f_last = nullptr;
    std::map<int, std::unique_ptr<SweepPlan>> sweeps;
    // root Robin work
    DevBuf<double> rootA, rootAi, rootgj, rootrhs;
    DevBuf<int> rootgji, rootinfo;
    // stage timers: event pairs for build, upwards, solve
    hipEvent_t ev[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    bool ev_recorded[3] = {false, false, false};
    // captured sweeps: one executable graph per (stage, k, device pointers, flags)
    struct SweepGraph {
        int kind = 0, k = 0, flags = 0;
        std::array<const void*, 4> ptr{};
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    std::vector<SweepGraph> graphs;
    unsigned long long graph_clock = 0;
    int autotune = 1;  // sweep GEMM tiles: 1 timed at the first sweep of a width, 0 fixed rule, 11/12/22 forced
    std::map<int, int> tuned;  // (depth, product, 16-column tiles of k) -> chosen tile
    // mixed precision: refinement steps per solve, uniform-mesh neighbour tables, work
    int refine_steps = 2;
    DevBuf<int> mix_nbr, mix_pos;
    DevBuf<double> mix_res, mix_del, mix_zero;
    // serving: slot events, leaf geometry and root boundary points of the load generator
    hipEvent_t slot_ev[EF_MAX_SLOTS] = {};
    bool slot_rec[EF_MAX_SLOTS] = {};
    DevBuf<double> loadxy, rootpts, loadfac;  // distinct cell-centre x then y values; factor work
    DevBuf<int> loadcol;                      // per leaf: offsets of its x and y centres in loadxy
    int load_nxv = 0, load_nyv = 0;
};

namespace {

void clear_graphs(ef_mesh_s* M);

const double* leaf_T_ptr(ef_mesh_s* M, int leaf) {
    const long long per = 16LL * M->P.N * M->P.N;
    if (M->variable) return M->leafT.p + (M->cache_operators ? 0 : (long long)leaf) * per;
    const int cls = M->cache_operators ? M->tleaf_class0 : M->leaf_class[leaf];
    return M->leafT.p + (long long)cls * per;
}

// DtN the parent merge reads: coarsened if tagged (coarsen_ overwrites it, :736-737)
const double* merge_T_ptr(ef_mesh_s* M, int id) {
    const NodeP& nd = M->P.nodes[id];
    if (nd.tag > 0) return M->Tar.p + nd.Tc_off.back();
    if (nd.leaf) return leaf_T_ptr(M, nd.leaf_index);
    return M->Tar.p + nd.T_off;
}
const double* full_T_ptr(ef_mesh_s* M, int id) {
    const NodeP& nd = M->P.nodes[id];
    if (nd.leaf) return leaf_T_ptr(M, nd.leaf_index);
    return M->Tar.p + nd.T_off;
}

GJWork gj_work(ef_mesh_s* M, int m, int B) {
    GJWork w;
    double* d = M->gjd.p;
    w.wb = d; d += (size_t)B * m * m;
    w.Pinv = d; d += (size_t)B * GJ_NB * GJ_NB;
    w.Mm = d;
    int* i = M->gji.p;
    w.cand = i; i += (size_t)B * GJ_MAXCH * GJ_NB;
    w.cfin = i; i += (size_t)B * GJ_NB;
    w.src = i; i += (size_t)B * m;
    w.ipiv = i; i += (size_t)B * m;
    w.prow = i;
    return w;
}

// ----- build ------------------------------------------------------------------
// True when some leaf's per-cell coefficients are not constant over the leaf (the
// fast-diagonalisation path needs c_W = c_E and c_S = c_N constant within a patch).
bool coefficients_vary(const Plan& P, const ef_coeffs& c) {
    if (c.solver_type != EF_SOLVER_FIVE_POINT || c.cell == nullptr) return false;
    const long long NN = (long long)P.N * P.N;
    for (size_t li = 0; li < P.leaves.size(); li++) {
        const double* base = c.cell + li * 6 * NN;
        const double a0 = base[0], b0 = base[NN], l0 = base[5 * NN];
        for (long long e = 0; e < NN; e++)
            if (base[e] != a0 || base[NN + e] != b0 || base[2 * NN + e] != b0 || base[3 * NN + e] != b0 ||
                base[4 * NN + e] != b0 || base[5 * NN + e] != l0)
                return true;
    }
    return false;
}

// fn(lo, hi) over [0, n) on up to 16 host threads (the calling thread takes the first chunk)
template <class F>
void parallel_chunks(size_t n, F&& fn) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), std::max<size_t>(1, n / 16));
    if (nt <= 1) {
        fn(size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 1; t < nt; t++) {
        const size_t lo = t * per, hi = std::min(n, lo + per);
        if (lo < hi) th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    fn(size_t(0), std::min(n, per));
    for (auto& x : th) x.join();
}

void setup_leaf_classes(ef_mesh_s* M, const ef_coeffs& c) {
    Plan& P = M->P;
    const int N = P.N;
    const long long NN = (long long)N * N;
    const size_t nl = P.leaves.size();
    M->classes_h.clear();
    M->leaf_class.assign(nl, -1);
    M->variable = coefficients_vary(P, c);
    if (M->variable) {
        if (N > BT_MAX_N) throw EfError(EF_ERR_ARG, "variable coefficients need N <= " + std::to_string(BT_MAX_N));
        // the samples go up in one asynchronous copy from page-locked memory, filled by host
        // threads; the symmetry and dominance scans run on the host meanwhile (the stream is
        // synchronised before the buffer is refilled, so no earlier copy still reads it)
        const size_t ncoef = nl * 6 * NN;
        M->vcoef.ensure(ncoef);
        if (M->coef_pin_n < ncoef) {
            if (M->coef_pin) EF_HIP(hipHostFree(M->coef_pin));
            M->coef_pin = nullptr;
            M->coef_pin_n = 0;
            EF_HIP(hipHostMalloc((void**)&M->coef_pin, ncoef * sizeof(double), hipHostMallocDefault));
            M->coef_pin_n = ncoef;
        }
        EF_HIP(hipStreamSynchronize(M->stream));  // no copy of an earlier build still reads the buffer
        parallel_chunks(nl, [&](size_t lo, size_t hi) {
            std::memcpy(M->coef_pin + lo * 6 * NN, c.cell + lo * 6 * NN, (hi - lo) * 6 * NN * sizeof(double));
        });
        EF_HIP(hipMemcpyAsync(M->vcoef.p, M->coef_pin, ncoef * sizeof(double), hipMemcpyHostToDevice, M->stream));
        M->vstencil.ensure(nl * 5 * NN);
        M->vSinv.ensure(nl * (size_t)N * NN);
        std::atomic<bool> sym(N == 8 || N == 16 || N == 32);
        if (sym)
            parallel_chunks(nl, [&](size_t lo, size_t hi) {
                for (size_t li = lo; li < hi && sym.load(std::memory_order_relaxed); li++)
                    if (!efk::cells_symmetric(c.cell + li * 6 * NN, N)) sym = false;
            });
        M->packed = sym;
        const long long NP = bt_packed_entries(N);
        if (M->packed) M->vSinvP.ensure(nl * (size_t)N * NP);
        M->leafT.ensure(nl * 16 * NN);
        std::vector<LeafDesc> ld(nl);
        for (size_t li = 0; li < nl; li++) {
            const NodeP& nd = P.nodes[P.leaves[li]];
            LeafDesc& L = ld[li];
            L.cls = -1;
            L.node = P.leaves[li];
            L.h_off = nd.h_off;
            for (int sd = 0; sd < 4; sd++) L.gs[sd] = nd.gsrc[sd];
            L.coef = M->vcoef.p + li * 6 * NN;
            L.stencil = M->vstencil.p + li * 5 * NN;
            L.Sinv = M->vSinv.p + li * (size_t)N * NN;
            L.T = M->leafT.p + li * 16 * NN;
            L.SinvP = M->packed ? M->vSinvP.p + li * NP : nullptr;  // block-major: block b at b * nl * NP
            L.pstride = (long long)nl * NP;
            L.dx = (nd.xu - nd.xl) / N;
            L.dy = (nd.yu - nd.yl) / N;
        }
        std::atomic<bool> dom(true);
        parallel_chunks(nl, [&](size_t lo, size_t hi) {
            for (size_t li = lo; li < hi && dom.load(std::memory_order_relaxed); li++)
                if (!efk::cells_dominant(c.cell + li * 6 * NN, N, ld[li].dx, ld[li].dy)) dom = false;
        });
        M->dominant = dom;
        M->leafdesc.upload(ld, M->stream);
        return;
    }
    std::map<std::vector<uint64_t>, int> seen;
    auto bits = [](double v) { uint64_t u; std::memcpy(&u, &v, 8); return u; };
    for (size_t li = 0; li < nl; li++) {
        const NodeP& nd = P.nodes[P.leaves[li]];
        const double dx = (nd.xu - nd.xl) / N, dy = (nd.yu - nd.yl) / N;
        double cx, cy, lam;
        if (c.solver_type == EF_SOLVER_FISHPACK) {
            cx = 1.0 / (dx * dx); cy = 1.0 / (dy * dy); lam = c.lambda0;
        } else if (c.cell == nullptr) {
            cx = c.alpha0 * c.beta0 / (dx * dx); cy = c.alpha0 * c.beta0 / (dy * dy); lam = c.lambda0;
        } else {
            // per-cell samples constant over this leaf
            const double* cc = c.cell + li * 6 * NN;
            cx = cc[0] * cc[NN] / (dx * dx); cy = cc[0] * cc[NN] / (dy * dy); lam = cc[5 * NN];
        }
        std::vector<uint64_t> key = {bits(cx), bits(cy), bits(lam), bits(dx), bits(dy)};
        auto it = seen.find(key);
        int id;
        if (it == seen.end()) {
            id = (int)M->classes_h.size();
            seen[key] = id;
            PatchCoef k;
            k.cx = cx; k.cy = cy; k.lam = lam; k.dx = dx; k.dy = dy;
            M->classes_h.push_back(k);
        } else id = it->second;
        M->leaf_class[li] = id;
    }
    M->tleaf_class0 = M->leaf_class[0];
    const int ncls = (int)M->classes_h.size();
    M->leafT.ensure((size_t)ncls * 16 * NN);
    M->classes.upload(M->classes_h, M->stream);
    M->leaf_cls.upload(M->leaf_class, M->stream);
    std::vector<long long> goff(4 * nl), hoff(nl);
    for (size_t li = 0; li < nl; li++) {
        for (int sd = 0; sd < 4; sd++) goff[4 * li + sd] = P.nodes[P.leaves[li]].gsrc[sd];
        hoff[li] = P.nodes[P.leaves[li]].h_off;
    }
    M->leaf_goff.upload(goff, M->stream);
    M->leaf_hoff.upload(hoff, M->stream);
    std::vector<double> QA, QB, mu, dinv;
    patch_fd_tables(N, N <= 16 ? 1 : 2, QA, QB, mu);
    patch_fd_dinv(N, N <= 16 ? 1 : 2, mu, M->classes_h, dinv);
    M->QA.upload(QA, M->stream);
    M->QB.upload(QB, M->stream);
    M->mu.upload(mu, M->stream);
    M->dinv.upload(dinv, M->stream);
}

// Arguments of patch_fd_kernel over the mesh's constant-coefficient leaves.
PatchFDArgs leaf_fd_args(ef_mesh_s* M, int op) {
    PatchFDArgs a{};
    a.N = M->P.N;
    a.op = op;
    a.coef = M->classes.p;
    a.QA = M->QA.p;
    a.QB = M->QB.p;
    a.mu = M->mu.p;
    a.dinv = M->dinv.p;
    a.vk = 1;
    return a;
}

// Descriptors of every merge group: coarsening steps, assembly, the S and T products.
void plan_build(ef_mesh_s* M) {
    Plan& P = M->P;
    hipStream_t st = M->stream;
    M->gplans.clear();
    M->bstore.clear();
    for (const Group& g : P.groups) {
        GroupPlan gp;
        const int n = g.n, B = (int)g.nodes.size();
        const long long n2 = (long long)n * n;
        // coarsening steps of tagged children (coarsen_, HPSAlgorithm.hpp:707-741)
        std::vector<std::vector<CoarsenDesc>> steps;
        std::vector<int> maxm;
        for (int id : g.nodes)
            for (int c = 0; c < 4; c++) {
                const NodeP& ch = P.nodes[P.nodes[id].child[c]];
                for (int t = 0; t < ch.tag; t++) {
                    if ((int)steps.size() <= t) { steps.resize(t + 1); maxm.resize(t + 1, 0); }
                    CoarsenDesc cd;
                    const int m = ch.s >> t;
                    cd.m = m;
                    cd.src = (t == 0) ? full_T_ptr(M, P.nodes[id].child[c]) : M->Tar.p + ch.Tc_off[t - 1];
                    cd.dst = M->Tar.p + ch.Tc_off[t];
                    check_range(cd.src, 16ULL * m * m * 8, "coarsen source");
                    check_range(cd.dst, 4ULL * m * m * 8, "coarsen destination");
                    steps[t].push_back(cd);
                    maxm[t] = std::max(maxm[t], m);
                }
            }
        for (size_t t = 0; t < steps.size(); t++) {
            gp.csteps.push_back(M->bstore.stage(steps[t]));
            gp.csteps_cnt.push_back((int)steps[t].size());
            gp.csteps_maxm.push_back(maxm[t]);
        }
        // assembly into the group's X work (slot b at b * 16 n^2) and each parent's T
        std::vector<AsmDesc> ad(B);
        std::vector<TreeGemm> pS, pT;
        for (int b = 0; b < B; b++) {
            const NodeP& nd = P.nodes[g.nodes[b]];
            AsmDesc& a = ad[b];
            a.n = n;
            a.pad = 0;
            for (int c = 0; c < 4; c++) {
                a.Tc[c] = merge_T_ptr(M, nd.child[c]);
                check_range(a.Tc[c], 16ULL * n2 * 8, "child DtN of a merge");
            }
            a.X = M->wX.p + (long long)b * 16 * n2;
            a.T = M->Tar.p + nd.T_off;
            check_range(a.X, 16ULL * n2 * 8, "merge X work");
            check_range(a.T, 64ULL * n2 * 8, "merge T");
            const double* Xi = M->Xiar.p + nd.Xi_off;
            double* Sm = M->Sar.p + nd.S_off;
            double* Tm = M->Tar.p + nd.T_off;
            // S (4n x 8n, WESN columns) = X^-1 S_RHS; column block q of S_RHS is nonzero in the
            // two interface rows r of child c = q_child(q): -T_c(side on r, side q) when c is r's
            // first child, +T_c(...) when its second (mergeS_, HPSAlgorithm.hpp:906-929)
            for (int q = 0; q < 8; q++) {
                const int c = q_child(q), sq = q_side(q);
                const double* Tc = a.Tc[c];
                TreeGemm p{};
                const int r0 = child_if(c, 0), r1 = child_if(c, 1);
                p.A0 = Xi + (long long)r0 * n; p.lda0 = 4 * n; p.K0 = n;
                p.B0 = Tc + (long long)side_on_if(c, r0) * n * 4 * n + (long long)sq * n; p.ldb0 = 4 * n;
                p.alpha0 = (if_c1(r0) == c) ? -1.0 : 1.0;
                p.A1 = Xi + (long long)r1 * n; p.lda1 = 4 * n; p.K1 = n;
                p.B1 = Tc + (long long)side_on_if(c, r1) * n * 4 * n + (long long)sq * n; p.ldb1 = 4 * n;
                p.alpha1 = (if_c1(r1) == c) ? -1.0 : 1.0;
                p.C = nullptr; p.beta = 0.0;
                p.D = Sm + (long long)q * n; p.ldd = 8 * n;
                p.M = 4 * n; p.N = n;
                pS.push_back(p);
            }
            // T (8n x 8n) += H S: block row q is child c's DtN rows of side q against its two
            // interface sides, times the matching rows of S (mergeT_, :940-968, WESN rows)
            for (int q = 0; q < 8; q++) {
                const int c = q_child(q), sq = q_side(q);
                const double* Tc = a.Tc[c];
                TreeGemm p{};
                const int r0 = child_if(c, 0), r1 = child_if(c, 1);
                p.A0 = Tc + (long long)sq * n * 4 * n + (long long)side_on_if(c, r0) * n; p.lda0 = 4 * n; p.K0 = n;
                p.B0 = Sm + (long long)r0 * n * 8 * n; p.ldb0 = 8 * n; p.alpha0 = 1.0;
                p.A1 = Tc + (long long)sq * n * 4 * n + (long long)side_on_if(c, r1) * n; p.lda1 = 4 * n; p.K1 = n;
                p.B1 = Sm + (long long)r1 * n * 8 * n; p.ldb1 = 8 * n; p.alpha1 = 1.0;
                p.D = Tm + (long long)q * n * 8 * n; p.ldd = 8 * n;
                p.C = p.D; p.ldc = 8 * n; p.beta = 1.0;
                p.M = n; p.N = 8 * n;
                pT.push_back(p);
            }
        }
        gp.asmd = M->bstore.stage(ad);
        const bool f32 = M->precision == EF_PRECISION_MIXED;  // mergeS_ / mergeT_ products on FP32 matrix cores
        gp.gS = plan_gemm(pS, n <= 16 ? 1 : 2, M->bstore, st, f32);
        gp.gT = plan_gemm(pT, 2, M->bstore, st, f32);
        M->gplans.push_back(std::move(gp));
    }
    M->bstore.flush(st);
}

void stage_begin(ef_mesh_s* M, int st) {
    if (!M->ev[st][0]) {
        EF_HIP(hipEventCreate(&M->ev[st][0]));
        EF_HIP(hipEventCreate(&M->ev[st][1]));
    }
    EF_HIP(hipEventRecord(M->ev[st][0], M->stream));
}
void stage_end(ef_mesh_s* M, int st) {
    EF_HIP(hipEventRecord(M->ev[st][1], M->stream));
    M->ev_recorded[st] = true;
}

void check_tree_kernels() {
    static bool done = false;
    if (done) return;
    check_kernel_resources((const void*)tree_gemm_kernel<1, 1, false>, "tree_gemm_kernel<1,1>");
    check_kernel_resources((const void*)tree_gemm_kernel<2, 2, false>, "tree_gemm_kernel<2,2>");
    check_kernel_resources((const void*)tree_gemm_kernel<1, 2, false>, "tree_gemm_kernel<1,2>");
    check_kernel_resources((const void*)tree_gemm_kernel<1, 2, true>, "tree_gemm_kernel<1,2,f32>");
    check_kernel_resources((const void*)tree_gemm_kernel<1, 1, true>, "tree_gemm_kernel<1,1,f32>");
    check_kernel_resources((const void*)tree_gemm_kernel<2, 2, true>, "tree_gemm_kernel<2,2,f32>");
    check_kernel_resources((const void*)tree_gemm_split_kernel, "tree_gemm_split_kernel");
    check_kernel_resources((const void*)gj_panel_kernel<0, 1, 1, true>, "gj_panel_kernel<0,1,1,fin>");
    check_kernel_resources((const void*)gj_panel_kernel<0, 1, 2, true>, "gj_panel_kernel<0,1,2,fin>");
    check_kernel_resources((const void*)gj_panel_kernel<0, 2, 2, true>, "gj_panel_kernel<0,2,2,fin>");
    check_kernel_resources((const void*)gj_panel_kernel<0, 4, 2, true>, "gj_panel_kernel<0,4,2,fin>");
    check_kernel_resources((const void*)gj_panel_kernel<0, 4, 2, false>, "gj_panel_kernel<0,4,2>");
    check_kernel_resources((const void*)gj_panel_kernel<1, 4, 1, true>, "gj_panel_kernel<1,4,1,fin>");
    check_kernel_resources((const void*)gj_mult_kernel, "gj_mult_kernel");
    check_kernel_resources((const void*)gj_update_kernel, "gj_update_kernel");
    check_kernel_resources((const void*)gj_perm_kernel, "gj_perm_kernel");
    check_kernel_resources((const void*)gj_gather_kernel, "gj_gather_kernel");
    check_kernel_resources((const void*)tree_assemble_kernel, "tree_assemble_kernel");
    check_kernel_resources((const void*)tree_coarsen_kernel, "tree_coarsen_kernel");
    check_kernel_resources((const void*)tree_up_gather_kernel, "tree_up_gather_kernel");
    check_kernel_resources((const void*)tree_uncoarsen_kernel, "tree_uncoarsen_kernel");
    check_kernel_resources((const void*)tree_side_copy_kernel, "tree_side_copy_kernel");
    check_kernel_resources((const void*)tree_side_interp_kernel, "tree_side_interp_kernel");
    check_kernel_resources((const void*)tree_root_dirichlet_kernel, "tree_root_dirichlet_kernel");
    check_kernel_resources((const void*)tree_root_robin_kernel, "tree_root_robin_kernel");
    done = true;
}

void do_build(ef_mesh_s* M, const ef_coeffs& c) {
    Plan& P = M->P;
    hipStream_t s = M->stream;
    const int N = P.N;
    check_tree_kernels();
    clear_graphs(M);  // captured sweeps read the operator arenas reallocated below
    M->solver_type = c.solver_type;
    setup_leaf_classes(M, c);
    // arenas and work (kept across builds of the mesh: their sizes depend on the tree only)
    M->Tar.ensure(std::max<long long>(P.T_elems, 1));
    M->Sar.ensure(std::max<long long>(P.S_elems, 1));
    M->Xiar.ensure(std::max<long long>(P.Xi_elems, 1));
    M->wX.ensure(std::max<long long>(P.work_X, 1));
    M->gjd.ensure(std::max<long long>(P.gj_d, 1));
    M->gji.ensure(std::max<long long>(P.gj_i, 1));
    M->info.ensure(std::max(P.ninternal, 1));
    // the merge descriptors depend only on these addresses and on the leaf DtN each leaf
    // reads: planned again only when one of them moved
    std::vector<const void*> sig = {M->Tar.p, M->Sar.p, M->Xiar.p, M->wX.p, M->gjd.p, M->gji.p, M->info.p,
                                    (const void*)(intptr_t)(M->precision + 1)};
    for (size_t li = 0; li < P.leaves.size(); li++) sig.push_back(leaf_T_ptr(M, (int)li));
    if (sig != M->plan_sig) {
        M->plan_sig.clear();
        plan_build(M);
        M->plan_sig = sig;
    }
    if (M->variable) {
        // block-tridiagonal factors, then the DtN of every leaf (buildD2N) in chunks of leaves
        const int nl = (int)P.leaves.size();
        launch_bt_factor(M->leafdesc.p, N, nl, s, M->dominant);
        EF_LAUNCH_CHECK();
        if (M->packed) {
            hipLaunchKernelGGL(bt_pack_kernel, dim3((unsigned)std::min<long long>(blocks_for((long long)nl * N * bt_packed_entries(N)), 65535)),
                               dim3(256), 0, s, M->leafdesc.p, nl, N);
            EF_LAUNCH_CHECK();
        }
        const long long per = (long long)N * N * 4 * N;
        const long long chunk = std::max<long long>(1, std::min<long long>(nl, (512LL << 20) / (per * 8)));
        M->vwork.ensure((size_t)(chunk * per));
        launch_bt_dtn(M->leafdesc.p, N, nl, M->vwork.p, chunk, s);
        EF_LAUNCH_CHECK();
    } else {
        // leaf DtN per class (buildD2N): one "patch" per class, T written class by class
        PatchFDArgs a = leaf_fd_args(M, PATCH_OP_DTN);
        a.npatches = (int)M->classes_h.size();
        a.nrhs = 1;
        a.out = M->leafT.p;
        launch_patch_fd(a, N <= 16 ? 1 : 2, patch_default_blocks(), s);
    }
    if (sync_check()) EF_HIP(hipStreamSynchronize(s));
    if (M->bstore.pending()) throw EfError(EF_ERR_INTERNAL, "merge descriptors staged but not uploaded");
    // merges, deepest level first (post-order dependencies)
    EF_HIP(hipMemsetAsync(M->info.p, 0, sizeof(int) * std::max(P.ninternal, 1), s));
    for (size_t gi = 0; gi < P.groups.size(); gi++) {
        const Group& g = P.groups[gi];
        const GroupPlan& gp = M->gplans[gi];
        const int n = g.n, B = (int)g.nodes.size(), m = 4 * n;
        for (size_t t = 0; t < gp.csteps.size(); t++) {
            const long long out = 4LL * gp.csteps_maxm[t] * gp.csteps_maxm[t];
            EF_LAUNCH("tree_coarsen_kernel", tree_coarsen_kernel, dim3(blocks_for(out), gp.csteps_cnt[t]), dim3(256), s,
                      gp.csteps[t]);
        }
        EF_LAUNCH("tree_assemble_kernel", tree_assemble_kernel, dim3(blocks_for(80LL * n * n), B), dim3(256), s, gp.asmd);
        run_gj_inverse(M->wX.p, M->Xiar.p + g.Xi0, m, B, gj_work(M, m, B), M->info.p + g.info0, s);
        run_gemm(gp.gS, s);
        run_gemm(gp.gT, s);
    }
    // zero pivots of any merge: one read at the end of the stage
    std::vector<int> info(std::max(P.ninternal, 1), 0);
    efk::copy_d2h(info.data(), M->info.p, sizeof(int) * info.size(), s);
    for (size_t gi = 0; gi < P.groups.size(); gi++) {
        const Group& g = P.groups[gi];
        for (size_t b = 0; b < g.nodes.size(); b++)
            if (info[g.info0 + b] != 0)
                throw EfError(EF_ERR_SINGULAR, "zero pivot in the interface system of a merge at depth " +
                                                   std::to_string(g.depth) + " (column " +
                                                   std::to_string(info[g.info0 + b]) + ")");
    }
    M->built = true;
}

// Mixed precision needs the five-point residual of the whole mesh, which this build forms on
// uniform meshes: every leaf at one level, neighbours found by position.
void setup_mixed(ef_mesh_s* M) {
    const Plan& P = M->P;
    if (M->homogeneous_rhs) throw EfError(EF_ERR_ARG, "mixed precision refines solves with a load (homogeneous-rhs is off)");
    const int nl = (int)P.leaves.size();
    const int L = P.nodes[P.leaves[0]].level;
    for (int li = 0; li < nl; li++)
        if (P.nodes[P.leaves[li]].level != L)
            throw EfError(EF_ERR_ARG, "mixed precision needs a uniform mesh (every leaf at one level)");
    const int side = 1 << L;
    std::vector<int> at((size_t)side * side, -1), pos(2 * (size_t)nl), nbr(4 * (size_t)nl, -1);
    for (int li = 0; li < nl; li++) {
        const NodeP& nd = P.nodes[P.leaves[li]];
        at[(size_t)nd.ix * side + nd.iy] = li;
        pos[2 * li] = nd.ix;
        pos[2 * li + 1] = nd.iy;
    }
    for (int li = 0; li < nl; li++) {
        const int ix = pos[2 * li], iy = pos[2 * li + 1];
        if (ix > 0) nbr[4 * li] = at[(size_t)(ix - 1) * side + iy];
        if (ix < side - 1) nbr[4 * li + 1] = at[(size_t)(ix + 1) * side + iy];
        if (iy > 0) nbr[4 * li + 2] = at[(size_t)ix * side + iy - 1];
        if (iy < side - 1) nbr[4 * li + 3] = at[(size_t)ix * side + iy + 1];
    }
    M->mix_nbr.upload(nbr, M->stream);
    M->mix_pos.upload(pos, M->stream);
}

// ----- sweeps ----------------------------------------------------------------
// The upwards and the Dirichlet solve sweeps are captured once per (k, device pointers) into
// a HIP graph and replayed with one hipGraphLaunch (EF_GRAPH=0 launches them one by one).
// Captured work is kernels only: every allocation, descriptor upload and host check happens
// before capture; graphs are dropped whenever the arenas or descriptor tables they point
// into are reallocated.
constexpr size_t MAX_GRAPHS = 64;

bool graphs_on() {  // read on every sweep, so a process can compare both ways
    const char* e = std::getenv("EF_GRAPH");
    return !(e && std::strcmp(e, "0") == 0) && !sync_check();
}

void clear_graphs(ef_mesh_s* M) {
    for (auto& g : M->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    M->graphs.clear();
}

template <class Body>
void run_sweep(ef_mesh_s* M, int kind, int k, const std::array<const void*, 4>& ptr, int flags, Body&& body) {
    if (!graphs_on()) {
        body();
        return;
    }
    hipStream_t s = M->stream;
    for (auto& g : M->graphs)
        if (g.kind == kind && g.k == k && g.flags == flags && g.ptr == ptr) {
            g.used = ++M->graph_clock;
            EF_HIP(hipGraphLaunch(g.exec, s));
            return;
        }
    EF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        body();
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        throw;
    }
    EF_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) throw EfError(EF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    if (M->graphs.size() >= MAX_GRAPHS) {
        auto lru = std::min_element(M->graphs.begin(), M->graphs.end(),
                                    [](const auto& a, const auto& b) { return a.used < b.used; });
        (void)hipGraphExecDestroy(lru->exec);
        M->graphs.erase(lru);
    }
    ef_mesh_s::SweepGraph g;
    g.kind = kind; g.k = k; g.flags = flags; g.ptr = ptr; g.exec = exec; g.used = ++M->graph_clock;
    M->graphs.push_back(g);
    EF_HIP(hipGraphLaunch(exec, s));
}

void ensure_vectors(ef_mesh_s* M, int k) {
    if (k <= M->kcap) return;
    clear_graphs(M);
    M->vec.alloc((size_t)M->P.vec_rows * k);
    EF_HIP(hipMemsetAsync(M->vec.p, 0, (size_t)M->P.vec_rows * k * sizeof(double), M->stream));
    M->kcap = k;
    M->sweeps.clear();  // descriptors point into the arena
    M->k_up = 0;        // and the upward state it held is gone
}

// Tile autotuner: at the first sweep of a width k, each level's w, h and u products are planned
// with 16x16, 16x32 and 32x32 blocks per wave, timed (HIP events on the mesh's stream, best of 3
// passes after one warm pass, over whatever the vector arena holds) and the fastest kept.  The
// choice cannot change a result: every output element sums its k range in the same order under
// any tile shape.  EF_AUTOTUNE=0 (or ef_mesh_set_autotune) keeps the fixed rule.
bool autotune_on() {
    const char* e = std::getenv("EF_AUTOTUNE");
    return !(e && std::strcmp(e, "0") == 0);
}

GemmLaunch tune_gemm(ef_mesh_s* M, const std::vector<TreeGemm>& probs, DevStore& store, int& tile) {
    hipStream_t s = M->stream;
    const int cand[3][2] = {{1, 1}, {1, 2}, {2, 2}};  // (rt, tn)
    GemmLaunch best;
    float best_ms = 0.0f;
    tile = 0;
    if (probs.empty()) return plan_gemm(probs, 1, store, s);
    hipEvent_t e0, e1;
    EF_HIP(hipEventCreate(&e0));
    EF_HIP(hipEventCreate(&e1));
    for (const auto& c : cand) {
        GemmLaunch g = plan_gemm(probs, c[1], store, s, false, c[0]);
        if (g.kw > 1) {  // split-K launches have one shape (16 x 16 tiles, a wave per chunk)
            tile = 11;
            best = g;
            break;
        }
        store.flush(s);
        run_gemm(g, s);  // warm
        float ms = 0.0f;
        for (int rep = 0; rep < 3; rep++) {
            EF_HIP(hipEventRecord(e0, s));
            run_gemm(g, s);
            EF_HIP(hipEventRecord(e1, s));
            EF_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            EF_HIP(hipEventElapsedTime(&t, e0, e1));
            ms = (rep == 0) ? t : std::min(ms, t);
        }
        if (tile == 0 || ms < best_ms) {
            best = g;
            best_ms = ms;
            tile = 10 * g.rt + g.tn;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

SweepPlan& sweep_plan(ef_mesh_s* M, int k) {
    auto it = M->sweeps.find(k);
    if (it != M->sweeps.end()) return *it->second;
    Plan& P = M->P;
    hipStream_t st = M->stream;
    auto sp = std::make_unique<SweepPlan>();
    double* V = M->vec.p;
    const long long vrows = P.vec_rows;
    if (vrows > 0x7fffffffLL) throw EfError(EF_ERR_ARG, "vector arena too large");
    auto vcheck = [&](long long off, long long rows, const char* what) {
        if (off < 0 || off + rows > vrows) throw EfError(EF_ERR_INTERNAL, std::string("vector arena range: ") + what);
    };
    const int D = P.maxdepth + 1;
    const int tn = k <= 16 ? 1 : 2;
    // split-K for the long products of the upper levels (few merges, large operators): chunks
    // of 128 once the concatenated k range reaches 512; the choice depends on the operator only,
    // never on k, so a column's result is the same in any batch
    auto split_kc = [](int K0, int K1) {
        if (K0 + K1 < 512 || K0 % 128 != 0 || (K1 > 0 && K1 % 128 != 0)) return 0;
        return 128;
    };
    sp->lv.resize(D);
    for (int dd = 0; dd < D; dd++) {
        LevelSweep& L = sp->lv[dd];
        std::vector<UpDesc> ud;
        std::vector<TreeGemm> pw, ph, pu;
        std::vector<SideCopy> cp, ip;
        int dn_cnt = 0;
        std::vector<std::vector<VecStepDesc>> us;
        for (int id : P.internal_at[dd]) {
            const NodeP& nd = P.nodes[id];
            const int n = nd.n;
            UpDesc u{};
            u.n = n;
            bool plain = (n % 4) == 0;  // no child coarsened: hDiff and tau's boundary h are read in place
            for (int c = 0; c < 4; c++) {
                const NodeP& ch = P.nodes[nd.child[c]];
                u.tag[c] = ch.tag;
                u.clen[c] = 4 * ch.s;
                u.ch_off[c] = ch.h_off;
                vcheck(ch.h_off, 4LL * ch.s, "child h");
                plain = plain && ch.tag == 0 && ch.s == n;
            }
            u.hd_off = nd.hd_off; u.w_off = nd.w_off; u.h_off = nd.h_off;
            vcheck(nd.hd_off, 4LL * n, "hDiff"); vcheck(nd.w_off, 4LL * n, "w"); vcheck(nd.h_off, 8LL * n, "h");
            if (!plain) {  // coarsenUpwards_ of a tagged child: gathered (with L21) into hDiff and tau's h
                ud.push_back(u);
                L.up_nmax = std::max(L.up_nmax, n);
            }
            L.nmax = std::max(L.nmax, n);
            const double* Xi = M->Xiar.p + nd.Xi_off;
            // w = X^-1 hDiff   (mergeW_ with the stored inverse instead of a re-factorisation)
            TreeGemm p{};
            p.A0 = Xi; p.lda0 = 4 * n; p.K0 = 4 * n; p.alpha0 = 1.0;
            if (plain) {
                // hDiff block r = h of child c2(r) on side s2(r) minus h of child c1(r) on side s1(r)
                p.B0 = V; p.ldb0 = 1; p.bcs = (int)vrows;
                p.dn = n;
                p.dsub = 1;
                for (int r = 0; r < 4; r++) {
                    p.d2[r] = u.ch_off[if_c2(r)] + (long long)if_s2(r) * n;
                    p.d1[r] = u.ch_off[if_c1(r)] + (long long)if_s1(r) * n;
                }
            } else {
                p.B0 = V + nd.hd_off; p.ldb0 = 1; p.bcs = (int)vrows;
            }
            p.D = V + nd.w_off; p.ldd = 1; p.dcs = (int)vrows;
            p.M = 4 * n; p.N = k;
            p.kc = split_kc(p.K0, 0);
            pw.push_back(p);
            // h(q) += T_c(side q, interface sides) w(interfaces of c)   (mergeH_, the H block of child c)
            for (int q = 0; q < 8; q++) {
                const int c = q_child(q), sq = q_side(q);
                const double* Tc = merge_T_ptr(M, nd.child[c]);
                const int r0 = child_if(c, 0), r1 = child_if(c, 1);
                TreeGemm h{};
                h.A0 = Tc + (long long)sq * n * 4 * n + (long long)side_on_if(c, r0) * n; h.lda0 = 4 * n; h.K0 = n;
                h.B0 = V + nd.w_off + (long long)r0 * n; h.ldb0 = 1; h.alpha0 = 1.0;
                h.A1 = Tc + (long long)sq * n * 4 * n + (long long)side_on_if(c, r1) * n; h.lda1 = 4 * n; h.K1 = n;
                h.B1 = V + nd.w_off + (long long)r1 * n; h.ldb1 = 1; h.alpha1 = 1.0;
                h.bcs = (int)vrows;
                h.D = V + nd.h_off + (long long)q * n; h.ldd = 1; h.dcs = (int)vrows;
                // the boundary part of tau's h: the child's own h on side sq (gathered into D when coarsened)
                h.C = plain ? V + u.ch_off[c] + (long long)sq * n : h.D;
                h.ldc = 1; h.ccs = (int)vrows; h.beta = 1.0;
                h.M = n; h.N = k;
                h.kc = split_kc(h.K0, h.K1);
                ph.push_back(h);
            }
            // downward: u_int = S g + w   (applyS_, HPSAlgorithm.hpp:1194-1230), g read side by
            // side where it lies; then the children's g written out where the sweep needs a copy
            // (split1to4: coarsened children, or every child without side reads)
            vcheck(nd.ui_off, 4LL * n, "interface potential");
            bool own_g = true;
            for (int sd = 0; sd < 4; sd++) {
                vcheck(nd.gsrc[sd], 2LL * n, "side of g");
                own_g = own_g && nd.gsrc[sd] == nd.g_off + 2LL * n * sd;
            }
            dn_cnt++;
            for (int c = 0; c < 4; c++) {
                const NodeP& ch = P.nodes[nd.child[c]];
                if (P.side_reads && ch.tag == 0) continue;
                SideCopy sc{};
                sc.len = n;
                for (int sd = 0; sd < 4; sd++) {
                    const int q = child_src_q(c, sd);
                    sc.src[sd] = q >= 0 ? nd.gsrc[q >> 1] + (long long)(q & 1) * n
                                        : nd.ui_off + (long long)child_src_r(c, sd) * n;
                    vcheck(sc.src[sd], n, "side source");
                }
                if (P.side_reads) {
                    // a coarsened child: its first uncoarsen_ step interpolates straight from the
                    // sources into that step's destination (g, or gt when more steps follow)
                    sc.dst = ((ch.tag - 1) % 2 == 0) ? ch.g_off : ch.gt_off;
                    vcheck(sc.dst, 8LL * n, "first uncoarsen step of a child");
                    ip.push_back(sc);
                } else {
                    sc.dst = ch.gc_off;
                    vcheck(sc.dst, 4LL * n, "child g");
                    cp.push_back(sc);
                }
            }
            TreeGemm pd{};
            pd.A0 = M->Sar.p + nd.S_off; pd.lda0 = 8 * n; pd.K0 = 8 * n; pd.alpha0 = 1.0;
            if (own_g) {
                pd.B0 = V + nd.g_off;
            } else {  // block-gather of the four sides (2n rows each)
                pd.B0 = V;
                pd.dn = 2 * n;
                pd.dsub = 0;
                for (int sd = 0; sd < 4; sd++) pd.d2[sd] = nd.gsrc[sd];
            }
            pd.ldb0 = 1; pd.bcs = (int)vrows;
            pd.D = V + nd.ui_off; pd.ldd = 1; pd.dcs = (int)vrows;
            if (!M->homogeneous_rhs) { pd.C = V + nd.w_off; pd.ldc = 1; pd.ccs = (int)vrows; pd.beta = 1.0; }
            pd.M = 4 * n; pd.N = k;
            pd.kc = split_kc(pd.K0, 0);
            pu.push_back(pd);
            // uncoarsen_ steps: coarse gc -> ... -> g, alternating through gt
            for (int stp = 1; stp <= nd.tag; stp++) {
                if (stp == 1 && P.side_reads) continue;  // done from the sources at the parent's depth
                if ((int)us.size() < stp) us.resize(stp);
                VecStepDesc v;
                const int nfine = nd.s >> (nd.tag - stp);
                v.nfine = nfine;
                const bool dst_is_g = ((nd.tag - stp) % 2) == 0;
                v.dst_off = dst_is_g ? nd.g_off : nd.gt_off;
                if (stp == 1) v.src_off = nd.gc_off;
                else v.src_off = (((nd.tag - (stp - 1)) % 2) == 0) ? nd.g_off : nd.gt_off;
                vcheck(v.dst_off, 4LL * nfine, "uncoarsen destination");
                vcheck(v.src_off, 2LL * nfine, "uncoarsen source");
                us[stp - 1].push_back(v);
            }
        }
        L.cnt = dn_cnt;
        L.copy_cnt = (int)cp.size();
        L.interp_cnt = (int)ip.size();
        L.up_cnt = (int)ud.size();
        check_range(V, (size_t)vrows * k * 8, "vector arena");
        L.up = sp->store.stage(ud);
        L.copies = sp->store.stage(cp);
        L.interps = sp->store.stage(ip);
        if (M->autotune == 1 && autotune_on()) {
            // widths with the same number of 16-column tiles share a choice: a tile's cost does
            // not depend on how many of its columns are live
            const std::vector<TreeGemm>* pr[3] = {&pw, &ph, &pu};
            GemmLaunch* gl[3] = {&L.gw, &L.gh, &L.gu};
            for (int t3 = 0; t3 < 3; t3++) {
                const int key = (dd * 3 + t3) * 4096 + (k + 15) / 16;
                auto hit = M->tuned.find(key);
                if (hit != M->tuned.end()) {
                    L.tile[t3] = hit->second;
                    const int tl = hit->second ? hit->second : 11;  // 0: no product at this depth
                    *gl[t3] = plan_gemm(*pr[t3], tl % 10, sp->store, st, false, tl / 10);
                } else {
                    M->k_up = 0;  // the timed passes overwrite the arena's w, h and interface rows
                    *gl[t3] = tune_gemm(M, *pr[t3], sp->store, L.tile[t3]);
                    M->tuned[key] = L.tile[t3];
                }
            }
        } else {
            const bool forced = M->autotune > 1;
            const int ftn = forced ? M->autotune % 10 : tn, frt = forced ? M->autotune / 10 : 0;
            L.gw = plan_gemm(pw, ftn, sp->store, st, false, frt);
            L.gh = plan_gemm(ph, ftn, sp->store, st, false, frt);
            L.gu = plan_gemm(pu, ftn, sp->store, st, false, frt);
            L.tile[0] = pw.empty() ? 0 : 10 * L.gw.rt + L.gw.tn;
            L.tile[1] = ph.empty() ? 0 : 10 * L.gh.rt + L.gh.tn;
            L.tile[2] = pu.empty() ? 0 : 10 * L.gu.rt + L.gu.tn;
        }
        for (auto& v : us) {
            if (v.empty()) continue;
            int nfmax = 0;
            for (auto& x : v) nfmax = std::max(nfmax, x.nfine);
            L.unc.push_back(sp->store.stage(v));
            L.unc_cnt.push_back((int)v.size());
            L.unc_nf.push_back(nfmax);
        }
    }
    sp->store.flush(st);
    SweepPlan& ref = *sp;
    M->sweeps[k] = std::move(sp);
    return ref;
}

// Leaf solves of every (leaf, rhs): mode 0 particular Neumann data, mode 1 interior solution.
void launch_leaf(ef_mesh_s* M, int mode, const double* f, int f_zero, double* u, int k) {
    const int N = M->P.N;
    const int nl = (int)M->P.leaves.size();
    if (M->variable) {
        if (M->packed && mode != 1 && M->vz.n < (size_t)nl * k * N * N)
            throw EfError(EF_ERR_INTERNAL, "leaf work buffer not sized before the sweep");
        launch_bt_solve(M->leafdesc.p, N, nl, k, mode, f, f_zero, (const double*)M->vec.p, M->vec.p, u, M->stream,
                        M->packed, M->vz.p, M->P.vec_rows);
        EF_LAUNCH_CHECK();
    } else {
        // mode 0: particular Neumann data (g = 0) into the h rows of every leaf;
        // mode 1: solve with the g rows of every leaf, u out
        PatchFDArgs a = leaf_fd_args(M, mode == 0 ? PATCH_OP_MAP_D2N : PATCH_OP_SOLVE);
        a.npatches = nl;
        a.nrhs = k;
        a.cls = M->leaf_cls.p;
        a.f = f_zero ? nullptr : f;
        a.vk = (int)M->P.vec_rows;  // arena column stride
        if (mode == 0) {
            a.g = nullptr;
            a.out = M->vec.p;
            a.h_off = M->leaf_hoff.p;
        } else {
            a.g = M->vec.p;
            a.g_off = M->leaf_goff.p;
            a.out = u;
        }
        launch_patch_fd(a, N <= 16 ? 1 : 2, patch_default_blocks(), M->stream);
    }
    if (sync_check()) EF_HIP(hipStreamSynchronize(M->stream));
}

void do_upwards(ef_mesh_s* M, const double* f_dev, int k) {
    if (!M->built) throw EfError(EF_ERR_STATE, "upwards stage called before the build stage");
    if (k < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    Plan& P = M->P;
    hipStream_t s = M->stream;
    ensure_vectors(M, k);
    SweepPlan& sp = sweep_plan(M, k);
    if (sp.store.pending()) throw EfError(EF_ERR_INTERNAL, "sweep descriptors staged but not uploaded");
    check_range(f_dev, (size_t)k * P.leaves.size() * P.N * P.N * 8, "upwards load f");
    if (M->variable && M->packed) {
        // the packed leaves' solve work grows with k; sweeps captured against the old buffer
        // (at a smaller k, while the vector arena was already large enough) must not replay
        const size_t need = (size_t)P.leaves.size() * k * P.N * P.N;
        if (M->vz.n < need) {
            clear_graphs(M);
            M->vz.ensure(need);
        }
    }
    run_sweep(M, 0, k, {f_dev, nullptr, nullptr, nullptr}, M->homogeneous_rhs ? 1 : 0, [&] {
        // leaf particular Neumann data (particularNeumannData, FiniteVolumeSolver.cpp:457-462)
        launch_leaf(M, 0, f_dev, 0, nullptr, k);
        if (!M->homogeneous_rhs) {  // upwards4to1 is skipped entirely for homogeneous-rhs (:532)
            for (int dd = P.maxdepth; dd >= 0; dd--) {
                const LevelSweep& L = sp.lv[dd];
                if (L.cnt == 0) continue;
                if (L.up_cnt > 0)
                    EF_LAUNCH("tree_up_gather_kernel", tree_up_gather_kernel,
                              dim3(vec_blocks(12LL * L.up_nmax * k, L.up_cnt), L.up_cnt), dim3(256), s, L.up, M->vec.p, k,
                              P.vec_rows);
                run_gemm(L.gw, s);
                run_gemm(L.gh, s);
            }
        }
    });
    M->k_up = k;
    M->f_last = f_dev;
}

void do_solve_once(ef_mesh_s* M, const double* f_dev, const double* r_dev, const double* a_dev, const double* b_dev,
                   bool robin, double* u_dev, int k) {
    if (!M->built) throw EfError(EF_ERR_STATE, "solve stage called before the build stage");
    if (k < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    Plan& P = M->P;
    hipStream_t s = M->stream;
    ensure_vectors(M, k);
    SweepPlan& sp = sweep_plan(M, k);  // (a new, tuned plan has overwritten any upward state)
    if (!M->homogeneous_rhs && M->k_up != k)
        throw EfError(EF_ERR_STATE, "solve stage needs a preceding upwards stage with the same number of right-hand sides");
    if (sp.store.pending()) throw EfError(EF_ERR_INTERNAL, "sweep descriptors staged but not uploaded");
    const NodeP& root = P.nodes[0];
    const int m = 4 * root.s;
    double* V = M->vec.p;
    check_range(r_dev, (size_t)m * k * 8, "root data r");
    check_range(a_dev, (size_t)m * 8, "root data a");
    check_range(u_dev, (size_t)k * P.leaves.size() * P.N * P.N * 8, "solution u");
    // downward sweep (solve4to1, HPSAlgorithm.hpp:562-577) and leaf solves (leafSolve, :584-596);
    // homogeneous-rhs solves with f = 0
    const bool f_zero = M->homogeneous_rhs || !f_dev;
    auto downward = [&] {
        for (int dd = 0; dd <= P.maxdepth; dd++) {
            const LevelSweep& L = sp.lv[dd];
            if (L.cnt == 0) continue;
            for (size_t stp = 0; stp < L.unc.size(); stp++)
                EF_LAUNCH("tree_uncoarsen_kernel", tree_uncoarsen_kernel,
                          dim3(vec_blocks(4LL * L.unc_nf[stp] * k, L.unc_cnt[stp]), L.unc_cnt[stp]), dim3(256), s, L.unc[stp], V, k,
                          P.vec_rows);
            run_gemm(L.gu, s);
            if (L.copy_cnt > 0)
                EF_LAUNCH("tree_side_copy_kernel", tree_side_copy_kernel, dim3(vec_blocks(4LL * L.nmax * k, L.copy_cnt), L.copy_cnt),
                          dim3(256), s, L.copies, V, k, P.vec_rows);
            if (L.interp_cnt > 0)
                EF_LAUNCH("tree_side_interp_kernel", tree_side_interp_kernel,
                          dim3(vec_blocks(8LL * L.nmax * k, L.interp_cnt), L.interp_cnt), dim3(256), s, L.interps, V, k, P.vec_rows);
        }
        launch_leaf(M, 1, f_dev ? f_dev : (const double*)V, f_zero ? 1 : 0, u_dev, k);
    };
    if (!robin) {
        run_sweep(M, 1, k, {f_dev, r_dev, a_dev, u_dev}, (M->homogeneous_rhs ? 1 : 0) | (f_zero ? 2 : 0), [&] {
            EF_LAUNCH("tree_root_dirichlet_kernel", tree_root_dirichlet_kernel, dim3(blocks_for((long long)m * k)),
                      dim3(256), s, r_dev, a_dev, V + root.g_off, m, k, P.vec_rows);
            downward();
        });
        return;
    } else {
        // g = (diag(a) + diag(b) T)^-1 (r - b h)
        check_range(b_dev, (size_t)m * 8, "root data b");
        if (m > GJ_MAXM) throw EfError(EF_ERR_ARG, "root boundary system larger than the inversion limit");
        M->rootA.ensure((size_t)m * m);
        M->rootAi.ensure((size_t)m * m);
        M->rootrhs.ensure((size_t)m * k);
        M->rootgj.ensure(gj_work_doubles(m, 1));
        M->rootgji.ensure(gj_work_ints(m, 1));
        M->rootinfo.ensure(1);
        const double* Troot = full_T_ptr(M, 0);
        const double* hroot = (root.leaf || M->homogeneous_rhs) ? nullptr : V + root.h_off;
        EF_LAUNCH("tree_root_robin_kernel", tree_root_robin_kernel, dim3(blocks_for((long long)m * m + (long long)m * k)),
                  dim3(256), s, Troot, a_dev, b_dev, r_dev, hroot, M->rootA.p, M->rootrhs.p, m, k, P.vec_rows);
        GJWork w;
        double* d = M->rootgj.p;
        w.wb = d; d += (size_t)m * m;
        w.Pinv = d; d += (size_t)GJ_NB * GJ_NB;
        w.Mm = d;
        int* ii = M->rootgji.p;
        w.cand = ii; ii += (size_t)GJ_MAXCH * GJ_NB;
        w.cfin = ii; ii += (size_t)GJ_NB;
        w.src = ii; ii += (size_t)m;
        w.ipiv = ii; ii += (size_t)m;
        w.prow = ii;
        EF_HIP(hipMemsetAsync(M->rootinfo.p, 0, sizeof(int), s));
        run_gj_inverse(M->rootA.p, M->rootAi.p, m, 1, w, M->rootinfo.p, s);
        TreeGemm p{};
        p.A0 = M->rootAi.p; p.lda0 = m; p.K0 = m; p.alpha0 = 1.0;
        p.B0 = M->rootrhs.p; p.ldb0 = 1; p.bcs = m;
        p.D = V + root.g_off; p.ldd = 1; p.dcs = (int)P.vec_rows;
        p.M = m; p.N = k;
        DevStore tmp;
        GemmLaunch g = plan_gemm(std::vector<TreeGemm>{p}, k <= 16 ? 1 : 2, tmp, s);
        tmp.flush(s);
        run_gemm(g, s);
        int info = 0;
        efk::copy_d2h(&info, M->rootinfo.p, sizeof(int), s);
        if (info) throw EfError(EF_ERR_SINGULAR, "zero pivot in the root boundary system");
    }
    downward();
}

// Mixed precision: the solve with the FP32-product operators, then refine_steps rounds of
// iterative refinement -- the FP64 residual of the five-point system (mixed_residual_kernel),
// the correction from an upwards + solve pass with that residual as the load and zero
// Dirichlet data, and u += correction.  The rounds reuse the upward state, so a mixed solve
// consumes its upwards stage (the next solve needs a new one).
void do_solve(ef_mesh_s* M, const double* f_dev, const double* r_dev, const double* a_dev, const double* b_dev,
              bool robin, double* u_dev, int k) {
    if (M->precision != EF_PRECISION_MIXED) {
        do_solve_once(M, f_dev, r_dev, a_dev, b_dev, robin, u_dev, k);
        return;
    }
    if (robin) throw EfError(EF_ERR_ARG, "mixed precision refines Dirichlet root conditions only (b = 0)");
    do_solve_once(M, f_dev, r_dev, a_dev, b_dev, false, u_dev, k);
    const Plan& P = M->P;
    const int nl = (int)P.leaves.size(), N = P.N, m = 4 * P.nodes[0].s;
    const long long cnt = (long long)k * nl * N * N;
    M->mix_res.ensure(cnt);
    M->mix_del.ensure(cnt);
    if (M->mix_zero.n < (size_t)k * m) {
        M->mix_zero.alloc((size_t)k * m);
        EF_LAUNCH("fill_kernel", fill_kernel, dim3(blocks_for((long long)k * m)), dim3(256), M->stream, M->mix_zero.p, 0.0,
                  (long long)k * m);
    }
    check_range(u_dev, (size_t)cnt * 8, "solution u");
    if (f_dev) check_range(f_dev, (size_t)cnt * 8, "load f");
    ResidualArgs ra{};
    ra.u = u_dev; ra.f = f_dev; ra.res = M->mix_res.p;
    ra.k = k; ra.nl = nl; ra.N = N; ra.M = P.nodes[0].s;
    ra.nbr = M->mix_nbr.p; ra.pos = M->mix_pos.p; ra.r = r_dev; ra.a = a_dev;
    ra.stencil = M->variable ? M->vstencil.p : nullptr;
    ra.coef = M->classes.p; ra.cls = M->leaf_cls.p;
    for (int it = 0; it < M->refine_steps; it++) {
        EF_LAUNCH("mixed_residual_kernel", mixed_residual_kernel, dim3(blocks_for(cnt)), dim3(256), M->stream, ra);
        do_upwards(M, M->mix_res.p, k);
        do_solve_once(M, M->mix_res.p, M->mix_zero.p, a_dev, nullptr, false, M->mix_del.p, k);
        EF_LAUNCH("axpy_kernel", axpy_kernel, dim3(blocks_for(cnt)), dim3(256), M->stream, u_dev,
                  (const double*)M->mix_del.p, cnt);
    }
    M->k_up = 0;
    M->f_last = nullptr;
}

}  // namespace

namespace efk {
void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {
std::mutex g_stage_mu;
void* g_stage = nullptr;  // kept until the process ends
constexpr size_t STAGE_BYTES = size_t(16) << 20;

void* stage_buffer() {
    if (!g_stage) EF_HIP(hipHostMalloc(&g_stage, STAGE_BYTES, hipHostMallocDefault));
    return g_stage;
}
}  // namespace

void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    void* st = stage_buffer();
    for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        std::memcpy(st, (const char*)src + off, n);
        EF_HIP(hipMemcpyAsync((char*)dst + off, st, n, hipMemcpyHostToDevice, s));
        EF_HIP(hipStreamSynchronize(s));
    }
}

void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    void* st = stage_buffer();
    for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        EF_HIP(hipMemcpyAsync(st, (const char*)src + off, n, hipMemcpyDeviceToHost, s));
        EF_HIP(hipStreamSynchronize(s));
        std::memcpy((char*)dst + off, st, n);
    }
}
}  // namespace efk

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* ef_last_error(void) { return g_last_error.c_str(); }
int ef_version(void) { return 1; }

int ef_ctx_create(int device, ef_ctx* out) {
    EF_API_BEGIN
    if (!out) throw EfError(EF_ERR_ARG, "null output handle");
    int ndev = 0;
    EF_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw EfError(EF_ERR_ARG, "no such device");
    EF_HIP(hipSetDevice(device));
    efk::noalloc_setup(device);
    auto* c = new ef_ctx_s();
    c->device = device;
    *out = c;
    EF_API_END
}

int ef_ctx_destroy(ef_ctx ctx) {
    EF_API_BEGIN
    delete ctx;
    EF_API_END
}

int ef_mesh_create(ef_ctx ctx, const ef_mesh_desc* desc, ef_mesh* out) {
    EF_API_BEGIN
    if (!ctx || !desc || !out) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ef_mesh_s> M(new ef_mesh_s());
    M->ctx = ctx;
    plan_tree(M->P, *desc);
    EF_HIP(hipStreamCreate(&M->own_stream));
    M->stream = M->own_stream;
    *out = M.release();
    EF_API_END
}

int ef_mesh_destroy(ef_mesh mesh) {
    EF_API_BEGIN
    if (!mesh) return EF_OK;
    (void)hipSetDevice(mesh->ctx->device);
    if (mesh->stream) (void)hipStreamSynchronize(mesh->stream);
    clear_graphs(mesh);
    for (auto& pr : mesh->ev)
        for (auto& e : pr)
            if (e) (void)hipEventDestroy(e);
    for (auto& e : mesh->slot_ev)
        if (e) (void)hipEventDestroy(e);
    hipStream_t s = mesh->own_stream;
    if (mesh->coef_pin) (void)hipHostFree(mesh->coef_pin);
    delete mesh;
    if (s) (void)hipStreamDestroy(s);
    EF_API_END
}

int ef_mesh_footprint(const ef_mesh_desc* desc, int precision, int max_rhs, size_t* bytes) {
    EF_API_BEGIN
    if (!desc || !bytes) throw EfError(EF_ERR_ARG, "null argument");
    Plan P;
    plan_tree(P, *desc);
    const long long N = P.N, nl = (long long)P.leaves.size(), k = std::max(max_rhs, 1);
    const bool variable = (precision & EF_FOOTPRINT_VARIABLE) != 0;
    long long elems = P.T_elems + P.S_elems + P.Xi_elems + P.work_X + P.gj_d;
    elems += (P.gj_i + P.ninternal + 1) / 2;                     // int work
    elems += nl * 16 * N * N;                                   // leaf DtN upper bound (one per leaf)
    elems += P.vec_rows * k;
    elems += 2LL * k * nl * N * N;                              // f and u staging
    if (variable)  // samples (6 N^2), stencil (5 N^2), pivot inverses (N^3), packed halves, solve work
        elems += nl * (11 * N * N + N * N * N + N * bt_packed_entries((int)N)) + nl * k * N * N;
    *bytes = (size_t)elems * sizeof(double);
    EF_API_END
}

int ef_mesh_info(ef_mesh mesh, int* info) {
    EF_API_BEGIN
    if (!mesh || !info) throw EfError(EF_ERR_ARG, "null argument");
    const Plan& P = mesh->P;
    int nint = 0;
    for (auto& nd : P.nodes) nint += nd.leaf ? 0 : 1;
    info[0] = (int)P.nodes.size();
    info[1] = (int)P.leaves.size();
    info[2] = P.nodes[0].s;
    info[3] = P.maxdepth;
    info[4] = P.N;
    info[5] = nint;
    info[6] = P.max_n;
    info[7] = 0;
    EF_API_END
}

int ef_mesh_nodes(ef_mesh mesh, int* info, double* bounds) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null argument");
    const Plan& P = mesh->P;
    for (size_t i = 0; i < P.nodes.size(); i++) {
        const NodeP& nd = P.nodes[i];
        if (info) {
            int v[12] = {nd.level, nd.ix, nd.iy, nd.leaf ? 1 : 0, nd.leaf_index, nd.tag, nd.s,
                         nd.child[0], nd.child[1], nd.child[2], nd.child[3], nd.parent};
            std::memcpy(info + 12 * i, v, sizeof(v));
        }
        if (bounds) {
            bounds[4 * i + 0] = nd.xl; bounds[4 * i + 1] = nd.xu;
            bounds[4 * i + 2] = nd.yl; bounds[4 * i + 3] = nd.yu;
        }
    }
    EF_API_END
}

int ef_mesh_set_stream(ef_mesh mesh, void* hip_stream) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    mesh->stream = hip_stream ? (hipStream_t)hip_stream : mesh->own_stream;
    EF_API_END
}

int ef_mesh_stream(ef_mesh mesh, void** hip_stream) {
    EF_API_BEGIN
    if (!mesh || !hip_stream) throw EfError(EF_ERR_ARG, "null argument");
    *hip_stream = (void*)mesh->stream;
    EF_API_END
}

int ef_mesh_set_options(ef_mesh mesh, int cache_operators, int homogeneous_rhs) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    if ((cache_operators != 0) != mesh->cache_operators && mesh->built)
        throw EfError(EF_ERR_STATE, "cache-operators must be set before the build stage");
    mesh->cache_operators = cache_operators != 0;
    if ((homogeneous_rhs != 0) != mesh->homogeneous_rhs) {
        clear_graphs(mesh);  // captured sweeps read the descriptors of the plans dropped here
        mesh->sweeps.clear();
        mesh->k_up = 0;
    }
    mesh->homogeneous_rhs = homogeneous_rhs != 0;
    EF_API_END
}

int ef_build(ef_mesh mesh, const ef_coeffs* coeffs, int precision) {
    EF_API_BEGIN
    if (!mesh || !coeffs) throw EfError(EF_ERR_ARG, "null argument");
    if (coeffs->solver_type != EF_SOLVER_FISHPACK && coeffs->solver_type != EF_SOLVER_FIVE_POINT)
        throw EfError(EF_ERR_ARG, "unknown solver type");
    if (precision != EF_PRECISION_FP64 && precision != EF_PRECISION_MIXED) throw EfError(EF_ERR_ARG, "unknown precision");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    if (precision == EF_PRECISION_MIXED) setup_mixed(mesh);
    mesh->built = false;
    mesh->sweeps.clear();
    mesh->k_up = 0;
    mesh->precision = precision;
    stage_begin(mesh, 0);
    do_build(mesh, *coeffs);
    stage_end(mesh, 0);
    EF_API_END
}

int ef_upwards_device(ef_mesh mesh, const double* f_dev, int nrhs) {
    EF_API_BEGIN
    if (!mesh || !f_dev) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    stage_begin(mesh, 1);
    do_upwards(mesh, f_dev, nrhs);
    stage_end(mesh, 1);
    EF_API_END
}

int ef_upwards(ef_mesh mesh, const double* f, int nrhs) {
    EF_API_BEGIN
    if (!mesh || !f) throw EfError(EF_ERR_ARG, "null argument");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    const size_t cnt = (size_t)nrhs * mesh->P.leaves.size() * mesh->P.N * mesh->P.N;
    mesh->fbuf.ensure(cnt);
    efk::copy_h2d(mesh->fbuf.p, f, cnt * sizeof(double), mesh->stream);
    stage_begin(mesh, 1);
    do_upwards(mesh, mesh->fbuf.p, nrhs);
    stage_end(mesh, 1);
    EF_HIP(hipStreamSynchronize(mesh->stream));
    EF_API_END
}

int ef_solve_device(ef_mesh mesh, const double* f_dev, const double* r_dev, const double* a_dev, const double* b_dev,
                    double* u_dev, int nrhs) {
    EF_API_BEGIN
    if (!mesh || !r_dev || !a_dev || !u_dev) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    stage_begin(mesh, 2);
    do_solve(mesh, f_dev, r_dev, a_dev, b_dev, b_dev != nullptr, u_dev, nrhs);
    stage_end(mesh, 2);
    EF_API_END
}

int ef_solve(ef_mesh mesh, const double* r, const double* a, const double* b, double* u, int nrhs) {
    EF_API_BEGIN
    if (!mesh || !r || !a || !u) throw EfError(EF_ERR_ARG, "null argument");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    const Plan& P = mesh->P;
    const int m = 4 * P.nodes[0].s;
    bool robin = false;
    if (b)
        for (int i = 0; i < m; i++)
            if (b[i] != 0.0) { robin = true; break; }
    mesh->rbuf.ensure((size_t)m * nrhs);
    mesh->abuf.ensure((size_t)m);
    mesh->bbuf.ensure((size_t)m);
    efk::copy_h2d(mesh->rbuf.p, r, sizeof(double) * m * nrhs, mesh->stream);
    efk::copy_h2d(mesh->abuf.p, a, sizeof(double) * m, mesh->stream);
    if (robin) efk::copy_h2d(mesh->bbuf.p, b, sizeof(double) * m, mesh->stream);
    const size_t cnt = (size_t)nrhs * P.leaves.size() * P.N * P.N;
    mesh->ubuf.ensure(cnt);
    const double* f = mesh->homogeneous_rhs ? nullptr : mesh->f_last;
    stage_begin(mesh, 2);
    do_solve(mesh, f, mesh->rbuf.p, mesh->abuf.p, robin ? mesh->bbuf.p : nullptr, robin, mesh->ubuf.p, nrhs);
    stage_end(mesh, 2);
    efk::copy_d2h(u, mesh->ubuf.p, cnt * sizeof(double), mesh->stream);
    EF_HIP(hipStreamSynchronize(mesh->stream));
    EF_API_END
}

int ef_mesh_sync(ef_mesh mesh) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    EF_HIP(hipStreamSynchronize(mesh->stream));
    EF_API_END
}

int ef_device_alloc(ef_ctx ctx, size_t bytes, void** ptr) {
    EF_API_BEGIN
    if (!ctx || !ptr) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(ctx->device));
    EF_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    EF_API_END
}

int ef_device_free(ef_ctx ctx, void* ptr) {
    EF_API_BEGIN
    if (!ctx) throw EfError(EF_ERR_ARG, "null context");
    EF_HIP(hipSetDevice(ctx->device));
    if (ptr) EF_HIP(hipFree(ptr));
    EF_API_END
}

int ef_memcpy(ef_ctx ctx, void* dst, const void* src, size_t bytes, int kind) {
    EF_API_BEGIN
    if (!ctx || (bytes && (!dst || !src))) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(ctx->device));
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) throw EfError(EF_ERR_ARG, "copy kind must be 0, 1 or 2");
    if (kind == 0) efk::copy_h2d(dst, src, bytes, nullptr);
    else if (kind == 1) efk::copy_d2h(dst, src, bytes, nullptr);
    else if (bytes % sizeof(double) == 0 && (uintptr_t)dst % 8 == 0 && (uintptr_t)src % 8 == 0) {
        // whole doubles: the library's copy kernel (one read and one write per element)
        const long long n = (long long)(bytes / sizeof(double));
        hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(n)), dim3(256), 0, nullptr, (const double*)src, (double*)dst, n);
        EF_LAUNCH_CHECK();
        EF_HIP(hipStreamSynchronize(nullptr));
    } else if (bytes) EF_HIP(hipMemcpy(dst, src, bytes, k));
    EF_API_END
}

int ef_mesh_time_sweep_gemms(ef_mesh mesh, int nrhs, int reps, double* ms_per_pass, int* launches) {
    EF_API_BEGIN
    if (!mesh || !ms_per_pass || !launches) throw EfError(EF_ERR_ARG, "null argument");
    if (!mesh->built) throw EfError(EF_ERR_STATE, "build stage first");
    if (nrhs < 1 || reps < 1) throw EfError(EF_ERR_ARG, "nrhs and reps must be >= 1");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    ensure_vectors(mesh, nrhs);
    SweepPlan& sp = sweep_plan(mesh, nrhs);
    const Plan& P = mesh->P;
    hipStream_t s = mesh->stream;
    int count = 0;
    auto pass = [&] {
        for (int dd = P.maxdepth; dd >= 0; dd--) {
            const LevelSweep& L = sp.lv[dd];
            if (L.cnt == 0) continue;
            if (!mesh->homogeneous_rhs) {
                run_gemm(L.gw, s);
                run_gemm(L.gh, s);
                count += (L.gw.nwaves > 0) + (L.gh.nwaves > 0);
            }
        }
        for (int dd = 0; dd <= P.maxdepth; dd++) {
            const LevelSweep& L = sp.lv[dd];
            if (L.cnt == 0) continue;
            run_gemm(L.gu, s);
            count += (L.gu.nwaves > 0);
        }
    };
    pass();  // warm
    hipEvent_t e0, e1;
    EF_HIP(hipEventCreate(&e0));
    EF_HIP(hipEventCreate(&e1));
    count = 0;
    EF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) pass();
    EF_HIP(hipEventRecord(e1, s));
    EF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    EF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_pass = ms / reps;
    *launches = count / reps;
    EF_API_END
}

// ----- serving data path -----------------------------------------------------------
int ef_host_alloc(ef_ctx ctx, size_t bytes, void** ptr) {
    EF_API_BEGIN
    if (!ctx || !ptr) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipSetDevice(ctx->device));
    EF_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    EF_API_END
}

int ef_host_free(ef_ctx ctx, void* ptr) {
    EF_API_BEGIN
    if (!ctx) throw EfError(EF_ERR_ARG, "null context");
    if (ptr) EF_HIP(hipHostFree(ptr));
    EF_API_END
}

int ef_mesh_set_autotune(ef_mesh mesh, int mode) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    if (mode != 0 && mode != 1 && mode != 11 && mode != 12 && mode != 22)
        throw EfError(EF_ERR_ARG, "autotune mode must be 0, 1, 11, 12 or 22");
    mesh->autotune = mode;
    mesh->tuned.clear();
    mesh->sweeps.clear();  // re-planned (and re-tuned) at the next sweep of each width,
    mesh->k_up = 0;        // so the next solve needs a new upwards stage
    clear_graphs(mesh);
    EF_API_END
}

int ef_mesh_sweep_tiles(ef_mesh mesh, int nrhs, int* tiles, int* levels) {
    EF_API_BEGIN
    if (!mesh || !levels) throw EfError(EF_ERR_ARG, "null argument");
    if (!mesh->built) throw EfError(EF_ERR_STATE, "build stage first");
    if (nrhs < 1) throw EfError(EF_ERR_ARG, "nrhs must be >= 1");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    ensure_vectors(mesh, nrhs);
    SweepPlan& sp = sweep_plan(mesh, nrhs);
    *levels = (int)sp.lv.size();
    if (tiles)
        for (size_t d = 0; d < sp.lv.size(); d++)
            for (int t3 = 0; t3 < 3; t3++) tiles[3 * d + t3] = sp.lv[d].tile[t3];
    EF_API_END
}

int ef_mesh_set_refinement(ef_mesh mesh, int steps) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    if (steps < 0 || steps > 16) throw EfError(EF_ERR_ARG, "refinement steps must be in 0..16");
    mesh->refine_steps = steps;
    EF_API_END
}

int ef_mesh_copy_async(ef_mesh mesh, void* dst, const void* src, size_t bytes, int kind) {
    EF_API_BEGIN
    if (!mesh || (bytes && (!dst || !src))) throw EfError(EF_ERR_ARG, "null argument");
    if (kind < 0 || kind > 2) throw EfError(EF_ERR_ARG, "copy kind must be 0, 1 or 2");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    if (bytes) EF_HIP(hipMemcpyAsync(dst, src, bytes, k, mesh->stream));
    EF_API_END
}

int ef_mesh_record(ef_mesh mesh, int slot) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    if (slot < 0 || slot >= EF_MAX_SLOTS) throw EfError(EF_ERR_ARG, "slot out of range");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    if (!mesh->slot_ev[slot]) EF_HIP(hipEventCreateWithFlags(&mesh->slot_ev[slot], hipEventDisableTiming));
    EF_HIP(hipEventRecord(mesh->slot_ev[slot], mesh->stream));
    mesh->slot_rec[slot] = true;
    EF_API_END
}

int ef_mesh_query(ef_mesh mesh, int slot, int* done) {
    EF_API_BEGIN
    if (!mesh || !done) throw EfError(EF_ERR_ARG, "null argument");
    if (slot < 0 || slot >= EF_MAX_SLOTS) throw EfError(EF_ERR_ARG, "slot out of range");
    *done = 1;
    if (!mesh->slot_rec[slot]) return EF_OK;
    const hipError_t e = hipEventQuery(mesh->slot_ev[slot]);
    if (e == hipErrorNotReady) { (void)hipGetLastError(); *done = 0; }
    else if (e != hipSuccess) throw EfError(EF_ERR_DEVICE, std::string("hipEventQuery: ") + hipGetErrorString(e));
    EF_API_END
}

int ef_mesh_wait(ef_mesh mesh, int slot) {
    EF_API_BEGIN
    if (!mesh) throw EfError(EF_ERR_ARG, "null mesh");
    if (slot < 0 || slot >= EF_MAX_SLOTS) throw EfError(EF_ERR_ARG, "slot out of range");
    if (mesh->slot_rec[slot]) EF_HIP(hipEventSynchronize(mesh->slot_ev[slot]));
    EF_API_END
}

int ef_mesh_eval_load(ef_mesh mesh, const double* params_dev, int nrhs, double* f_dev, double* r_dev) {
    EF_API_BEGIN
    if (!mesh || !params_dev || !f_dev || !r_dev) throw EfError(EF_ERR_ARG, "null argument");
    if (nrhs < 1 || nrhs > 65535) throw EfError(EF_ERR_ARG, "nrhs must be in 1..65535");
    EF_HIP(hipSetDevice(mesh->ctx->device));
    const Plan& P = mesh->P;
    const int nl = (int)P.leaves.size(), N = P.N, m = 4 * P.nodes[0].s;
    if (!mesh->loadxy.p) {
        // distinct (xl, dx) and (yl, dy) of the leaves: their N cell centres are computed once
        std::map<std::pair<double, double>, int> xcols, ycols;
        std::vector<int> col(2 * (size_t)nl);
        std::vector<double> xs, ys, pts(2 * (size_t)m);
        for (int li = 0; li < nl; li++) {
            const NodeP& nd = P.nodes[P.leaves[li]];
            const double dx = (nd.xu - nd.xl) / N, dy = (nd.yu - nd.yl) / N;
            auto ix = xcols.emplace(std::make_pair(nd.xl, dx), (int)xs.size());
            if (ix.second)
                for (int i = 0; i < N; i++) xs.push_back((nd.xl + dx / 2) + i * dx);
            auto iy = ycols.emplace(std::make_pair(nd.yl, dy), (int)ys.size());
            if (iy.second)
                for (int j = 0; j < N; j++) ys.push_back((nd.yl + dy / 2) + j * dy);
            col[2 * li] = ix.first->second;
            col[2 * li + 1] = iy.first->second;
        }
        // root boundary points, sides W, E, S, N (HPSAlgorithm.hpp:375-400)
        const NodeP& root = P.nodes[0];
        const int M = root.s;
        const double dx = (root.xu - root.xl) / M, dy = (root.yu - root.yl) / M;
        for (int i = 0; i < M; i++) {
            const double ys0 = (root.yl + dy / 2) + i * dy, xs0 = (root.xl + dx / 2) + i * dx;
            pts[i] = root.xl;         pts[m + i] = ys0;
            pts[M + i] = root.xu;     pts[m + M + i] = ys0;
            pts[2 * M + i] = xs0;     pts[m + 2 * M + i] = root.yl;
            pts[3 * M + i] = xs0;     pts[m + 3 * M + i] = root.yu;
        }
        mesh->load_nxv = (int)xs.size();
        mesh->load_nyv = (int)ys.size();
        xs.insert(xs.end(), ys.begin(), ys.end());
        mesh->loadxy.upload(xs, mesh->stream);
        mesh->loadcol.upload(col, mesh->stream);
        mesh->rootpts.upload(pts, mesh->stream);
    }
    const int nxv = mesh->load_nxv, nyv = mesh->load_nyv;
    check_range(params_dev, (size_t)nrhs * 64 * 8, "load parameters");
    check_range(f_dev, (size_t)nrhs * nl * N * N * 8, "load f");
    check_range(r_dev, (size_t)nrhs * m * 8, "load r");
    mesh->loadfac.ensure((size_t)nrhs * 8 * (nxv + nyv));
    EF_LAUNCH("serve_factor_kernel", serve_factor_kernel, dim3(blocks_for(8LL * (nxv + nyv)), nrhs), dim3(256),
              mesh->stream, (const double*)mesh->loadxy.p, nxv, nyv, params_dev, mesh->loadfac.p);
    EF_LAUNCH("serve_cells_kernel", serve_cells_kernel, dim3((unsigned)std::min(nl, 512), nrhs), dim3(256),
              mesh->stream, (const int*)mesh->loadcol.p, nl, N, nxv, nyv, params_dev, (const double*)mesh->loadfac.p,
              f_dev);
    EF_LAUNCH("serve_bdata_kernel", serve_bdata_kernel, dim3((unsigned)std::min<long long>(blocks_for((long long)nrhs * m), 1024)),
              dim3(256), mesh->stream, params_dev, nrhs, (const double*)mesh->rootpts.p, m, r_dev);
    EF_API_END
}

int ef_debug_invert(ef_ctx ctx, int m, int B, const double* A, double* Ainv, int* ipiv) {
    EF_API_BEGIN
    if (!ctx || !A || !Ainv) throw EfError(EF_ERR_ARG, "null argument");
    if (m < 1 || B < 1) throw EfError(EF_ERR_ARG, "m and B must be >= 1");
    EF_HIP(hipSetDevice(ctx->device));
    check_tree_kernels();
    const size_t mm = (size_t)m * m * B;
    DevBuf<double> wa, xi, d;
    DevBuf<int> ii, info;
    wa.alloc(mm);
    xi.alloc(mm);
    d.alloc(gj_work_doubles(m, B));
    ii.alloc(gj_work_ints(m, B));
    info.alloc(B);
    GJWork w;
    double* dp = d.p;
    w.wb = dp; dp += mm;
    w.Pinv = dp; dp += (size_t)B * GJ_NB * GJ_NB;
    w.Mm = dp;
    int* ip = ii.p;
    w.cand = ip; ip += (size_t)B * GJ_MAXCH * GJ_NB;
    w.cfin = ip; ip += (size_t)B * GJ_NB;
    w.src = ip; ip += (size_t)B * m;
    w.ipiv = ip; ip += (size_t)B * m;
    w.prow = ip;
    hipStream_t s = ctx->stream;
    efk::copy_h2d(wa.p, A, mm * sizeof(double), s);
    EF_HIP(hipMemsetAsync(info.p, 0, sizeof(int) * B, s));
    run_gj_inverse(wa.p, xi.p, m, B, w, info.p, s);
    efk::copy_d2h(Ainv, xi.p, mm * sizeof(double), s);
    if (ipiv) efk::copy_d2h(ipiv, w.ipiv, sizeof(int) * B * m, s);
    std::vector<int> inf(B);
    efk::copy_d2h(inf.data(), info.p, sizeof(int) * B, s);
    for (int b = 0; b < B; b++)
        if (inf[b]) throw EfError(EF_ERR_SINGULAR, "zero pivot in matrix " + std::to_string(b));
    EF_API_END
}

int ef_mesh_stage_ms(ef_mesh mesh, double* ms) {
    EF_API_BEGIN
    if (!mesh || !ms) throw EfError(EF_ERR_ARG, "null argument");
    EF_HIP(hipStreamSynchronize(mesh->stream));
    for (int st = 0; st < 3; st++) {
        ms[st] = -1.0;
        if (!mesh->ev_recorded[st]) continue;
        float t = 0.0f;
        EF_HIP(hipEventElapsedTime(&t, mesh->ev[st][0], mesh->ev[st][1]));
        ms[st] = t;
    }
    EF_API_END
}

int ef_get_operator(ef_mesh mesh, int node, int which, int rhs, double* out, int* rows, int* cols) {
    EF_API_BEGIN
    if (!mesh || !rows || !cols) throw EfError(EF_ERR_ARG, "null argument");
    ef_mesh_s* M = mesh;
    Plan& P = M->P;
    if (node < 0 || node >= (int)P.nodes.size()) throw EfError(EF_ERR_ARG, "node index out of range");
    const NodeP& nd = P.nodes[node];
    EF_HIP(hipSetDevice(M->ctx->device));
    EF_HIP(hipStreamSynchronize(M->stream));
    const int n = nd.n;
    auto copy_mat = [&](const double* src, int r, int c) {
        *rows = r; *cols = c;
        if (out) {
            efk::copy_d2h(out, src, sizeof(double) * r * c, M->stream);
            EF_HIP(hipStreamSynchronize(M->stream));
        }
    };
    auto need_built = [&]() { if (!M->built) throw EfError(EF_ERR_STATE, "operators exist after the build stage"); };
    auto need_internal = [&]() { if (nd.leaf) throw EfError(EF_ERR_ARG, "leaf nodes carry no S, X, H or w"); };
    auto copy_vec = [&](long long off, int len) {
        *rows = len; *cols = 1;
        if (!out) return;
        if (M->kcap == 0) throw EfError(EF_ERR_STATE, "no vectors yet");
        if (rhs < 0 || rhs >= M->kcap) throw EfError(EF_ERR_ARG, "rhs out of range");
        efk::copy_d2h(out, M->vec.p + (long long)rhs * P.vec_rows + off, sizeof(double) * len, M->stream);
    };
    switch (which) {
        case 0: {  // T (coarsened if the node was tagged, as coarsen_ leaves it)
            need_built();
            const int m = 4 * (nd.s >> nd.tag);
            copy_mat(merge_T_ptr(M, node), m, m);
            break;
        }
        case 1: need_built(); need_internal(); copy_mat(M->Sar.p + nd.S_off, 4 * n, 8 * n); break;
        case 2: {  // X: re-assembled from the children's stored DtN (mergeX_)
            need_built(); need_internal();
            DevBuf<double> tX, tT;
            tX.alloc(16LL * n * n);
            tT.alloc(64LL * n * n);
            AsmDesc ad{};
            ad.n = n;
            for (int c = 0; c < 4; c++) ad.Tc[c] = merge_T_ptr(M, nd.child[c]);
            ad.X = tX.p; ad.T = tT.p;
            DevBuf<AsmDesc> dad;
            dad.upload(std::vector<AsmDesc>{ad}, M->stream);
            EF_LAUNCH("tree_assemble_kernel", tree_assemble_kernel, dim3(blocks_for(80LL * n * n), 1), dim3(256), M->stream,
                      dad.p);
            EF_HIP(hipStreamSynchronize(M->stream));
            copy_mat(tX.p, 4 * n, 4 * n);
            break;
        }
        case 3: {  // H (8n x 4n of mergeT_): child c's tau-side rows against its interface sides
            need_built(); need_internal();
            *rows = 8 * n; *cols = 4 * n;
            if (out) {
                std::fill(out, out + 32LL * n * n, 0.0);
                std::vector<double> Tc(16LL * n * n);
                for (int c = 0; c < 4; c++) {
                    efk::copy_d2h(Tc.data(), merge_T_ptr(M, nd.child[c]), sizeof(double) * Tc.size(), M->stream);
                    // tau sides of child c in the reference's row order: (W or E), then (S or N)
                    const int ts[2] = {(c & 1) ? (int)E : (int)W, (c >> 1) ? (int)N : (int)S};
                    for (int rr = 0; rr < 2 * n; rr++)
                        for (int slot = 0; slot < 2; slot++) {
                            const int r = child_if(c, slot), side = side_on_if(c, r);
                            for (int j = 0; j < n; j++)
                                out[(long long)(c * 2 * n + rr) * 4 * n + (long long)r * n + j] =
                                    Tc[(long long)(ts[rr / n] * n + rr % n) * 4 * n + (long long)side * n + j];
                        }
                }
            }
            break;
        }
        case 5: {  // g, side by side from where the downward sweep reads it
            const int len = nd.s;
            double* o = out;
            for (int sd = 0; sd < 4; sd++) {
                copy_vec(nd.gsrc[sd], len);
                if (out) out += len;
            }
            out = o;
            *rows = 4 * len;
            break;
        }
        case 7: {  // h, coarsened in place by coarsenUpwards_ for tagged nodes
            const int len = 4 * nd.s;
            if (!out) { *rows = len >> nd.tag; *cols = 1; break; }
            std::vector<double> full(len);
            copy_vec(nd.h_off, len);
            std::copy(out, out + len, full.begin());
            int cur = len;
            for (int t = 0; t < nd.tag; t++) {
                const int side = cur / 4;
                std::vector<double> nx(cur / 2);
                for (int sd = 0; sd < 4; sd++)
                    for (int i = 0; i < side / 2; i++)
                        nx[sd * side / 2 + i] = 0.5 * full[sd * side + 2 * i] + 0.5 * full[sd * side + 2 * i + 1];
                full = nx;
                cur /= 2;
            }
            std::copy(full.begin(), full.end(), out);
            *rows = cur; *cols = 1;
            break;
        }
        case 8: need_internal(); copy_vec(nd.w_off, 4 * n); break;
        case 4: case 6: throw EfError(EF_ERR_ARG, "leaf u and f are returned by ef_solve / given to ef_upwards");
        default: throw EfError(EF_ERR_ARG, "unknown operator code");
    }
    EF_API_END
}

}  // extern "C"
