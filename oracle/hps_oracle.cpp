// hps_oracle.cpp -- CPU FP64 restatement of EllipticForest's HPS factor-and-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the MI355X library is
// compared against (parity tests, smoke check, and the timed CPU baseline of the
// benchmark).  Nothing in the product links, loads or calls it.
//
// It restates, step for step and in plain C++17 (no MPI, p4est, PETSc, FISHPACK or
// LAPACK), the reference algorithm:
//   * FiniteVolumeGrid            src/Patches/FiniteVolume/FiniteVolumeGrid.cpp:9-36
//   * FiniteVolumeSolver::solve   src/Patches/FiniteVolume/FiniteVolumeSolver.cpp:25-294
//       - FivePointStencil: the dense 5-point system of :64-184 (ghost = 2g - u),
//         solved exactly (band LU with partial pivoting == dense getrf pivots)
//       - FISHPACK90 (:224-290): the same staggered 5-point Helmholtz system with
//         alpha = beta = 1 and lambda = lambda(0,0)  (assumption: hstcrt_ with
//         MBDCND = NBDCND = 1 is an exact direct solver of that system)
//   * mapD2N :296-353, buildD2N :355-455 (live loop :446-451), particularNeumannData :457-462
//   * HPSAlgorithm::buildStage/upwardsStage/solveStage  src/HPSAlgorithm.hpp:120-445
//     with every family step: coarsen_ :707, createIndexSets_ :752,
//     createMatrixBlocks_ :799, mergeX_ :870, mergeS_ :906, mergeT_ :940,
//     reorderOperators_ :979, mergePatch_ :1004, coarsenUpwards_ :1024,
//     mergeW_ :1059 (re-factorises X on every upwards pass), mergeH_ :1098,
//     reorderOperatorsUpwards_ :1128, uncoarsen_ :1165, applyS_ :1194
//   * L21 / L12 interpolation     src/SpecialMatrices.hpp:93-173
//   * Matrix::blockPermute        src/Matrix.hpp:430-472, Vector::blockPermute Vector.hpp:432-460
//   * dgesv / dgemm / dgemv       src/Matrix.hpp:789-953 (own kernels; partial pivoting)
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -shared).  C entry points at the end.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Dense row-major matrix (Matrix.hpp:711-725 stores A(i,j) at i*ncols + j).
// ---------------------------------------------------------------------------
struct Mat {
    int r = 0, c = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(int rows, int cols, double fill = 0.0) : r(rows), c(cols), v((size_t)rows * cols, fill) {}
    double& operator()(int i, int j) { return v[(size_t)i * c + j]; }
    double operator()(int i, int j) const { return v[(size_t)i * c + j]; }
};
using Vec = std::vector<double>;

static int g_threads = 1;

// C = A * B   (dgemm wrapper semantics, Matrix.hpp:827-856)
static Mat matmul(const Mat& A, const Mat& B) {
    if (A.c != B.r) throw std::invalid_argument("matmul: inner dimensions differ");
    Mat C(A.r, B.c);
    const int M = A.r, K = A.c, N = B.c;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (M * (long)N * K > 200000)
    for (int i = 0; i < M; i++) {
        double* ci = &C.v[(size_t)i * N];
        const double* ai = &A.v[(size_t)i * K];
        for (int k = 0; k < K; k++) {
            const double a = ai[k];
            if (a == 0.0) continue;
            const double* bk = &B.v[(size_t)k * N];
            for (int j = 0; j < N; j++) ci[j] += a * bk[j];
        }
    }
    return C;
}

// y = A * x  (dgemv wrapper, Matrix.hpp:789-816)
static Vec matvec(const Mat& A, const Vec& x) {
    if ((int)x.size() != A.c) throw std::invalid_argument("matvec: size mismatch");
    Vec y(A.r, 0.0);
    for (int i = 0; i < A.r; i++) {
        double s = 0.0;
        const double* ai = &A.v[(size_t)i * A.c];
        for (int j = 0; j < A.c; j++) s += ai[j] * x[j];
        y[i] = s;
    }
    return y;
}

// Solve A X = B in place of B with LU and partial pivoting (dgesv semantics,
// Matrix.hpp:866-953: pivot = first row of largest magnitude in the column).
// B is row-major n x nrhs.  Returns LAPACK-style info (0 ok, k>0 zero pivot).
static int lu_solve_inplace(int n, std::vector<double> A, int nrhs, double* B) {
    std::vector<int> piv(n);
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = std::fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = std::fabs(A[(size_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (best == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
        }
        const double inv = 1.0 / A[(size_t)k * n + k];
        const double* ak = &A[(size_t)k * n];
#pragma omp parallel for schedule(static) num_threads(g_threads) if ((n - k) * (long)(n - k) > 40000)
        for (int i = k + 1; i < n; i++) {
            double* ai = &A[(size_t)i * n];
            const double l = ai[k] * inv;
            ai[k] = l;
            if (l == 0.0) continue;
            for (int j = k + 1; j < n; j++) ai[j] -= l * ak[j];
        }
    }
    // Row interchanges on B, then forward (unit L) and backward (U) substitution.
    for (int k = 0; k < n; k++) {
        if (piv[k] != k) {
            for (int j = 0; j < nrhs; j++) std::swap(B[(size_t)k * nrhs + j], B[(size_t)piv[k] * nrhs + j]);
        }
    }
    // Columns of B are independent: split them over threads.
    const int nchunk = std::max(1, std::min(g_threads, nrhs));
#pragma omp parallel for schedule(static) num_threads(g_threads) if (nchunk > 1)
    for (int t = 0; t < nchunk; t++) {
        const int j0 = (int)((long)nrhs * t / nchunk), j1 = (int)((long)nrhs * (t + 1) / nchunk);
        for (int i = 0; i < n; i++) {
            double* bi = &B[(size_t)i * nrhs];
            const double* ai = &A[(size_t)i * n];
            for (int k = 0; k < i; k++) {
                const double l = ai[k];
                if (l == 0.0) continue;
                const double* bk = &B[(size_t)k * nrhs];
                for (int j = j0; j < j1; j++) bi[j] -= l * bk[j];
            }
        }
        for (int i = n - 1; i >= 0; i--) {
            double* bi = &B[(size_t)i * nrhs];
            const double* ai = &A[(size_t)i * n];
            for (int k = i + 1; k < n; k++) {
                const double u = ai[k];
                if (u == 0.0) continue;
                const double* bk = &B[(size_t)k * nrhs];
                for (int j = j0; j < j1; j++) bi[j] -= u * bk[j];
            }
            const double inv = 1.0 / ai[i];
            for (int j = j0; j < j1; j++) bi[j] *= inv;
        }
    }
    return info;
}

static Mat solve(const Mat& A, const Mat& B) {
    if (A.r != A.c || A.r != B.r) throw std::invalid_argument("solve: shape mismatch");
    Mat X = B;
    lu_solve_inplace(A.r, A.v, B.c, X.v.data());
    return X;
}
static Vec solve(const Mat& A, const Vec& b) {
    if (A.r != A.c || A.r != (int)b.size()) throw std::invalid_argument("solve: shape mismatch");
    Vec x = b;
    lu_solve_inplace(A.r, A.v, 1, x.data());
    return x;
}

// Matrix::getFromIndexSet (Matrix.hpp:316-350)
static Mat gather(const Mat& A, const std::vector<int>& I, const std::vector<int>& J) {
    Mat R((int)I.size(), (int)J.size());
    for (size_t i = 0; i < I.size(); i++)
        for (size_t j = 0; j < J.size(); j++) R((int)i, (int)j) = A(I[i], J[j]);
    return R;
}
static Vec gather(const Vec& v, const std::vector<int>& I) {
    Vec r(I.size());
    for (size_t i = 0; i < I.size(); i++) r[i] = v[I[i]];
    return r;
}
static Mat negate(Mat A) { for (auto& x : A.v) x = -x; return A; }
static Mat sub(const Mat& A, const Mat& B) { Mat C = A; for (size_t i = 0; i < C.v.size(); i++) C.v[i] -= B.v[i]; return C; }
static void add_inplace(Mat& A, const Mat& B) { for (size_t i = 0; i < A.v.size(); i++) A.v[i] += B.v[i]; }
static void set_block(Mat& A, int r0, int c0, const Mat& B) {
    for (int i = 0; i < B.r; i++)
        for (int j = 0; j < B.c; j++) A(r0 + i, c0 + j) = B(i, j);
}
static Mat block_diag(const std::vector<Mat>& D) {
    int R = 0, C = 0;
    for (auto& d : D) { R += d.r; C += d.c; }
    Mat A(R, C);
    int r0 = 0, c0 = 0;
    for (auto& d : D) { set_block(A, r0, c0, d); r0 += d.r; c0 += d.c; }
    return A;
}
static std::vector<int> range(int a, int b) {  // [a, b)
    std::vector<int> r;
    for (int i = a; i < b; i++) r.push_back(i);
    return r;
}
static std::vector<int> cat(const std::vector<int>& a, const std::vector<int>& b) {
    std::vector<int> r = a;
    r.insert(r.end(), b.begin(), b.end());
    return r;
}

// Matrix::blockPermute (Matrix.hpp:430-472): new block k = old block pi[k].
static std::vector<int> block_index(const std::vector<int>& pi, const std::vector<int>& sizes) {
    std::vector<int> idx;
    for (int p : pi) {
        int start = 0;
        for (int q = 0; q < p; q++) start += sizes[q];
        for (int t = 0; t < sizes[p]; t++) idx.push_back(start + t);
    }
    return idx;
}
static Mat block_permute(const Mat& A, const std::vector<int>& pr, const std::vector<int>& pc,
                         const std::vector<int>& R, const std::vector<int>& C) {
    int rs = 0, cs = 0;
    for (int x : R) rs += x;
    for (int x : C) cs += x;
    if (rs != A.r || cs != A.c) throw std::invalid_argument("blockPermute: block sizes do not add up");
    return gather(A, block_index(pr, R), block_index(pc, C));
}
static Vec block_permute(const Vec& v, const std::vector<int>& pi, const std::vector<int>& S) {
    int s = 0;
    for (int x : S) s += x;
    if (s != (int)v.size()) throw std::invalid_argument("blockPermute: block sizes do not add up");
    return gather(v, block_index(pi, S));
}

// InterpolationMatrixFine2Coarse(ncoarse): ncoarse x 2ncoarse, averaging (SpecialMatrices.hpp:93-116)
static Mat L21(int nc) {
    Mat L(nc, 2 * nc);
    for (int i = 0; i < nc; i++) { L(i, 2 * i) = 0.5; L(i, 2 * i + 1) = 0.5; }
    return L;
}
// InterpolationMatrixCoarse2Fine(nfine): nfine x nfine/2 (SpecialMatrices.hpp:136-173):
// linear 0.75/0.25 between coarse neighbours, quadratic extrapolation at both ends.
static Mat L12(int nf) {
    const int nc = nf / 2;
    Mat L(nf, nc);
    L(0, 0) = 1.40625; L(0, 1) = -0.5625; L(0, 2) = 0.15625;
    L(nf - 1, nc - 3) = 0.15625; L(nf - 1, nc - 2) = -0.5625; L(nf - 1, nc - 1) = 1.40625;
    for (int j = 0; 2 * j + 1 < nf - 1; j++) {
        const int i = 2 * j + 1;
        L(i, j) = 0.75; L(i, j + 1) = 0.25;
        L(i + 1, j) = 0.25; L(i + 1, j + 1) = 0.75;
    }
    return L;
}
static Mat block_diag4(const Mat& B) { return block_diag({B, B, B, B}); }

// ---------------------------------------------------------------------------
// FiniteVolumeGrid (FiniteVolumeGrid.cpp:9-36): cell centred, dx = (xu - xl)/nx.
// ---------------------------------------------------------------------------
struct Grid {
    int nx = 0, ny = 0;
    double xl = 0, xu = 0, yl = 0, yu = 0, dx = 0, dy = 0;
    Grid() = default;
    Grid(int nx_, double xl_, double xu_, int ny_, double yl_, double yu_)
        : nx(nx_), ny(ny_), xl(xl_), xu(xu_), yl(yl_), yu(yu_), dx((xu_ - xl_) / nx_), dy((yu_ - yl_) / ny_) {}
    double x(int i) const { return (xl + dx / 2) + i * dx; }
    double y(int j) const { return (yl + dy / 2) + j * dy; }
};

// Coefficients of one leaf, already sampled the way FiniteVolumeSolver.cpp:64-78 samples
// them: alpha(x_i,y_j), beta at the four faces of cell (i,j), lambda(x_i,y_j).
// Arrays are indexed like a leaf cell vector: j + i*ny.
struct LeafCoef {
    int fishpack = 1;        // 1: FISHPACK90 semantics (alpha=beta=1, lambda = lambda0)
    double lambda0 = 0.0;
    Vec alpha, bW, bE, bS, bN, lam;
};

// Band LU with partial pivoting of the leaf 5-point matrix.  Row I = j + i*ny has
// neighbours I +- 1 (y) and I +- ny (x), so kl = ku = ny; storage keeps the
// kl extra super-diagonals that row exchanges can fill.  The pivot sequence equals
// that of a dense getrf on the same matrix, because column k is zero below row k+kl.
struct BandLU {
    int n = 0, kl = 0, ku = 0, w = 0;
    std::vector<double> ab;
    std::vector<int> piv;
    double& at(int i, int j) { return ab[(size_t)i * w + (j - i + kl)]; }
    double at(int i, int j) const { return ab[(size_t)i * w + (j - i + kl)]; }
    void init(int n_, int kl_, int ku_) {
        n = n_; kl = kl_; ku = ku_; w = 2 * kl + ku + 1;
        ab.assign((size_t)n * w, 0.0);
        piv.assign(n, 0);
    }
    void factor() {
        for (int k = 0; k < n; k++) {
            const int iend = std::min(n - 1, k + kl);
            int p = k;
            double best = std::fabs(at(k, k));
            for (int i = k + 1; i <= iend; i++) {
                double v = std::fabs(at(i, k));
                if (v > best) { best = v; p = i; }
            }
            piv[k] = p;
            if (best == 0.0) throw std::runtime_error("leaf 5-point matrix is singular");
            const int jend = std::min(n - 1, k + ku + kl);
            if (p != k)
                for (int j = k; j <= jend; j++) std::swap(at(k, j), at(p, j));
            const double inv = 1.0 / at(k, k);
            for (int i = k + 1; i <= iend; i++) {
                const double l = at(i, k) * inv;
                at(i, k) = l;
                if (l == 0.0) continue;
                for (int j = k + 1; j <= jend; j++) at(i, j) -= l * at(k, j);
            }
        }
    }
    void solve(double* b) const {
        for (int k = 0; k < n; k++) {
            if (piv[k] != k) std::swap(b[k], b[piv[k]]);
            const int iend = std::min(n - 1, k + kl);
            for (int i = k + 1; i <= iend; i++) b[i] -= at(i, k) * b[k];
        }
        for (int i = n - 1; i >= 0; i--) {
            const int jend = std::min(n - 1, i + ku + kl);
            double s = b[i];
            for (int j = i + 1; j <= jend; j++) s -= at(i, j) * b[j];
            b[i] = s / at(i, i);
        }
    }
};

// Stencil coefficients of cell (i,j): FiniteVolumeSolver.cpp:64-85.
struct Stencil { double cW, cE, cS, cN, cC; };
static Stencil stencil(const Grid& g, const LeafCoef& k, int i, int j) {
    const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy;
    Stencil s;
    if (k.fishpack) {
        // hstcrt_ solves u_xx + u_yy + lambda u = f; alpha and beta are not used (:224-290).
        s.cW = 1.0 / dx2; s.cE = 1.0 / dx2; s.cS = 1.0 / dy2; s.cN = 1.0 / dy2;
        s.cC = -(2.0 / dx2 + 2.0 / dy2) + k.lambda0;
    } else {
        const int I = j + i * g.ny;
        const double a = k.alpha[I];
        s.cW = a * k.bW[I] / dx2; s.cE = a * k.bE[I] / dx2;
        s.cS = a * k.bS[I] / dy2; s.cN = a * k.bN[I] / dy2;
        s.cC = -a * ((k.bE[I] + k.bW[I]) / dx2 + (k.bN[I] + k.bS[I]) / dy2) + k.lam[I];
    }
    return s;
}

// Leaf patch solver: factorisation of the 5-point matrix of one leaf, plus the
// boundary-data terms of FiniteVolumeSolver.cpp:100-184 (ghost value 2g - u).
struct LeafSolver {
    Grid grid;
    LeafCoef coef;
    BandLU lu;
    std::vector<Stencil> st;
    void setup(const Grid& g, const LeafCoef& k) {
        grid = g; coef = k;
        const int nx = g.nx, ny = g.ny, n = nx * ny;
        st.resize(n);
        lu.init(n, ny, ny);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                const int I = j + i * ny;
                Stencil s = stencil(g, k, i, j);
                st[I] = s;
                double diag = s.cC;
                if (i > 0) lu.at(I, I - ny) = s.cW; else diag -= s.cW;
                if (i < nx - 1) lu.at(I, I + ny) = s.cE; else diag -= s.cE;
                if (j > 0) lu.at(I, I - 1) = s.cS; else diag -= s.cS;
                if (j < ny - 1) lu.at(I, I + 1) = s.cN; else diag -= s.cN;
                lu.at(I, I) = diag;
            }
        lu.factor();
    }
    // u = solve(grid, g, f) with g in WESN order (segments of length nx, as the reference uses nside = nx)
    Vec solve(const Vec& g, const Vec& f) const {
        const int nx = grid.nx, ny = grid.ny, ns = grid.nx;
        Vec b = f;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                const int I = j + i * ny;
                const Stencil& s = st[I];
                double add = 0.0;
                if (i == 0) add += s.cW * g[0 * ns + j];
                if (i == nx - 1) add += s.cE * g[1 * ns + j];
                if (j == 0) add += s.cS * g[2 * ns + i];
                if (j == ny - 1) add += s.cN * g[3 * ns + i];
                if (i == 0 || j == 0 || i == nx - 1 || j == ny - 1) b[I] += -2.0 * add;
            }
        lu.solve(b.data());
        return b;
    }
    // FiniteVolumeSolver::mapD2N (:296-353): h = +-(2/d)(u_edge - g), WESN.
    Vec mapD2N(const Vec& g, const Vec& f) const {
        const int ns = grid.nx;
        Vec u = solve(g, f);
        Vec h(4 * ns);
        const double tx = 2.0 / grid.dx, ty = 2.0 / grid.dy;
        for (int t = 0; t < ns; t++) {
            const double uW = u[t], uE = u[(ns - 1) * ns + t];
            const double uS = u[t * ns], uN = u[(t + 1) * ns - 1];
            h[0 * ns + t] = tx * (uW - g[0 * ns + t]);
            h[1 * ns + t] = -tx * (uE - g[1 * ns + t]);
            h[2 * ns + t] = ty * (uS - g[2 * ns + t]);
            h[3 * ns + t] = -ty * (uN - g[3 * ns + t]);
        }
        return h;
    }
    // FiniteVolumeSolver::buildD2N live loop (:446-451): column j = mapD2N(e_j, 0).
    Mat buildD2N() const {
        const int M = 4 * grid.nx;
        Mat T(M, M);
        Vec e(M, 0.0), fz((size_t)grid.nx * grid.ny, 0.0);
        for (int j = 0; j < M; j++) {
            e[j] = 1.0;
            Vec col = mapD2N(e, fz);
            for (int i = 0; i < M; i++) T(i, j) = col[i];
            e[j] = 0.0;
        }
        return T;
    }
};

// ---------------------------------------------------------------------------
// Quadtree: nodes in pre-order (root first, children 0..3 = SW, SE, NW, NE as
// FiniteVolumeNodeFactory.cpp:15-67 lays them out and p4est numbers siblings).
// ---------------------------------------------------------------------------
struct Node {
    int level = 0, ix = 0, iy = 0;
    int parent = -1, child[4] = {-1, -1, -1, -1};
    bool leaf = false;
    int leaf_index = -1;
    Grid grid;
    int n_coarsens = 0;
    Mat T, S, X, H;
    Vec u, g, f, h, w;
    int size() const { return grid.nx >> n_coarsens; }  // PatchBase::size (Patch.hpp:99-101)
};

struct Forest {
    int N = 0;
    int solver_type = 1;  // 0 five-point, 1 FISHPACK
    double lambda0 = 0.0;
    bool cache_operators = false, homogeneous_rhs = false;
    bool built = false;
    std::vector<Node> nodes;
    std::vector<int> leaves;      // node index per leaf, pre-order
    std::vector<int> postorder;   // node indices, children before parents
    std::vector<LeafCoef> coef;   // per leaf (five-point)

    int make(int parent, int level, int ix, int iy, const Grid& g) {
        Node nd;
        nd.level = level; nd.ix = ix; nd.iy = iy; nd.parent = parent; nd.grid = g;
        nodes.push_back(nd);
        return (int)nodes.size() - 1;
    }
};

// Pre-order construction from a pre-order leaf list: a node is a leaf iff the next
// unconsumed leaf has the same (level, ix, iy); otherwise it has four children.
// Child grids are quadrants of the parent with the same nx (FiniteVolumeNodeFactory.cpp:15-67).
static int build_node(Forest& F, int parent, int level, int ix, int iy, const Grid& g,
                      const int* L, const int* X, const int* Y, int nleaves, int& cursor) {
    const int id = F.make(parent, level, ix, iy, g);
    if (cursor >= nleaves) throw std::invalid_argument("leaf list ended early");
    if (L[cursor] == level) {
        if (X[cursor] != ix || Y[cursor] != iy) throw std::invalid_argument("leaf list is not a pre-order quadtree leaf list");
        F.nodes[id].leaf = true;
        F.nodes[id].leaf_index = (int)F.leaves.size();
        F.leaves.push_back(id);
        cursor++;
        return id;
    }
    if (L[cursor] < level) throw std::invalid_argument("leaf list is not a pre-order quadtree leaf list");
    const double xm = (g.xl + g.xu) / 2.0, ym = (g.yl + g.yu) / 2.0;
    const Grid cg[4] = {Grid(g.nx, g.xl, xm, g.ny, g.yl, ym), Grid(g.nx, xm, g.xu, g.ny, g.yl, ym),
                        Grid(g.nx, g.xl, xm, g.ny, ym, g.yu), Grid(g.nx, xm, g.xu, g.ny, ym, g.yu)};
    for (int c = 0; c < 4; c++) {
        int cid = build_node(F, id, level + 1, 2 * ix + (c & 1), 2 * iy + (c >> 1), cg[c], L, X, Y, nleaves, cursor);
        F.nodes[id].child[c] = cid;
    }
    return id;
}

static void make_postorder(Forest& F) {
    F.postorder.clear();
    std::vector<std::pair<int, int>> stack = {{0, 0}};
    while (!stack.empty()) {
        auto& top = stack.back();
        Node& nd = F.nodes[top.first];
        if (!nd.leaf && top.second < 4) {
            int c = nd.child[top.second++];
            stack.push_back({c, 0});
        } else {
            F.postorder.push_back(top.first);
            stack.pop_back();
        }
    }
}

static LeafSolver leaf_solver(const Forest& F, const Node& nd) {
    LeafSolver s;
    LeafCoef k;
    if (F.solver_type == 1) { k.fishpack = 1; k.lambda0 = F.lambda0; }
    else k = F.coef[nd.leaf_index];
    s.setup(nd.grid, k);
    return s;
}

// ----- merge4to1 steps (HPSAlgorithm.hpp:497-518, :676-1009) -----
struct Blocks {
    int n = 0;
    std::vector<int> ab, ag, at, ba, bo, bt, ga, go, gt, ob, og, ot;  // index sets
};

// createIndexSets_ (:752-788)
static Blocks index_sets(Node* c[4]) {
    const int s0 = c[0]->size();
    for (int i = 1; i < 4; i++)
        if (c[i]->size() != s0)
            throw std::invalid_argument("Size of children patches are not the same; something probably went wrong with the coarsening...");
    Blocks b;
    const int n = c[0]->T.r / 4;
    b.n = n;
    auto W = range(0, n), E = range(n, 2 * n), S = range(2 * n, 3 * n), N = range(3 * n, 4 * n);
    b.ab = E; b.ag = N; b.at = cat(W, S);
    b.ba = W; b.bo = N; b.bt = cat(E, S);
    b.ga = S; b.go = E; b.gt = cat(W, N);
    b.ob = S; b.og = W; b.ot = cat(E, N);
    return b;
}

// tagPatchesForCoarsening_ (:676-696) + coarsen_ (:707-741)
static void coarsen(Node* c[4]) {
    int smin = c[0]->size();
    for (int i = 1; i < 4; i++) smin = std::min(smin, c[i]->size());
    for (int i = 0; i < 4; i++) {
        const int tag = (int)std::log2((double)c[i]->size()) - (int)std::log2((double)smin);
        for (int n = 0; n < tag; n++) {
            const int nfine = c[i]->grid.nx >> n;
            const int ncoarse = nfine / 2;
            Mat A = block_diag4(L21(ncoarse)), B = block_diag4(L12(nfine));
            c[i]->T = matmul(A, c[i]->T);
            c[i]->T = matmul(c[i]->T, B);
            c[i]->n_coarsens++;
        }
    }
}

// mergeX_ (:870-895) from the blocks of createMatrixBlocks_ (:799-859)
static Mat merge_X(Node* c[4], const Blocks& b) {
    const Mat &Ta = c[0]->T, &Tb = c[1]->T, &Tg = c[2]->T, &To = c[3]->T;
    Mat T_ag_ag = gather(Ta, b.ag, b.ag), T_ga_ga = gather(Tg, b.ga, b.ga);
    Mat T_ag_ab = gather(Ta, b.ag, b.ab), T_ga_go = negate(gather(Tg, b.ga, b.go));
    Mat T_bo_bo = gather(Tb, b.bo, b.bo), T_ob_ob = gather(To, b.ob, b.ob);
    Mat T_bo_ba = gather(Tb, b.bo, b.ba), T_ob_og = negate(gather(To, b.ob, b.og));
    Mat T_ab_ag = gather(Ta, b.ab, b.ag), T_ba_bo = negate(gather(Tb, b.ba, b.bo));
    Mat T_ab_ab = gather(Ta, b.ab, b.ab), T_ba_ba = gather(Tb, b.ba, b.ba);
    Mat T_go_ga = gather(Tg, b.go, b.ga), T_og_ob = negate(gather(To, b.og, b.ob));
    Mat T_go_go = gather(Tg, b.go, b.go), T_og_og = gather(To, b.og, b.og);
    const int n = b.n;
    Mat X = block_diag({sub(T_ag_ag, T_ga_ga), sub(T_bo_bo, T_ob_ob), sub(T_ab_ab, T_ba_ba), sub(T_go_go, T_og_og)});
    set_block(X, 0 * n, 2 * n, T_ag_ab);
    set_block(X, 0 * n, 3 * n, T_ga_go);
    set_block(X, 1 * n, 2 * n, T_bo_ba);
    set_block(X, 1 * n, 3 * n, T_ob_og);
    set_block(X, 2 * n, 0 * n, T_ab_ag);
    set_block(X, 2 * n, 1 * n, T_ba_bo);
    set_block(X, 3 * n, 0 * n, T_go_ga);
    set_block(X, 3 * n, 1 * n, T_og_ob);
    return X;
}

static const std::vector<int> PI_WESN = {0, 4, 2, 6, 1, 3, 5, 7};

static void merge4to1(Node& tau, Node* c[4]) {
    coarsen(c);
    Blocks b = index_sets(c);
    const int n = b.n;
    const Mat &Ta = c[0]->T, &Tb = c[1]->T, &Tg = c[2]->T, &To = c[3]->T;
    tau.X = merge_X(c, b);
    // mergeS_ (:906-929)
    Mat R(4 * n, 8 * n);
    set_block(R, 0 * n, 0 * n, negate(gather(Ta, b.ag, b.at)));
    set_block(R, 0 * n, 4 * n, gather(Tg, b.ga, b.gt));
    set_block(R, 1 * n, 2 * n, negate(gather(Tb, b.bo, b.bt)));
    set_block(R, 1 * n, 6 * n, gather(To, b.ob, b.ot));
    set_block(R, 2 * n, 0 * n, negate(gather(Ta, b.ab, b.at)));
    set_block(R, 2 * n, 2 * n, gather(Tb, b.ba, b.bt));
    set_block(R, 3 * n, 4 * n, negate(gather(Tg, b.go, b.gt)));
    set_block(R, 3 * n, 6 * n, gather(To, b.og, b.ot));
    tau.S = solve(tau.X, R);
    // mergeT_ (:940-968)
    Mat TL = block_diag({gather(Ta, b.at, b.at), gather(Tb, b.bt, b.bt), gather(Tg, b.gt, b.gt), gather(To, b.ot, b.ot)});
    Mat H(8 * n, 4 * n);
    set_block(H, 0 * n, 0 * n, gather(Ta, b.at, b.ag));
    set_block(H, 0 * n, 2 * n, gather(Ta, b.at, b.ab));
    set_block(H, 2 * n, 1 * n, gather(Tb, b.bt, b.bo));
    set_block(H, 2 * n, 2 * n, gather(Tb, b.bt, b.ba));
    set_block(H, 4 * n, 0 * n, gather(Tg, b.gt, b.ga));
    set_block(H, 4 * n, 3 * n, gather(Tg, b.gt, b.go));
    set_block(H, 6 * n, 1 * n, gather(To, b.ot, b.ob));
    set_block(H, 6 * n, 3 * n, gather(To, b.ot, b.og));
    tau.H = H;
    Mat T = matmul(H, tau.S);
    add_inplace(T, TL);
    // reorderOperators_ (:979-993)
    std::vector<int> s4(4, n), s8(8, n);
    tau.S = block_permute(tau.S, {0, 1, 2, 3}, PI_WESN, s4, s8);
    tau.T = block_permute(T, PI_WESN, PI_WESN, s8, s8);
    // mergePatch_ (:1004-1009)
    tau.grid = Grid(c[0]->size() + c[1]->size(), c[0]->grid.xl, c[1]->grid.xu,
                    c[0]->size() + c[2]->size(), c[0]->grid.yl, c[2]->grid.yu);
}

static void upwards4to1(Forest& F, Node& tau, Node* c[4]) {
    if (F.homogeneous_rhs) return;  // :532
    // coarsenUpwards_ (:1024-1048)
    for (int i = 0; i < 4; i++)
        for (int n = 0; n < c[i]->n_coarsens; n++) {
            const int nfine = c[i]->grid.nx >> n;
            c[i]->h = matvec(block_diag4(L21(nfine / 2)), c[i]->h);
        }
    Blocks b = index_sets(c);
    const int n = b.n;
    Mat X = merge_X(c, b);  // rebuilt, as :544 does
    // mergeW_ (:1059-1087)
    Vec hd(4 * n);
    auto diff = [&](int off, const Vec& a, const std::vector<int>& Ia, const Vec& bb, const std::vector<int>& Ib) {
        for (int t = 0; t < n; t++) hd[off + t] = a[Ia[t]] - bb[Ib[t]];
    };
    diff(0 * n, c[2]->h, b.ga, c[0]->h, b.ag);
    diff(1 * n, c[3]->h, b.ob, c[1]->h, b.bo);
    diff(2 * n, c[1]->h, b.ba, c[0]->h, b.ab);
    diff(3 * n, c[3]->h, b.og, c[2]->h, b.go);
    tau.w = solve(X, hd);
    // mergeH_ (:1098-1117)
    Vec h = matvec(tau.H, tau.w);
    Vec upd = gather(c[0]->h, b.at);
    Vec t1 = gather(c[1]->h, b.bt), t2 = gather(c[2]->h, b.gt), t3 = gather(c[3]->h, b.ot);
    upd.insert(upd.end(), t1.begin(), t1.end());
    upd.insert(upd.end(), t2.begin(), t2.end());
    upd.insert(upd.end(), t3.begin(), t3.end());
    for (size_t i = 0; i < h.size(); i++) h[i] += upd[i];
    // reorderOperatorsUpwards_ (:1128-1150)
    tau.h = block_permute(h, PI_WESN, std::vector<int>(8, n));
}

static void split1to4(Forest& F, Node& tau, Node* c[4]) {
    // uncoarsen_ (:1165-1183)
    for (int n = 0; n < tau.n_coarsens; n++) {
        const int nfine = tau.grid.nx >> (tau.n_coarsens - (n + 1));
        tau.g = matvec(block_diag4(L12(nfine)), tau.g);
    }
    // applyS_ (:1194-1230)
    Vec ui = matvec(tau.S, tau.g);
    if (!F.homogeneous_rhs)
        for (size_t i = 0; i < ui.size(); i++) ui[i] += tau.w[i];
    const int n = c[0]->size();
    auto seg = [&](const Vec& v, int k) { return Vec(v.begin() + k * n, v.begin() + (k + 1) * n); };
    Vec g_ag = seg(ui, 0), g_bo = seg(ui, 1), g_ab = seg(ui, 2), g_go = seg(ui, 3);
    Vec aW = seg(tau.g, 0), gW = seg(tau.g, 1), bE = seg(tau.g, 2), oE = seg(tau.g, 3);
    Vec aS = seg(tau.g, 4), bS = seg(tau.g, 5), gN = seg(tau.g, 6), oN = seg(tau.g, 7);
    auto join = [](std::initializer_list<const Vec*> parts) {
        Vec r;
        for (auto* p : parts) r.insert(r.end(), p->begin(), p->end());
        return r;
    };
    c[0]->g = join({&aW, &g_ab, &aS, &g_ag});
    c[1]->g = join({&g_ab, &bE, &bS, &g_bo});
    c[2]->g = join({&gW, &g_go, &g_ag, &gN});
    c[3]->g = join({&g_go, &oE, &g_bo, &oN});
}

static void children(Forest& F, Node& nd, Node* c[4]) {
    for (int i = 0; i < 4; i++) c[i] = &F.nodes[nd.child[i]];
}

// buildStage (:120-161)
static void build(Forest& F) {
    Mat cached;
    bool have_cached = false;
    for (int id : F.postorder) {
        Node& nd = F.nodes[id];
        if (nd.leaf) {
            if (F.cache_operators) {
                if (!have_cached) { cached = leaf_solver(F, nd).buildD2N(); have_cached = true; }
                nd.T = cached;
            } else {
                nd.T = leaf_solver(F, nd).buildD2N();
            }
        } else {
            Node* c[4];
            children(F, nd, c);
            merge4to1(nd, c);
        }
    }
    F.built = true;
}

// upwardsStage (:225-272) with f given per leaf (pre-order leaf order, j + i*ny)
static void upwards(Forest& F, const double* f_all) {
    const size_t nc = (size_t)F.N * F.N;
    for (int id : F.postorder) {
        Node& nd = F.nodes[id];
        if (nd.leaf) {
            nd.f.assign(f_all + nd.leaf_index * nc, f_all + (nd.leaf_index + 1) * nc);
            nd.h = leaf_solver(F, nd).mapD2N(Vec(4 * nd.grid.nx, 0.0), nd.f);  // particularNeumannData
        } else {
            Node* c[4];
            children(F, nd, c);
            upwards4to1(F, nd, c);
        }
    }
}

// solveStage (:343-445): root boundary data a u + b du/dn = r at root face points
static void solve_stage(Forest& F, const double* r, const double* a, const double* b, double* u_all) {
    Node& root = F.nodes[0];
    const int M = root.grid.nx;
    Mat A(4 * M, 4 * M), B(4 * M, 4 * M);
    Vec rhs(4 * M);
    for (int i = 0; i < 4 * M; i++) { A(i, i) = a[i]; B(i, i) = b[i]; }
    Vec h = root.h;
    if (h.empty()) h.assign(4 * M, 0.0);
    Vec bh = matvec(B, h);
    for (int i = 0; i < 4 * M; i++) rhs[i] = r[i] - bh[i];
    Mat BT = matmul(B, root.T);
    add_inplace(A, BT);
    root.g = solve(A, rhs);
    // pre-order split then leaf solves (Quadtree.hpp:345-360)
    std::vector<int> stack = {0};
    const size_t nc = (size_t)F.N * F.N;
    while (!stack.empty()) {
        int id = stack.back();
        stack.pop_back();
        Node& nd = F.nodes[id];
        if (nd.leaf) {
            if (F.homogeneous_rhs) nd.f.assign(nc, 0.0);  // leafSolve (:584-596)
            nd.u = leaf_solver(F, nd).solve(nd.g, nd.f);
            std::copy(nd.u.begin(), nd.u.end(), u_all + nd.leaf_index * nc);
        } else {
            Node* c[4];
            children(F, nd, c);
            split1to4(F, nd, c);
            for (int i = 3; i >= 0; i--) stack.push_back(nd.child[i]);
        }
    }
}

}  // namespace orc

// ===========================================================================
// C entry points (ctypes).  Status: 0 ok, nonzero error; message via orc_last_error.
// ===========================================================================
using namespace orc;
static thread_local std::string g_err;

#define ORC_TRY(...)                                   \
    try { __VA_ARGS__; return 0; }                     \
    catch (const std::exception& e) { g_err = e.what(); return 1; }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_set_threads(int t) { g_threads = std::max(1, t); }

// Forest over leaves given in pre-order as (level, ix, iy); N cells per leaf side.
int orc_create(int nleaves, const int* level, const int* ix, const int* iy, int N, double xl, double xu,
               double yl, double yu, int solver_type, void** out) {
    ORC_TRY({
        auto* F = new Forest();
        F->N = N;
        F->solver_type = solver_type;
        int cursor = 0;
        try {
            build_node(*F, -1, 0, 0, 0, Grid(N, xl, xu, N, yl, yu), level, ix, iy, nleaves, cursor);
            if (cursor != nleaves) throw std::invalid_argument("leaf list has extra entries");
        } catch (...) { delete F; throw; }
        make_postorder(*F);
        F->coef.resize(F->leaves.size());
        *out = F;
    })
}
void orc_destroy(void* f) { delete (Forest*)f; }
int orc_num_nodes(void* f) { return (int)((Forest*)f)->nodes.size(); }
int orc_num_leaves(void* f) { return (int)((Forest*)f)->leaves.size(); }

int orc_set_options(void* f, int cache_operators, int homogeneous_rhs, double lambda0) {
    auto* F = (Forest*)f;
    F->cache_operators = cache_operators != 0;
    F->homogeneous_rhs = homogeneous_rhs != 0;
    F->lambda0 = lambda0;
    return 0;
}

// Five-point coefficients of one leaf: alpha, beta at W/E/S/N faces, lambda (N*N each).
int orc_set_leaf_coeffs(void* f, int leaf, const double* alpha, const double* bW, const double* bE,
                        const double* bS, const double* bN, const double* lam) {
    ORC_TRY({
        auto* F = (Forest*)f;
        if (leaf < 0 || leaf >= (int)F->leaves.size()) throw std::out_of_range("leaf index");
        const size_t n = (size_t)F->N * F->N;
        LeafCoef& k = F->coef[leaf];
        k.fishpack = 0;
        k.alpha.assign(alpha, alpha + n); k.bW.assign(bW, bW + n); k.bE.assign(bE, bE + n);
        k.bS.assign(bS, bS + n); k.bN.assign(bN, bN + n); k.lam.assign(lam, lam + n);
    })
}

// info[0..11] = level, ix, iy, leaf, leaf_index, n_coarsens, nx, child0..3, parent
int orc_node_info(void* f, int node, int* info, double* bounds) {
    ORC_TRY({
        auto* F = (Forest*)f;
        const Node& nd = F->nodes.at(node);
        int v[12] = {nd.level, nd.ix, nd.iy, nd.leaf ? 1 : 0, nd.leaf_index, nd.n_coarsens, nd.grid.nx,
                     nd.child[0], nd.child[1], nd.child[2], nd.child[3], nd.parent};
        std::memcpy(info, v, sizeof(v));
        if (bounds) { bounds[0] = nd.grid.xl; bounds[1] = nd.grid.xu; bounds[2] = nd.grid.yl; bounds[3] = nd.grid.yu; }
    })
}

int orc_build(void* f) { ORC_TRY(build(*(Forest*)f)) }
int orc_upwards(void* f, const double* f_all) { ORC_TRY(upwards(*(Forest*)f, f_all)) }
int orc_solve(void* f, const double* r, const double* a, const double* b, double* u_all) {
    ORC_TRY(solve_stage(*(Forest*)f, r, a, b, u_all))
}
int orc_root_size(void* f) { return ((Forest*)f)->nodes[0].grid.nx; }

// which: 0 T, 1 S, 2 X, 3 H, 4 u, 5 g, 6 f, 7 h, 8 w.  With out == NULL only the shape is returned.
int orc_get(void* f, int node, int which, double* out, int* rows, int* cols) {
    ORC_TRY({
        auto* F = (Forest*)f;
        const Node& nd = F->nodes.at(node);
        const Mat* m = nullptr;
        const Vec* v = nullptr;
        switch (which) {
            case 0: m = &nd.T; break;
            case 1: m = &nd.S; break;
            case 2: m = &nd.X; break;
            case 3: m = &nd.H; break;
            case 4: v = &nd.u; break;
            case 5: v = &nd.g; break;
            case 6: v = &nd.f; break;
            case 7: v = &nd.h; break;
            case 8: v = &nd.w; break;
            default: throw std::invalid_argument("which");
        }
        if (m) { *rows = m->r; *cols = m->c; if (out) std::memcpy(out, m->v.data(), m->v.size() * sizeof(double)); }
        else { *rows = (int)v->size(); *cols = 1; if (out) std::memcpy(out, v->data(), v->size() * sizeof(double)); }
    })
}

// ----- building blocks, exposed for known-answer tests -----
// dgesv: A row-major n x n, B row-major n x nrhs overwritten by the solution.
int orc_dgesv(int n, int nrhs, const double* A, double* B) {
    ORC_TRY({
        std::vector<double> a(A, A + (size_t)n * n);
        if (lu_solve_inplace(n, a, nrhs, B) != 0) throw std::runtime_error("singular matrix");
    })
}
int orc_dgemm(int m, int k, int n, const double* A, const double* B, double* C) {
    ORC_TRY({
        Mat a(m, k), b(k, n);
        a.v.assign(A, A + (size_t)m * k);
        b.v.assign(B, B + (size_t)k * n);
        Mat c = matmul(a, b);
        std::memcpy(C, c.v.data(), c.v.size() * sizeof(double));
    })
}
int orc_block_permute(int nr, int nc, const double* A, int nbr, const int* pr, const int* R, int nbc,
                      const int* pc, const int* C, double* out) {
    ORC_TRY({
        Mat a(nr, nc);
        a.v.assign(A, A + (size_t)nr * nc);
        Mat b = block_permute(a, std::vector<int>(pr, pr + nbr), std::vector<int>(pc, pc + nbc),
                              std::vector<int>(R, R + nbr), std::vector<int>(C, C + nbc));
        std::memcpy(out, b.v.data(), b.v.size() * sizeof(double));
    })
}
int orc_interp(int kind, int n, double* out) {  // kind 0: L21(ncoarse=n), 1: L12(nfine=n)
    ORC_TRY({
        Mat L = kind == 0 ? L21(n) : L12(n);
        std::memcpy(out, L.v.data(), L.v.size() * sizeof(double));
    })
}

// Leaf-level operations on one patch.  coeffs == NULL selects FISHPACK semantics with lambda0;
// otherwise coeffs points at 6 consecutive N*N arrays: alpha, bW, bE, bS, bN, lambda.
static LeafSolver make_leaf(int N, double xl, double xu, double yl, double yu, const double* coeffs, double lambda0) {
    LeafSolver s;
    LeafCoef k;
    const size_t n = (size_t)N * N;
    if (!coeffs) { k.fishpack = 1; k.lambda0 = lambda0; }
    else {
        k.fishpack = 0;
        k.alpha.assign(coeffs, coeffs + n); k.bW.assign(coeffs + n, coeffs + 2 * n);
        k.bE.assign(coeffs + 2 * n, coeffs + 3 * n); k.bS.assign(coeffs + 3 * n, coeffs + 4 * n);
        k.bN.assign(coeffs + 4 * n, coeffs + 5 * n); k.lam.assign(coeffs + 5 * n, coeffs + 6 * n);
    }
    s.setup(Grid(N, xl, xu, N, yl, yu), k);
    return s;
}
int orc_leaf_solve(int N, double xl, double xu, double yl, double yu, const double* coeffs, double lambda0,
                   const double* g, const double* f, double* u) {
    ORC_TRY({
        LeafSolver s = make_leaf(N, xl, xu, yl, yu, coeffs, lambda0);
        Vec r = s.solve(Vec(g, g + 4 * N), Vec(f, f + (size_t)N * N));
        std::memcpy(u, r.data(), r.size() * sizeof(double));
    })
}
int orc_leaf_map_d2n(int N, double xl, double xu, double yl, double yu, const double* coeffs, double lambda0,
                     const double* g, const double* f, double* h) {
    ORC_TRY({
        LeafSolver s = make_leaf(N, xl, xu, yl, yu, coeffs, lambda0);
        Vec r = s.mapD2N(Vec(g, g + 4 * N), Vec(f, f + (size_t)N * N));
        std::memcpy(h, r.data(), r.size() * sizeof(double));
    })
}
int orc_leaf_dtn(int N, double xl, double xu, double yl, double yu, const double* coeffs, double lambda0, double* T) {
    ORC_TRY({
        LeafSolver s = make_leaf(N, xl, xu, yl, yu, coeffs, lambda0);
        Mat t = s.buildD2N();
        std::memcpy(T, t.v.data(), t.v.size() * sizeof(double));
    })
}

}  // extern "C"
