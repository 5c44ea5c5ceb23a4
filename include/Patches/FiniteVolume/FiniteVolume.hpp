// FiniteVolume.hpp -- drop-in FiniteVolumeGrid / FiniteVolumeSolver / FiniteVolumePatch /
// FiniteVolumeNodeFactory (src/Patches/FiniteVolume/*) over the MI355X C ABI.
#ifndef EF_FINITE_VOLUME_HPP_
#define EF_FINITE_VOLUME_HPP_

#include <cmath>
#include <functional>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../EllipticForest.hpp"

namespace EllipticForest {

enum DimensionIndex { X = 0, Y = 1 };
#ifndef XDIM
#define XDIM 0
#define YDIM 1
#endif

// Cell-centred patch grid (FiniteVolumeGrid.cpp:9-36): dx = (xu - xl)/nx, point(i) = xl + dx/2 + i dx.
class FiniteVolumeGrid : public MPI::MPIObject {
public:
    FiniteVolumeGrid() = default;
    FiniteVolumeGrid(MPI::Communicator comm, int nx, double x_lower, double x_upper, int ny, double y_lower,
                     double y_upper)
        : MPIObject(comm), nx_(nx), ny_(ny), xl_(x_lower), xu_(x_upper), yl_(y_lower), yu_(y_upper),
          dx_((x_upper - x_lower) / nx), dy_((y_upper - y_lower) / ny) {}
    double point(DimensionIndex dim, int index) const {
        if (dim == DimensionIndex::X) return (xl_ + dx_ / 2) + index * dx_;
        if (dim == DimensionIndex::Y) return (yl_ + dy_ / 2) + index * dy_;
        throw std::invalid_argument("Invalid `dim` argument");
    }
    std::string name() const { return "FiniteVolumeGrid"; }
    std::size_t nx() const { return (std::size_t)nx_; }
    std::size_t ny() const { return (std::size_t)ny_; }
    double xLower() const { return xl_; }
    double xUpper() const { return xu_; }
    double yLower() const { return yl_; }
    double yUpper() const { return yu_; }
    double dx() const { return dx_; }
    double dy() const { return dy_; }
    double operator()(std::size_t dim, std::size_t index) const { return point((DimensionIndex)dim, (int)index); }
private:
    int nx_ = 0, ny_ = 0;
    double xl_ = 0, xu_ = 0, yl_ = 0, yu_ = 0, dx_ = 0, dy_ = 0;
};

enum FiniteVolumeSolverType { FivePointStencil, FISHPACK90 };
using Analytical2DFunction = std::function<double(double, double)>;

// Per-cell coefficient samples of one leaf, as FiniteVolumeSolver.cpp:64-78 takes them:
// alpha(x_i,y_j), beta at the W/E/S/N faces, lambda(x_i,y_j); index j + i*ny.
inline void sampleCellCoefficients(const FiniteVolumeGrid& g, const Analytical2DFunction& alpha,
                                   const Analytical2DFunction& beta, const Analytical2DFunction& lambda,
                                   double* out /* 6 * nx * ny */) {
    const int nx = (int)g.nx(), ny = (int)g.ny();
    const std::size_t n = (std::size_t)nx * ny;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const double x = g(0, i), y = g(1, j);
            const std::size_t I = (std::size_t)j + (std::size_t)i * ny;
            out[I] = alpha(x, y);
            out[n + I] = beta(x - g.dx() / 2.0, y);
            out[2 * n + I] = beta(x + g.dx() / 2.0, y);
            out[3 * n + I] = beta(x, y - g.dy() / 2.0);
            out[4 * n + I] = beta(x, y + g.dy() / 2.0);
            out[5 * n + I] = lambda(x, y);
        }
}

class FiniteVolumePatch;

// Leaf patch solver (FiniteVolumeSolver.hpp:53-162).  The single-patch operations run a one-leaf
// mesh on the device: solve = leaf solve with Dirichlet data g and load f, buildD2N = its DtN
// operator T, particularNeumannData = its h, mapD2N = T g + h.
class FiniteVolumeSolver : public MPI::MPIObject {
public:
    FiniteVolumeSolverType solver_type = FiniteVolumeSolverType::FivePointStencil;
    Analytical2DFunction alpha_function = [](double, double) { return 1.0; };
    Analytical2DFunction beta_function = [](double, double) { return 1.0; };
    Analytical2DFunction lambda_function = [](double, double) { return 0.0; };
    Analytical2DFunction rhs_function;

    FiniteVolumeSolver() = default;
    FiniteVolumeSolver(MPI::Communicator comm, Analytical2DFunction alpha, Analytical2DFunction beta,
                       Analytical2DFunction lambda, Analytical2DFunction rhs)
        : MPIObject(comm), alpha_function(alpha), beta_function(beta), lambda_function(lambda), rhs_function(rhs) {}
    std::string name() const { return "FiniteVolumeSolver"; }
    int gridIndex2MatrixIndex(int i, int j, int nx, int ny) const { (void)nx; return j + i * ny; }

    // Coefficients of ef_build for leaves with the given grids (z-order).
    void fillCoefficients(const std::vector<const FiniteVolumeGrid*>& leaf_grids, ef_coeffs& c,
                          std::vector<double>& storage) const {
        c.solver_type = solver_type == FISHPACK90 ? EF_SOLVER_FISHPACK : EF_SOLVER_FIVE_POINT;
        c.alpha0 = 1.0; c.beta0 = 1.0;
        c.lambda0 = lambda_function(0, 0);  // FISHPACK uses lambda(0,0) (FiniteVolumeSolver.cpp:262)
        c.cell = nullptr;
        if (c.solver_type == EF_SOLVER_FIVE_POINT) {
            const std::size_t per = 6 * leaf_grids.front()->nx() * leaf_grids.front()->ny();
            storage.assign(per * leaf_grids.size(), 0.0);
            for (std::size_t k = 0; k < leaf_grids.size(); k++)
                sampleCellCoefficients(*leaf_grids[k], alpha_function, beta_function, lambda_function,
                                       storage.data() + k * per);
            c.cell = storage.data();
        }
    }

    // The single-grid operations run as a batch of one patch (ef_patches_*).
    Vector<double> solve(FiniteVolumeGrid& grid, Vector<double>& dirichlet_data, Vector<double>& rhs_data) {
        OnePatch p(*this, grid);
        p.checkSizes(&dirichlet_data, &rhs_data);
        Vector<double> u((std::size_t)p.N * p.N);
        efCheck(ef_patches_apply(p.h, EF_PATCH_SOLVE, dirichlet_data.dataPointer(), rhs_data.dataPointer(),
                                 u.dataPointer(), 1), "FiniteVolumeSolver::solve");
        return u;
    }
    Vector<double> mapD2N(FiniteVolumeGrid& grid, Vector<double>& dirichlet_data, Vector<double>& rhs_data) {
        OnePatch p(*this, grid);
        p.checkSizes(&dirichlet_data, &rhs_data);
        Vector<double> h((std::size_t)4 * p.N);
        efCheck(ef_patches_apply(p.h, EF_PATCH_MAP_D2N, dirichlet_data.dataPointer(), rhs_data.dataPointer(),
                                 h.dataPointer(), 1), "FiniteVolumeSolver::mapD2N");
        return h;
    }
    Matrix<double> buildD2N(FiniteVolumeGrid& grid) {
        OnePatch p(*this, grid);
        Matrix<double> T((std::size_t)4 * p.N, (std::size_t)4 * p.N);
        efCheck(ef_patches_dtn(p.h, T.dataPointer()), "FiniteVolumeSolver::buildD2N");
        return T;
    }
    Vector<double> particularNeumannData(FiniteVolumeGrid& grid, Vector<double>& rhs_data) {
        OnePatch p(*this, grid);
        p.checkSizes(nullptr, &rhs_data);
        Vector<double> h((std::size_t)4 * p.N);
        efCheck(ef_patches_apply(p.h, EF_PATCH_MAP_D2N, nullptr, rhs_data.dataPointer(), h.dataPointer(), 1),
                "FiniteVolumeSolver::particularNeumannData");
        return h;
    }

private:
    struct OnePatch {
        ef_patches h = nullptr;
        int N = 0;
        OnePatch(const FiniteVolumeSolver& s, const FiniteVolumeGrid& g) {
            if (g.nx() != g.ny()) throw std::invalid_argument("[EllipticForest::FiniteVolumeSolver] square patches only");
            N = (int)g.nx();
            const double bounds[4] = {g.xLower(), g.xUpper(), g.yLower(), g.yUpper()};
            ef_patch_batch b{1, N, bounds};
            ef_coeffs c{};
            std::vector<double> store;
            s.fillCoefficients({&g}, c, store);
            efCheck(ef_patches_create(defaultContext(), &b, &c, &h), "FiniteVolumeSolver");
        }
        ~OnePatch() { ef_patches_destroy(h); }
        OnePatch(const OnePatch&) = delete;
        OnePatch& operator=(const OnePatch&) = delete;
        void checkSizes(const Vector<double>* g, const Vector<double>* f) const {
            if (g && g->size() != (std::size_t)4 * N)
                throw std::invalid_argument("[EllipticForest::FiniteVolumeSolver] dirichlet_data must have 4N entries");
            if (f && f->size() != (std::size_t)N * N)
                throw std::invalid_argument("[EllipticForest::FiniteVolumeSolver] rhs_data must have N*N entries");
        }
    };
};

// Per-node patch (FiniteVolumePatch.hpp:11-261).  After the stages the operators and vectors
// live on the device; the accessors copy a node's data to the host on first use.
class FiniteVolumePatch : public MPI::MPIObject {
public:
    int n_coarsens = 0;
    bool leaf = false;

    FiniteVolumePatch() = default;
    FiniteVolumePatch(MPI::Communicator comm, FiniteVolumeGrid grid) : MPIObject(comm), grid_(grid) {}
    std::string name() const { return "FiniteVolumePatch"; }
    FiniteVolumeGrid& grid() { return grid_; }
    int size() { return (int)(grid_.nx() / (std::size_t)std::pow(2, n_coarsens)); }  // Patch.hpp:99-101

    Matrix<double>& matrixX() { return fetchM(2, X_); }
    Matrix<double>& matrixH() { return fetchM(3, H_); }
    Matrix<double>& matrixS() { return fetchM(1, S_); }
    Matrix<double>& matrixT() { return fetchM(0, T_); }
    Vector<double>& vectorU() { return u_; }
    Vector<double>& vectorG() { return fetchV(5, g_); }
    Vector<double>& vectorV() { return v_; }
    Vector<double>& vectorF() { return f_; }
    Vector<double>& vectorH() { return fetchV(7, h_); }
    Vector<double>& vectorW() { return fetchV(8, w_); }

    double dataSize() {  // FiniteVolumePatch.cpp:70-87, in MB
        const double MB = 1024.0 * 1024.0;
        double s = (4 * sizeof(int) + sizeof(bool)) / MB;
        for (auto* m : {&T_, &H_, &S_, &X_}) s += m->nRows() * m->nCols() * sizeof(double) / MB;
        for (auto* v : {&u_, &g_, &v_, &f_, &h_, &w_}) s += v->size() * sizeof(double) / MB;
        return s;
    }
    std::string str() {
        std::ostringstream o;
        o << "FiniteVolumePatch: nx=" << grid_.nx() << " [" << grid_.xLower() << "," << grid_.xUpper() << "]x["
          << grid_.yLower() << "," << grid_.yUpper() << "] n_coarsens=" << n_coarsens;
        return o.str();
    }

    // set by HPSAlgorithm: where this patch's data lives, and the stage counter it belongs to
    void attach(ef_mesh m, int node, const long* stage) { mesh_ = m; node_ = node; stage_ = stage; }
    void setGrid(const FiniteVolumeGrid& g) { grid_ = g; }

private:
    Matrix<double>& fetchM(int which, Matrix<double>& dst) {
        if (mesh_ && stage_ && fetched_[which] != *stage_) {
            int r = 0, c = 0;
            if (ef_get_operator(mesh_, node_, which, 0, nullptr, &r, &c) == EF_OK) {
                dst = Matrix<double>(r, c);
                efCheck(ef_get_operator(mesh_, node_, which, 0, dst.dataPointer(), &r, &c), "FiniteVolumePatch");
            }
            fetched_[which] = *stage_;
        }
        return dst;
    }
    Vector<double>& fetchV(int which, Vector<double>& dst) {
        if (mesh_ && stage_ && fetched_[which] != *stage_) {
            int r = 0, c = 0;
            if (ef_get_operator(mesh_, node_, which, 0, nullptr, &r, &c) == EF_OK) {
                std::vector<double> buf((std::size_t)r * c);
                if (ef_get_operator(mesh_, node_, which, 0, buf.data(), &r, &c) == EF_OK) dst = Vector<double>(buf);
            }
            fetched_[which] = *stage_;
        }
        return dst;
    }
    FiniteVolumeGrid grid_;
    Matrix<double> X_, H_, S_, T_;
    Vector<double> u_, g_, v_, f_, h_, w_;
    ef_mesh mesh_ = nullptr;
    int node_ = -1;
    const long* stage_ = nullptr;
    long fetched_[9] = {-1, -1, -1, -1, -1, -1, -1, -1, -1};
};

// Node factory (FiniteVolumeNodeFactory.cpp:11-93): children are the parent's quadrants
// with the same nx, sibling 0 = SW, 1 = SE, 2 = NW, 3 = NE.
class FiniteVolumeNodeFactory : public MPI::MPIObject {
public:
    FiniteVolumeNodeFactory() = default;
    explicit FiniteVolumeNodeFactory(MPI::Communicator comm) : MPIObject(comm) {}
    FiniteVolumePatch createChildPatch(FiniteVolumePatch& parent, int sibling_id) {
        auto& g = parent.grid();
        const double xm = (g.xLower() + g.xUpper()) / 2.0, ym = (g.yLower() + g.yUpper()) / 2.0;
        const int nx = (int)g.nx(), ny = (int)g.ny();
        FiniteVolumeGrid cg;
        switch (sibling_id) {
            case 0: cg = FiniteVolumeGrid(getComm(), nx, g.xLower(), xm, ny, g.yLower(), ym); break;
            case 1: cg = FiniteVolumeGrid(getComm(), nx, xm, g.xUpper(), ny, g.yLower(), ym); break;
            case 2: cg = FiniteVolumeGrid(getComm(), nx, g.xLower(), xm, ny, ym, g.yUpper()); break;
            case 3: cg = FiniteVolumeGrid(getComm(), nx, xm, g.xUpper(), ny, ym, g.yUpper()); break;
            default: throw std::invalid_argument("[EllipticForest::FiniteVolumeNodeFactory] sibling id 0..3");
        }
        return FiniteVolumePatch(getComm(), cg);
    }
    Node<FiniteVolumePatch>* createNode(FiniteVolumePatch data, std::string path, int level, int, int) {
        auto* n = new Node<FiniteVolumePatch>();
        n->data = data; n->path = path; n->level = level;
        return n;
    }
    Node<FiniteVolumePatch>* createChildNode(Node<FiniteVolumePatch>* parent, int sibling_id, int, int) {
        auto* n = new Node<FiniteVolumePatch>();
        n->data = createChildPatch(parent->data, sibling_id);
        n->path = parent->path + std::to_string(sibling_id);
        n->level = parent->level + 1;
        return n;
    }
    Node<FiniteVolumePatch>* createParentNode(std::vector<Node<FiniteVolumePatch>*> children, int, int) {
        auto& a = children[0]->data.grid();
        FiniteVolumeGrid pg(getComm(), (int)a.nx(), a.xLower(), children[1]->data.grid().xUpper(), (int)a.ny(),
                            a.yLower(), children[2]->data.grid().yUpper());
        auto* n = new Node<FiniteVolumePatch>();
        n->data = FiniteVolumePatch(getComm(), pg);
        n->path = children[0]->path.substr(0, children[0]->path.size() - 1);
        n->level = children[0]->level - 1;
        return n;
    }
};

}  // namespace EllipticForest

#endif  // EF_FINITE_VOLUME_HPP_
