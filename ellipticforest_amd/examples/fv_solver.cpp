// fv_solver.cpp -- the single-grid FiniteVolumeSolver operations through the drop-in C++ API,
// in the shape of the reference's test_finite_volume_solver.cpp: u = sin x + sin y on
// [-1,1]^2, Dirichlet data at the faces, f = -(sin x + sin y), FISHPACK semantics; the
// max-norm error for n = 8, 16, ..., M and the convergence order of the last two.  Then, on an
// 8 x 8 grid over [0,1] x [0.5,1.25], the DtN matrix, the Neumann data of the same problem
// and its particular part, printed for comparison with other implementations.
//   fv_solver [--max-n=M]   (n = 8, 16, ..., M; default 32, the reference's test runs to 512)
// Output lines: "err n linf", "order p", "T i j v" (a few entries), "h i v", "hp i v".
//   fv_solver --five-point [--max-n=M]   the same with the five-point stencil and variable
// coefficients alpha = 1, beta = 1 + 0.5 sin(2 pi x) cos(2 pi y), lambda = -1 (M <= 32), the
// manufactured solution u = sin(3x) cosh(y) and f = div(beta grad u) - u.
#include <EllipticForest.hpp>
#include <Patches/FiniteVolume/FiniteVolume.hpp>

#include <cmath>
#include <cstdio>
#include <vector>

using namespace EllipticForest;

static bool fivePoint = false;
static double uPoisson(double x, double y) { return std::sin(x) + std::sin(y); }
static double uVar(double x, double y) { return std::sin(3.0 * x) * std::cosh(y); }
static double betaVar(double x, double y) { return 1.0 + 0.5 * std::sin(2.0 * M_PI * x) * std::cos(2.0 * M_PI * y); }
static double uExact(double x, double y) { return fivePoint ? uVar(x, y) : uPoisson(x, y); }
// f = div(beta grad u) + lambda u = beta lap(u) + grad(beta) . grad(u) - u for the variable problem
static double fRHS(double x, double y) {
    if (!fivePoint) return -uPoisson(x, y);
    const double ux = 3.0 * std::cos(3.0 * x) * std::cosh(y), uy = std::sin(3.0 * x) * std::sinh(y);
    const double lap = -8.0 * std::sin(3.0 * x) * std::cosh(y);
    const double bx = M_PI * std::cos(2.0 * M_PI * x) * std::cos(2.0 * M_PI * y);
    const double by = -M_PI * std::sin(2.0 * M_PI * x) * std::sin(2.0 * M_PI * y);
    return betaVar(x, y) * lap + bx * ux + by * uy - uVar(x, y);
}

static void problem(FiniteVolumeGrid& grid, int n, Vector<double>& g, Vector<double>& f, Vector<double>& ue) {
    g = Vector<double>(4 * n);
    f = Vector<double>(n * n);
    ue = Vector<double>(n * n);
    for (int j = 0; j < n; j++) {
        g[j] = uExact(grid.xLower(), grid(YDIM, j));
        g[n + j] = uExact(grid.xUpper(), grid(YDIM, j));
    }
    for (int i = 0; i < n; i++) {
        g[2 * n + i] = uExact(grid(XDIM, i), grid.yLower());
        g[3 * n + i] = uExact(grid(XDIM, i), grid.yUpper());
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            f[j + i * n] = fRHS(grid(XDIM, i), grid(YDIM, j));
            ue[j + i * n] = uExact(grid(XDIM, i), grid(YDIM, j));
        }
}

int main(int argc, char** argv) {
    EllipticForestApp app(&argc, &argv);
    fivePoint = app.options.optionExists("five-point");
    FiniteVolumeSolver solver;
    solver.solver_type = fivePoint ? FiniteVolumeSolverType::FivePointStencil : FiniteVolumeSolverType::FISHPACK90;
    solver.alpha_function = [](double, double) { return 1.0; };
    if (fivePoint) {
        solver.beta_function = betaVar;
        solver.lambda_function = [](double, double) { return -1.0; };
    } else {
        solver.beta_function = [](double, double) { return 1.0; };
        solver.lambda_function = [](double, double) { return 0.0; };
    }

    const int max_n = app.options.optionExists("max-n") ? std::get<int>(app.options["max-n"]) : 32;
    std::vector<int> ns;
    for (int n = 8; n <= max_n; n *= 2) ns.push_back(n);
    std::vector<double> errors(ns.size());
    for (std::size_t t = 0; t < ns.size(); t++) {
        const int n = ns[t];
        FiniteVolumeGrid grid(MPI_COMM_WORLD, n, -1.0, 1.0, n, -1.0, 1.0);
        Vector<double> g, f, ue;
        problem(grid, n, g, f, ue);
        Vector<double> u = solver.solve(grid, g, f);
        double e = 0.0;
        for (int k = 0; k < n * n; k++) e = std::fmax(e, std::fabs(u[k] - ue[k]));
        errors[t] = e;
        std::printf("err %d %.17g\n", n, e);
    }
    const std::size_t L = ns.size();
    std::printf("order %.17g\n", std::log(errors[L - 1] / errors[L - 2]) / std::log((double)ns[L - 2] / ns[L - 1]));

    const int n = 8;
    FiniteVolumeGrid grid(MPI_COMM_WORLD, n, 0.0, 1.0, n, 0.5, 1.25);
    Matrix<double> T = solver.buildD2N(grid);
    const int picks[6][2] = {{0, 0}, {0, 1}, {3, 17}, {8, 8}, {31, 31}, {20, 5}};
    for (auto& p : picks) std::printf("T %d %d %.17g\n", p[0], p[1], T(p[0], p[1]));
    Vector<double> g, f, ue;
    problem(grid, n, g, f, ue);
    Vector<double> h = solver.mapD2N(grid, g, f);
    Vector<double> hp = solver.particularNeumannData(grid, f);
    for (int i = 0; i < 4 * n; i++) std::printf("h %d %.17g\n", i, h[i]);
    for (int i = 0; i < 4 * n; i++) std::printf("hp %d %.17g\n", i, hp[i]);
    return 0;
}
