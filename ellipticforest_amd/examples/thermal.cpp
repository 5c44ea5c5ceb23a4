// thermal.cpp -- implicit time stepping through the drop-in C++ API, in the shape of the
// reference's examples/thermal/main.cpp:272-355: one build stage, then per time step an
// upwards stage whose callback reads the previous step's vectorU of every leaf
// ((I/dt - Laplacian) u^{n+1} = u^n/dt + s as the five-point problem
// Laplacian u - u/dt = -u^n/dt - s), and a solve stage with Dirichlet data on the west and
// east sides and Neumann data (a = 0, b = 1) on the south and north sides.
//   thermal [--level=L] [--nx=N] [--steps=K] [--dt=t]
// Prints "u leaf cell value" for every cell of the last step and "max |u|".
#include <EllipticForest.hpp>
#include <Patches/FiniteVolume/FiniteVolume.hpp>

#include <cmath>
#include <cstdio>

using namespace EllipticForest;

static double sources(double x, double y, double t) { return std::exp(-t) * std::cos(x) * std::sin(1.5 * y); }
static double uWest(double, double y, double t) { return std::sin(y + t); }
static double uEast(double, double y, double t) { return std::cos(y - t); }
static double dudnSouth(double x, double, double t) { return 0.5 * x * std::exp(-t); }
static double dudnNorth(double x, double, double t) { return -0.25 * std::cos(x + t); }

int main(int argc, char** argv) {
    EllipticForestApp app(&argc, &argv);
    auto opt = [&](const char* k, int d) { return app.options.optionExists(k) ? std::get<int>(app.options[k]) : d; };
    const int level = opt("level", 2), nx = opt("nx", 8), nt = opt("steps", 3);
    const double dt = app.options.optionExists("dt") ? std::get<double>(app.options["dt"]) : 0.1;

    FiniteVolumeGrid grid(MPI_COMM_WORLD, nx, -1.0, 1.0, nx, -1.0, 1.0);
    FiniteVolumePatch root_patch(MPI_COMM_WORLD, grid);
    FiniteVolumeNodeFactory node_factory(MPI_COMM_WORLD);
    Mesh<FiniteVolumePatch> mesh{};
    mesh.uniform(level, root_patch, node_factory);

    FiniteVolumeSolver solver{};
    solver.solver_type = FiniteVolumeSolverType::FivePointStencil;
    solver.alpha_function = [](double, double) { return 1.0; };
    solver.beta_function = [](double, double) { return 1.0; };
    solver.lambda_function = [&](double, double) { return -1.0 / dt; };

    HPSAlgorithm<FiniteVolumeGrid, FiniteVolumeSolver, FiniteVolumePatch, double> HPS(MPI_COMM_WORLD, mesh, solver);
    HPS.setupStage();
    HPS.buildStage();
    for (int n = 0; n <= nt; n++) {
        const double time = n * dt;
        HPS.upwardsStage([&](FiniteVolumePatch& patch) {
            auto& g = patch.grid();
            const int px = (int)g.nx(), py = (int)g.ny();
            patch.vectorF() = Vector<double>(px * py);
            auto& u = patch.vectorU();
            auto& f = patch.vectorF();
            for (int i = 0; i < px; i++)
                for (int j = 0; j < py; j++) {
                    const int I = j + i * py;
                    f[I] = (n == 0) ? 0.0 : -(1.0 / dt) * u[I] - sources(g(0, i), g(1, j), time);
                }
        });
        HPS.solveStage([&](int side, double x, double y, double* a, double* b) {
            switch (side) {
                case 0: *a = 1.0; *b = 0.0; return uWest(x, y, time);
                case 1: *a = 1.0; *b = 0.0; return uEast(x, y, time);
                case 2: *a = 0.0; *b = 1.0; return dudnSouth(x, y, time);
                default: *a = 0.0; *b = 1.0; return dudnNorth(x, y, time);
            }
        });
    }
    double umax = 0.0;
    int leaf = 0;
    mesh.quadtree.traversePreOrder([&](Node<FiniteVolumePatch>* node) {
        if (!node->leaf) return 1;
        auto& u = node->data.vectorU();
        for (std::size_t k = 0; k < u.size(); k++) {
            std::printf("u %d %zu %.17g\n", leaf, k, u[k]);
            umax = std::fmax(umax, std::fabs(u[k]));
        }
        leaf++;
        return 1;
    });
    std::printf("max %.17g\n", umax);
    return 0;
}
