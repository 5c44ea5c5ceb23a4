// elliptic_single.cpp -- Poisson on [-10,10]^2 with u = sin x + sin y through the drop-in
// C++ API: the driver workflow of the reference's examples/elliptic-single (options,
// refineByFunction, FISHPACK leaf solver, setup/build/upwards/solve, errors at cell centres).
//   elliptic_single [--max-level=L] [--nx=N] [--min-level=L0] [--threshold=t] [--uniform=1]
// Prints "errors l1 l2 linf", the leaf count and the stage times.
#include <EllipticForest.hpp>
#include <Patches/FiniteVolume/FiniteVolume.hpp>

#include <cmath>
#include <cstdio>

using namespace EllipticForest;

static double uFunction(double x, double y) { return std::sin(x) + std::sin(y); }
static double fFunction(double x, double y) { return -uFunction(x, y); }

static int opt(EllipticForestApp& app, const char* k, int d) {
    return app.options.optionExists(k) ? std::get<int>(app.options[k]) : d;
}

int main(int argc, char** argv) {
    EllipticForestApp app(&argc, &argv);
    const double threshold = app.options.optionExists("threshold") ? std::get<double>(app.options["threshold"]) : 1.2;
    const int min_level = opt(app, "min-level", 0);
    const int max_level = opt(app, "max-level", 4);
    const int nx = opt(app, "nx", 8);
    const bool uniform = opt(app, "uniform", 0) != 0;
    const double xl = -10, xu = 10, yl = -10, yu = 10;

    FiniteVolumeGrid grid(MPI_COMM_WORLD, nx, xl, xu, nx, yl, yu);
    FiniteVolumePatch root_patch(MPI_COMM_WORLD, grid);
    FiniteVolumeNodeFactory node_factory(MPI_COMM_WORLD);
    Mesh<FiniteVolumePatch> mesh{};
    if (uniform) mesh.uniform(max_level, root_patch, node_factory);
    else
        mesh.refineByFunction([&](double x, double y) { return std::fabs(fFunction(x, y)) > threshold; }, threshold,
                              min_level, max_level, root_patch, node_factory);

    FiniteVolumeSolver solver{};
    solver.solver_type = FiniteVolumeSolverType::FISHPACK90;
    solver.alpha_function = [](double, double) { return 1.0; };
    solver.beta_function = [](double, double) { return 1.0; };
    solver.lambda_function = [](double, double) { return 0.0; };

    HPSAlgorithm<FiniteVolumeGrid, FiniteVolumeSolver, FiniteVolumePatch, double> HPS(MPI_COMM_WORLD, mesh, solver);
    HPS.setupStage();
    HPS.buildStage();
    HPS.upwardsStage([&](double x, double y) { return fFunction(x, y); });
    HPS.solveStage([&](int side, double x, double y, double* a, double* b) {
        (void)side;
        *a = 1.0;
        *b = 0.0;
        return uFunction(x, y);
    });

    // errors at cell centres, area-weighted as examples/elliptic-multiple/main.cpp:347-374
    double l1 = 0, l2 = 0, li = 0;
    int leaves = 0;
    mesh.quadtree.traversePostOrder([&](Node<FiniteVolumePatch>* node) {
        if (!node->leaf) return 1;
        auto& patch = node->data;
        auto& g = patch.grid();
        for (std::size_t i = 0; i < g.nx(); i++)
            for (std::size_t j = 0; j < g.ny(); j++) {
                const double d = patch.vectorU()[j + i * g.ny()] - uFunction(g(XDIM, i), g(YDIM, j));
                l1 += g.dx() * g.dy() * std::fabs(d);
                l2 += g.dx() * g.dy() * d * d;
                li = std::fmax(li, std::fabs(d));
            }
        leaves++;
        return 1;
    });
    const double area = (xu - xl) * (yu - yl);
    std::printf("leaves %d\n", leaves);
    std::printf("errors %.17g %.17g %.17g\n", l1 / area, std::sqrt(l2 / area), li);
    std::printf("root_T_rows %zu\n", mesh.quadtree.root().matrixT().nRows());
    for (const char* t : {"setup-stage", "build-stage", "upwards-stage", "solve-stage"})
        std::printf("timer %s %.6f\n", t, app.timers[t].time());
    return 0;
}
