// vtk_mesh.cpp -- VTK output of a mesh through the drop-in C++ API, in the shape of the
// reference's examples/elliptic-single/main.cpp:225-273: mesh functions from vectors over the
// leaf cells and from functions sampled at the cell centres, then Mesh::toVTK.
//   vtk_mesh --out=<base> [--nx=N] [--max-level=L]
// Builds refineByFunction(|f| > 1.2, levels 0..L) on [-10,10]^2 with N x N cells per patch and
// writes <base>-mesh.pvtu / _0000.vtu and <base>-quadtree.pvtu / _0000.vtu, then the same with
// number 7 (<base>-mesh-0007...).  Prints "leaves n" and "cells n".
#include <EllipticForest.hpp>
#include <Patches/FiniteVolume/FiniteVolume.hpp>

#include <cmath>
#include <cstdio>

using namespace EllipticForest;

int main(int argc, char** argv) {
    EllipticForestApp app(&argc, &argv);
    const std::string base = std::get<std::string>(app.options["out"]);
    const int nx = app.options.optionExists("nx") ? std::get<int>(app.options["nx"]) : 4;
    const int max_level = app.options.optionExists("max-level") ? std::get<int>(app.options["max-level"]) : 3;
    auto f = [](double x, double y) { return -(std::sin(x) + std::sin(y)); };

    FiniteVolumeGrid grid(MPI_COMM_WORLD, nx, -10.0, 10.0, nx, -10.0, 10.0);
    FiniteVolumePatch root_patch(MPI_COMM_WORLD, grid);
    FiniteVolumeNodeFactory node_factory(MPI_COMM_WORLD);
    Mesh<FiniteVolumePatch> mesh{};
    mesh.refineByFunction([&](double x, double y) { return std::fabs(f(x, y)) > 1.2; }, 1.2, 0, max_level, root_patch,
                          node_factory);

    // a vector over every leaf cell (the cell's x centre), in the order the leaves are visited
    Vector<double> xMesh{};
    xMesh.name() = "x_centre";
    int leaves = 0;
    mesh.quadtree.traversePreOrder([&](Node<FiniteVolumePatch>* node) {
        if (!node->leaf) return 1;
        auto& g = node->data.grid();
        for (std::size_t i = 0; i < g.nx(); i++)
            for (std::size_t j = 0; j < g.ny(); j++) xMesh.append(g(0, i));
        leaves++;
        return 1;
    });
    mesh.addMeshFunction(xMesh);
    mesh.addMeshFunction(f, "f_rhs");
    mesh.addMeshFunction([](double x, double y) { return x * y; }, "xy");
    mesh.toVTK(base);
    mesh.toVTK(base, 7);
    std::printf("leaves %d\ncells %zu\n", leaves, xMesh.size());
    return 0;
}
