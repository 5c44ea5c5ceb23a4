// EllipticForest.hpp -- drop-in C++ surface of the MI355X HPS path.
//
// A driver written against the reference (e.g. examples/elliptic-single/main.cpp) includes
// <EllipticForest.hpp> and <Patches/FiniteVolume/FiniteVolume.hpp> and keeps using
//   EllipticForestApp (options "cache-operators", "homogeneous-rhs"; timers
//                     "setup-stage", "build-stage", "upwards-stage", "solve-stage"),
//   Mesh<Patch>::refineByFunction / iteratePatches / iterateCells, quadtree traversals,
//   HPSAlgorithm<FiniteVolumeGrid, FiniteVolumeSolver, FiniteVolumePatch, double>
//     setupStage / buildStage / upwardsStage / solveStage,
//   FiniteVolumeGrid / FiniteVolumeSolver / FiniteVolumePatch / FiniteVolumeNodeFactory.
// The numerics run in libefhps.so (include/ef_hps.h) on one GPU; these headers sample the
// user's functions on the host, move data through the C ABI and turn status codes back
// into the reference's exception types.  Single rank: MPI communicators are accepted and
// ignored.  Link with -lefhps.
#ifndef ELLIPTIC_FOREST_HPP_
#define ELLIPTIC_FOREST_HPP_

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <algorithm>
#include <cstdio>
#include <deque>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "ef_hps.h"

#ifndef MPI_COMM_WORLD
#define EF_NO_MPI_HEADER 1
typedef int MPI_Comm;
#define MPI_COMM_WORLD 0
#define MPI_COMM_SELF 1
#endif

namespace EllipticForest {

// ---------------------------------------------------------------------------
// errors: C ABI status -> reference exception types (std::invalid_argument for shape and
// argument errors, std::runtime_error otherwise)
// ---------------------------------------------------------------------------
inline void efCheck(int rc, const char* where) {
    if (rc == EF_OK) return;
    std::string msg = std::string("[EllipticForest::") + where + "] " + ef_last_error();
    if (rc == EF_ERR_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

namespace MPI {
using Communicator = MPI_Comm;
class MPIObject {
public:
    MPIObject() = default;
    explicit MPIObject(Communicator comm) : comm_(comm) {}
    Communicator getComm() const { return comm_; }
    int getRank() const { return 0; }
    int getSize() const { return 1; }
protected:
    Communicator comm_ = MPI_COMM_WORLD;
};
}  // namespace MPI

// ---------------------------------------------------------------------------
// Vector / Matrix: the subset of src/Vector.hpp and src/Matrix.hpp that drivers touch
// (row-major storage, Matrix.hpp:711-725).
// ---------------------------------------------------------------------------
template <typename T>
class Vector {
public:
    Vector() = default;
    explicit Vector(std::size_t n, T v = T(0)) : d_(n, v) {}
    Vector(std::initializer_list<T> l) : d_(l) {}
    explicit Vector(std::vector<T> v) : d_(std::move(v)) {}
    std::size_t size() const { return d_.size(); }
    T& operator[](std::size_t i) { return d_[i]; }
    const T& operator[](std::size_t i) const { return d_[i]; }
    T& operator()(std::size_t i) { return d_.at(i); }
    std::vector<T>& data() { return d_; }
    const std::vector<T>& data() const { return d_; }
    T* dataPointer() { return d_.data(); }
    std::string& name() { return name_; }
    void append(const Vector<T>& v) { d_.insert(d_.end(), v.d_.begin(), v.d_.end()); }
    void append(T v) { d_.push_back(v); }
    // VTK data-array view (Vector.hpp:670-757): type, name, components, ASCII data written with
    // std::to_string and followed by a space, range from std::to_string of the extremes
    void setType(std::string t) { vtkType_ = std::move(t); }
    std::string getType() const { return vtkType_; }
    std::string getName() const { return name_; }
    void setNumberOfComponents(std::string n) { vtkComponents_ = std::move(n); }
    std::string getNumberOfComponents() const { return vtkComponents_; }
    std::string getFormat() const { return "ascii"; }
    std::string getRangeMin() const { return std::to_string(d_.empty() ? T(0) : *std::min_element(d_.begin(), d_.end())); }
    std::string getRangeMax() const { return std::to_string(d_.empty() ? T(0) : *std::max_element(d_.begin(), d_.end())); }
    std::string getData() const {
        std::string str;
        for (const T& v : d_) str += std::to_string(v) + " ";
        return str;
    }
    Vector<T> getSegment(std::size_t start, std::size_t len) const {
        if (start + len > d_.size()) throw std::out_of_range("[EllipticForest::Vector::getSegment] out of range");
        return Vector<T>(std::vector<T>(d_.begin() + start, d_.begin() + start + len));
    }
    void setSegment(std::size_t start, const Vector<T>& v) {
        if (start + v.size() > d_.size()) throw std::out_of_range("[EllipticForest::Vector::setSegment] out of range");
        std::copy(v.d_.begin(), v.d_.end(), d_.begin() + start);
    }
private:
    std::vector<T> d_;
    std::string name_ = "Vector";
    std::string vtkType_ = "Float32";
    std::string vtkComponents_ = "1";
};

// ---------------------------------------------------------------------------
// VTK output: an XML tree written as XMLTree.cpp:27-66 writes it (two spaces of indent per
// level, attributes in insertion order, a node's data on its own line before its children).
// ---------------------------------------------------------------------------
namespace vtk {
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attributes;
    std::string data;
    std::vector<XmlNode> children;
};
inline void writeNode(std::ofstream& file, const XmlNode& node, const std::string& prefix) {
    std::string header = prefix + "<" + node.name + " ";
    for (auto& a : node.attributes) header += a.first + "=\"" + a.second + "\" ";
    header.pop_back();
    file << header << ">" << std::endl;
    if (!node.data.empty()) file << node.data << std::endl;
    for (auto& c : node.children) writeNode(file, c, prefix + "  ");
    file << prefix << "</" << node.name << ">" << std::endl;
}
inline void writeTree(const std::string& filename, const XmlNode& root) {
    std::ofstream file(filename);
    if (!file) throw std::runtime_error("[EllipticForest::vtk] cannot open " + filename);
    file << "<?xml version=\"1.0\"?>" << std::endl;
    writeNode(file, root, "");
}
// DataArray / PDataArray nodes of a vector (VTK.cpp:5-24)
template <typename T>
inline XmlNode dataArray(const Vector<T>& v) {
    return XmlNode{"DataArray",
                   {{"type", v.getType()}, {"Name", v.getName()}, {"NumberOfComponents", v.getNumberOfComponents()},
                    {"format", v.getFormat()}, {"RangeMin", v.getRangeMin()}, {"RangeMax", v.getRangeMax()}},
                   v.getData(), {}};
}
template <typename T>
inline XmlNode pDataArray(const Vector<T>& v) {
    return XmlNode{"PDataArray",
                   {{"type", v.getType()}, {"Name", v.getName()}, {"NumberOfComponents", v.getNumberOfComponents()},
                    {"format", v.getFormat()}},
                   "", {}};
}
inline XmlNode vtkFile() {
    return XmlNode{"VTKFile", {{"type", "UnstructuredGrid"}, {"version", "0.1"}, {"byte_order", "LittleEndian"}}, "", {}};
}
}  // namespace vtk

template <typename T>
inline double vectorInfNorm(const Vector<T>& a, const Vector<T>& b) {
    double m = 0.0;
    for (std::size_t i = 0; i < a.size(); i++) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return m;
}

template <typename T>
class Matrix {
public:
    Matrix() = default;
    Matrix(std::size_t r, std::size_t c, T v = T(0)) : r_(r), c_(c), d_(r * c, v) {}
    std::size_t nRows() const { return r_; }
    std::size_t nCols() const { return c_; }
    T& operator()(std::size_t i, std::size_t j) { return d_[i * c_ + j]; }
    const T& operator()(std::size_t i, std::size_t j) const { return d_[i * c_ + j]; }
    std::vector<T>& data() { return d_; }
    T* dataPointer() { return d_.data(); }
private:
    std::size_t r_ = 0, c_ = 0;
    std::vector<T> d_;
};

// ---------------------------------------------------------------------------
// App: options and timers (EllipticForestApp.hpp:39-137, Options.hpp:27-33, Timer.cpp:8-23)
// ---------------------------------------------------------------------------
class Timer {
public:
    void start() { t0_ = std::chrono::steady_clock::now(); running_ = true; }
    double stop() {
        if (running_) acc_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        running_ = false;
        return acc_;
    }
    double time() const { return acc_; }
    void clear() { acc_ = 0.0; }
private:
    std::chrono::steady_clock::time_point t0_{};
    double acc_ = 0.0;
    bool running_ = false;
};

class Options {
public:
    using OptionTypes = std::variant<std::string, bool, int, double>;
    OptionTypes& operator[](const std::string& key) { return map_[key]; }
    template <typename T>
    void setOption(const std::string& key, T value) { map_[key] = value; }
    void setOption(const std::string& key, const char* value) { map_[key] = std::string(value); }
    bool optionExists(const std::string& key) const { return map_.count(key) != 0; }
private:
    std::map<std::string, OptionTypes> map_;
};

class EllipticForestApp {
public:
    Options options;
    std::map<std::string, Timer> timers;

    EllipticForestApp() { instancePtr() = this; defaults(); }
    EllipticForestApp(int* argc, char*** argv) {
        instancePtr() = this;
        defaults();
        for (int i = 1; argc && argv && i < *argc; i++) parseArg((*argv)[i]);
    }
    ~EllipticForestApp() {
        if (instancePtr() == this) instancePtr() = nullptr;
    }
    static EllipticForestApp& getInstance() {
        static EllipticForestApp fallback(0);
        return instancePtr() ? *instancePtr() : fallback;
    }
    void addTimer(const std::string& name) { timers[name]; }
    int log(const char* fmt, ...) {
        va_list a; va_start(a, fmt);
        std::printf("[EllipticForest 0] "); std::vprintf(fmt, a); std::printf("\n");
        va_end(a); return 0;
    }
    int logHead(const char* fmt, ...) {
        if (!verbose_) return 0;
        va_list a; va_start(a, fmt);
        std::printf("[EllipticForest] "); std::vprintf(fmt, a); std::printf("\n");
        va_end(a); return 0;
    }
    void setVerbose(bool v) { verbose_ = v; }
private:
    explicit EllipticForestApp(int) { defaults(); }
    static EllipticForestApp*& instancePtr() { static EllipticForestApp* p = nullptr; return p; }
    void defaults() {
        options.setOption("cache-operators", false);
        options.setOption("homogeneous-rhs", false);
    }
    // --key=value with type inference (InputParser.cpp:9-38)
    void parseArg(const std::string& s) {
        if (s.rfind("--", 0) != 0) return;
        const auto eq = s.find('=');
        const std::string key = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
        if (eq == std::string::npos) { options.setOption(key, true); return; }
        const std::string v = s.substr(eq + 1);
        if (v == "true" || v == "false") { options.setOption(key, v == "true"); return; }
        char* end = nullptr;
        const long iv = std::strtol(v.c_str(), &end, 10);
        if (end && *end == '\0') { options.setOption(key, (int)iv); return; }
        const double dv = std::strtod(v.c_str(), &end);
        if (end && *end == '\0') { options.setOption(key, dv); return; }
        options.setOption(key, v);
    }
    bool verbose_ = false;
};

// One device context per process, created on first use.
inline ef_ctx defaultContext() {
    static std::shared_ptr<ef_ctx_s> ctx;
    if (!ctx) {
        ef_ctx c = nullptr;
        efCheck(ef_ctx_create(0, &c), "defaultContext");
        ctx.reset(c, [](ef_ctx p) { ef_ctx_destroy(p); });
    }
    return ctx.get();
}

// ---------------------------------------------------------------------------
// Quadtree nodes and traversals (Quadtree.hpp:41-513, QuadNode.hpp:10-164): nodes in
// pre-order, children 0..3 = SW, SE, NW, NE, path = "0" + sibling ids.
// ---------------------------------------------------------------------------
template <typename PatchType>
struct Node {
    PatchType data;
    std::string path;
    int level = 0;
    bool leaf = false;
    int index = -1;       // pre-order node index
    int children[4] = {-1, -1, -1, -1};
};

template <typename PatchType>
class Quadtree {
public:
    std::vector<Node<PatchType>> nodes;  // pre-order
    PatchType& root() { return nodes.at(0).data; }
    void traversePreOrder(std::function<int(Node<PatchType>*)> fn) {
        for (auto& nd : nodes) if (!fn(&nd)) return;
    }
    void traversePreOrder(std::function<void(PatchType&)> fn) {
        for (auto& nd : nodes) fn(nd.data);
    }
    void traversePostOrder(std::function<int(Node<PatchType>*)> fn) { post(0, fn); }
    void traversePostOrder(std::function<void(PatchType&)> fn) {
        post(0, [&](Node<PatchType>* n) { fn(n->data); return 1; });
    }
private:
    void post(int i, const std::function<int(Node<PatchType>*)>& fn) {
        if (i < 0 || nodes.empty()) return;
        for (int c : nodes[i].children) post(c, fn);
        fn(&nodes[i]);
    }
};

template <typename PatchType> class AbstractNodeFactory;

// Mesh: refineByFunction (Mesh.hpp:111-180) through ef_refine_by_function, which restates
// p4est's uniform-refine-balance sequence; iteratePatches/iterateCells in pre-order.
template <typename PatchType>
class Mesh {
public:
    Quadtree<PatchType> quadtree;
    std::vector<int> leaf_level, leaf_ix, leaf_iy;  // leaves in z-order

    Mesh() = default;

    template <typename Factory>
    void refineByFunction(std::function<bool(double x, double y)> fn, double threshold, int min_level, int max_level,
                          PatchType& root_patch, Factory& node_factory) {
        (void)threshold;
        auto& g = root_patch.grid();
        struct Ctx { std::function<bool(double, double)>* fn; };
        Ctx c{&fn};
        auto cb = [](const double* x, const double* y, int n, void* user) -> int {
            auto* cc = (Ctx*)user;
            for (int i = 0; i < n; i++) if ((*cc->fn)(x[i], y[i])) return 1;
            return 0;
        };
        int nleaves = 0;
        efCheck(ef_refine_by_function(cb, &c, (int)g.nx(), (int)g.ny(), g.xLower(), g.xUpper(), g.yLower(), g.yUpper(),
                                      min_level, max_level, 0, nullptr, nullptr, nullptr, &nleaves),
                "Mesh::refineByFunction");
        leaf_level.assign(nleaves, 0); leaf_ix.assign(nleaves, 0); leaf_iy.assign(nleaves, 0);
        efCheck(ef_refine_by_function(cb, &c, (int)g.nx(), (int)g.ny(), g.xLower(), g.xUpper(), g.yLower(), g.yUpper(),
                                      min_level, max_level, nleaves, leaf_level.data(), leaf_ix.data(), leaf_iy.data(),
                                      &nleaves),
                "Mesh::refineByFunction");
        buildTree(root_patch, node_factory);
    }

    // uniform mesh at `level` (the quadrant list p4est_new_ext(level, fill_uniform) makes)
    template <typename Factory>
    void uniform(int level, PatchType& root_patch, Factory& node_factory) {
        leaf_level.clear(); leaf_ix.clear(); leaf_iy.clear();
        zorderUniform(0, 0, 0, level);
        buildTree(root_patch, node_factory);
    }

    void iteratePatches(std::function<void(PatchType&)> fn) {
        for (auto& nd : quadtree.nodes) if (nd.leaf) fn(nd.data);
    }
    void iterateCells(std::function<void(double, double)> fn) {
        for (auto& nd : quadtree.nodes) {
            if (!nd.leaf) continue;
            auto& g = nd.data.grid();
            for (std::size_t i = 0; i < g.nx(); i++)
                for (std::size_t j = 0; j < g.ny(); j++) fn(g(0, i), g(1, j));
        }
    }
    int nLeaves() const { return (int)leaf_level.size(); }

    // Mesh functions written as cell data by toVTK (Mesh.hpp:352-385): a vector over all leaf
    // cells in iterateCells order, or a function sampled at the cell centres
    void addMeshFunction(Vector<double>& meshFunction) { mesh_functions.push_back(&meshFunction); }
    void addMeshFunction(std::function<double(double x, double y)> meshFunction, std::string meshFunctionName) {
        owned_functions_.emplace_back();
        Vector<double>& v = owned_functions_.back();
        v.name() = meshFunctionName;
        iterateCells([&](double xc, double yc) { v.append(meshFunction(xc, yc)); });
        mesh_functions.push_back(&v);
    }
    void clearMeshFunctions() {
        mesh_functions.clear();
        owned_functions_.clear();
    }

    // Mesh.hpp:393-424: "<base>-mesh[-NNNN].pvtu" with "<base>-mesh[-NNNN]_0000.vtu" (every leaf
    // cell as its own quadrilateral, the mesh functions as cell data; the same text as the
    // reference's writer), and "<base>-quadtree[-NNNN].pvtu" with its _0000.vtu (one pixel per
    // leaf patch: the geometry of p4est's quadtree file, written here in ASCII)
    void toVTK(std::string baseFilename, int number = -1) {
        auto numbered = [&](std::string f) {
            if (number != -1) {
                char buffer[32];
                std::snprintf(buffer, 32, "%04i", number);
                f += "-" + std::string(buffer);
            }
            return f;
        };
        writeMeshVTK(numbered(baseFilename + "-mesh"));
        writeQuadtreeVTK(numbered(baseFilename + "-quadtree"));
    }

    std::vector<Vector<double>*> mesh_functions;

private:
    std::deque<Vector<double>> owned_functions_;

    // points, connectivity, offsets and types as Mesh.hpp:186-265 builds them: leaves in
    // pre-order, cells i-major, four corners per cell counter-clockwise from (x_i, y_j)
    void writeMeshVTK(const std::string& name) {
        Vector<double> points;
        points.name() = "points";
        points.setNumberOfComponents("3");
        Vector<int> connectivity, offsets, types;
        connectivity.name() = "connectivity";
        connectivity.setType("Int64");
        offsets.name() = "offsets";
        offsets.setType("Int64");
        types.name() = "types";
        types.setType("Int64");
        int index = 0, npoints = 0, ncells = 0;
        for (auto& nd : quadtree.nodes) {
            if (!nd.leaf) continue;
            auto& g = nd.data.grid();
            const double xLower = g.xLower(), yLower = g.yLower(), dx = g.dx(), dy = g.dy();
            const int nx = (int)g.nx(), ny = (int)g.ny();
            for (int i = 0; i < nx; i++)
                for (int j = 0; j < ny; j++) {
                    const double x0 = xLower + i * dx, y0 = yLower + j * dy;
                    const double x1 = xLower + (i + 1) * dx, y1 = yLower + (j + 1) * dy;
                    points.append(Vector<double>{x0, y0, 0.0, x1, y0, 0.0, x1, y1, 0.0, x0, y1, 0.0});
                    connectivity.append(Vector<int>{index, index + 1, index + 2, index + 3});
                    offsets.append(Vector<int>{(index + 1) * 4, (index + 2) * 4, (index + 3) * 4, (index + 4) * 4});
                    types.append(Vector<int>{9, 9, 9, 9});
                    index += 4;
                    npoints += 4;
                    ncells += 1;
                }
        }
        std::string names;
        for (auto* f : mesh_functions) names += f->getName() + " ";
        // the piece (VTK.cpp:160-207)
        vtk::XmlNode vtu = vtk::vtkFile();
        vtk::XmlNode piece{"Piece", {{"NumberOfPoints", std::to_string(npoints)}, {"NumberOfCells", std::to_string(ncells)}}, "", {}};
        piece.children.push_back(vtk::XmlNode{"Points", {}, "", {vtk::dataArray(points)}});
        piece.children.push_back(vtk::XmlNode{"Cells", {}, "", {vtk::dataArray(connectivity), vtk::dataArray(offsets), vtk::dataArray(types)}});
        if (!mesh_functions.empty()) {
            vtk::XmlNode cd{"CellData", {{"Scalars", names}}, "", {}};
            for (auto* f : mesh_functions) cd.children.push_back(vtk::dataArray(*f));
            piece.children.push_back(cd);
        }
        vtu.children.push_back(vtk::XmlNode{"UnstructuredGrid", {}, "", {piece}});
        char buffer[256];
        std::snprintf(buffer, 256, (name + "_%04i.vtu").c_str(), 0);
        const std::string source(buffer);
        vtk::writeTree(source, vtu);
        // the parallel header (VTK.cpp:239-300), one piece per rank
        vtk::XmlNode pvtu = vtk::vtkFile();
        vtk::XmlNode pgrid{"PUnstructuredGrid", {{"GhostLevel", "0"}}, "", {}};
        pgrid.children.push_back(vtk::XmlNode{"PPoints", {}, "", {vtk::pDataArray(points)}});
        pgrid.children.push_back(vtk::XmlNode{"PPointData", {}, "", {}});
        vtk::XmlNode pcd{"PCellData", {}, "", {}};
        if (!mesh_functions.empty()) pcd.attributes.push_back({"Scalars", names});
        for (auto* f : mesh_functions) pcd.children.push_back(vtk::pDataArray(*f));
        pgrid.children.push_back(pcd);
        pgrid.children.push_back(vtk::XmlNode{"Piece", {{"Source", source}}, "", {}});
        pvtu.children.push_back(pgrid);
        vtk::writeTree(name + ".pvtu", pvtu);
    }

    // one VTK_PIXEL (type 8, corners in z-order) per leaf patch, as p4est_vtk_write_header writes
    // the quadrants with scale 1
    void writeQuadtreeVTK(const std::string& name) {
        Vector<double> points;
        points.name() = "Position";
        points.setNumberOfComponents("3");
        Vector<int> connectivity, offsets, types;
        connectivity.name() = "connectivity";
        connectivity.setType("Int32");
        offsets.name() = "offsets";
        offsets.setType("Int32");
        types.name() = "types";
        types.setType("UInt8");
        int k = 0;
        for (auto& nd : quadtree.nodes) {
            if (!nd.leaf) continue;
            auto& g = nd.data.grid();
            const double xl = g.xLower(), xu = g.xUpper(), yl = g.yLower(), yu = g.yUpper();
            points.append(Vector<double>{xl, yl, 0.0, xu, yl, 0.0, xl, yu, 0.0, xu, yu, 0.0});
            connectivity.append(Vector<int>{4 * k, 4 * k + 1, 4 * k + 2, 4 * k + 3});
            offsets.append(4 * (k + 1));
            types.append(8);
            k++;
        }
        vtk::XmlNode vtu = vtk::vtkFile();
        vtk::XmlNode piece{"Piece", {{"NumberOfPoints", std::to_string(4 * k)}, {"NumberOfCells", std::to_string(k)}}, "", {}};
        piece.children.push_back(vtk::XmlNode{"Points", {}, "", {vtk::dataArray(points)}});
        piece.children.push_back(vtk::XmlNode{"Cells", {}, "", {vtk::dataArray(connectivity), vtk::dataArray(offsets), vtk::dataArray(types)}});
        vtu.children.push_back(vtk::XmlNode{"UnstructuredGrid", {}, "", {piece}});
        const std::string source = name + "_0000.vtu";
        vtk::writeTree(source, vtu);
        vtk::XmlNode pvtu = vtk::vtkFile();
        vtk::XmlNode pgrid{"PUnstructuredGrid", {{"GhostLevel", "0"}}, "", {}};
        pgrid.children.push_back(vtk::XmlNode{"PPoints", {}, "", {vtk::pDataArray(points)}});
        pgrid.children.push_back(vtk::XmlNode{"Piece", {{"Source", source}}, "", {}});
        pvtu.children.push_back(pgrid);
        vtk::writeTree(name + ".pvtu", pvtu);
    }

    void zorderUniform(int l, int x, int y, int level) {
        if (l == level) { leaf_level.push_back(l); leaf_ix.push_back(x); leaf_iy.push_back(y); return; }
        for (int c = 0; c < 4; c++) zorderUniform(l + 1, 2 * x + (c & 1), 2 * y + (c >> 1), level);
    }
    template <typename Factory>
    void buildTree(PatchType& root_patch, Factory& factory) {
        quadtree.nodes.clear();
        std::size_t cursor = 0;
        Node<PatchType> root;
        root.data = root_patch;
        root.path = "0";
        quadtree.nodes.push_back(root);
        grow(0, 0, 0, 0, cursor, factory);
        if (cursor != leaf_level.size()) throw std::invalid_argument("[EllipticForest::Mesh] leaf list is not a quadtree");
    }
    template <typename Factory>
    void grow(int id, int l, int x, int y, std::size_t& cursor, Factory& factory) {
        quadtree.nodes[id].level = l;
        quadtree.nodes[id].index = id;
        if (cursor < leaf_level.size() && leaf_level[cursor] == l) {
            if (leaf_ix[cursor] != x || leaf_iy[cursor] != y) throw std::invalid_argument("[EllipticForest::Mesh] bad leaf order");
            quadtree.nodes[id].leaf = true;
            cursor++;
            return;
        }
        for (int c = 0; c < 4; c++) {
            Node<PatchType> child;
            child.data = factory.createChildPatch(quadtree.nodes[id].data, c);
            child.path = quadtree.nodes[id].path + std::to_string(c);
            quadtree.nodes.push_back(child);
            const int cid = (int)quadtree.nodes.size() - 1;
            quadtree.nodes[id].children[c] = cid;
            grow(cid, l + 1, 2 * x + (c & 1), 2 * y + (c >> 1), cursor, factory);
        }
    }
};

}  // namespace EllipticForest

#include "HPSAlgorithm.hpp"

#endif  // ELLIPTIC_FOREST_HPP_
