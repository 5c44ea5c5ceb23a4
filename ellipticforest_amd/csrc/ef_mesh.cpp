// ef_mesh.cpp -- quadtree refinement for the drop-in Mesh::refineByFunction.
//
// Reproduces the forest that src/Mesh.hpp:111-180 obtains from p4est on one tree:
//   p4est_new_ext(min_level, fill_uniform)        -> uniform quadrants at min_level
//   p4est_refine(recursive, cb)                   -> split while level < max_level and
//                                                    the callback marks a cell centre
//   p4est_balance(P4EST_CONNECT_CORNER)           -> 2:1 balance over faces and corners
// and returns the leaves in z-order (children SW, SE, NW, NE), the order p4est stores
// local quadrants in and Mesh::iteratePatches visits them.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "ef_hps.h"

namespace efk {
void set_last_error(const std::string& msg);  // ef_api.hip
}

namespace {

inline uint64_t key(int l, long long x, long long y) {
    return ((uint64_t)l << 58) | ((uint64_t)x << 29) | (uint64_t)y;
}

struct Forest {
    std::unordered_set<uint64_t> leaves;
    bool has(int l, long long x, long long y) const { return leaves.count(key(l, x, y)) != 0; }
    // leaf covering the level-l quadrant (x, y), as its level, or -1
    int cover(int l, long long x, long long y) const {
        for (int a = l; a >= 0; a--)
            if (has(a, x >> (l - a), y >> (l - a))) return a;
        return -1;
    }
};

void refine(Forest& F, int l, long long x, long long y, ef_refine_fn fn, void* user, int nx, int ny, double xl,
            double xu, double yl, double yu, int max_level, std::vector<double>& px, std::vector<double>& py) {
    if (l < max_level) {
        // quadrant corners through the affine map of the single square tree (P4est.cpp:7-33)
        const double h = 1.0 / (double)(1LL << l);
        const double x0 = xl + (xu - xl) * (x * h), x1 = xl + (xu - xl) * ((x + 1) * h);
        const double y0 = yl + (yu - yl) * (y * h), y1 = yl + (yu - yl) * ((y + 1) * h);
        const double dx = (x1 - x0) / nx, dy = (y1 - y0) / ny;
        int t = 0;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++, t++) {
                px[t] = x0 + (i + 0.5) * dx;
                py[t] = y0 + (j + 0.5) * dy;
            }
        if (fn(px.data(), py.data(), nx * ny, user)) {
            for (int c = 0; c < 4; c++)
                refine(F, l + 1, 2 * x + (c & 1), 2 * y + (c >> 1), fn, user, nx, ny, xl, xu, yl, yu, max_level, px, py);
            return;
        }
    }
    F.leaves.insert(key(l, x, y));
}

// Ripple balance: a leaf at level l whose face or corner neighbour region is covered
// by a leaf coarser than l-1 forces that leaf to split; repeat to a fixed point.
void balance(Forest& F) {
    for (;;) {
        std::vector<uint64_t> split;
        for (uint64_t kk : F.leaves) {
            const int l = (int)(kk >> 58);
            const long long x = (long long)((kk >> 29) & ((1ULL << 29) - 1)), y = (long long)(kk & ((1ULL << 29) - 1));
            if (l < 2) continue;
            const long long n = 1LL << l;
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++) {
                    if (!dx && !dy) continue;
                    const long long qx = x + dx, qy = y + dy;
                    if (qx < 0 || qy < 0 || qx >= n || qy >= n) continue;
                    const int a = F.cover(l, qx, qy);
                    if (a >= 0 && a < l - 1) split.push_back(key(a, qx >> (l - a), qy >> (l - a)));
                }
        }
        if (split.empty()) return;
        std::sort(split.begin(), split.end());
        split.erase(std::unique(split.begin(), split.end()), split.end());
        for (uint64_t kk : split) {
            if (!F.leaves.erase(kk)) continue;
            const int l = (int)(kk >> 58);
            const long long x = (long long)((kk >> 29) & ((1ULL << 29) - 1)), y = (long long)(kk & ((1ULL << 29) - 1));
            for (int c = 0; c < 4; c++) F.leaves.insert(key(l + 1, 2 * x + (c & 1), 2 * y + (c >> 1)));
        }
    }
}

void zorder(const Forest& F, int l, long long x, long long y, std::vector<int>& L, std::vector<int>& X, std::vector<int>& Y) {
    if (F.has(l, x, y)) {
        L.push_back(l); X.push_back((int)x); Y.push_back((int)y);
        return;
    }
    if (l >= 28) throw std::runtime_error("leaf set does not tile the domain");
    for (int c = 0; c < 4; c++) zorder(F, l + 1, 2 * x + (c & 1), 2 * y + (c >> 1), L, X, Y);
}

}  // namespace

extern "C" int ef_refine_by_function(ef_refine_fn fn, void* user, int nx, int ny, double xl, double xu, double yl,
                                     double yu, int min_level, int max_level, int capacity, int* level, int* ix,
                                     int* iy, int* nleaves) {
    try {
        if (!fn || !nleaves || nx < 1 || ny < 1 || min_level < 0 || max_level < min_level || max_level > 24)
            throw std::invalid_argument("invalid refinement arguments");
        Forest F;
        std::vector<double> px((size_t)nx * ny), py((size_t)nx * ny);
        const long long n0 = 1LL << min_level;
        for (long long x = 0; x < n0; x++)
            for (long long y = 0; y < n0; y++) refine(F, min_level, x, y, fn, user, nx, ny, xl, xu, yl, yu, max_level, px, py);
        balance(F);
        std::vector<int> L, X, Y;
        zorder(F, 0, 0, 0, L, X, Y);
        *nleaves = (int)L.size();
        if (capacity < (int)L.size()) return capacity == 0 ? EF_OK : EF_ERR_ARG;
        std::copy(L.begin(), L.end(), level);
        std::copy(X.begin(), X.end(), ix);
        std::copy(Y.begin(), Y.end(), iy);
        return EF_OK;
    } catch (const std::exception& e) {
        efk::set_last_error(e.what());
        return EF_ERR_ARG;
    }
}
