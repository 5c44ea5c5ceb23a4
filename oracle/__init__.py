"""CPU FP64 oracle of EllipticForest's HPS factor-and-solve path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()`` and
by ``bench.py``'s CPU-baseline leg, and only as the checker or the timed CPU
baseline.  The product (``ellipticforest_amd``) never imports it.

``libhps_oracle.so`` (built from ``hps_oracle.cpp`` by ``oracle/Makefile``) restates
src/HPSAlgorithm.hpp and src/Patches/FiniteVolume/FiniteVolumeSolver.cpp step by
step; ``mesh.py`` restates Mesh::refineByFunction.  This module is the ctypes view
of the library plus the host-side sampling of the reference drivers:

* f at leaf cell centres, index ``j + i*ny``   (HPSAlgorithm.hpp:241-249)
* root boundary data at root-grid face points  (HPSAlgorithm.hpp:375-400)
* alpha / beta-at-faces / lambda per cell      (FiniteVolumeSolver.cpp:64-78)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import mesh  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhps_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

FIVE_POINT = 0
FISHPACK = 1


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s"] + (["-B"] if force else []), check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_create.argtypes = [ctypes.c_int, _ip, _ip, _ip, ctypes.c_int] + [ctypes.c_double] * 4 + [
            ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        for name in ("orc_num_nodes", "orc_num_leaves", "orc_root_size", "orc_build"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_destroy.restype = None
        L.orc_set_options.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_set_leaf_coeffs.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 6
        L.orc_node_info.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp]
        L.orc_upwards.argtypes = [ctypes.c_void_p, _dp]
        L.orc_solve.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip]
        L.orc_dgesv.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_dgemm.argtypes = [ctypes.c_int] * 3 + [_dp] * 3
        L.orc_block_permute.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip,
                                        ctypes.c_int, _ip, _ip, _dp]
        L.orc_interp.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        leaf_args = [ctypes.c_int] + [ctypes.c_double] * 4 + [_dp, ctypes.c_double]
        L.orc_leaf_solve.argtypes = leaf_args + [_dp, _dp, _dp]
        L.orc_leaf_map_d2n.argtypes = leaf_args + [_dp, _dp, _dp]
        L.orc_leaf_dtn.argtypes = leaf_args + [_dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().orc_set_threads(int(n))


# ----- building blocks -----------------------------------------------------------
def dgesv(A, B):
    A = _f64(A)
    B = _f64(B).copy()
    vec = B.ndim == 1
    B2 = B.reshape(A.shape[0], -1).copy()
    _check(lib().orc_dgesv(A.shape[0], B2.shape[1], _d(A), _d(B2)))
    return B2.ravel() if vec else B2


def dgemm(A, B):
    A, B = _f64(A), _f64(B)
    C = np.zeros((A.shape[0], B.shape[1]))
    _check(lib().orc_dgemm(A.shape[0], A.shape[1], B.shape[1], _d(A), _d(B), _d(C)))
    return C


def block_permute(A, pi_rows, pi_cols, R, C):
    A = _f64(A)
    pr, pc = np.asarray(pi_rows, np.int32), np.asarray(pi_cols, np.int32)
    Rr, Cc = np.asarray(R, np.int32), np.asarray(C, np.int32)
    out = np.zeros_like(A)
    _check(lib().orc_block_permute(A.shape[0], A.shape[1], _d(A), len(pr), _i(pr), _i(Rr), len(pc), _i(pc),
                                   _i(Cc), _d(out)))
    return out


def interp_fine2coarse(ncoarse):
    out = np.zeros((ncoarse, 2 * ncoarse))
    _check(lib().orc_interp(0, ncoarse, _d(out)))
    return out


def interp_coarse2fine(nfine):
    out = np.zeros((nfine, nfine // 2))
    _check(lib().orc_interp(1, nfine, _d(out)))
    return out


def _coeff_ptr(coeffs):
    if coeffs is None:
        return None, ctypes.cast(None, _dp)
    c = _f64(coeffs)
    return c, _d(c)


def leaf_solve(N, bounds, g, f, coeffs=None, lambda0=0.0):
    keep, cp = _coeff_ptr(coeffs)
    g, f = _f64(g), _f64(f)
    u = np.zeros(N * N)
    _check(lib().orc_leaf_solve(N, *bounds, cp, lambda0, _d(g), _d(f), _d(u)))
    return u


def leaf_map_d2n(N, bounds, g, f, coeffs=None, lambda0=0.0):
    keep, cp = _coeff_ptr(coeffs)
    g, f = _f64(g), _f64(f)
    h = np.zeros(4 * N)
    _check(lib().orc_leaf_map_d2n(N, *bounds, cp, lambda0, _d(g), _d(f), _d(h)))
    return h


def leaf_dtn(N, bounds, coeffs=None, lambda0=0.0):
    keep, cp = _coeff_ptr(coeffs)
    T = np.zeros((4 * N, 4 * N))
    _check(lib().orc_leaf_dtn(N, *bounds, cp, lambda0, _d(T)))
    return T


# ----- host-side sampling, as the reference drivers do it -------------------------
def cell_centres(N, bounds):
    """FiniteVolumeGrid::point (FiniteVolumeGrid.cpp:24-36): xl + dx/2 + i*dx."""
    xl, xu, yl, yu = bounds
    dx, dy = (xu - xl) / N, (yu - yl) / N
    xs = (xl + dx / 2) + np.arange(N) * dx
    ys = (yl + dy / 2) + np.arange(N) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")  # index j + i*ny
    return X.ravel(), Y.ravel(), dx, dy


def sample_coeffs(N, bounds, alpha, beta, lam):
    """alpha(x_i,y_j), beta at the four faces, lambda(x_i,y_j) (FiniteVolumeSolver.cpp:64-78)."""
    X, Y, dx, dy = cell_centres(N, bounds)
    out = np.empty((6, N * N))
    out[0] = alpha(X, Y)
    out[1] = beta(X - dx / 2.0, Y)
    out[2] = beta(X + dx / 2.0, Y)
    out[3] = beta(X, Y - dy / 2.0)
    out[4] = beta(X, Y + dy / 2.0)
    out[5] = lam(X, Y)
    return out


def root_face_points(M, bounds):
    """Root-grid face points, side n = 0..3 (W, E, S, N), HPSAlgorithm.hpp:375-400."""
    xl, xu, yl, yu = bounds
    dx, dy = (xu - xl) / M, (yu - yl) / M
    ys = (yl + dy / 2) + np.arange(M) * dy
    xs = (xl + dx / 2) + np.arange(M) * dx
    px = np.concatenate([np.full(M, xl), np.full(M, xu), xs, xs])
    py = np.concatenate([ys, ys, np.full(M, yl), np.full(M, yu)])
    side = np.repeat(np.arange(4), M)
    return side, px, py


class Forest:
    """One HPS forest in the oracle: leaves (level, ix, iy) in z-order, N x N cells each."""

    def __init__(self, leaves, N, bounds, solver=FISHPACK, lambda0=0.0, coeff_fns=None,
                 cache_operators=False, homogeneous_rhs=False):
        L = lib()
        leaves = list(leaves)
        lv = np.array([l[0] for l in leaves], np.int32)
        ix = np.array([l[1] for l in leaves], np.int32)
        iy = np.array([l[2] for l in leaves], np.int32)
        h = ctypes.c_void_p()
        _check(L.orc_create(len(leaves), _i(lv), _i(ix), _i(iy), N, *bounds, solver, ctypes.byref(h)))
        self.h = h
        self.N = N
        self.bounds = tuple(bounds)
        self.solver = solver
        self.leaves = leaves
        _check(L.orc_set_options(h, int(cache_operators), int(homogeneous_rhs), float(lambda0)))
        self.leaf_bounds = [self.node_info(n)[1] for n in self.leaf_nodes()]
        if solver == FIVE_POINT:
            alpha, beta, lam = coeff_fns
            for k, b in enumerate(self.leaf_bounds):
                c = sample_coeffs(N, b, alpha, beta, lam)
                _check(L.orc_set_leaf_coeffs(h, k, *[_d(c[t]) for t in range(6)]))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_destroy(self.h)
            self.h = None

    def num_nodes(self):
        return lib().orc_num_nodes(self.h)

    def node_info(self, node):
        info = np.zeros(12, np.int32)
        b = np.zeros(4)
        _check(lib().orc_node_info(self.h, node, _i(info), _d(b)))
        d = dict(level=int(info[0]), ix=int(info[1]), iy=int(info[2]), leaf=bool(info[3]),
                 leaf_index=int(info[4]), n_coarsens=int(info[5]), nx=int(info[6]),
                 children=[int(c) for c in info[7:11]], parent=int(info[11]))
        return d, tuple(b)

    def leaf_nodes(self):
        return [n for n in range(self.num_nodes()) if self.node_info(n)[0]["leaf"]]

    def root_size(self):
        return lib().orc_root_size(self.h)

    def build(self):
        _check(lib().orc_build(self.h))

    def sample_rhs(self, fn):
        """f at leaf cell centres, leaves in z-order (HPSAlgorithm.hpp:241-249)."""
        out = np.empty((len(self.leaf_bounds), self.N * self.N))
        for k, b in enumerate(self.leaf_bounds):
            X, Y, _, _ = cell_centres(self.N, b)
            out[k] = fn(X, Y)
        return out

    def upwards(self, f_all):
        f_all = _f64(f_all)
        _check(lib().orc_upwards(self.h, _d(f_all)))

    def boundary_data(self, bc_fn):
        """bc_fn(side, x, y) -> (r, a, b) arrays at root face points."""
        M = self.root_size()
        side, px, py = root_face_points(M, self.bounds)
        r, a, b = bc_fn(side, px, py)
        return (_f64(np.broadcast_to(r, px.shape)), _f64(np.broadcast_to(a, px.shape)),
                _f64(np.broadcast_to(b, px.shape)))

    def solve(self, r, a, b):
        r, a, b = _f64(r), _f64(a), _f64(b)
        u = np.zeros((len(self.leaf_bounds), self.N * self.N))
        _check(lib().orc_solve(self.h, _d(r), _d(a), _d(b), _d(u)))
        return u

    def get(self, node, which):
        codes = dict(T=0, S=1, X=2, H=3, u=4, g=5, f=6, h=7, w=8)
        r, c = ctypes.c_int(), ctypes.c_int()
        L = lib()
        _check(L.orc_get(self.h, node, codes[which], ctypes.cast(None, _dp), ctypes.byref(r), ctypes.byref(c)))
        out = np.zeros((r.value, c.value))
        _check(L.orc_get(self.h, node, codes[which], _d(out), ctypes.byref(r), ctypes.byref(c)))
        return out if which in ("T", "S", "X", "H") else out.ravel()
