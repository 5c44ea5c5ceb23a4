"""ctypes binding of libefhps.so (declared in include/ef_hps.h).

There is no CPU fallback: if the library or a GPU is missing, loading or creating a
context raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefhps.so")

EF_OK = 0
EF_ERR_ARG = 1
EF_ERR_STATE = 2
EF_ERR_DEVICE = 3
EF_ERR_SINGULAR = 4
EF_ERR_INTERNAL = 5

EF_SOLVER_FIVE_POINT = 0
EF_SOLVER_FISHPACK = 1
EF_PRECISION_FP64 = 0
EF_PRECISION_MIXED = 1
EF_FOOTPRINT_VARIABLE = 0x100

# Every symbol include/ef_hps.h declares.
EXPORTS = [
    "ef_ctx_create", "ef_ctx_destroy", "ef_last_error", "ef_version", "ef_refine_by_function",
    "ef_mesh_create", "ef_mesh_destroy", "ef_mesh_footprint", "ef_mesh_info", "ef_mesh_nodes",
    "ef_mesh_set_stream", "ef_mesh_stream", "ef_mesh_set_options", "ef_build", "ef_upwards", "ef_upwards_device",
    "ef_solve", "ef_solve_device", "ef_get_operator", "ef_mesh_sync", "ef_device_alloc", "ef_device_free",
    "ef_memcpy", "ef_debug_invert", "ef_host_alloc", "ef_host_free", "ef_mesh_copy_async", "ef_mesh_record",
    "ef_mesh_query", "ef_mesh_wait", "ef_mesh_eval_load", "ef_mesh_time_sweep_gemms", "ef_mesh_set_refinement",
    "ef_mesh_set_autotune", "ef_mesh_sweep_tiles",
    "ef_mesh_stage_ms", "ef_patches_create", "ef_patches_destroy", "ef_patches_set_stream",
    "ef_patches_info", "ef_patches_apply", "ef_patches_apply_device", "ef_patches_apply_device_noalloc",
    "ef_patches_dtn", "ef_patches_sync",
    "ef_patches_last_ms", "ef_patches_stream", "ef_timer_create", "ef_timer_destroy", "ef_timer_record",
    "ef_timer_elapsed_ms",
]

EF_PATCH_SOLVE = 0
EF_PATCH_MAP_D2N = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

REFINE_FN = ctypes.CFUNCTYPE(ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_void_p)


class MeshDesc(ctypes.Structure):
    _fields_ = [("nleaves", ctypes.c_int), ("level", _ip), ("ix", _ip), ("iy", _ip), ("N", ctypes.c_int),
                ("xl", ctypes.c_double), ("xu", ctypes.c_double), ("yl", ctypes.c_double), ("yu", ctypes.c_double)]


class Coeffs(ctypes.Structure):
    _fields_ = [("solver_type", ctypes.c_int), ("alpha0", ctypes.c_double), ("beta0", ctypes.c_double),
                ("lambda0", ctypes.c_double), ("cell", _dp)]


class PatchBatch(ctypes.Structure):
    _fields_ = [("npatches", ctypes.c_int), ("N", ctypes.c_int), ("bounds", _dp)]


class EfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[ef error {code}] {msg}")
        self.code = code


_lib = None


def load_library(path):
    """Open one build of the C ABI and declare its signatures."""
    L = ctypes.CDLL(path)
    L.ef_last_error.restype = ctypes.c_char_p
    L.ef_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.ef_ctx_destroy.argtypes = [_vp]
    L.ef_refine_by_function.argtypes = [REFINE_FN, _vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip]
    L.ef_mesh_create.argtypes = [_vp, ctypes.POINTER(MeshDesc), ctypes.POINTER(_vp)]
    L.ef_mesh_destroy.argtypes = [_vp]
    L.ef_mesh_footprint.argtypes = [ctypes.POINTER(MeshDesc), ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_size_t)]
    L.ef_mesh_info.argtypes = [_vp, _ip]
    L.ef_mesh_nodes.argtypes = [_vp, _ip, _dp]
    L.ef_mesh_set_stream.argtypes = [_vp, _vp]
    L.ef_mesh_stream.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.ef_mesh_set_options.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.ef_build.argtypes = [_vp, ctypes.POINTER(Coeffs), ctypes.c_int]
    L.ef_upwards.argtypes = [_vp, _dp, ctypes.c_int]
    L.ef_upwards_device.argtypes = [_vp, _vp, ctypes.c_int]
    L.ef_solve.argtypes = [_vp, _dp, _dp, _dp, _dp, ctypes.c_int]
    L.ef_solve_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
    L.ef_get_operator.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip, _ip]
    L.ef_mesh_sync.argtypes = [_vp]
    L.ef_device_alloc.argtypes = [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.ef_device_free.argtypes = [_vp, _vp]
    L.ef_memcpy.argtypes = [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]
    L.ef_debug_invert.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]
    L.ef_host_alloc.argtypes = [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.ef_host_free.argtypes = [_vp, _vp]
    L.ef_mesh_copy_async.argtypes = [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]
    L.ef_mesh_record.argtypes = [_vp, ctypes.c_int]
    L.ef_mesh_query.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.ef_mesh_wait.argtypes = [_vp, ctypes.c_int]
    L.ef_mesh_eval_load.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
    L.ef_mesh_time_sweep_gemms.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip]
    L.ef_mesh_set_refinement.argtypes = [_vp, ctypes.c_int]
    L.ef_mesh_set_autotune.argtypes = [_vp, ctypes.c_int]
    L.ef_mesh_sweep_tiles.argtypes = [_vp, ctypes.c_int, _ip, _ip]
    L.ef_mesh_stage_ms.argtypes = [_vp, _dp]
    L.ef_patches_create.argtypes = [_vp, ctypes.POINTER(PatchBatch), ctypes.POINTER(Coeffs), ctypes.POINTER(_vp)]
    L.ef_patches_destroy.argtypes = [_vp]
    L.ef_patches_set_stream.argtypes = [_vp, _vp]
    L.ef_patches_info.argtypes = [_vp, _ip]
    L.ef_patches_apply.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    L.ef_patches_apply_device.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int]
    L.ef_patches_apply_device_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int]
    L.ef_patches_dtn.argtypes = [_vp, _dp]
    L.ef_patches_sync.argtypes = [_vp]
    L.ef_patches_last_ms.argtypes = [_vp, _dp]
    L.ef_patches_stream.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.ef_timer_create.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.ef_timer_destroy.argtypes = [_vp]
    L.ef_timer_record.argtypes = [_vp, _vp, ctypes.c_int]
    L.ef_timer_elapsed_ms.argtypes = [_vp, _dp]
    return L


def lib():
    """The product library, libefhps.so next to this file."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        _lib = load_library(LIB_PATH)
    return _lib


def check(rc, L=None):
    if rc != EF_OK:
        raise EfError(rc, (L or lib()).ef_last_error().decode())


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def refine_by_function(fn, nx, ny, bounds, min_level, max_level):
    """Leaves (level, ix, iy) in z-order; ``fn(x, y)`` is vectorised over one quadrant's cell centres."""
    L = lib()
    err = []

    def cb(px, py, npts, user):
        try:
            x = np.ctypeslib.as_array(px, shape=(npts,))
            y = np.ctypeslib.as_array(py, shape=(npts,))
            return 1 if bool(np.any(fn(x, y))) else 0
        except Exception as e:  # propagate after the C call returns
            err.append(e)
            return 0

    cfn = REFINE_FN(cb)
    n = ctypes.c_int(0)
    check(L.ef_refine_by_function(cfn, None, nx, ny, *bounds, min_level, max_level, 0, None, None, None,
                                  ctypes.byref(n)))
    if err:
        raise err[0]
    cap = n.value
    lv, ix, iy = (np.zeros(cap, np.int32) for _ in range(3))
    check(L.ef_refine_by_function(cfn, None, nx, ny, *bounds, min_level, max_level, cap, iptr(lv), iptr(ix),
                                  iptr(iy), ctypes.byref(n)))
    if err:
        raise err[0]
    return [(int(a), int(b), int(c)) for a, b, c in zip(lv, ix, iy)]


class Context:
    """A device context.  ``library`` selects a build of the C ABI (default: the product
    library); the test suite passes its CPU-emulation build here."""

    def __init__(self, device=0, library=None):
        self.L = library if library is not None else lib()
        h = _vp()
        check(self.L.ef_ctx_create(device, ctypes.byref(h)), self.L)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None:
            self.L.ef_ctx_destroy(self.h)
            self.h = None


class DeviceArray:
    """A float64 array in the context's device memory (ef_device_alloc)."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.size = int(np.prod(self.shape))
        p = _vp()
        check(ctx.L.ef_device_alloc(ctx.h, self.size * 8, ctypes.byref(p)), ctx.L)
        self.ptr = p.value

    def upload(self, host, offset=0):
        host = np.ascontiguousarray(host, dtype=np.float64)
        if offset + host.size > self.size:
            raise ValueError("upload past the end of the device array")
        check(self.ctx.L.ef_memcpy(self.ctx.h, self.ptr + 8 * offset, host.ctypes.data, host.size * 8, 0), self.ctx.L)

    def download(self, count=None):
        count = self.size if count is None else count
        out = np.empty(count)
        check(self.ctx.L.ef_memcpy(self.ctx.h, out.ctypes.data, self.ptr, count * 8, 1), self.ctx.L)
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            self.ctx.L.ef_device_free(self.ctx.h, self.ptr)
            self.ptr = None


class DeviceMesh:
    """One factorisable quadtree mesh on the GPU: the C-ABI handle plus host-side helpers."""

    def __init__(self, ctx, leaves, N, bounds):
        self.ctx = ctx
        self.N = N
        self.bounds = tuple(float(b) for b in bounds)
        self.leaves = list(leaves)
        self._lv = np.array([l[0] for l in leaves], np.int32)
        self._ix = np.array([l[1] for l in leaves], np.int32)
        self._iy = np.array([l[2] for l in leaves], np.int32)
        self.L = ctx.L
        d = MeshDesc(len(leaves), iptr(self._lv), iptr(self._ix), iptr(self._iy), N, *self.bounds)
        h = _vp()
        self._check(self.L.ef_mesh_create(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        info = np.zeros(8, np.int32)
        self._check(self.L.ef_mesh_info(h, iptr(info)))
        self.nnodes, self.nleaves, self.M, self.maxdepth = (int(v) for v in info[:4])
        self.n_internal, self.max_n = int(info[5]), int(info[6])
        ni = np.zeros(12 * self.nnodes, np.int32)
        nb = np.zeros(4 * self.nnodes)
        self._check(self.L.ef_mesh_nodes(h, iptr(ni), dptr(nb)))
        self.node_info = ni.reshape(-1, 12)
        self.node_bounds = nb.reshape(-1, 4)
        self.leaf_nodes = [i for i in range(self.nnodes) if self.node_info[i, 3]]

    def _check(self, rc):
        check(rc, self.L)

    def __del__(self):
        if getattr(self, "h", None) is not None:
            self.L.ef_mesh_destroy(self.h)
            self.h = None

    def set_options(self, cache_operators=False, homogeneous_rhs=False):
        self._check(self.L.ef_mesh_set_options(self.h, int(cache_operators), int(homogeneous_rhs)))

    def set_stream(self, stream_ptr):
        self._check(self.L.ef_mesh_set_stream(self.h, stream_ptr))

    @property
    def stream(self):
        """The hipStream_t (as an integer) the mesh's stages run on."""
        s = _vp()
        self._check(self.L.ef_mesh_stream(self.h, ctypes.byref(s)))
        return s.value

    def build(self, solver_type=EF_SOLVER_FISHPACK, lambda0=0.0, alpha0=1.0, beta0=1.0, cell=None,
              precision=EF_PRECISION_FP64):
        keep = None
        cp = ctypes.cast(None, _dp)
        if cell is not None:
            keep = f64(cell)
            cp = dptr(keep)
        c = Coeffs(solver_type, alpha0, beta0, lambda0, cp)
        self._check(self.L.ef_build(self.h, ctypes.byref(c), precision))
        del keep

    def set_autotune(self, mode):
        """Sweep GEMM tiles: 1/True timed (default), 0/False fixed rule, 11/12/22 one forced tile."""
        self._check(self.L.ef_mesh_set_autotune(self.h, int(mode)))

    def sweep_tiles(self, k):
        """Chosen tiles per depth (root first): [[w, h, u], ...] as rows x 10 + columns (units of 16)."""
        n = ctypes.c_int(0)
        self._check(self.L.ef_mesh_sweep_tiles(self.h, int(k), ctypes.cast(None, _ip), ctypes.byref(n)))
        out = np.zeros(3 * n.value, np.int32)
        self._check(self.L.ef_mesh_sweep_tiles(self.h, int(k), out.ctypes.data_as(_ip), ctypes.byref(n)))
        return out.reshape(-1, 3)

    def set_refinement(self, steps):
        """Iterative-refinement rounds of each solve of a mixed-precision build."""
        self._check(self.L.ef_mesh_set_refinement(self.h, int(steps)))

    def upwards(self, f):
        """f: [nrhs][nleaves][N*N] (or [nleaves][N*N] for one right-hand side)."""
        f = f64(f)
        k = 1 if f.ndim == 2 else f.shape[0]
        self._check(self.L.ef_upwards(self.h, dptr(f), k))

    def solve(self, r, a, b=None):
        """r: [nrhs][4M] (or [4M]); a, b: [4M].  Returns u [nrhs][nleaves][N*N] (or [nleaves][N*N])."""
        r = f64(r)
        single = r.ndim == 1
        k = 1 if single else r.shape[0]
        a = f64(np.broadcast_to(a, (4 * self.M,)))
        bp = ctypes.cast(None, _dp)
        if b is not None:
            b = f64(np.broadcast_to(b, (4 * self.M,)))
            bp = dptr(b)
        u = np.zeros((k, self.nleaves, self.N * self.N))
        self._check(self.L.ef_solve(self.h, dptr(r), dptr(a), bp, dptr(u), k))
        return u[0] if single else u

    def upwards_device(self, f_ptr, k):
        self._check(self.L.ef_upwards_device(self.h, f_ptr, k))

    def solve_device(self, f_ptr, r_ptr, a_ptr, b_ptr, u_ptr, k):
        self._check(self.L.ef_solve_device(self.h, f_ptr, r_ptr, a_ptr, b_ptr, u_ptr, k))

    def sync(self):
        self._check(self.L.ef_mesh_sync(self.h))

    def stage_ms(self):
        """HIP-event times of the last build, upwards and solve calls (ms; -1 if not run)."""
        out = np.zeros(3)
        self._check(self.L.ef_mesh_stage_ms(self.h, dptr(out)))
        return dict(zip(("build-stage", "upwards-stage", "solve-stage"), out.tolist()))

    CODES = dict(T=0, S=1, X=2, H=3, u=4, g=5, f=6, h=7, w=8)

    def get(self, node, which, rhs=0):
        code = self.CODES[which]
        r, c = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.ef_get_operator(self.h, node, code, rhs, ctypes.cast(None, _dp), ctypes.byref(r),
                                    ctypes.byref(c)))
        out = np.zeros((r.value, c.value))
        self._check(self.L.ef_get_operator(self.h, node, code, rhs, dptr(out), ctypes.byref(r), ctypes.byref(c)))
        return out if which in ("T", "S", "X", "H") else out.ravel()

    @staticmethod
    def footprint(leaves, N, bounds, precision=EF_PRECISION_FP64, max_rhs=1, variable=False):
        lv = np.array([l[0] for l in leaves], np.int32)
        ix = np.array([l[1] for l in leaves], np.int32)
        iy = np.array([l[2] for l in leaves], np.int32)
        d = MeshDesc(len(leaves), iptr(lv), iptr(ix), iptr(iy), N, *[float(b) for b in bounds])
        out = ctypes.c_size_t()
        flags = precision | (EF_FOOTPRINT_VARIABLE if variable else 0)
        check(lib().ef_mesh_footprint(ctypes.byref(d), flags, max_rhs, ctypes.byref(out)))
        return int(out.value)


class Patches:
    """A batch of independent N x N leaf patches (FiniteVolumeSolver on one grid each,
    FiniteVolumeSolver.hpp:122-149) on the device.

    ``bounds`` is [npatches][4] (xl, xu, yl, yu).  ``cell`` (five-point only) holds per-cell
    samples [npatches][6][N*N] (alpha, beta W/E/S/N, lambda); without it alpha0, beta0
    and lambda0 apply to every cell."""

    def __init__(self, ctx, bounds, N, solver_type=EF_SOLVER_FISHPACK, lambda0=0.0, alpha0=1.0, beta0=1.0,
                 cell=None):
        self.ctx, self.L = ctx, ctx.L
        self.bounds = f64(np.atleast_2d(bounds))
        self.P, self.N = self.bounds.shape[0], int(N)
        self._cell = None if cell is None else f64(cell)
        b = PatchBatch(self.P, self.N, dptr(self.bounds))
        c = Coeffs(solver_type, alpha0, beta0, lambda0, dptr(self._cell) if self._cell is not None else None)
        h = _vp()
        check(self.L.ef_patches_create(ctx.h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(h)), self.L)
        self.h = h
        info = np.zeros(3, np.int32)
        check(self.L.ef_patches_info(self.h, iptr(info)), self.L)
        self.variable = bool(info[2])
        self.packed = info[2] == 2  # symmetric: packed inverse pivot blocks

    def __del__(self):
        if getattr(self, "h", None) is not None:
            self.L.ef_patches_destroy(self.h)
            self.h = None

    def _apply(self, op, g, f, nrhs):
        P, N = self.P, self.N
        g = None if g is None else f64(g).reshape(nrhs, P, 4 * N)
        f = None if f is None else f64(f).reshape(nrhs, P, N * N)
        out = np.empty((nrhs, P, N * N if op == EF_PATCH_SOLVE else 4 * N))
        check(self.L.ef_patches_apply(self.h, op, None if g is None else dptr(g), None if f is None else dptr(f),
                                      dptr(out), nrhs), self.L)
        return out

    def solve(self, g, f, nrhs=1):
        """u [nrhs][npatches][N*N] from Dirichlet data g and load f (either may be None: zero)."""
        return self._apply(EF_PATCH_SOLVE, g, f, nrhs)

    def map_d2n(self, g, f, nrhs=1):
        """Neumann data h [nrhs][npatches][4N]; g = None gives particularNeumannData."""
        return self._apply(EF_PATCH_MAP_D2N, g, f, nrhs)

    def dtn(self):
        """DtN operator of every patch, [npatches][4N][4N]."""
        T = np.empty((self.P, 4 * self.N, 4 * self.N))
        check(self.L.ef_patches_dtn(self.h, dptr(T)), self.L)
        return T

    def apply_device(self, op, g_ptr, f_ptr, out_ptr, nrhs):
        check(self.L.ef_patches_apply_device(self.h, op, g_ptr, f_ptr, out_ptr, nrhs), self.L)

    def sync(self):
        check(self.L.ef_patches_sync(self.h), self.L)

    def last_ms(self):
        ms = ctypes.c_double(0.0)
        check(self.L.ef_patches_last_ms(self.h, ctypes.byref(ms)), self.L)
        return ms.value

    @property
    def stream(self):
        s = _vp()
        check(self.L.ef_patches_stream(self.h, ctypes.byref(s)), self.L)
        return s.value


class Timer:
    """Two HIP events on a stream: GPU time of everything enqueued between start and stop."""

    def __init__(self, ctx):
        self.ctx, self.L = ctx, ctx.L
        h = _vp()
        check(self.L.ef_timer_create(ctx.h, ctypes.byref(h)), self.L)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None:
            self.L.ef_timer_destroy(self.h)
            self.h = None

    def start(self, stream=None):
        check(self.L.ef_timer_record(self.h, stream, 0), self.L)

    def stop(self, stream=None):
        check(self.L.ef_timer_record(self.h, stream, 1), self.L)

    def elapsed_ms(self):
        ms = ctypes.c_double(0.0)
        check(self.L.ef_timer_elapsed_ms(self.h, ctypes.byref(ms)), self.L)
        return ms.value
