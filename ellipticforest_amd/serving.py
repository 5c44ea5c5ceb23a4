"""Serving loop over factorised meshes (continuous batching of solve requests).

* Meshes are built once and stay resident; a new mesh is admitted when its device
  footprint (``ef_mesh_footprint`` plus the serving slots) fits the HBM budget, evicting the
  least recently used idle meshes if needed.  A mesh registered with ``register_mesh`` is
  admitted (built) when a request names it while it is not resident; if the meshes holding
  the budget are busy, the loop first drains the batches in flight.
* A request names a resident mesh and carries either the 64 parameters of a serving-workload
  load (``submit_load``: 8 Gaussians for f, 8 Fourier modes for the Dirichlet data, evaluated
  on the device by ``ef_mesh_eval_load``) or explicit host arrays f ([nleaves][N*N]) and r
  ([4M]) (``submit``).
* The loop takes the mesh with the oldest waiting request and packs up to ``max_batch`` of
  that mesh's waiting requests of one kind as right-hand-side columns into a free slot.  A
  slot owns page-locked host buffers and device buffers; its inputs go up with
  ``ef_mesh_copy_async`` on the mesh's stream, then the load evaluation, upwards and solve
  stages and the copies of requested results are queued behind them and an event is
  recorded.  The host returns at once to admit and pack the next batch while the device
  works (up to ``nslots`` batches in flight); a batch completes when its event has fired.
  Requests that arrive while others are on the GPU join the next batch.
* With ``buckets`` the batch width is rounded up to the next bucket (padding columns are
  zero loads), so the sweep replays one of a few captured graphs; without it the width is
  exact and the stages are launched one by one.
* Every right-hand side is reduced in an order that does not depend on the batch width, so
  a request's result is the same alone and in a batch.
* Results stay in the slot's device buffer; a request that wants its solution on the host
  (``want_u``) has it copied into one of the slot's ``fetch_rows`` page-locked rows, so at
  most that many such requests share a batch.
* ``stats()`` reports solves/s and p50/p99 latency (arrival to completion).
"""
from __future__ import annotations

import collections
import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

NPARAM = 64  # doubles per serving-workload request


class PinnedArray:
    """Page-locked host memory (``ef_host_alloc``) viewed as a float64 numpy array."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        n = int(np.prod(self.shape))
        p = _lib._vp()
        _lib.check(ctx.L.ef_host_alloc(ctx.h, max(n, 1) * 8, ctypes.byref(p)), ctx.L)
        self.ptr = p.value
        self.a = np.ctypeslib.as_array((ctypes.c_double * max(n, 1)).from_address(self.ptr))[:n].reshape(self.shape)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.ctx.L.ef_host_free(self.ctx.h, self.ptr)
            self.ptr = None


def workload_params(rng, bounds, count):
    """Parameters of ``count`` serving-workload requests on the domain ``bounds``: K ~ U{1..8}
    Gaussians with centres uniform in the domain, widths log-uniform in [0.2, 2] and
    amplitudes N(0, 1) (amp 0 marks an unused Gaussian), and 8 Dirichlet modes a cos(kx x +
    ky y + phi) with a ~ N(0, 1), integer kx, ky in [-4, 4] scaled to the domain and phi ~
    U[0, 2 pi)."""
    xl, xu, yl, yu = bounds
    P = np.zeros((count, NPARAM))
    for i in range(count):
        K = int(rng.integers(1, 9))
        for g in range(K):
            P[i, 4 * g:4 * g + 4] = (rng.standard_normal(), rng.uniform(xl, xu), rng.uniform(yl, yu),
                                     np.exp(rng.uniform(np.log(0.2), np.log(2.0))))
        for q in range(8):
            P[i, 32 + 4 * q:36 + 4 * q] = (rng.standard_normal(), rng.integers(-4, 5) * np.pi / (xu - xl),
                                           rng.integers(-4, 5) * np.pi / (yu - yl), rng.uniform(0, 2 * np.pi))
    return P


def workload_load(params, leaf_bounds, N, root_points):
    """Host evaluation of a request's load (the oracle of ``ef_mesh_eval_load``): f per leaf at
    cell centres, r at the root boundary points."""
    f = np.zeros((len(leaf_bounds), N * N))
    for k, (xl, xu, yl, yu) in enumerate(leaf_bounds):
        dx, dy = (xu - xl) / N, (yu - yl) / N
        X, Y = np.meshgrid((xl + dx / 2) + np.arange(N) * dx, (yl + dy / 2) + np.arange(N) * dy, indexing="ij")
        X, Y = X.ravel(), Y.ravel()
        for g in range(8):
            a, cx, cy, w = params[4 * g:4 * g + 4]
            if a != 0.0:
                f[k] += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * w * w))
    px, py = root_points
    r = np.zeros(len(px))
    for q in range(8):
        a, kx, ky, phi = params[32 + 4 * q:36 + 4 * q]
        r += a * np.cos(kx * px + ky * py + phi)
    return f, r


_DONE = threading.Condition()  # shared by every request's completion flag


class Done:
    """A request's completion: ``set`` / ``is_set`` / ``wait`` like threading.Event, on one
    condition shared by all requests (an Event per request cost more than the rest of a
    submission)."""
    __slots__ = ("_flag",)

    def __init__(self):
        self._flag = False

    def set(self):
        with _DONE:
            self._flag = True
            _DONE.notify_all()

    def is_set(self):
        return self._flag

    def wait(self, timeout=None):
        with _DONE:
            return _DONE.wait_for(lambda: self._flag, timeout)


def _set_done(reqs):
    """Mark a finished batch's requests complete under one acquisition of the condition."""
    with _DONE:
        for q in reqs:
            q.done._flag = True
        _DONE.notify_all()


@dataclass
class Request:
    mesh: str
    params: np.ndarray | None = None
    f: np.ndarray | None = None
    r: np.ndarray | None = None
    want_u: bool = True
    rid: int = 0
    t_arrival: float = field(default_factory=time.perf_counter)
    t_done: float = 0.0
    u: np.ndarray | None = None
    done: Done = field(default_factory=Done)

    @property
    def kind(self):
        return "load" if self.params is not None else "host"


class _Slot:
    def __init__(self, ctx, nl, NN, M4, max_batch, fetch_rows):
        self.params = PinnedArray(ctx, (max_batch, NPARAM))
        self.dparams = _lib.DeviceArray(ctx, (max_batch, NPARAM))
        self.f = _lib.DeviceArray(ctx, (max_batch, nl, NN))
        self.r = _lib.DeviceArray(ctx, (max_batch, M4))
        self.u = _lib.DeviceArray(ctx, (max_batch, nl, NN))
        self.hu = PinnedArray(ctx, (fetch_rows, nl, NN))  # results fetched to the host
        self.hf = None  # page-locked staging of host-array requests, made on first use
        self.hr = None
        self.batch = None


class _Resident:
    def __init__(self, ctx, dm, footprint, max_batch, nslots, fetch_rows):
        self.dm = dm
        self.footprint = footprint
        self.last_used = time.perf_counter()
        nl, NN, M4 = dm.nleaves, dm.N * dm.N, 4 * dm.M
        self.slots = [_Slot(ctx, nl, NN, M4, max_batch, fetch_rows) for _ in range(nslots)]
        self.next_slot = 0
        self.a = _lib.DeviceArray(ctx, (M4,))
        self.a.upload(np.ones(M4))


class Server:
    def __init__(self, ctx, hbm_budget_bytes=200 << 30, max_batch=32, nslots=3, buckets=(1, 2, 4, 8, 16, 32, 64),
                 fetch_rows=8):
        if not 1 <= nslots <= 8:
            raise ValueError("nslots must be in 1..8")
        self.ctx = ctx
        self.fetch_rows = max(1, min(fetch_rows, max_batch))  # results a batch can return to the host
        self.budget = hbm_budget_bytes
        self.max_batch = max_batch
        self.nslots = nslots
        self.buckets = sorted(b for b in buckets if b <= max_batch) if buckets else None
        if self.buckets and self.buckets[-1] != max_batch:
            self.buckets.append(max_batch)
        self.meshes = collections.OrderedDict()
        self.recipes = {}
        self.admissions = []  # (key, build seconds) of every mesh built
        self.evictions = 0
        self.pending = collections.deque()
        self.lock = threading.Condition()
        self.completed = []
        self.inflight = collections.deque()
        self._next_id = 0
        self._stop = False

    # ----- meshes ---------------------------------------------------------------------
    def used_bytes(self):
        return sum(m.footprint for m in self.meshes.values())

    def slot_bytes(self, nleaves, N, M):
        per = 2 * nleaves * N * N + 4 * M + NPARAM  # f, u, r, parameters (device)
        return self.nslots * self.max_batch * per * 8

    def add_mesh(self, key, leaves, N, bounds, **build_args):
        """Build and keep a factorised mesh; evict idle meshes (LRU) to fit the budget."""
        probe = _lib.DeviceMesh.footprint(leaves, N, bounds, max_rhs=self.max_batch,
                                          variable=build_args.get("cell") is not None)
        fp = probe + self.slot_bytes(len(leaves), N, 0)
        if fp > self.budget:
            raise MemoryError(f"mesh {key} needs {fp} bytes, budget is {self.budget}")
        with self.lock:
            busy = {r.mesh for r in self.pending} | {res_key for res_key, _ in self._inflight_keys()}
            while self.used_bytes() + fp > self.budget:
                victim = next((k for k in self.meshes if k not in busy), None)
                if victim is None:
                    raise MemoryError("no idle mesh to evict")
                del self.meshes[victim]
                self.evictions += 1
        t0 = time.perf_counter()
        dm = _lib.DeviceMesh(self.ctx, leaves, N, bounds)
        dm.build(**build_args)
        dm.sync()
        self.admissions.append((key, time.perf_counter() - t0))
        # plan (and time the GEMM tiles of) every batch width before requests arrive: the widest
        # first, so the vector arena is sized once
        for w in sorted(self.buckets or range(1, self.max_batch + 1), reverse=True):
            dm.sweep_tiles(w)
        with self.lock:
            self.meshes[key] = _Resident(self.ctx, dm, fp, self.max_batch, self.nslots, self.fetch_rows)
        return dm

    def register_mesh(self, key, leaves, N, bounds, **build_args):
        """Keep how to build a mesh; it is admitted when a request names it and it is not resident."""
        self.recipes[key] = (leaves, N, bounds, build_args)

    def _admit(self, key):
        """Build a registered mesh for a request that names it (the submitting thread drives the
        loop meanwhile: call it from the thread that runs ``run``)."""
        leaves, N, bounds, build_args = self.recipes[key]
        try:
            self.add_mesh(key, leaves, N, bounds, **build_args)
        except MemoryError:
            self.run()  # the meshes holding the budget were busy: finish their batches, then evict
            self.add_mesh(key, leaves, N, bounds, **build_args)

    def _inflight_keys(self):
        return [(k, s) for k, s, _ in self.inflight]

    # ----- requests -------------------------------------------------------------------
    def _enqueue(self, req):
        if req.mesh not in self.meshes and req.mesh in self.recipes:
            self._admit(req.mesh)
        with self.lock:
            if req.mesh not in self.meshes:
                raise KeyError(f"mesh {req.mesh} is not resident")
            req.rid = self._next_id
            self._next_id += 1
            self.pending.append(req)
            self.lock.notify()
        return req

    def submit_load(self, mesh, params, want_u=True):
        p = np.ascontiguousarray(params, dtype=np.float64).ravel()
        if p.size != NPARAM:
            raise ValueError(f"a load request has {NPARAM} parameters")
        return self._enqueue(Request(mesh=mesh, params=p, want_u=want_u))

    def submit(self, mesh, f, r, want_u=True):
        return self._enqueue(Request(mesh=mesh, f=np.ascontiguousarray(f, dtype=np.float64),
                                     r=np.ascontiguousarray(r, dtype=np.float64), want_u=want_u))

    def _take_batch(self):
        with self.lock:
            if not self.pending:
                return None, None, []
            key, kind = self.pending[0].mesh, self.pending[0].kind
            batch, rest = [], []
            fetch = 0
            # scan only until the batch is full: the skipped requests go back in front, in order
            # (a scan of the whole queue per batch made a deep queue cost O(queue) per batch)
            while self.pending and len(batch) < self.max_batch:
                q = self.pending.popleft()
                if q.mesh == key and q.kind == kind and (not q.want_u or fetch < self.fetch_rows):
                    batch.append(q)
                    fetch += q.want_u
                else:
                    rest.append(q)
            self.pending.extendleft(reversed(rest))
            self.meshes.move_to_end(key)
            return key, self.meshes[key], batch

    # ----- device work ----------------------------------------------------------------
    def _width(self, n):
        if not self.buckets:
            return n
        return next(b for b in self.buckets if b >= n)

    def _launch(self, key, res, si, batch):
        dm, L, s = res.dm, self.ctx.L, res.slots[si]
        n, k = len(batch), self._width(len(batch))
        nl, NN, M4 = dm.nleaves, dm.N * dm.N, 4 * dm.M
        if batch[0].kind == "load":
            s.params.a[:n] = np.stack([q.params for q in batch])
            s.params.a[n:k] = 0.0  # padding columns: zero load, zero boundary data
            _lib.check(L.ef_mesh_copy_async(dm.h, s.dparams.ptr, s.params.ptr, k * NPARAM * 8, 0), L)
            _lib.check(L.ef_mesh_eval_load(dm.h, s.dparams.ptr, k, s.f.ptr, s.r.ptr), L)
        else:
            if s.hf is None:
                s.hf = PinnedArray(self.ctx, (self.max_batch, nl, NN))
                s.hr = PinnedArray(self.ctx, (self.max_batch, M4))
            for j, q in enumerate(batch):
                s.hf.a[j] = q.f.reshape(nl, NN)
                s.hr.a[j] = q.r
            s.hf.a[n:k] = 0.0
            s.hr.a[n:k] = 0.0
            _lib.check(L.ef_mesh_copy_async(dm.h, s.f.ptr, s.hf.ptr, k * nl * NN * 8, 0), L)
            _lib.check(L.ef_mesh_copy_async(dm.h, s.r.ptr, s.hr.ptr, k * M4 * 8, 0), L)
        dm.upwards_device(s.f.ptr, k)
        dm.solve_device(s.f.ptr, s.r.ptr, res.a.ptr, None, s.u.ptr, k)
        per = nl * NN * 8
        s.rows = {}
        for j, q in enumerate(batch):
            if q.want_u:
                row = len(s.rows)
                s.rows[j] = row
                _lib.check(L.ef_mesh_copy_async(dm.h, s.hu.ptr + row * per, s.u.ptr + j * per, per, 1), L)
        _lib.check(L.ef_mesh_record(dm.h, si), L)
        s.batch = batch
        self.inflight.append((key, si, res))

    def _done(self, res, si):
        flag = ctypes.c_int(0)
        _lib.check(self.ctx.L.ef_mesh_query(res.dm.h, si, ctypes.byref(flag)), self.ctx.L)
        return flag.value != 0

    def _complete(self, block):
        """Finish the oldest in-flight batch (waiting for it when ``block``)."""
        key, si, res = self.inflight[0]
        if not block and not self._done(res, si):
            return False
        if block:
            _lib.check(self.ctx.L.ef_mesh_wait(res.dm.h, si), self.ctx.L)
        self.inflight.popleft()
        s = res.slots[si]
        now = time.perf_counter()
        for j, q in enumerate(s.batch):
            if q.want_u:
                q.u = s.hu.a[s.rows[j]].copy()
            q.t_done = now
        _set_done(s.batch)
        self.completed.extend(s.batch)
        s.batch = None
        res.last_used = now
        return True

    def run(self, stop_when_idle=True, poll_s=5e-5):
        """Continuous batching: admit and queue batches while earlier ones run on the device."""
        while not self._stop:
            while self.inflight and self._complete(block=False):
                pass
            key, res, batch = self._take_batch()
            if batch:
                si = res.next_slot
                res.next_slot = (si + 1) % self.nslots
                while any(k == key and s == si for k, s, _ in self.inflight):
                    self._complete(block=True)  # the slot's previous batch must be finished
                self._launch(key, res, si, batch)
                continue
            if self.inflight:
                if not self._complete(block=False):
                    with self.lock:
                        if not self.pending:
                            self.lock.wait(timeout=poll_s)
                continue
            if stop_when_idle:
                break
            with self.lock:
                if not self.pending:
                    self.lock.wait(timeout=poll_s)

    def stop(self):
        self._stop = True
        with self.lock:
            self.lock.notify_all()

    def stats(self):
        if not self.completed:
            return {}
        lat = np.array([(q.t_done - q.t_arrival) * 1e3 for q in self.completed])
        t0 = min(q.t_arrival for q in self.completed)
        t1 = max(q.t_done for q in self.completed)
        return dict(requests=len(lat), solves_per_s=len(lat) / max(t1 - t0, 1e-12),
                    p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)))


def drive_poisson(server, mesh, params, rate_hz, want_u_every=0, seed=0):
    """Submit ``params`` (one request per row) at Poisson arrival times of rate ``rate_hz``
    from a feeder thread while ``server.run`` serves them; returns the requests."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(params))
    reqs = [None] * len(params)

    def feed():
        t = time.perf_counter()
        for i, p in enumerate(params):
            t += gaps[i]
            while True:
                d = t - time.perf_counter()
                if d <= 0:
                    break
                time.sleep(min(d, 1e-3) if d > 2e-4 else 0)
            reqs[i] = server.submit_load(mesh, p, want_u=bool(want_u_every) and i % want_u_every == 0)
        server.stop()

    th = threading.Thread(target=feed, daemon=True)
    server._stop = False
    th.start()
    server.run(stop_when_idle=False)
    th.join()
    server._stop = False
    server.run(stop_when_idle=True)  # drain
    return reqs
