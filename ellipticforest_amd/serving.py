"""Serving loop over factorised meshes (continuous batching of solve requests).

* Meshes are built once and stay resident; a new mesh is admitted when its device
  footprint (``ef_mesh_footprint``) fits the HBM budget, evicting the least recently
  used idle meshes if needed.
* A request carries the load f ([nleaves][N*N]) and the Dirichlet data r ([4M]) for one
  resident mesh.  The loop takes the mesh with the oldest waiting request, packs up to
  ``max_batch`` of that mesh's waiting requests as right-hand-side columns into one of
  two staging slots, and runs upwards + solve on the device for the whole batch.  While
  one slot runs on the device the host packs the next batch into the other slot, so
  requests that arrive while others are on the GPU join the next launch.
* Each right-hand side is reduced in an order that does not depend on the batch width,
  so a request's result is the same alone and in a batch.
* ``stats()`` reports solves/s and p50/p99 latency (arrival to result on the host).

Device memory comes from a buffer provider: ``DeviceBuffers`` allocates HBM and moves
bytes through the library's own entry points (``ef_device_alloc``, ``ef_memcpy``), so the
process holds a single HIP runtime.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class DeviceBuffers:
    """HBM buffers of the context's GPU (``_lib.DeviceArray``)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def empty(self, shape):
        return _lib.DeviceArray(self.ctx, shape)

    def ptr(self, buf):
        return buf.ptr

    def upload(self, buf, host, offset=0):
        buf.upload(np.ascontiguousarray(host).ravel(), offset)

    def download(self, buf, count):
        return buf.download(count)


@dataclass
class Request:
    mesh: str
    f: np.ndarray
    r: np.ndarray
    rid: int = 0
    t_arrival: float = field(default_factory=time.perf_counter)
    t_done: float = 0.0
    u: np.ndarray | None = None
    done: threading.Event = field(default_factory=threading.Event)


class _Resident:
    def __init__(self, dm, footprint, max_batch, buffers):
        self.dm = dm
        self.footprint = footprint
        self.last_used = time.perf_counter()
        nl, NN, M4 = dm.nleaves, dm.N * dm.N, 4 * dm.M
        # two staging slots: the device works on one while the host packs the other
        self.slots = [dict(f=buffers.empty((max_batch, nl, NN)), r=buffers.empty((max_batch, M4)),
                           u=buffers.empty((max_batch, nl, NN))) for _ in range(2)]
        self.a = buffers.empty((M4,))
        buffers.upload(self.a, np.ones(M4))


class Server:
    def __init__(self, ctx, buffers, hbm_budget_bytes=200 << 30, max_batch=32):
        self.ctx = ctx
        self.buffers = buffers
        self.budget = hbm_budget_bytes
        self.max_batch = max_batch
        self.meshes = collections.OrderedDict()
        self.pending = collections.deque()
        self.lock = threading.Condition()
        self.completed = []
        self._next_id = 0
        self._stop = False

    # ----- meshes ---------------------------------------------------------------------
    def used_bytes(self):
        return sum(m.footprint for m in self.meshes.values())

    def add_mesh(self, key, leaves, N, bounds, **build_args):
        """Build and keep a factorised mesh; evict idle meshes (LRU) to fit the budget."""
        fp = _lib.DeviceMesh.footprint(leaves, N, bounds, max_rhs=self.max_batch)
        fp += 3 * 2 * self.max_batch * len(leaves) * N * N * 8  # staging slots
        if fp > self.budget:
            raise MemoryError(f"mesh {key} needs {fp} bytes, budget is {self.budget}")
        with self.lock:
            busy = {r.mesh for r in self.pending}
            while self.used_bytes() + fp > self.budget:
                victim = next((k for k in self.meshes if k not in busy), None)
                if victim is None:
                    raise MemoryError("no idle mesh to evict")
                del self.meshes[victim]
        dm = _lib.DeviceMesh(self.ctx, leaves, N, bounds)
        dm.build(**build_args)
        with self.lock:
            self.meshes[key] = _Resident(dm, fp, self.max_batch, self.buffers)
        return dm

    # ----- requests -------------------------------------------------------------------
    def submit(self, mesh, f, r):
        req = Request(mesh=mesh, f=np.ascontiguousarray(f, dtype=np.float64),
                      r=np.ascontiguousarray(r, dtype=np.float64))
        with self.lock:
            if mesh not in self.meshes:
                raise KeyError(f"mesh {mesh} is not resident")
            req.rid = self._next_id
            self._next_id += 1
            self.pending.append(req)
            self.lock.notify()
        return req

    def _take_batch(self):
        with self.lock:
            if not self.pending:
                return None, []
            key = self.pending[0].mesh
            batch, rest = [], collections.deque()
            while self.pending:
                q = self.pending.popleft()
                if q.mesh == key and len(batch) < self.max_batch:
                    batch.append(q)
                else:
                    rest.append(q)
            self.pending = rest
            self.meshes.move_to_end(key)
            return self.meshes[key], batch

    def _pack(self, res, slot, batch):
        B = self.buffers
        for j, q in enumerate(batch):
            B.upload(slot["f"], q.f, offset=j * q.f.size)
            B.upload(slot["r"], q.r, offset=j * q.r.size)

    def _launch(self, res, slot, k):
        dm, B = res.dm, self.buffers
        dm.upwards_device(B.ptr(slot["f"]), k)
        dm.solve_device(B.ptr(slot["f"]), B.ptr(slot["r"]), B.ptr(res.a), None, B.ptr(slot["u"]), k)

    def _finish(self, res, slot, batch):
        res.dm.sync()
        per = res.dm.nleaves * res.dm.N * res.dm.N
        u = self.buffers.download(slot["u"], len(batch) * per).reshape(len(batch), res.dm.nleaves, -1)
        now = time.perf_counter()
        for j, q in enumerate(batch):
            q.u = u[j].copy()
            q.t_done = now
            q.done.set()
        self.completed.extend(batch)
        res.last_used = now

    def run(self, stop_when_idle=True, poll_s=1e-4):
        """Continuous batching: launch batch i, pack batch i+1 while i runs, then finish i."""
        inflight = None
        slot_id = 0
        while not self._stop:
            res, batch = self._take_batch()
            if batch:
                slot = res.slots[slot_id]
                self._pack(res, slot, batch)
                if inflight is not None:
                    self._finish(*inflight)
                self._launch(res, slot, len(batch))
                inflight = (res, slot, batch)
                slot_id ^= 1
                continue
            if inflight is not None:
                self._finish(*inflight)
                inflight = None
                continue
            if stop_when_idle:
                break
            with self.lock:
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
