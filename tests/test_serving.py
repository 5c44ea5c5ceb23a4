"""Serving loop (ellipticforest_amd/serving.py) driven through the CPU-emulation build of
the library: buffers come from the library's own ef_device_alloc / ef_memcpy (host memory
in the emulation), so the device-pointer entry points run unchanged.  Checks: a request's result is the same alone and in a batch (bitwise: every
right-hand side is reduced in a batch-independent order), results match the oracle, and
meshes are admitted / evicted by footprint."""
import numpy as np
import pytest

import oracle
from oracle import mesh as omesh
from ellipticforest_amd import serving


def reference(leaves, N, bounds, f, r):
    O = oracle.Forest(leaves, N, bounds, solver=oracle.FISHPACK)
    O.build()
    O.upwards(f)
    return O.solve(r, np.ones_like(r), np.zeros_like(r))


def test_batch_invariance_and_parity(emu_ctx):
    leaves, N, bounds = omesh.uniform(1), 8, (0.0, 1.0, 0.0, 1.0)
    srv = serving.Server(emu_ctx, serving.DeviceBuffers(emu_ctx), max_batch=4)
    dm = srv.add_mesh("m", leaves, N, bounds)
    rng = np.random.default_rng(7)
    reqs = [(rng.standard_normal((len(leaves), N * N)), rng.standard_normal(4 * dm.M)) for _ in range(6)]
    batched = [srv.submit("m", f, r) for f, r in reqs]
    srv.run()
    assert all(q.done.is_set() for q in batched)
    for (f, r), q in zip(reqs, batched):
        alone = srv.submit("m", f, r)
        srv.run()
        assert np.array_equal(alone.u, q.u)
        ref = reference(leaves, N, bounds, f, r)
        assert np.abs(q.u - ref).max() / np.abs(ref).max() < 1e-10
    st = srv.stats()
    assert st["requests"] == 12 and st["p99_ms"] >= st["p50_ms"] > 0


def test_admission_by_footprint(emu_ctx):
    leaves, N, bounds = omesh.uniform(1), 8, (0.0, 1.0, 0.0, 1.0)
    srv = serving.Server(emu_ctx, serving.DeviceBuffers(emu_ctx), max_batch=2)
    srv.add_mesh("a", leaves, N, bounds)
    one = srv.used_bytes()
    srv.budget = int(one * 1.5)  # room for one mesh only
    srv.add_mesh("b", leaves, N, bounds)
    assert list(srv.meshes) == ["b"]
    with pytest.raises(KeyError):
        srv.submit("a", np.zeros((4, 64)), np.zeros(4 * 16))
    srv.budget = one // 2
    with pytest.raises(MemoryError):
        srv.add_mesh("c", leaves, N, bounds)
