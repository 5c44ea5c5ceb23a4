"""Serving loop (ellipticforest_amd/serving.py).  The CPU tests drive it through the
CPU-emulation build of the library (page-locked buffers, asynchronous copies, slot events and
the load evaluation map to host memory there, so the same code path runs); the GPU tests run
it on the device.  Checks: a request's result is the same alone and in a batch (bitwise:
every right-hand side is reduced in a batch-independent order), results match the oracle,
device-evaluated loads match their host evaluation, meshes are admitted / evicted by
footprint, and the Poisson-arrival driver serves every request."""
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


def leaf_bounds(dm):
    return [tuple(dm.node_bounds[n]) for n in dm.leaf_nodes]


def host_array_case(ctx, leaves, N, bounds, nreq, max_batch):
    srv = serving.Server(ctx, max_batch=max_batch)
    dm = srv.add_mesh("m", leaves, N, bounds)
    rng = np.random.default_rng(7)
    reqs = [(rng.standard_normal((len(leaves), N * N)), rng.standard_normal(4 * dm.M)) for _ in range(nreq)]
    batched = [srv.submit("m", f, r) for f, r in reqs]
    srv.run()
    assert all(q.done.is_set() for q in batched)
    for i in (0, nreq - 1):
        f, r = reqs[i]
        alone = srv.submit("m", f, r)
        srv.run()
        assert np.array_equal(alone.u, batched[i].u)
        ref = reference(leaves, N, bounds, f, r)
        assert np.abs(batched[i].u - ref).max() / np.abs(ref).max() < 1e-10
    st = srv.stats()
    assert st["requests"] == nreq + 2 and st["p99_ms"] >= st["p50_ms"] > 0


def load_case(ctx, leaves, N, bounds, nreq, max_batch, buckets=(1, 2, 4, 8, 16, 32)):
    srv = serving.Server(ctx, max_batch=max_batch, buckets=buckets)
    dm = srv.add_mesh("m", leaves, N, bounds)
    P = serving.workload_params(np.random.default_rng(3), bounds, nreq)
    reqs = [srv.submit_load("m", p) for p in P]
    srv.run()
    import ellipticforest_amd as ef
    side, px, py = oracle.root_face_points(dm.M, bounds)
    for i in (0, nreq // 2, nreq - 1):
        f, r = serving.workload_load(P[i], leaf_bounds(dm), N, (px, py))
        alone = srv.submit_load("m", P[i])
        srv.run()
        assert np.array_equal(alone.u, reqs[i].u)
        ref = reference(leaves, N, bounds, f, r)
        assert np.abs(reqs[i].u - ref).max() / np.abs(ref).max() < 1e-10
    return srv, dm, P, reqs


def test_batch_invariance_and_parity(emu_ctx):
    host_array_case(emu_ctx, omesh.uniform(1), 8, (0.0, 1.0, 0.0, 1.0), nreq=6, max_batch=4)


def test_load_requests(emu_ctx):
    """Requests as workload parameters: device-evaluated loads against the host evaluation
    (through the solution), bucketed batch widths, invariance alone vs in a batch."""
    load_case(emu_ctx, omesh.uniform(1), 8, (-1.0, 1.0, -1.0, 1.0), nreq=7, max_batch=4, buckets=(1, 2, 4))


def test_admission_by_footprint(emu_ctx):
    leaves, N, bounds = omesh.uniform(1), 8, (0.0, 1.0, 0.0, 1.0)
    srv = serving.Server(emu_ctx, max_batch=2)
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


def test_admission_on_demand(emu_ctx):
    """Registered meshes (different N, variable coefficients) are built when a request names
    them; with room for two, a third evicts the least recently used idle one, and a request for
    an evicted mesh builds it again.  Results match a fresh single-mesh server bitwise."""
    import ellipticforest_amd as ef
    bounds = (-1.0, 1.0, -1.0, 1.0)
    specs = {"a": (omesh.uniform(1), 8), "b": (omesh.uniform(1), 4), "c": (omesh.uniform(1), 8)}

    def cells(leaves, N, lam):
        dm = ef.DeviceMesh(emu_ctx, leaves, N, bounds)
        beta = lambda x, y: 1.0 + 0.5 * np.sin(x + 2 * y)
        return np.concatenate([oracle.sample_coeffs(N, tuple(dm.node_bounds[n]), lambda x, y: np.ones_like(x), beta,
                                                    lambda x, y: lam + 0 * x).ravel() for n in dm.leaf_nodes])

    srv = serving.Server(emu_ctx, max_batch=2, buckets=None)
    args = {}
    for i, (key, (leaves, N)) in enumerate(specs.items()):
        args[key] = dict(solver_type=ef._lib.EF_SOLVER_FIVE_POINT, cell=cells(leaves, N, -0.5 * i))
        srv.register_mesh(key, leaves, N, bounds, **args[key])
    fp = {k: ef._lib.DeviceMesh.footprint(l, N, bounds, max_rhs=2, variable=True) + srv.slot_bytes(len(l), N, 0)
          for k, (l, N) in specs.items()}
    srv.budget = fp["a"] + fp["c"] + fp["b"] // 2  # two meshes fit, not three
    P = serving.workload_params(np.random.default_rng(4), bounds, 4)
    reqs = []
    for key, p in zip(("a", "b", "c", "a"), P):
        reqs.append(srv.submit_load(key, p))
        srv.run()
    assert [k for k, _ in srv.admissions] == ["a", "b", "c", "a"] and srv.evictions >= 2
    assert len(srv.meshes) <= 2
    for key, p, q in zip(("a", "b", "c", "a"), P, reqs):
        one = serving.Server(emu_ctx, max_batch=2, buckets=None)
        one.add_mesh(key, specs[key][0], specs[key][1], bounds, **args[key])
        alone = one.submit_load(key, p)
        one.run()
        assert np.array_equal(alone.u, q.u)


def test_poisson_driver(emu_ctx):
    leaves, N, bounds = omesh.uniform(1), 8, (-1.0, 1.0, -1.0, 1.0)
    srv = serving.Server(emu_ctx, max_batch=4, buckets=None)
    srv.add_mesh("m", leaves, N, bounds)
    P = serving.workload_params(np.random.default_rng(5), bounds, 12)
    reqs = serving.drive_poisson(srv, "m", P, rate_hz=2000.0, want_u_every=4)
    assert all(q is not None and q.done.is_set() for q in reqs)
    assert sum(q.u is not None for q in reqs) == 3
    assert srv.stats()["requests"] == 12


def test_workload_params_shape():
    P = serving.workload_params(np.random.default_rng(0), (-10, 10, -10, 10), 50)
    K = (P[:, 0:32:4] != 0).sum(1)
    assert P.shape == (50, serving.NPARAM) and K.min() >= 1 and K.max() <= 8
    w = P[:, 3:32:4][P[:, 0:32:4] != 0]
    assert w.min() >= 0.2 and w.max() <= 2.0


@pytest.mark.gpu
def test_gpu_serving_host_arrays(ctx):
    host_array_case(ctx, omesh.refine_by_function(lambda x, y: np.abs(np.sin(x) + np.sin(y)) > 1.2, 8, 8,
                                                  -10, 10, -10, 10, 0, 4), 8, (-10, 10, -10, 10), nreq=40,
                    max_batch=32)


@pytest.mark.gpu
def test_gpu_serving_loads_and_poisson(ctx):
    """Loads evaluated on the device, widths past 16 columns (two-tile GEMMs), bitwise
    invariance alone vs in a batch, oracle parity, then a Poisson-arrival run."""
    leaves = omesh.refine_by_function(lambda x, y: np.abs(np.sin(x) + np.sin(y)) > 1.2, 16, 16, -10, 10, -10, 10, 0, 4)
    srv, dm, P, reqs = load_case(ctx, leaves, 16, (-10.0, 10.0, -10.0, 10.0), nreq=45, max_batch=32)
    P2 = serving.workload_params(np.random.default_rng(9), (-10.0, 10.0, -10.0, 10.0), 300)
    srv.completed.clear()
    out = serving.drive_poisson(srv, "m", P2, rate_hz=5000.0, want_u_every=50)
    assert all(q.done.is_set() for q in out)
    st = srv.stats()
    assert st["requests"] == 300 and st["p99_ms"] >= st["p50_ms"] > 0
    alone = srv.submit_load("m", P2[100])
    srv.run()
    assert np.array_equal(alone.u, out[100].u)


@pytest.mark.gpu
def test_gpu_multi_mesh_admission(ctx):
    """Three adaptive variable-coefficient meshes (N = 8, 16, 8) registered with a budget for
    two: requests build meshes on demand and evict idle ones; every result equals the same
    request on a fresh single-mesh server, bitwise, and one matches the oracle."""
    import ellipticforest_amd as ef
    bounds = (-10.0, 10.0, -10.0, 10.0)
    F = lambda x, y: -(np.sin(x) + np.sin(y))
    specs = {}
    for key, (N, maxl, a, lam) in {"a": (8, 4, 1.0, -0.5), "b": (16, 3, 2.5, -1.5), "c": (8, 5, 0.5, 0.0)}.items():
        leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, N, N, *bounds, 0, maxl)
        dm = ef.DeviceMesh(ctx, leaves, N, bounds)
        beta = (lambda aa: (lambda x, y: 1.0 + 0.5 * np.sin(aa * x + 0.7 * y)))(a)
        cell = np.concatenate([oracle.sample_coeffs(N, tuple(dm.node_bounds[n]), lambda x, y: np.ones_like(x), beta,
                                                    (lambda l: (lambda x, y: l + 0 * x))(lam)).ravel()
                               for n in dm.leaf_nodes])
        specs[key] = (leaves, N, dict(solver_type=ef._lib.EF_SOLVER_FIVE_POINT, cell=cell), beta, lam)
    srv = serving.Server(ctx, max_batch=8, buckets=(1, 2, 4, 8))
    fp = {}
    for key, (leaves, N, args, _, _) in specs.items():
        srv.register_mesh(key, leaves, N, bounds, **args)
        fp[key] = ef._lib.DeviceMesh.footprint(leaves, N, bounds, max_rhs=8, variable=True) + \
            srv.slot_bytes(len(leaves), N, 0)
    srv.budget = sorted(fp.values())[-1] + sorted(fp.values())[-2] + 1
    P = serving.workload_params(np.random.default_rng(8), bounds, 12)
    order = ["a", "b", "c", "a", "c", "b"] * 2
    reqs = []
    for i in range(0, 12, 3):  # ticks of three requests
        for key, p in zip(order[i:i + 3], P[i:i + 3]):
            reqs.append((key, p, srv.submit_load(key, p)))
        srv.run()
    assert srv.evictions >= 1 and len(srv.admissions) >= 4
    for key, p, q in reqs[:4]:
        leaves, N, args, _, _ = specs[key]
        one = serving.Server(ctx, max_batch=8, buckets=(1, 2, 4, 8))
        one.add_mesh(key, leaves, N, bounds, **args)
        alone = one.submit_load(key, p)
        one.run()
        assert np.array_equal(alone.u, q.u), key
    key, p, q = reqs[1]
    leaves, N, args, beta, lam = specs[key]
    O = oracle.Forest(leaves, N, bounds, solver=oracle.FIVE_POINT,
                      coeff_fns=(lambda x, y: np.ones_like(x), beta, lambda x, y: lam + 0 * x))
    O.build()
    dm = srv.meshes[key].dm if key in srv.meshes else ef.DeviceMesh(ctx, leaves, N, bounds)
    side, px, py = oracle.root_face_points(dm.M, bounds)
    f, r = serving.workload_load(p, [tuple(dm.node_bounds[n]) for n in dm.leaf_nodes], N, (px, py))
    O.upwards(f)
    ref = O.solve(r, np.ones_like(r), np.zeros_like(r))
    assert np.abs(q.u - ref).max() / np.abs(ref).max() < 1e-10
