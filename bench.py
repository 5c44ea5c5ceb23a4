"""Benchmark of the HPS factor-and-solve path on one MI355X.

Workload (default): adaptive Poisson on [-10,10]^2, f = -(sin x + sin y), Dirichlet
u = sin x + sin y, refineByFunction(|f| > 1.2, min level 0, max level 7), 16x16 cells per
patch (7,750 leaves, root side 512), FISHPACK semantics, FP64.

A step is one upwards + solve pass of a batch of --nrhs (default 64) right-hand sides whose loads and
boundary data are already resident in HBM (HPSAlgorithm::upwardsStage + solveStage over the
whole tree); value = solves per second over the K timed steps, queued back to back on the
mesh's stream between one pair of HIP events.  Reported beside it: buildStage seconds (HIP
events around ef_build, steady state and first call), per-request latency percentiles (one
right-hand side per request, host clock around a synchronised upwards + solve), the relative
error of the result against the FP64 oracle, the sweep's HBM roofline (algorithmic operator
and vector bytes per step over the step time) and the build's matrix-core rate.

The CPU baseline is the oracle restatement of the reference algorithm (oracle/), timed on
this host: one buildStage with all host threads, upwards + solve passes at one thread and at
all threads.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

U = lambda x, y: np.sin(x) + np.sin(y)
F = lambda x, y: -(np.sin(x) + np.sin(y))
BOUNDS = (-10.0, 10.0, -10.0, 10.0)
METRIC = "solves/s + p50/p99 latency at 1e-8 rel err vs FP64; buildStage seconds"
# HBM bytes per launch measured by two rocprofv3 counter passes (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, factors from the copy kernel); profiles/patch_fd_kernel_adaptive_poisson_leafset.md and
# profiles/bt_solve_sym_kernel_variable_beta_leafset.md have the commands and the raw values.
TRAFFIC_PER_LAUNCH = 572.9e6          # leaf solves of this workload, 16 right-hand sides
TRAFFIC_MAP_D2N = 319.0e6             # particular Neumann data, same leaf set and k
TRAFFIC_VAR_BETA = 1329.8e6           # variable-beta leaf solves, packed symmetric blocks
TRAFFIC_GEMM_PER_LAUNCH = 135.36e6    # tree_gemm_kernel in the sweep at 64 columns, per launch: profiles/tree_gemm_kernel_adaptive_poisson_sweep.md


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(max_level, N):
    import ellipticforest_amd as ef
    leaves = ef.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, N, N, BOUNDS, 0, max_level)
    return leaves


def cpu_baseline(leaves, N, nsolve, threads):
    """Oracle (reference algorithm restated in C++) on this host: buildStage with all threads,
    upwards + solve at one thread (nsolve passes) and at all threads (one pass)."""
    import oracle
    oracle.set_threads(threads)
    O = oracle.Forest(leaves, N, BOUNDS, solver=oracle.FISHPACK)
    t0 = time.perf_counter()
    O.build()
    t_build = time.perf_counter() - t0
    f = O.sample_rhs(F)
    r, a, b = O.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    t0 = time.perf_counter()
    O.upwards(f)
    O.solve(r, a, b)
    t_all = time.perf_counter() - t0
    oracle.set_threads(1)
    times = []
    u = None
    for _ in range(nsolve):
        t0 = time.perf_counter()
        O.upwards(f)
        u = O.solve(r, a, b)
        times.append(time.perf_counter() - t0)
    err = 0.0
    for k, bd in enumerate(O.leaf_bounds):
        X, Y, _, _ = oracle.cell_centres(N, bd)
        err = max(err, float(np.abs(u[k] - U(X, Y)).max()))
    return dict(build_s=t_build, solve_s=float(np.mean(times)), solves_per_s=1.0 / float(np.mean(times)),
                solves_per_s_all=1.0 / t_all, build_threads=threads, solve_threads=1, disc_err=err, u=u, f=f, r=r,
                a=a)


def tree_counts(D):
    """Interface sizes n of every merge and the coarsening steps (fine side m) of the mesh."""
    info = D.node_info
    ns, coarsen = [], []
    for i in range(D.nnodes):
        level, is_leaf, tag, s = info[i, 0], info[i, 3], info[i, 5], info[i, 6]
        if not is_leaf:
            ns.append(s // 2)
        m = s
        for _ in range(tag):
            coarsen.append(m)
            m //= 2
    return np.array(ns, np.float64), np.array(coarsen, np.float64)


def sweep_bytes(D, N, k):
    """Algorithmic HBM bytes of one upwards + solve step with k right-hand sides: every merge's
    X^-1 (16 n^2), S (32 n^2) and the 16 n^2 nonzeros of H (its children's DtN blocks) read
    once, and per right-hand side each leaf's load f read and solution u written (FP64).  The
    constant-coefficient leaves share their sine-transform tables, so leaf operators add nothing."""
    ns, _ = tree_counts(D)
    return 8.0 * (64.0 * float(np.sum(ns ** 2)) + k * D.nleaves * 2.0 * N * N)


def plain_merges(D):
    """Per merge (in node order): True when no child is coarsened, so its w product reads hDiff as
    differences of the children's h in place (8 n k reads) instead of a gathered hDiff (4 n k)."""
    info = D.node_info
    return np.array([all(info[info[i, 7 + c], 5] == 0 for c in range(4))
                     for i in range(D.nnodes) if not info[i, 3]], bool)


def gemm_bytes(D, k):
    """Algorithmic HBM bytes of the tree GEMMs of one step: per merge the operators X^-1 (16 n^2),
    the 16 n^2 nonzeros of H and S (32 n^2) read once, and the vectors of its products
    (w = X^-1 hDiff: 12 n k with hDiff formed from the children's h in place, 8 n k from a gathered
    hDiff; h += H w: 32 n k; u = S g + w: 16 n k, g read side by side where it lies), FP64."""
    ns, _ = tree_counts(D)
    vec = np.where(plain_merges(D), 60.0, 56.0)
    return 8.0 * float(np.sum(64.0 * ns ** 2 + vec * ns * k))


def build_flops(D, N, nclasses):
    """Matrix-core work of buildStage.  dense: the reference's count (dgesv of X against the dense
    S_RHS, dense H S, L21 T L12, leaf DtN by 32 N^4), sum 810.67 n^3 + 48 m^3 + 32 N^4 per leaf
    class.  executed: what the kernels do (X^-1 by Gauss-Jordan 2 (4n)^3, S from the nonzero
    half of S_RHS 128 n^3, H S from the nonzero half of H 256 n^3; coarsening as sparse reads)."""
    ns, cm = tree_counts(D)
    dense = float(np.sum(810.67 * ns ** 3) + np.sum(48.0 * cm ** 3) + 32.0 * N ** 4 * nclasses)
    executed = float(np.sum(512.0 * ns ** 3))
    return dense, executed


def gpu_run(leaves, N, k, steps, warmup, ref, nlat):
    """Device path through the C ABI only (device memory, streams and HIP-event timers of
    libefhps.so; no second HIP runtime in the process)."""
    import ellipticforest_amd as ef
    ctx = ef.Context(0)
    D = ef.DeviceMesh(ctx, leaves, N, BOUNDS)
    builds = []
    for _ in range(3):  # the first call includes allocations and kernel loading
        t0 = time.perf_counter()
        D.build()
        D.sync()
        builds.append(dict(wall_s=time.perf_counter() - t0, event_s=D.stage_ms()["build-stage"] / 1e3))
    log("buildStage: " + ", ".join(f"{b['event_s'] * 1e3:.2f} ms" for b in builds))
    nl = len(leaves)
    f = ef.DeviceArray(ctx, (k, nl, N * N))
    r = ef.DeviceArray(ctx, (k, 4 * D.M))
    a = ef.DeviceArray(ctx, (4 * D.M,))
    u = ef.DeviceArray(ctx, (k, nl, N * N))
    f.upload(np.broadcast_to(ref["f"], (k,) + ref["f"].shape))
    r.upload(np.broadcast_to(ref["r"], (k,) + ref["r"].shape))
    a.upload(ref["a"])

    def step(kk, uu):
        D.upwards_device(f.ptr, kk)
        D.solve_device(f.ptr, r.ptr, a.ptr, None, uu.ptr, kk)

    for _ in range(warmup):
        step(k, u)
    D.sync()
    # timed region: K steps queued back to back between two HIP events on the mesh's stream
    timer = ef.Timer(ctx)
    t0 = time.perf_counter()
    timer.start(D.stream)
    for _ in range(steps):
        step(k, u)
    timer.stop(D.stream)
    ms_step = timer.elapsed_ms() / steps
    D.sync()
    wall_step = (time.perf_counter() - t0) / steps * 1e3
    st = D.stage_ms()  # the last step's stages
    uu = u.download().reshape(k, nl, N * N)
    rel = float(max(np.abs(uu[j] - ref["u"]).max() for j in range(k)) / np.abs(ref["u"]).max())
    # per-request latency: one right-hand side at a time, host clock around a synchronised step
    u1 = ef.DeviceArray(ctx, (1, nl, N * N))
    for _ in range(3):
        step(1, u1)
    D.sync()
    lat = []
    for _ in range(nlat):
        t1 = time.perf_counter()
        step(1, u1)
        D.sync()
        lat.append((time.perf_counter() - t1) * 1e3)
    rel1 = float(np.abs(u1.download().reshape(nl, N * N) - ref["u"]).max() / np.abs(ref["u"]).max())
    # the sweep's dominant kernel: its GEMM launches alone, between HIP events on the mesh's stream
    import ctypes
    gms, glaunch = ctypes.c_double(0.0), ctypes.c_int(0)
    ef._lib.check(ctx.L.ef_mesh_time_sweep_gemms(D.h, k, max(steps, 5), ctypes.byref(gms), ctypes.byref(glaunch)), ctx.L)
    gbytes = gemm_bytes(D, k)
    nbytes = sweep_bytes(D, N, k)
    dense, executed = build_flops(D, N, len({lv for lv, _, _ in leaves}))  # FISHPACK leaf classes: one per leaf level
    b_s = builds[-1]["event_s"]
    return dict(builds=builds, build_s=b_s, ms_per_step=ms_step, wall_ms_per_step=wall_step, value=k / (ms_step / 1e3),
                rel_err=max(rel, rel1), upwards_ms=st["upwards-stage"], solve_ms=st["solve-stage"],
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)), nlat=nlat,
                sweep_bytes=nbytes, sweep_gbs=nbytes / (ms_step / 1e3) / 1e9,
                gemm_ms=gms.value, gemm_launches=glaunch.value, gemm_bytes=gbytes, gemm_gbs=gbytes / (gms.value / 1e3) / 1e9,
                build_tflops_dense=dense / b_s / 1e12, build_tflops_executed=executed / b_s / 1e12,
                dense_flops=dense, executed_flops=executed, nnodes=D.nnodes, M=D.M)


def serving_run(leaves, N, nreq, max_batch, seed=11):
    """The serving workload on the factorised mesh: ``nreq`` requests, each f = K ~ U{1..8}
    Gaussians and Dirichlet data of 8 Fourier modes (serving.workload_params), evaluated on the
    device.  saturated: all requests queued at once (capacity); poisson: the same count at
    Poisson arrivals of a fixed rate (80 % of capacity), latency = arrival -> result complete in
    HBM; plain: one request at a time through the same path, k = 1, waiting for each.  A sample
    of results is fetched and compared with a one-request solve (bitwise) and its max |u|."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import serving
    ctx = ef.Context(0)
    srv = serving.Server(ctx, max_batch=max_batch)
    dm = srv.add_mesh("cfg4", leaves, N, BOUNDS)
    P = serving.workload_params(np.random.default_rng(seed), BOUNDS, nreq)
    for p in P[:2 * max_batch]:
        srv.submit_load("cfg4", p, want_u=False)
    srv.run()
    srv.completed.clear()
    t0 = time.perf_counter()
    for p in P:
        srv.submit_load("cfg4", p, want_u=False)
    srv.run()
    cap = nreq / (time.perf_counter() - t0)
    sat = srv.stats()
    srv.completed.clear()
    rate = 0.8 * cap
    reqs = serving.drive_poisson(srv, "cfg4", P, rate_hz=rate, want_u_every=max(nreq // 20, 1), seed=seed)
    poi = srv.stats()
    sample = [q for q in reqs if q.u is not None]
    same = True
    for q in sample[:3]:
        alone = srv.submit_load("cfg4", q.params)
        srv.run()
        same = same and bool(np.array_equal(alone.u, q.u))
    srv.completed.clear()
    plain = serving.Server(ctx, max_batch=1, nslots=1, buckets=None)
    plain.meshes["cfg4"] = srv.meshes["cfg4"]  # same factorised mesh
    t0 = time.perf_counter()
    nplain = min(300, nreq)
    for p in P[:nplain]:
        plain.submit_load("cfg4", p, want_u=False)
        plain.run()
    plain_rate = nplain / (time.perf_counter() - t0)
    pst = plain.stats()
    return dict(workload="serving: %d requests on the adaptive Poisson mesh (%d leaves, %dx%d cells), each f = "
                "K~U{1..8} Gaussians (centres uniform, widths log-uniform [0.2,2], amplitudes N(0,1)), Dirichlet = "
                "8 Fourier modes; loads evaluated on the device from 64 parameters per request" % (
                    nreq, len(leaves), N, N),
                max_batch=max_batch, slots=srv.nslots, bucketed_widths=list(srv.buckets),
                saturated_solves_per_s=cap, saturated_p50_ms=sat["p50_ms"], saturated_p99_ms=sat["p99_ms"],
                poisson_rate_hz=rate, solves_per_s=poi["solves_per_s"], p50_ms=poi["p50_ms"], p99_ms=poi["p99_ms"],
                plain_per_request_solves_per_s=plain_rate, plain_p50_ms=pst["p50_ms"], plain_p99_ms=pst["p99_ms"],
                served_over_plain=cap / plain_rate, batch_invariant_sample=same, results_fetched=len(sample),
                note="latency: request arrival to its solution complete in HBM (results stay on the device; a "
                     "fetched 16 MB field adds its PCIe copy)")


def leaf_bounds(leaves, N):
    out = []
    for lv, ix, iy in leaves:
        w = (BOUNDS[1] - BOUNDS[0]) / (1 << lv)
        h = (BOUNDS[3] - BOUNDS[2]) / (1 << lv)
        out.append((BOUNDS[0] + ix * w, BOUNDS[0] + (ix + 1) * w, BOUNDS[2] + iy * h, BOUNDS[2] + (iy + 1) * h))
    return np.array(out)


def patch_inputs(P, N, k, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((k, P, 4 * N)), rng.standard_normal((k, P, N * N))


def patch_check_items(P, k, count=24, seed=1):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(k)), int(p)) for p in rng.choice(P, count, replace=False)]


def cpu_patch_baseline(bounds, N, k, sample, seconds=10.0):
    """Oracle leaf solves (the reference's FiniteVolumeSolver::solve restated), one thread,
    on the first patches of the workload; also the oracle's answers for the items the GPU
    result is checked on."""
    import oracle
    oracle.set_threads(1)
    g, f = patch_inputs(len(bounds), N, k)
    n, t0 = 0, time.perf_counter()
    while n < sample and time.perf_counter() - t0 < seconds:
        p = n % len(bounds)
        oracle.leaf_solve(N, bounds[p], g[0, p], f[0, p])
        n += 1
    dt = time.perf_counter() - t0
    check = {(c, p): oracle.leaf_solve(N, bounds[p], g[c, p], f[c, p]) for c, p in patch_check_items(len(bounds), k)}
    return dict(value=n / dt, patches=n, seconds=dt, check=check)


def cpu_leaf_op_baseline(op, bounds, N, cells=None, seconds=5.0, sample=2000):
    """Oracle leaf operations at one thread on the first patches of a leaf set (a bounded
    sample): "map" = particular Neumann data of f, "solve" = the leaf solve with g and f."""
    import oracle
    oracle.set_threads(1)
    g, f = patch_inputs(len(bounds), N, 1)
    n, t0 = 0, time.perf_counter()
    while n < sample and time.perf_counter() - t0 < seconds:
        p = n % len(bounds)
        kw = dict(coeffs=cells[p]) if cells is not None else {}
        if op == "map":
            oracle.leaf_map_d2n(N, bounds[p], np.zeros(4 * N), f[0, p], **kw)
        else:
            oracle.leaf_solve(N, bounds[p], g[0, p], f[0, p], **kw)
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=n / dt, patches=n, seconds=dt)


def patch_bytes_flops(P, N, k):
    """Algorithmic HBM bytes and flops of one batched leaf-solve launch: read f (N^2) and
    g (4N), write u (N^2) per patch and right-hand side; four N x N x N products."""
    NB = 1 if N <= 16 else 2
    Np = 16 * NB
    return P * k * (2 * N * N + 4 * N) * 8, P * k * 4 * 2 * Np ** 3


def gpu_patch_run(bounds, N, k, steps, warmup, check=None):
    """Leaf solves of the solve stage as a batch (ef_patches_apply_device, SOLVE): every
    leaf patch of the workload with k right-hand sides resident in HBM."""
    import ellipticforest_amd as ef
    ctx = ef.Context(0)
    P = len(bounds)
    pt = ef.Patches(ctx, bounds, N)
    g_h, f_h = patch_inputs(P, N, k)
    g, f, u = ef.DeviceArray(ctx, g_h.shape), ef.DeviceArray(ctx, f_h.shape), ef.DeviceArray(ctx, f_h.shape)
    g.upload(g_h)
    f.upload(f_h)
    for _ in range(warmup):
        pt.apply_device(ef._lib.EF_PATCH_SOLVE, g.ptr, f.ptr, u.ptr, k)
    pt.sync()
    # the timed region: `steps` launches queued back to back on the batch's stream between
    # two HIP events on that stream
    timer = ef.Timer(ctx)
    t0 = time.perf_counter()
    timer.start(pt.stream)
    for _ in range(steps):
        pt.apply_device(ef._lib.EF_PATCH_SOLVE, g.ptr, f.ptr, u.ptr, k)
    timer.stop(pt.stream)
    ms = timer.elapsed_ms() / steps
    wall = time.perf_counter() - t0
    # single launches, each between its own event pair (shows the per-launch spread)
    kern_ms = []
    for _ in range(steps):
        pt.apply_device(ef._lib.EF_PATCH_SOLVE, g.ptr, f.ptr, u.ptr, k)
        kern_ms.append(pt.last_ms())
    err = None
    if check:
        uu = u.download().reshape(k, P, N * N)
        err = max(float(np.abs(uu[c, p] - uo).max() / np.abs(uo).max()) for (c, p), uo in check.items())
    nbytes, nflops = patch_bytes_flops(P, N, k)
    return dict(ms=ms, ms_min=float(np.min(kern_ms)), ms_single_mean=float(np.mean(kern_ms)), wall_ms=wall * 1e3 / steps, value=P * k / (ms / 1e3),
                rel_err=err, bytes=nbytes, flops=nflops, gbs=nbytes / (ms / 1e3) / 1e9,
                tflops=nflops / (ms / 1e3) / 1e12)


def time_launches(ef, ctx, pt, op, g_ptr, f_ptr, out_ptr, k, steps, warmup):
    """Mean GPU time (ms) of `steps` queued launches between one pair of HIP events."""
    for _ in range(warmup):
        pt.apply_device(op, g_ptr, f_ptr, out_ptr, k)
    pt.sync()
    timer = ef.Timer(ctx)
    timer.start(pt.stream)
    for _ in range(steps):
        pt.apply_device(op, g_ptr, f_ptr, out_ptr, k)
    timer.stop(pt.stream)
    return timer.elapsed_ms() / steps


def gpu_particular_run(bounds, N, k, steps, warmup):
    """Particular Neumann data of every leaf (the leaf step of the upwards stage,
    FiniteVolumeSolver::particularNeumannData): f in, h out."""
    import ellipticforest_amd as ef
    ctx = ef.Context(0)
    P = len(bounds)
    pt = ef.Patches(ctx, bounds, N)
    _, f_h = patch_inputs(P, N, k)
    f, h = ef.DeviceArray(ctx, f_h.shape), ef.DeviceArray(ctx, (k, P, 4 * N))
    f.upload(f_h)
    ms = time_launches(ef, ctx, pt, ef._lib.EF_PATCH_MAP_D2N, None, f.ptr, h.ptr, k, steps, warmup)
    nbytes = P * k * (N * N + 4 * N) * 8
    return dict(ms=ms, value=P * k / (ms / 1e3), bytes=nbytes, gbs=nbytes / (ms / 1e3) / 1e9)


def var_beta_cells(bounds, N, beta=None, lam=-1.0):
    """alpha, beta at the four faces and lambda per cell (FiniteVolumeSolver.cpp:64-78) for
    alpha = 1 and the given beta (default 1 + 0.5 sin(2 pi x) cos(2 pi y)) and constant lambda."""
    if beta is None:
        beta = lambda x, y: 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    out = np.empty((len(bounds), 6, N * N))
    for p, (xl, xu, yl, yu) in enumerate(bounds):
        dx, dy = (xu - xl) / N, (yu - yl) / N
        X, Y = np.meshgrid(xl + (np.arange(N) + 0.5) * dx, yl + (np.arange(N) + 0.5) * dy, indexing="ij")
        X, Y = X.ravel(), Y.ravel()
        out[p] = [np.ones_like(X), beta(X - dx / 2, Y), beta(X + dx / 2, Y), beta(X, Y - dy / 2),
                  beta(X, Y + dy / 2), np.full_like(X, lam)]
    return out


def multi_mesh_run(nmeshes, nreq, max_batch=32, tick=64, seed=1, budget_frac=0.6):
    """Multi-mesh serving with admission by HBM footprint: ``nmeshes`` meshes drawn with seed 1
    -- max level ~ U{4..8}, N ~ {8, 16, 32}, alpha = 1, beta = 1 + 0.5 sin(a x + b y) with a, b ~
    U[0, 4], lambda ~ U[-2, 0], on [-10,10]^2 refined where |sin x + sin y| > 1.2, five-point
    leaves, FP64 -- registered with one server whose budget holds ``budget_frac`` of their
    footprints.  ``nreq`` cfg4-style load requests (Gaussians + Fourier Dirichlet modes) arrive
    in ticks of ``tick`` requests for one mesh each, the mesh drawn with probability 1/(i + 1)
    (normalised); a request for a mesh that is not resident builds it (evicting least recently
    used idle meshes).  Latency: arrival to result in HBM, builds included."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import serving, _lib
    ctx = ef.Context(0)
    rng = np.random.default_rng(seed)
    srv = serving.Server(ctx, max_batch=max_batch, buckets=(1, 2, 4, 8, 16, 32))
    specs = []
    for i in range(nmeshes):
        maxl, N = int(rng.integers(4, 9)), int(rng.choice([8, 16, 32]))
        a, b, lam = float(rng.uniform(0, 4)), float(rng.uniform(0, 4)), float(rng.uniform(-2, 0))
        leaves = ef.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, N, N, BOUNDS, 0, maxl)
        fp = _lib.DeviceMesh.footprint(leaves, N, BOUNDS, max_rhs=max_batch, variable=True) + \
            srv.slot_bytes(len(leaves), N, 0)
        probe = ef.DeviceMesh(ctx, leaves, N, BOUNDS)
        lb = np.array([tuple(probe.node_bounds[n]) for n in probe.leaf_nodes])
        del probe
        beta = (lambda aa, bb: (lambda x, y: 1.0 + 0.5 * np.sin(aa * x + bb * y)))(a, b)
        cell = var_beta_cells(lb, N, beta, lam)
        srv.register_mesh(f"m{i}", leaves, N, BOUNDS, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
        specs.append(dict(key=f"m{i}", max_level=maxl, N=N, leaves=len(leaves), a=a, b=b, lam=lam,
                          footprint_gb=fp / 1e9))
        log(f"multi-mesh: registered m{i}: max level {maxl}, N {N}, {len(leaves)} leaves, {fp / 1e9:.2f} GB")
    fps = [sp["footprint_gb"] * 1e9 for sp in specs]
    srv.budget = int(max(budget_frac * sum(fps), 1.05 * max(fps)))
    P = serving.workload_params(rng, BOUNDS, nreq)
    pop = 1.0 / np.arange(1, nmeshes + 1)
    ticks = rng.choice(nmeshes, size=(nreq + tick - 1) // tick, p=pop / pop.sum())
    t0 = time.perf_counter()
    for t, s0 in enumerate(range(0, nreq, tick)):
        for i in range(s0, min(nreq, s0 + tick)):
            srv.submit_load(specs[ticks[t]]["key"], P[i], want_u=False)
        srv.run()
        if t % 8 == 7:
            log(f"multi-mesh: tick {t + 1}/{len(ticks)}, {len(srv.admissions)} builds, {srv.evictions} evictions, "
                f"{time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    st = srv.stats()
    bt = np.array([t for _, t in srv.admissions])
    out = dict(workload="multi-mesh serving: %d meshes (seed %d: max level U{4..8}, N in {8,16,32}, beta = 1 + 0.5 "
               "sin(a x + b y), a, b ~ U[0,4], lambda ~ U[-2,0], alpha = 1, five-point, FP64) on [-10,10]^2 "
               "refined where |sin x + sin y| > 1.2; %d load requests in ticks of %d for one mesh each, drawn with "
               "probability proportional to 1/(i+1); "
               "HBM budget %.1f GB = %.0f%% of the meshes' footprints" % (nmeshes, seed, nreq, tick,
                                                                         srv.budget / 1e9, 100 * budget_frac),
               meshes=specs, solves_per_s=nreq / wall, p50_ms=st["p50_ms"], p99_ms=st["p99_ms"],
               admissions=len(bt), evictions=srv.evictions, build_s_mean=float(bt.mean()),
               build_s_max=float(bt.max()), build_s_total=float(bt.sum()),
               note="precision FP64: mixed-precision builds refine on uniform meshes only")
    log(f"multi-mesh: {out['solves_per_s']:.0f} solves/s, p50 {out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms, "
        f"{len(bt)} builds ({bt.sum():.2f} s), {srv.evictions} evictions")
    return out


def variable_path_run(steps, level=6, N=32, nrhs=4):
    """The whole path on the variable-coefficient workload ([-1,1]^2, uniform level 6 = 4,096
    leaves, 32x32 cells, alpha = 1, beta = 1 + 0.5 sin(2 pi x) cos(2 pi y), lambda = -1,
    manufactured u = sin(3x) cosh(y)), built in FP64 and in mixed precision (FP32 S and T
    products, FP64 X^-1, FP64 residual + 2 rounds of iterative refinement per solve): buildStage
    seconds (HIP events, steady state), upwards + solve per step of nrhs right-hand sides, the
    discretisation error against u, and the mixed result's error against the FP64 one."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    ctx = ef.Context(0)
    uex = lambda x, y: np.sin(3 * x) * np.cosh(y)
    beta = lambda x, y: 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)

    def fex(x, y):
        lap = -8.0 * np.sin(3 * x) * np.cosh(y)
        ux, uy = 3 * np.cos(3 * x) * np.cosh(y), np.sin(3 * x) * np.sinh(y)
        bx = np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
        by = -np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
        return beta(x, y) * lap + bx * ux + by * uy - uex(x, y)

    B = (-1.0, 1.0, -1.0, 1.0)
    leaves = ef.refine_by_function(lambda x, y: np.ones_like(x, dtype=bool), N, N, B, level, level)
    out = dict(workload="variable-coefficient elliptic on [-1,1]^2, uniform level %d (%d leaves), %dx%d cells, "
               "alpha=1, beta=1+0.5 sin(2 pi x) cos(2 pi y), lambda=-1, manufactured u=sin(3x)cosh(y), %d "
               "right-hand sides per step" % (level, len(leaves), N, N, nrhs))
    u64 = None
    for name, prec in (("fp64", _lib.EF_PRECISION_FP64), ("mixed", _lib.EF_PRECISION_MIXED)):
        D = ef.DeviceMesh(ctx, leaves, N, B)
        lb = [tuple(D.node_bounds[n]) for n in D.leaf_nodes]
        cell = var_beta_cells(np.array(lb), N)
        builds = []
        for _ in range(2):
            D.build(solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell, precision=prec)
            D.sync()
            builds.append(D.stage_ms()["build-stage"] / 1e3)
        centres = []
        for (xl, xu, yl, yu) in lb:
            dx, dy = (xu - xl) / N, (yu - yl) / N
            X, Y = np.meshgrid(xl + dx / 2 + np.arange(N) * dx, yl + dy / 2 + np.arange(N) * dy, indexing="ij")
            centres.append((X.ravel(), Y.ravel()))
        f1 = np.stack([fex(X, Y) for X, Y in centres])
        ue = np.stack([uex(X, Y) for X, Y in centres])
        M = D.M
        dxr = 2.0 / M
        t = -1 + dxr / 2 + np.arange(M) * dxr
        px = np.concatenate([np.full(M, -1.0), np.full(M, 1.0), t, t])
        py = np.concatenate([t, t, np.full(M, -1.0), np.full(M, 1.0)])
        nl = len(leaves)
        f = ef.DeviceArray(ctx, (nrhs, nl, N * N))
        r = ef.DeviceArray(ctx, (nrhs, 4 * M))
        a = ef.DeviceArray(ctx, (4 * M,))
        u = ef.DeviceArray(ctx, (nrhs, nl, N * N))
        f.upload(np.broadcast_to(f1, (nrhs,) + f1.shape))
        r.upload(np.broadcast_to(uex(px, py), (nrhs, 4 * M)))
        a.upload(np.ones(4 * M))
        D.upwards_device(f.ptr, nrhs)
        D.solve_device(f.ptr, r.ptr, a.ptr, None, u.ptr, nrhs)
        D.sync()
        timer = ef.Timer(ctx)
        timer.start(D.stream)
        for _ in range(steps):
            D.upwards_device(f.ptr, nrhs)
            D.solve_device(f.ptr, r.ptr, a.ptr, None, u.ptr, nrhs)
        timer.stop(D.stream)
        ms = timer.elapsed_ms() / steps
        uu = u.download().reshape(nrhs, nl, N * N)[0]
        res = dict(build_s=builds[-1], build_first_s=builds[0], ms_per_step=ms, solves_per_s=nrhs / (ms / 1e3),
                   disc_error=float(np.abs(uu - ue).max()))
        if u64 is None:
            u64 = uu
        else:
            res["rel_err_vs_fp64"] = float(np.abs(uu - u64).max() / np.abs(u64).max())
            res["refinement_rounds"] = 2
        out[name] = res
        log(f"variable path {name}: build {builds[-1] * 1e3:.1f} ms, step {ms:.3f} ms, disc err {res['disc_error']:.3e}"
            + (f", vs fp64 {res['rel_err_vs_fp64']:.2e}" if "rel_err_vs_fp64" in res else ""))
        del D
    out["mixed_build_speedup"] = out["fp64"]["build_s"] / out["mixed"]["build_s"]
    return out


def gpu_variable_run(steps, warmup, level=6, N=32):
    """Leaf solves of the variable-beta leaf set ([-1,1]^2, uniform level 6, N = 32: 4,096 patches,
    five-point stencil with variable beta), one right-hand side per launch."""
    import ellipticforest_amd as ef
    ctx = ef.Context(0)
    w = 2.0 / (1 << level)
    bounds = np.array([(-1 + i * w, -1 + (i + 1) * w, -1 + j * w, -1 + (j + 1) * w)
                       for i in range(1 << level) for j in range(1 << level)])
    P = len(bounds)
    pt = ef.Patches(ctx, bounds, N, solver_type=ef._lib.EF_SOLVER_FIVE_POINT, cell=var_beta_cells(bounds, N))
    g_h, f_h = patch_inputs(P, N, 1)
    g, f, u = ef.DeviceArray(ctx, g_h.shape), ef.DeviceArray(ctx, f_h.shape), ef.DeviceArray(ctx, f_h.shape)
    g.upload(g_h)
    f.upload(f_h)
    ms = time_launches(ef, ctx, pt, ef._lib.EF_PATCH_SOLVE, g.ptr, f.ptr, u.ptr, 1, steps, warmup)
    # the leaf DtN of the set (buildD2N), once, outside the timed region: its kernel shows in a
    # kernel trace of this run; the wall time includes copying T (537 MB) to the host
    t0 = time.perf_counter()
    pt.dtn()
    dtn_wall_ms = (time.perf_counter() - t0) * 1e3
    dtn_ms = pt.last_ms()  # HIP events around the DtN launches on the batch's stream
    # per patch: the inverse pivot blocks, read by the forward and again by the backward sweep
    # (N blocks of N (N + 1) / 2 packed entries for a symmetric matrix, N^2 otherwise), the
    # stencil's cW and cE (N^2 each) and cS, cN on the two edge columns, f and u (N^2 each), g (4N)
    block = N * (N + 1) // 2 if pt.packed else N * N
    nbytes = P * (2 * N * block + 2 * N * N + 2 * N + 2 * N * N + 4 * N) * 8
    return dict(ms=ms, value=P / (ms / 1e3), bytes=nbytes, gbs=nbytes / (ms / 1e3) / 1e9, patches=P,
                packed=pt.packed, dtn_ms=dtn_ms, dtn_wall_ms=dtn_wall_ms, bounds=bounds,
                cells=pt_cells_sample(bounds, N))


def pt_cells_sample(bounds, N, n=512):
    """Per-cell coefficients of the first n patches (the CPU baseline's sample)."""
    return var_beta_cells(bounds[:n], N)


def traffic_probe(bounds, N, k, copy_doubles=32 << 20, level=6, NV=32):
    """Launches for the byte counters: 5 leaf solves and 5 particular-Neumann-data maps of the
    leaf set, 5 leaf solves of the variable-beta leaf set, then 3 plain copies of known size."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    ctx = ef.Context(0)
    P = len(bounds)
    g_h, f_h = patch_inputs(P, N, k)
    g, f, u = ef.DeviceArray(ctx, g_h.shape), ef.DeviceArray(ctx, f_h.shape), ef.DeviceArray(ctx, f_h.shape)
    h = ef.DeviceArray(ctx, g_h.shape)
    g.upload(g_h)
    f.upload(f_h)
    pt = ef.Patches(ctx, bounds, N)
    for _ in range(5):
        pt.apply_device(_lib.EF_PATCH_SOLVE, g.ptr, f.ptr, u.ptr, k)
    for _ in range(5):
        pt.apply_device(_lib.EF_PATCH_MAP_D2N, None, f.ptr, h.ptr, k)
    pt.sync()
    w = 2.0 / (1 << level)
    vb = np.array([(-1 + i * w, -1 + (i + 1) * w, -1 + j * w, -1 + (j + 1) * w)
                   for i in range(1 << level) for j in range(1 << level)])
    pv = ef.Patches(ctx, vb, NV, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=var_beta_cells(vb, NV))
    vg_h, vf_h = patch_inputs(len(vb), NV, 1)
    vg, vf, vu = ef.DeviceArray(ctx, vg_h.shape), ef.DeviceArray(ctx, vf_h.shape), ef.DeviceArray(ctx, vf_h.shape)
    vg.upload(vg_h)
    vf.upload(vf_h)
    for _ in range(5):
        pv.apply_device(_lib.EF_PATCH_SOLVE, vg.ptr, vf.ptr, vu.ptr, 1)
    pv.sync()
    a, b = ef.DeviceArray(ctx, (copy_doubles,)), ef.DeviceArray(ctx, (copy_doubles,))
    for _ in range(3):
        _lib.check(ctx.L.ef_memcpy(ctx.h, b.ptr, a.ptr, copy_doubles * 8, 2), ctx.L)
    log(f"traffic probe: leaf solve {patch_bytes_flops(P, N, k)[0]} algorithmic bytes per launch; "
        f"copy reads and writes {copy_doubles * 8} bytes each")


def gemm_traffic_probe(leaves, N, k, copy_doubles=32 << 20):
    """Counter-pass workload: the step's tree GEMMs of the default mesh at k right-hand sides
    (3 passes through ef_mesh_time_sweep_gemms, after one warm pass) and 3 device-to-device
    copies of 256 MiB on the library's copy kernel (the calibration of the byte counters)."""
    import ctypes
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    ctx = ef.Context(0)
    D = ef.DeviceMesh(ctx, leaves, N, BOUNDS)
    D.build()
    ms, nl = ctypes.c_double(0.0), ctypes.c_int(0)
    _lib.check(ctx.L.ef_mesh_time_sweep_gemms(D.h, k, 3, ctypes.byref(ms), ctypes.byref(nl)), ctx.L)
    a, b = ef.DeviceArray(ctx, (copy_doubles,)), ef.DeviceArray(ctx, (copy_doubles,))
    for _ in range(3):
        _lib.check(ctx.L.ef_memcpy(ctx.h, b.ptr, a.ptr, copy_doubles * 8, 2), ctx.L)
    log(f"gemm traffic probe: {nl.value} GEMM launches per pass, {ms.value:.3f} ms per pass, "
        f"algorithmic {gemm_bytes(D, k) / 1e6:.1f} MB per pass")


def copy_bandwidth(doubles=64 << 20, reps=5):
    """Achievable HBM bandwidth of a plain copy on this GPU (the library's copy kernel through
    ef_memcpy, 512 MiB read + 512 MiB written), best of `reps`, in GB/s."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    ctx = ef.Context(0)
    a, b = ef.DeviceArray(ctx, (doubles,)), ef.DeviceArray(ctx, (doubles,))
    t = ef.Timer(ctx)
    best = None
    for _ in range(reps):
        t.start(None)
        _lib.check(ctx.L.ef_memcpy(ctx.h, b.ptr, a.ptr, doubles * 8, 2), ctx.L)
        t.stop(None)
        ms = t.elapsed_ms()
        best = ms if best is None else min(best, ms)
    return 2 * doubles * 8 / (best / 1e3) / 1e9


def leaf_components(out, leaves, args, cpu_patch):
    """Leaf-patch kernels on their own (the leaf stage of the path as one batch)."""
    bounds = leaf_bounds(leaves, args.N)
    k = args.patch_nrhs
    pr = gpu_patch_run(bounds, args.N, k, args.steps, args.warmup, cpu_patch["check"])
    log(f"leaf solves: {pr['ms']:.4f} ms per launch of {len(bounds)} x {k}, {pr['gbs']:.0f} GB/s")
    comps = [dict(
        name="leaf solve stage (FiniteVolumeSolver::solve on every leaf, ef_patches_apply_device)",
        kernel="patch_fd_kernel<1, true, 0>", value=pr["value"], unit="leaf solves/s", ms_per_launch=pr["ms"],
        patches=len(bounds), nrhs=k, rel_err_vs_fp64_oracle=pr["rel_err"], achieved_gbs=pr["gbs"],
        algorithmic_bytes_per_launch=pr["bytes"],
        traffic_from_profiles=TRAFFIC_PER_LAUNCH if k == 16 else None)]
    pp = gpu_particular_run(bounds, args.N, k, args.steps, args.warmup)
    comps.append(dict(
        name="upwards leaf stage (particularNeumannData on every leaf, ef_patches_apply_device MAP_D2N)",
        kernel="patch_fd_kernel<1, true, 1>", value=pp["value"], unit="leaf maps/s", ms_per_launch=pp["ms"],
        patches=len(bounds), nrhs=k, achieved_gbs=pp["gbs"], algorithmic_bytes_per_launch=pp["bytes"]))
    pv = gpu_variable_run(args.steps, args.warmup)
    comps.append(dict(
        name="variable-beta leaf solves (five-point, [-1,1]^2 uniform level 6: 4,096 patches, N = 32, 1 right-hand side)",
        kernel="bt_solve_sym_kernel<32>" if pv["packed"] else "bt_solve_wave_kernel<32>", value=pv["value"],
        unit="leaf solves/s", ms_per_launch=pv["ms"], patches=pv["patches"], nrhs=1, achieved_gbs=pv["gbs"],
        algorithmic_bytes_per_launch=pv["bytes"], leaf_dtn_ms=pv["dtn_ms"]))
    out["components"] = comps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--max-level", type=int, default=7)
    ap.add_argument("--N", type=int, default=16)
    ap.add_argument("--cpu-solves", type=int, default=3)
    ap.add_argument("--latency-requests", type=int, default=200)
    ap.add_argument("--patch-nrhs", type=int, default=16)
    ap.add_argument("--leaf-components", action="store_true", help="also time the leaf-patch kernels on their own")
    ap.add_argument("--gemm-traffic-probe", action="store_true",
                    help="only the step's tree GEMMs (1 + 3 passes) and 3 copies of 256 MiB on the library's copy "
                         "kernel (the calibration of HBM byte counters), then exit")
    ap.add_argument("--serving-requests", type=int, default=10000, help="requests of the serving run (0: skip)")
    ap.add_argument("--no-variable-path", action="store_true", help="skip the variable-coefficient whole path")
    ap.add_argument("--multi-mesh-requests", type=int, default=1536, help="requests of the multi-mesh run (0: skip)")
    ap.add_argument("--multi-mesh-count", type=int, default=6)
    ap.add_argument("--serving-batch", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the oracle timings (profiling runs)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only 5 leaf-solve, 5 particular-Neumann-data and 5 variable-beta leaf-solve launches and 3 "
                         "device-to-device copies of 256 MiB on the library's copy kernel (the calibration of "
                         "HBM byte counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the HPS path of this build runs on a single MI355X; "
              f"--gpus {args.gpus} is not supported (multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    leaves = workload(args.max_level, args.N)
    log(f"workload: {len(leaves)} leaves, N={args.N}")
    if args.traffic_probe:
        traffic_probe(leaf_bounds(leaves, args.N), args.N, args.patch_nrhs)
        return
    if args.gemm_traffic_probe:
        gemm_traffic_probe(leaves, args.N, args.nrhs)
        return
    threads = os.cpu_count() or 1
    if args.no_cpu_baseline:
        import oracle
        O = oracle.Forest(leaves, args.N, BOUNDS, solver=oracle.FISHPACK)
        f = O.sample_rhs(F)
        r, a, b = O.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
        ref = dict(solves_per_s=None, build_s=None, disc_err=None, u=None, f=f, r=r, a=a)
    else:
        log(f"cpu baseline: oracle build with {threads} threads, {args.cpu_solves} single-thread solves")
        ref = cpu_baseline(leaves, args.N, args.cpu_solves, threads)
        log(f"cpu baseline: build {ref['build_s']:.2f}s, solve {ref['solve_s']:.2f}s (1 thread), "
            f"{1.0 / ref['solves_per_s_all']:.2f}s ({threads} threads), disc err {ref['disc_err']:.3e}")
    out = dict(metric=METRIC, value=None, unit="solves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=None, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic (analytic f and Dirichlet data sampled at cell centres / root face points)",
               config=dict(workload="adaptive Poisson [-10,10]^2, |f|>1.2 refinement, levels 0..%d, %dx%d cells per "
                           "patch, FISHPACK semantics, %d right-hand sides per step" % (
                               args.max_level, args.N, args.N, args.nrhs),
                           leaves=len(leaves), nrhs_per_step=args.nrhs, parallelism="single GPU"),
               build_s=None, p50_ms=None, p99_ms=None, rel_err_vs_fp64_oracle=None, roofline=None,
               cpu_baseline=None if args.no_cpu_baseline else dict(
                   value=ref["solves_per_s"], unit="solves/s", cores=1, kind="port",
                   sample="oracle (reference algorithm restated in C++, -O3 -march=x86-64-v3): upwards + solve of "
                          "1 right-hand side on the full workload, mean of %d passes at 1 thread" % args.cpu_solves,
                   value_all_cores=ref["solves_per_s_all"], host_threads=threads,
                   build_s=ref["build_s"], build_threads=threads, discretization_error=ref["disc_err"]))
    if ref["u"] is None:
        ref["u"] = np.zeros((len(leaves), args.N * args.N))
    try:
        g = gpu_run(leaves, args.N, args.nrhs, args.steps, args.warmup, ref, args.latency_requests)
        log(f"step {g['ms_per_step']:.4f} ms ({g['value']:.0f} solves/s), sweep {g['sweep_gbs']:.0f} GB/s, "
            f"build {g['build_s'] * 1e3:.2f} ms, p50 {g['p50_ms']:.3f} ms p99 {g['p99_ms']:.3f} ms, rel err {g['rel_err']:.2e}")
        out.update(value=g["value"], ms_per_step=g["ms_per_step"], build_s=g["build_s"], p50_ms=g["p50_ms"],
                   p99_ms=g["p99_ms"], rel_err_vs_fp64_oracle=None if args.no_cpu_baseline else g["rel_err"],
                   wall_ms_per_step=g["wall_ms_per_step"], upwards_ms=g["upwards_ms"], solve_ms=g["solve_ms"],
                   build_calls=g["builds"], latency_requests=g["nlat"])
        per_launch = g["gemm_bytes"] / max(g["gemm_launches"], 1)
        out["roofline"] = dict(
            bound="hbm", achieved=g["gemm_gbs"], peak=8000.0, unit="GB/s", frac=g["gemm_gbs"] / 8000.0,
            traffic=TRAFFIC_GEMM_PER_LAUNCH, kernel="tree_gemm_kernel (the step's tree GEMMs: w = X^-1 hDiff, "
            "h += H w, u = S g + w; %d launches per step, split-K chunks summed inside the workgroup)" % g["gemm_launches"],
            algorithmic_bytes_per_launch=per_launch, gemm_ms_per_step=g["gemm_ms"],
            note="achieved = algorithmic bytes of the step's GEMM launches / their HIP-event time (GEMMs alone, "
                 "ef_mesh_time_sweep_gemms); traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE / "
                 "WRITE_SIZE passes scaled by the copy-kernel factors (profiles/)")
        out["step_roofline"] = dict(
            bound="hbm", achieved=g["sweep_gbs"], peak=8000.0, unit="GB/s", frac=g["sweep_gbs"] / 8000.0,
            algorithmic_bytes_per_step=g["sweep_bytes"],
            note="whole step: operator bytes + leaf loads and solutions over the HIP-event step time")
        out["build_roofline"] = dict(
            bound="mfma", peak=78.6, unit="TFLOP/s", achieved_dense_count=g["build_tflops_dense"],
            achieved_executed=g["build_tflops_executed"], frac_executed=g["build_tflops_executed"] / 78.6,
            dense_flops=g["dense_flops"], executed_flops=g["executed_flops"],
            speedup_vs_cpu_build=None if args.no_cpu_baseline else ref["build_s"] / g["build_s"])
    except Exception as e:  # the device path reports why it did not run
        out["error"] = f"{type(e).__name__}: {e}"
        log(out["error"])
    if args.serving_requests > 0 and "error" not in out:
        try:
            sv = serving_run(leaves, args.N, args.serving_requests, args.serving_batch)
            log(f"serving: {sv['saturated_solves_per_s']:.0f} solves/s saturated, Poisson {sv['solves_per_s']:.0f}/s "
                f"p50 {sv['p50_ms']:.2f} ms p99 {sv['p99_ms']:.2f} ms; plain {sv['plain_per_request_solves_per_s']:.0f}/s")
            out["serving"] = sv
        except Exception as e:
            out["serving_error"] = f"{type(e).__name__}: {e}"
            log(out["serving_error"])
    if args.multi_mesh_requests > 0 and "error" not in out:
        try:
            out["multi_mesh"] = multi_mesh_run(args.multi_mesh_count, args.multi_mesh_requests)
        except Exception as e:
            out["multi_mesh_error"] = f"{type(e).__name__}: {e}"
            log(out["multi_mesh_error"])
    if not args.no_variable_path and "error" not in out:
        try:
            out["variable_path"] = variable_path_run(max(args.steps // 4, 3))
        except Exception as e:
            out["variable_path_error"] = f"{type(e).__name__}: {e}"
            log(out["variable_path_error"])
    if args.leaf_components:
        try:
            cpu_patch = dict(check=None) if args.no_cpu_baseline else cpu_patch_baseline(
                leaf_bounds(leaves, args.N), args.N, args.patch_nrhs, 20000)
            leaf_components(out, leaves, args, cpu_patch)
        except Exception as e:
            out["components_error"] = f"{type(e).__name__}: {e}"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
