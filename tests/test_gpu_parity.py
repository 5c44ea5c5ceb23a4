"""GPU parity: the HIP library, called through the C ABI, against the oracle restatement
and the independent global-solve fixtures.

Tolerances (FP64 path): relative max-norm 1e-10 against the oracle on the solution and
1e-11 on stored operators; the discretisation error must equal the oracle's to 1e-10
relative.  Integer/index work (mesh, node tables) is compared exactly.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import mesh as omesh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U = lambda x, y: np.sin(x) + np.sin(y)
F = lambda x, y: -(np.sin(x) + np.sin(y))
TOL_U = 1e-10
TOL_OP = 1e-11


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def solve_both(ctx, leaves, N, bounds, lam=0.0, check_ops=False, nodes=None):
    import ellipticforest_amd as ef
    O = oracle.Forest(leaves, N, bounds, solver=oracle.FISHPACK, lambda0=lam)
    O.build()
    f = O.sample_rhs(F)
    O.upwards(f)
    r, a, b = O.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    uo = O.solve(r, a, b)
    D = ef.DeviceMesh(ctx, leaves, N, bounds)
    D.build(lambda0=lam)
    D.upwards(f)
    ug = D.solve(r, a)
    if check_ops:
        for node in (nodes if nodes is not None else range(D.nnodes)):
            assert rel(D.get(node, "T"), O.get(node, "T")) < TOL_OP, ("T", node)
            if not D.node_info[node, 3]:
                for w in ("S", "X", "H"):
                    assert rel(D.get(node, w), O.get(node, w)) < TOL_OP, (w, node)
                # w is part of the solve (X^-1 hDiff): the library applies the stored inverse, the
                # oracle re-solves with LU (mergeW_), so it carries cond(X) eps like the solution
                assert rel(D.get(node, "w"), O.get(node, "w")) < TOL_U, ("w", node)
            assert rel(D.get(node, "g"), O.get(node, "g")) < TOL_U, ("g", node)
    return O, D, uo, ug


def test_uniform_L1_N8(ctx):
    say("uniform level 1, N=8, [-1,1]^2, FISHPACK")
    O, D, uo, ug = solve_both(ctx, omesh.uniform(1), 8, (-1, 1, -1, 1), check_ops=True)
    assert rel(ug, uo) < TOL_U


@pytest.mark.parametrize("name", ["global_5pt_L2N8_pm1", "global_5pt_L2N8_pm10", "global_5pt_L2N8_helmholtz",
                                  "global_5pt_L3N16_pm10"])
def test_uniform_equals_global_solve(ctx, name):
    import ellipticforest_amd as ef
    say(f"golden {name}")
    d = np.load(os.path.join(GOLD, name + ".npz"))
    level, N, bounds, lam = int(d["level"]), int(d["N"]), tuple(d["bounds"]), float(d["lam"])
    O = oracle.Forest(omesh.uniform(level), N, bounds, solver=oracle.FISHPACK, lambda0=lam)
    f = O.sample_rhs(F)
    D = ef.DeviceMesh(ctx, omesh.uniform(level), N, bounds)
    side, px, py = oracle.root_face_points(D.M, bounds)
    D.build(lambda0=lam)
    D.upwards(f)
    u = D.solve(U(px, py), np.ones_like(px))
    assert rel(u, d["u"]) < 1e-11


def test_uniform_level3_operators_and_error(ctx):
    """[-10,10]^2, uniform level 3, 16x16 cells per patch, FISHPACK Poisson."""
    say("uniform L3 N16 [-10,10]^2, operators of every node")
    O, D, uo, ug = solve_both(ctx, omesh.uniform(3), 16, (-10, 10, -10, 10), check_ops=True)
    assert rel(ug, uo) < TOL_U
    eo = eg = 0.0
    for k, bd in enumerate(O.leaf_bounds):
        X, Y, _, _ = oracle.cell_centres(16, bd)
        eo = max(eo, np.abs(uo[k] - U(X, Y)).max())
        eg = max(eg, np.abs(ug[k] - U(X, Y)).max())
    assert abs(eg - eo) <= 1e-10 * eo


def test_adaptive_fixture(ctx):
    import ellipticforest_amd as ef
    say("adaptive fixture: |f|>1.2, levels 0..3, N=8")
    d = np.load(os.path.join(GOLD, "oracle_fixtures.npz"))
    leaves = [tuple(int(v) for v in l) for l in d["adaptive_leaves"]]
    O = oracle.Forest(leaves, 8, (-10, 10, -10, 10), solver=oracle.FISHPACK)
    f = O.sample_rhs(F)
    D = ef.DeviceMesh(ctx, leaves, 8, (-10, 10, -10, 10))
    side, px, py = oracle.root_face_points(D.M, (-10, 10, -10, 10))
    D.build()
    D.upwards(f)
    assert rel(D.solve(U(px, py), np.ones_like(px)), d["adaptive_u"]) < TOL_U


@pytest.mark.parametrize("maxl,N,lam", [(4, 8, 0.0), (5, 16, -1.0), (6, 8, 0.0)])
def test_adaptive_vs_oracle(ctx, maxl, N, lam):
    say(f"adaptive |f|>1.2 levels 0..{maxl}, N={N}, lambda={lam}")
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, N, N, -10, 10, -10, 10, 0, maxl)
    O, D, uo, ug = solve_both(ctx, leaves, N, (-10, 10, -10, 10), lam=lam, check_ops=(maxl == 4))
    assert rel(ug, uo) < TOL_U


def test_adaptive_linear_exactness(ctx):
    import ellipticforest_amd as ef
    say("adaptive levels 0..5 N=8: linear solution reproduced")
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, 5)
    D = ef.DeviceMesh(ctx, leaves, 8, (-10, 10, -10, 10))
    D.build()
    D.upwards(np.zeros((len(leaves), 64)))
    lin = lambda x, y: 0.3 * x - 0.7 * y + 2.0
    side, px, py = oracle.root_face_points(D.M, (-10, 10, -10, 10))
    u = D.solve(lin(px, py), np.ones_like(px))
    for k, node in enumerate(D.leaf_nodes):
        X, Y, _, _ = oracle.cell_centres(8, tuple(D.node_bounds[node]))
        assert np.abs(u[k] - lin(X, Y)).max() < 1e-10


def test_multi_rhs_batch_equals_single(ctx, k=5, maxl=4, cols=None):
    import ellipticforest_amd as ef
    say(f"multi-RHS: batch of {k} vs one at a time, adaptive levels 0..{maxl} N=8")
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, maxl)
    D = ef.DeviceMesh(ctx, leaves, 8, (-10, 10, -10, 10))
    D.build()
    rng = np.random.default_rng(0)
    f = rng.standard_normal((k, len(leaves), 64))
    r = rng.standard_normal((k, 4 * D.M))
    a = np.ones(4 * D.M)
    D.upwards(f)
    ub = D.solve(r, a)
    for j in (range(k) if cols is None else cols):
        D.upwards(f[j])
        us = D.solve(r[j], a)
        assert np.array_equal(ub[j], us), j  # every column reduced in a batch-independent order


def test_multi_rhs_wide_batch(ctx):
    """More than 16 columns: two-column-tile GEMMs (and 16-row tiles where M <= 16)."""
    test_multi_rhs_batch_equals_single(ctx, k=40, maxl=4, cols=(0, 17, 39))


def tile_case(ctx, k, maxl, N):
    """Sweep GEMM tile shapes: the fixed rule, each forced tile (16x16, 16x32, 32x32) and the
    timed choice give the same bits (every output element sums its k range in one order)."""
    import ellipticforest_amd as ef
    say(f"sweep tiles: fixed rule vs forced 11/12/22 vs timed, k={k} levels 0..{maxl} N={N}")
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, N, N, -10, 10, -10, 10, 0, maxl)
    D = ef.DeviceMesh(ctx, leaves, N, (-10, 10, -10, 10))
    D.build()
    rng = np.random.default_rng(4)
    f = rng.standard_normal((k, len(leaves), N * N))
    r = rng.standard_normal((k, 4 * D.M))
    a = np.ones(4 * D.M)
    out = {}
    for mode in (0, 11, 12, 22, 1):
        D.set_autotune(mode)
        tiles = D.sweep_tiles(k)
        assert tiles.shape == (D.maxdepth + 1, 3)
        used = set(tiles.ravel().tolist()) - {0}
        assert used <= {11, 12, 22} and used, tiles
        # (split-K launches, the upper levels' long products, always run 16 x 16 tiles: 11)
        if mode > 1:
            assert used <= {mode, 11}
        if mode == 0:
            assert used <= ({11} if k <= 16 else {11, 12, 22})
        D.upwards(f)
        out[mode] = D.solve(r, a)
    for mode in (11, 12, 22, 1):
        assert np.array_equal(out[mode], out[0]), mode
    # a tuned width's solve needs its own upwards stage: the timed passes overwrite the state
    from ellipticforest_amd._lib import EfError
    D.upwards(f)
    D.sweep_tiles(k + 1)
    with pytest.raises(EfError, match="preceding upwards"):
        D.solve(r, a)
    return out[0]


def test_sweep_tiles_bitwise(ctx):
    tile_case(ctx, 40, 5, 8)
    tile_case(ctx, 7, 4, 16)


def test_homogeneous_rhs_option(ctx, level=2):
    import ellipticforest_amd as ef
    say("homogeneous-rhs option: leaves solve with f = 0, w skipped")
    leaves = omesh.uniform(level)
    O = oracle.Forest(leaves, 8, (0, 1, 0, 1), solver=oracle.FISHPACK, homogeneous_rhs=True)
    O.build()
    f = O.sample_rhs(F)
    O.upwards(f)
    r, a, b = O.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    uo = O.solve(r, a, b)
    D = ef.DeviceMesh(ctx, leaves, 8, (0, 1, 0, 1))
    D.set_options(homogeneous_rhs=True)
    D.build()
    D.upwards(f)
    assert rel(D.solve(r, a), uo) < TOL_U


def test_robin_root_condition(ctx, level=2):
    import ellipticforest_amd as ef
    say(f"root condition a u + b du/dn = r with b != 0, uniform level {level} N=8")
    leaves = omesh.uniform(level)
    O = oracle.Forest(leaves, 8, (0, 1, 0, 1), solver=oracle.FISHPACK)
    O.build()
    f = O.sample_rhs(F)
    O.upwards(f)
    r, a, b = O.boundary_data(lambda s, x, y: (U(x, y), 1.0 + 0 * x, 0.25 + 0.1 * np.cos(x + y)))
    uo = O.solve(r, a, b)
    D = ef.DeviceMesh(ctx, leaves, 8, (0, 1, 0, 1))
    D.build()
    D.upwards(f)
    assert rel(D.solve(r, a, b), uo) < TOL_U


def test_adaptive_poisson_full_size(ctx):
    """The adaptive Poisson mesh at full size (|f| > 1.2, levels 0..7): 7,750 leaves, 16x16 cells, against the oracle."""
    import ellipticforest_amd as ef
    say("adaptive Poisson full size: |f|>1.2 levels 0..7, N=16, 7750 leaves")
    oracle.set_threads(min(16, os.cpu_count() or 1))
    leaves = ef.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 16, 16, (-10, 10, -10, 10), 0, 7)
    assert len(leaves) == 7750
    O, D, uo, ug = solve_both(ctx, leaves, 16, (-10, 10, -10, 10), check_ops=True, nodes=[0, 1, 2, 5, 21])
    assert rel(ug, uo) < TOL_U


# ----- variable coefficients (the variable-beta operator at reduced sizes) ------------
ALPHA = lambda x, y: np.ones_like(x)
BETA = lambda x, y: 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
LAMBDA = lambda x, y: -np.ones_like(x)
U3 = lambda x, y: np.sin(3 * x) * np.cosh(y)


def F3(x, y):
    """f = div(beta grad u) - u for u = sin(3x) cosh(y) (alpha = 1, lambda = -1)."""
    lap = -8.0 * np.sin(3 * x) * np.cosh(y)
    ux, uy = 3 * np.cos(3 * x) * np.cosh(y), np.sin(3 * x) * np.sinh(y)
    bx = np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
    by = -np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    return BETA(x, y) * lap + bx * ux + by * uy - U3(x, y)


def variable_both(ctx, leaves, N, bounds=(-1.0, 1.0, -1.0, 1.0), check_ops=False):
    import ellipticforest_amd as ef
    O = oracle.Forest(leaves, N, bounds, solver=oracle.FIVE_POINT, coeff_fns=(ALPHA, BETA, LAMBDA))
    O.build()
    f = O.sample_rhs(F3)
    O.upwards(f)
    r, a, b = O.boundary_data(lambda s, x, y: (U3(x, y), 1.0, 0.0))
    uo = O.solve(r, a, b)
    D = ef.DeviceMesh(ctx, leaves, N, bounds)
    cell = np.concatenate([oracle.sample_coeffs(N, tuple(D.node_bounds[n]), ALPHA, BETA, LAMBDA).ravel()
                           for n in D.leaf_nodes])
    D.build(solver_type=ef._lib.EF_SOLVER_FIVE_POINT, cell=cell)
    D.upwards(f)
    ug = D.solve(r, a)
    if check_ops:
        for node in range(D.nnodes):
            assert rel(D.get(node, "T"), O.get(node, "T")) < TOL_OP, ("T", node)
            if not D.node_info[node, 3]:
                assert rel(D.get(node, "S"), O.get(node, "S")) < TOL_OP, ("S", node)
    err = max(np.abs(ug[k] - U3(*oracle.cell_centres(N, tuple(D.node_bounds[n]))[:2])).max()
              for k, n in enumerate(D.leaf_nodes))
    return uo, ug, err


def test_variable_coefficients_uniform(ctx):
    say("variable beta, lambda=-1, uniform level 2, N=16, [-1,1]^2")
    uo, ug, err = variable_both(ctx, omesh.uniform(2), 16, check_ops=True)
    assert rel(ug, uo) < TOL_U


def test_variable_coefficients_adaptive(ctx, maxl=4, N=8, check_ops=True):
    """Variable beta on an adaptive mesh (coarsened children, leaves of several sizes), every
    operator and the solution against the oracle."""
    say(f"variable beta, lambda=-1, adaptive levels 0..{maxl}, N={N}, [-1,1]^2")
    leaves = omesh.refine_by_function(lambda x, y: (x < -0.4) & (y < 0.1), N, N, -1, 1, -1, 1, 0, maxl)
    assert len({lv for lv, _, _ in leaves}) > 1
    uo, ug, err = variable_both(ctx, leaves, N, check_ops=check_ops)
    assert rel(ug, uo) < TOL_U


@pytest.mark.parametrize("name", ["global_5pt_varbeta_L2N8", "global_5pt_varbeta_L3N16"])
def test_variable_equals_global_solve(ctx, name):
    """The variable-beta operator on a uniform mesh against the independent global 5-point
    solve with face-sampled beta (tests/golden/make_golden.py)."""
    import ellipticforest_amd as ef
    say(f"golden {name}")
    d = np.load(os.path.join(GOLD, name + ".npz"))
    level, N, bounds = int(d["level"]), int(d["N"]), tuple(d["bounds"])
    leaves = omesh.uniform(level)
    D = ef.DeviceMesh(ctx, leaves, N, bounds)
    cell = np.concatenate([oracle.sample_coeffs(N, tuple(D.node_bounds[n]), ALPHA, BETA, LAMBDA).ravel()
                           for n in D.leaf_nodes])
    f = np.stack([F3(*oracle.cell_centres(N, tuple(D.node_bounds[n]))[:2]) for n in D.leaf_nodes])
    side, px, py = oracle.root_face_points(D.M, bounds)
    D.build(solver_type=ef._lib.EF_SOLVER_FIVE_POINT, cell=cell)
    D.upwards(f)
    u = D.solve(U3(px, py), np.ones_like(px))
    assert rel(u, d["u"]) < 1e-11


def test_variable_coefficients_N32_shape(ctx):
    """The variable-beta operator on N = 32 patches at level 3: parity and second-order error."""
    say("variable beta: uniform level 3, N=32")
    uo, ug, err = variable_both(ctx, omesh.uniform(3), 32)
    assert rel(ug, uo) < TOL_U
    assert err < 2e-4


def graph_replay_case(ctx, leaves, N, bounds, k=3):
    """Upwards + Dirichlet solve through device pointers with the captured sweep against the
    same calls launched kernel by kernel (EF_GRAPH=0): bitwise equal on every replay, with
    other work on the mesh between replays, when the inputs change under the same pointers,
    and after the vector arena has been reallocated for a larger k."""
    from ellipticforest_amd import _lib
    D = _lib.DeviceMesh(ctx, leaves, N, bounds)
    D.build()
    nl, m = len(leaves), 4 * D.M
    rng = np.random.default_rng(11)
    f = rng.standard_normal((2, k, nl, N * N))
    r = rng.standard_normal((2, k, m))
    a = np.ones(m)
    fd, rd, ad, ud = (_lib.DeviceArray(ctx, s) for s in (f[0].shape, r[0].shape, a.shape, f[0].shape))
    ad.upload(a)

    def sweep(i):
        fd.upload(f[i])
        rd.upload(r[i])
        D.upwards_device(fd.ptr, k)
        D.solve_device(fd.ptr, rd.ptr, ad.ptr, None, ud.ptr, k)
        D.sync()
        return ud.download().reshape(f[i].shape)

    os.environ["EF_GRAPH"] = "0"
    try:
        eager = [sweep(0), sweep(1)]
    finally:
        del os.environ["EF_GRAPH"]
    for rep in range(3):
        for i in (0, 1):
            say(f"graph replay {rep} input {i}")
            assert np.array_equal(sweep(i), eager[i]), (rep, i)
        D.upwards(f[1][:1])  # other work on the mesh between replays (k = 1, host buffers)
        D.solve(r[1][:1], a)
    big, ub = _lib.DeviceArray(ctx, (k + 2, nl, N * N)), _lib.DeviceArray(ctx, (k + 2, nl, N * N))
    rb = _lib.DeviceArray(ctx, (k + 2, m))
    big.upload(np.zeros((k + 2, nl, N * N)))
    rb.upload(np.zeros((k + 2, m)))
    D.upwards_device(big.ptr, k + 2)  # grows the arena: every captured sweep is dropped
    D.solve_device(big.ptr, rb.ptr, ad.ptr, None, ub.ptr, k + 2)
    D.sync()
    assert np.array_equal(sweep(0), eager[0])
    return eager


def variable_graph_widths_case(ctx, level=1, N=8):
    """Captured sweeps of a variable-coefficient mesh whose vector arena was sized up front
    (as the serving loop plans every bucket width): widths 2, 4, then 2 again, each against the
    same calls launched kernel by kernel.  The solve work buffer of the packed five-point leaves
    grows with the width; a sweep captured before it grew must not be replayed."""
    from ellipticforest_amd import _lib
    say(f"variable-coefficient graph replay across widths, uniform level {level} N={N}")
    leaves = omesh.uniform(level)
    bounds = (-1.0, 1.0, -1.0, 1.0)
    D = _lib.DeviceMesh(ctx, leaves, N, bounds)
    cell = np.concatenate([oracle.sample_coeffs(N, tuple(D.node_bounds[n]), ALPHA, BETA, LAMBDA).ravel()
                           for n in D.leaf_nodes])
    D.build(solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
    D.sweep_tiles(8)  # sizes the vector arena for 8 right-hand sides
    nl, m = len(leaves), 4 * D.M
    rng = np.random.default_rng(5)
    a = np.ones(m)
    ad = _lib.DeviceArray(ctx, a.shape)
    ad.upload(a)

    bufs = {}  # one set of device buffers per width, reused (as a serving slot reuses its own)

    def sweep(k, seed):
        r2 = np.random.default_rng(seed)
        f, r = r2.standard_normal((k, nl, N * N)), r2.standard_normal((k, m))
        if k not in bufs:
            bufs[k] = (_lib.DeviceArray(ctx, f.shape), _lib.DeviceArray(ctx, r.shape), _lib.DeviceArray(ctx, f.shape))
        fd, rd, ud = bufs[k]
        fd.upload(f)
        rd.upload(r)
        D.upwards_device(fd.ptr, k)
        D.solve_device(fd.ptr, rd.ptr, ad.ptr, None, ud.ptr, k)
        D.sync()
        return ud.download().reshape(f.shape)

    os.environ["EF_GRAPH"] = "0"
    try:
        e2, e4 = sweep(2, 1), sweep(4, 2)
    finally:
        del os.environ["EF_GRAPH"]
    del D  # a fresh mesh: its first width-2 sweep is captured while the work buffer fits 2
    D = _lib.DeviceMesh(ctx, leaves, N, bounds)
    D.build(solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
    D.sweep_tiles(8)
    for k, seed, ref in ((2, 1, e2), (4, 2, e4), (2, 1, e2), (4, 2, e4)):
        assert np.array_equal(sweep(k, seed), ref), k


def test_variable_graph_widths(ctx):
    variable_graph_widths_case(ctx)


def test_graph_replay_matches_eager(ctx):
    say("graph replay, adaptive N=8 mesh with coarsening")
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, 5)
    graph_replay_case(ctx, leaves, 8, (-10, 10, -10, 10))


def inverse_case(ctx, m, B, seed, integer=False, check_pivots=True, sparse=False):
    """ef_debug_invert (the merges' batched Gauss-Jordan inversion) on random matrices: the
    inverse against numpy's, and the row interchanges against LAPACK dgetrf's (scipy
    lu_factor), index for index, when no step spans more than one 512-row chunk.  Integer
    matrices make ties in |pivot| common: the first maximum of the interchanged column wins,
    as in idamax.  Sparse integer matrices keep ties alive after elimination, including ties
    with rows that an earlier interchange of the same step moved."""
    import scipy.linalg as sla
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(seed)
    if sparse:
        A = np.empty((B, m, m))
        for b in range(B):
            while True:
                A[b] = rng.integers(-3, 4, (m, m)) * (rng.random((m, m)) < 0.15)
                if np.linalg.matrix_rank(A[b]) == m:
                    break
    else:
        A = (rng.integers(-3, 4, (B, m, m)).astype(np.float64) if integer else rng.standard_normal((B, m, m)))
    Ainv = np.empty_like(A)
    ipiv = np.empty((B, m), np.int32)
    _lib.check(ctx.L.ef_debug_invert(ctx.h, m, B, A.ctypes.data_as(_lib._dp), Ainv.ctypes.data_as(_lib._dp),
                                     ipiv.ctypes.data_as(_lib._ip)), ctx.L)
    for b in range(B):
        ref = np.linalg.inv(A[b])
        cond = np.linalg.cond(A[b])
        assert rel(Ainv[b], ref) < 1e-15 * cond * m, (m, b, rel(Ainv[b], ref), cond)
        if check_pivots:
            assert np.array_equal(ipiv[b], sla.lu_factor(A[b])[1]), (m, b)


@pytest.mark.parametrize("m,B,integer", [(16, 5, False), (64, 40, False), (64, 7, True), (96, 3, True),
                                         (200, 2, False), (512, 1, False), (512, 2, True)])
def test_batched_inverse_pivots(ctx, m, B, integer):
    say(f"batched inverse m={m} B={B} integer={integer}")
    inverse_case(ctx, m, B, seed=m + B, integer=integer)


@pytest.mark.parametrize("m,B", [(64, 12), (128, 4), (256, 2), (480, 1)])
def test_batched_inverse_pivot_ties(ctx, m, B):
    say(f"batched inverse with pivot ties m={m} B={B}")
    inverse_case(ctx, m, B, seed=7 * m + B, sparse=True)


@pytest.mark.parametrize("m,B", [(1024, 1), (1056, 2), (2048, 1)])
def test_batched_inverse_tournament(ctx, m, B):
    """Steps over more than one chunk choose pivots by tournament: the inverse is checked, the
    interchanges are not LAPACK's."""
    say(f"batched inverse (tournament) m={m} B={B}")
    inverse_case(ctx, m, B, seed=m, check_pivots=False)


def mixed_case(ctx, level, N, variable):
    """Mixed precision (FP32 S and T products, FP64 X^-1, FP64 residual + iterative refinement)
    against the FP64 build of the same mesh: the operators differ at FP32 rounding (neither
    equal nor wrong), the unrefined solve is off by FP32-level error, and refinement brings the
    solution to the FP64 one (<= 1e-8 is the target; 1e-10 is asserted after two rounds)."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    leaves, bounds = omesh.uniform(level), (-1.0, 1.0, -1.0, 1.0)
    uf, ff = (U3, F3) if variable else (U, F)

    def make(prec):
        D = ef.DeviceMesh(ctx, leaves, N, bounds)
        if variable:
            cell = np.concatenate([oracle.sample_coeffs(N, tuple(D.node_bounds[n]), ALPHA, BETA, LAMBDA).ravel()
                                   for n in D.leaf_nodes])
            D.build(solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell, precision=prec)
        else:
            D.build(precision=prec)
        return D

    D64, Dm = make(_lib.EF_PRECISION_FP64), make(_lib.EF_PRECISION_MIXED)
    f = np.stack([ff(*oracle.cell_centres(N, tuple(D64.node_bounds[n]))[:2]) for n in D64.leaf_nodes])
    side, px, py = oracle.root_face_points(D64.M, bounds)
    r, a = uf(px, py), np.ones_like(px)
    D64.upwards(f)
    u64 = D64.solve(r, a)
    for w in ("S", "T"):
        d = rel(Dm.get(0, w), D64.get(0, w))
        assert 1e-13 < d < 1e-4, (w, d)
    errs = {}
    for steps in (0, 1, 2):
        Dm.set_refinement(steps)
        Dm.upwards(f)
        errs[steps] = rel(Dm.solve(r, a), u64)
    say(f"mixed precision vs FP64: {errs}")
    assert 1e-12 < errs[0] < 1e-3
    assert errs[1] < 1e-8 and errs[2] < 1e-10
    with pytest.raises(Exception):
        Dm.solve(r, a)  # the refinement consumed the upwards stage
    return errs


def test_mixed_precision_poisson(ctx):
    say("mixed precision: Poisson, uniform level 3, N=16")
    mixed_case(ctx, 3, 16, variable=False)


def test_mixed_precision_variable_beta(ctx):
    say("mixed precision: variable beta, uniform level 3, N=16")
    mixed_case(ctx, 3, 16, variable=True)


def test_mixed_precision_refuses_adaptive(ctx):
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    leaves = omesh.refine_by_function(lambda x, y: (x < 0.2) & (y < 0.2), 8, 8, 0, 1, 0, 1, 0, 3)
    assert len({l[0] for l in leaves}) > 1
    D = ef.DeviceMesh(ctx, leaves, 8, (0, 1, 0, 1))
    with pytest.raises(Exception):
        D.build(precision=_lib.EF_PRECISION_MIXED)
