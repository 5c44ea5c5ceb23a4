"""Batched leaf-patch operations (ef_patches_*): FiniteVolumeSolver::solve, mapD2N,
particularNeumannData and buildD2N (FiniteVolumeSolver.cpp:25-462) on many patches.

The oracle's leaf functions are the reference's algorithm restated (pinned in
test_oracle.py by the second-order convergence the reference asserts, the DtN
invariants and the global 5-point solves).  Tolerances (FP64): relative max-norm 1e-12
on u, h and T against the oracle for every patch and right-hand side.

The emulator cases run on the CPU (the HIP sources compiled against tests/cpu_emu);
the gpu cases call the product library through the C ABI on an MI355X, at the oracle's
sizes and at the full leaf set of the adaptive Poisson mesh (|f| > 1.2, levels 0..7: 7,750 patches, N = 16), where
size-independent properties are checked: linearity in (g, f), agreement of mapD2N with
T g + particular data, and the exact Neumann data of a linear solution.
"""
import os

import numpy as np
import pytest

import oracle
from oracle import mesh as omesh

TOL = 1e-12
F = lambda x, y: -(np.sin(x) + np.sin(y))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def random_bounds(rng, P):
    lo = rng.uniform(-2, 2, (P, 2))
    ext = rng.uniform(0.05, 1.5, (P, 2))
    return np.stack([lo[:, 0], lo[:, 0] + ext[:, 0], lo[:, 1], lo[:, 1] + ext[:, 1]], axis=1)


def check_against_oracle(pt, bounds, N, rng, nrhs=2, coeffs=None, lam=0.0, dtn=True):
    P = len(bounds)
    g = rng.standard_normal((nrhs, P, 4 * N))
    f = rng.standard_normal((nrhs, P, N * N))
    u = pt.solve(g, f, nrhs)
    h = pt.map_d2n(g, f, nrhs)
    hp = pt.map_d2n(None, f, nrhs)
    u0 = pt.solve(g, None, nrhs)
    for c in range(nrhs):
        for p in range(P):
            kw = dict(coeffs=None if coeffs is None else coeffs[p], lambda0=lam)
            assert rel(u[c, p], oracle.leaf_solve(N, bounds[p], g[c, p], f[c, p], **kw)) < TOL, ("u", c, p)
            assert rel(h[c, p], oracle.leaf_map_d2n(N, bounds[p], g[c, p], f[c, p], **kw)) < TOL, ("h", c, p)
            assert rel(hp[c, p], oracle.leaf_map_d2n(N, bounds[p], np.zeros(4 * N), f[c, p], **kw)) < TOL, ("hp", c, p)
            assert rel(u0[c, p], oracle.leaf_solve(N, bounds[p], g[c, p], np.zeros(N * N), **kw)) < TOL, ("u0", c, p)
    if dtn:
        T = pt.dtn()
        for p in range(P):
            kw = dict(coeffs=None if coeffs is None else coeffs[p], lambda0=lam)
            assert rel(T[p], oracle.leaf_dtn(N, bounds[p], **kw)) < TOL, ("T", p)


# ----- CPU: the kernels through the emulation build ----------------------------------
@pytest.mark.parametrize("N", [4, 8, 12, 16])
@pytest.mark.parametrize("lam", [0.0, -1.0])
def test_emu_fishpack_patches(emu_ctx, N, lam):
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(N)
    bounds = random_bounds(rng, 3)
    pt = _lib.Patches(emu_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FISHPACK, lambda0=lam)
    assert not pt.variable
    check_against_oracle(pt, bounds, N, rng, lam=lam)


def test_emu_fishpack_patches_N32(emu_ctx):
    """Two-by-two MFMA block grid (N > 16)."""
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(32)
    bounds = random_bounds(rng, 2)
    pt = _lib.Patches(emu_ctx, bounds, 32, lambda0=-0.5)
    check_against_oracle(pt, bounds, 32, rng, nrhs=1, lam=-0.5, dtn=False)


@pytest.mark.parametrize("N", [8, pytest.param(16, marks=pytest.mark.skipif(
    not __import__("os").environ.get("EF_SLOW_TESTS"), reason="a minute on the emulator; the GPU runs N = 16"))])
def test_emu_five_point_constant_and_variable(emu_ctx, N):
    """Constant coefficients (fast diagonalisation), variable beta (a symmetric matrix: packed
    inverse pivot blocks) and variable alpha (rows scaled by alpha, not symmetric: full blocks)."""
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(5 + N)
    bounds = random_bounds(rng, 3)
    const = np.stack([oracle.sample_coeffs(N, b, lambda x, y: 2.0 + 0 * x, lambda x, y: 0.5 + 0 * x,
                                           lambda x, y: -0.7 + 0 * x) for b in bounds])
    pt = _lib.Patches(emu_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=const)
    assert not pt.variable
    check_against_oracle(pt, bounds, N, rng, coeffs=const)
    beta = lambda x, y: 1 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    var = np.stack([oracle.sample_coeffs(N, b, lambda x, y: 1.0 + 0 * x, beta, lambda x, y: -1.0 + 0 * x)
                    for b in bounds])
    pv = _lib.Patches(emu_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=var)
    assert pv.variable and pv.packed
    check_against_oracle(pv, bounds, N, rng, coeffs=var)
    nonsym = alpha_cells(bounds, N)
    pn = _lib.Patches(emu_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=nonsym)
    assert pn.variable and not pn.packed
    check_against_oracle(pn, bounds, N, rng, coeffs=nonsym)


def test_emu_patch_argument_errors(emu_ctx):
    from ellipticforest_amd import _lib
    with pytest.raises(_lib.EfError) as e:
        _lib.Patches(emu_ctx, [[0, 1, 0, 1]], 2048)
    assert e.value.code == _lib.EF_ERR_ARG
    with pytest.raises(_lib.EfError) as e:
        _lib.Patches(emu_ctx, [[0, 1, 0, 1]], 2)
    assert e.value.code == _lib.EF_ERR_ARG
    with pytest.raises(_lib.EfError) as e:
        _lib.Patches(emu_ctx, [[0, 0, 0, 1]], 8)
    assert e.value.code == _lib.EF_ERR_ARG
    pt = _lib.Patches(emu_ctx, [[0, 1, 0, 1]], 8)
    with pytest.raises(_lib.EfError):
        pt.apply_device(7, None, None, None, 1)
    # empty batch and no right-hand sides: argument errors, not empty results
    with pytest.raises(_lib.EfError) as e:
        _lib.Patches(emu_ctx, np.zeros((0, 4)), 16)
    assert e.value.code == _lib.EF_ERR_ARG
    with pytest.raises(_lib.EfError) as e:
        pt.solve(np.zeros((0, 1, 32)), np.zeros((0, 1, 64)), 0)
    assert e.value.code == _lib.EF_ERR_ARG
    # no load and no boundary data (both NULL): the zero solution and zero Neumann data
    p16 = _lib.Patches(emu_ctx, [[0, 1, 0, 1], [1, 2, 0, 1]], 16)
    assert np.abs(p16.solve(None, None, 1)).max() == 0.0
    assert np.abs(p16.map_d2n(None, None, 1)).max() == 0.0


# ----- GPU: the product library through the C ABI -----------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("N", [4, 8, 12, 16, 24, 32])
def test_gpu_fishpack_patches_vs_oracle(patch_ctx, N):
    """Every register-resident variant: one MFMA block (N <= 16) or two by two (N <= 32), each
    with N filling the blocks (16, 32) or padded (4, 8, 12, 24)."""
    from ellipticforest_amd import _lib
    say(f"patches FISHPACK N={N}, 5 patches, 3 rhs, lambda=-1")
    rng = np.random.default_rng(100 + N)
    bounds = random_bounds(rng, 5)
    pt = _lib.Patches(patch_ctx, bounds, N, lambda0=-1.0)
    check_against_oracle(pt, bounds, N, rng, nrhs=3, lam=-1.0)


@pytest.mark.gpu
def test_gpu_five_point_constant_vs_oracle(patch_ctx):
    from ellipticforest_amd import _lib
    say("patches five-point constant coefficients N=16")
    rng = np.random.default_rng(7)
    N = 16
    bounds = random_bounds(rng, 4)
    const = np.stack([oracle.sample_coeffs(N, b, lambda x, y: 1.5 + 0 * x, lambda x, y: 0.8 + 0 * x,
                                           lambda x, y: -2.0 + 0 * x) for b in bounds])
    pt = _lib.Patches(patch_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=const)
    assert not pt.variable
    check_against_oracle(pt, bounds, N, rng, coeffs=const)


def adaptive_poisson_leaf_bounds():
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 16, 16, -10, 10, -10, 10, 0, 7)
    out = []
    for lv, ix, iy in leaves:
        w = 20.0 / (1 << lv)
        out.append((-10 + ix * w, -10 + (ix + 1) * w, -10 + iy * w, -10 + (iy + 1) * w))
    return np.array(out)


@pytest.mark.gpu
def test_gpu_adaptive_poisson_leaf_set_properties(patch_ctx):
    """All 7,750 leaf patches of the adaptive Poisson mesh, 4 right-hand sides: a sample of patches
    against the oracle, linearity, mapD2N = T g + particular data, and the exact Neumann
    data of u = 3x - 2y + 1 (the 5-point scheme is exact on linear functions)."""
    from ellipticforest_amd import _lib
    bounds = adaptive_poisson_leaf_bounds()
    P, N, k = len(bounds), 16, 4
    say(f"adaptive Poisson leaf set: {P} patches, N={N}, {k} rhs")
    assert P == 7750
    pt = _lib.Patches(patch_ctx, bounds, N)
    rng = np.random.default_rng(11)
    g1, g2 = rng.standard_normal((2, k, P, 4 * N))
    f1, f2 = rng.standard_normal((2, k, P, N * N))
    u1, u2 = pt.solve(g1, f1, k), pt.solve(g2, f2, k)
    u12 = pt.solve(2.0 * g1 - g2, 2.0 * f1 - f2, k)
    assert rel(u12, 2.0 * u1 - u2) < 1e-12
    for p in rng.choice(P, 12, replace=False):
        c = int(rng.integers(k))
        assert rel(u1[c, p], oracle.leaf_solve(N, bounds[p], g1[c, p], f1[c, p])) < TOL, p
    h = pt.map_d2n(g1, f1, k)
    hp = pt.map_d2n(None, f1, k)
    sample = rng.choice(P, 16, replace=False)
    T = _lib.Patches(patch_ctx, bounds[sample], N).dtn()
    for i, p in enumerate(sample):
        assert rel(h[0, p], T[i] @ g1[0, p] + hp[0, p]) < 1e-11, p
    # linear solution: g at boundary faces, h = the exact derivative (d/dx, d/dy)
    lin = lambda x, y: 3.0 * x - 2.0 * y + 1.0
    g = np.empty((1, P, 4 * N))
    for p, (xl, xu, yl, yu) in enumerate(bounds):
        xs = xl + (np.arange(N) + 0.5) * (xu - xl) / N
        ys = yl + (np.arange(N) + 0.5) * (yu - yl) / N
        g[0, p] = np.concatenate([lin(xl, ys), lin(xu, ys), lin(xs, yl), lin(xs, yu)])
    hl = pt.map_d2n(g, None, 1)[0]
    expect = np.concatenate([np.full(N, 3.0), np.full(N, 3.0), np.full(N, -2.0), np.full(N, -2.0)])
    assert np.abs(hl - expect).max() < 1e-8


@pytest.mark.parametrize("N", [24, 32])
def test_emu_grid_stride_loop_two_by_two_blocks(emu_ctx, monkeypatch, N):
    """One workgroup (four waves) for 2 patches x 3 right-hand sides (and, N = 24, 2 x 4N DtN
    columns) with the two-by-two MFMA block grid: each wave fetches and solves several items in
    turn."""
    from ellipticforest_amd import _lib
    monkeypatch.setenv("EF_PATCH_MAX_BLOCKS", "1")
    rng = np.random.default_rng(N + 5)
    bounds = random_bounds(rng, 2)
    pt = _lib.Patches(emu_ctx, bounds, N, lambda0=-0.75)
    check_against_oracle(pt, bounds, N, rng, nrhs=3, lam=-0.75, dtn=(N == 24))


@pytest.mark.parametrize("N", [8, 16])
def test_emu_grid_stride_loop(emu_ctx, monkeypatch, N):
    """Two workgroups for 3 patches x 5 right-hand sides and a DtN of 3 x 4N columns: every
    wave loops over several items (N = 16: the unpadded single-block variant)."""
    from ellipticforest_amd import _lib
    monkeypatch.setenv("EF_PATCH_MAX_BLOCKS", "2")
    rng = np.random.default_rng(21)
    bounds = random_bounds(rng, 3)
    pt = _lib.Patches(emu_ctx, bounds, N, lambda0=-0.25)
    check_against_oracle(pt, bounds, N, rng, nrhs=5, lam=-0.25)


def test_emu_timer_and_stream(emu_ctx):
    from ellipticforest_amd import _lib
    pt = _lib.Patches(emu_ctx, [[0, 1, 0, 1]], 8)
    t = _lib.Timer(emu_ctx)
    with pytest.raises(_lib.EfError):
        t.elapsed_ms()  # nothing recorded yet
    t.start(pt.stream)
    pt.solve(np.zeros((1, 1, 32)), np.ones((1, 1, 64)))
    t.stop(pt.stream)
    assert t.elapsed_ms() >= 0.0


BETA3 = lambda x, y: 1 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)


def variable_cells(bounds, N, beta=BETA3, lam=-1.0):
    return np.stack([oracle.sample_coeffs(N, b, lambda x, y: 1.0 + 0 * x, beta, lambda x, y: lam + 0 * x)
                     for b in bounds])


def alpha_cells(bounds, N, lam=1.0):
    """alpha = 1 + 0.3 sin(x + 2y) scales the rows of the five-point matrix: not symmetric.  With
    lambda = +1 the rows are not diagonally dominant either, so the block LU pivots (lambda <= 0
    takes the factorisation without pivoting)."""
    return np.stack([oracle.sample_coeffs(N, b, lambda x, y: 1.0 + 0.3 * np.sin(x + 2 * y), BETA3,
                                          lambda x, y: lam + 0 * x) for b in bounds])


@pytest.mark.skipif(not os.environ.get("EF_EMU_SLOW"),
                    reason="about 2.5 minutes on 8 cores (the N = 32 factorisation's 1,024 cross-lane "
                           "reads per block, each a barrier of 64 emulated lanes); set EF_EMU_SLOW=1. "
                           "The same kernels run at N = 32 in the GPU tests and at N = 8, 16 here.")
def test_emu_variable_patches_N32(emu_ctx):
    """Largest block-tridiagonal size (N = 32) on the emulation build: packed symmetric blocks,
    lambda = +1 so the factorisation pivots (the pivot-free variant is emulated at N = 8 and 16
    and runs at N = 32 in the GPU tests)."""
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(33)
    bounds = random_bounds(rng, 1)
    cell = variable_cells(bounds, 32, lam=1.0)
    pv = _lib.Patches(emu_ctx, bounds, 32, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
    assert pv.variable and pv.packed
    check_against_oracle(pv, bounds, 32, rng, nrhs=1, coeffs=cell, dtn=False)  # N = 32 DtN: GPU test


@pytest.mark.gpu
@pytest.mark.parametrize("N", [4, 8, 12, 16, 24, 32])
@pytest.mark.parametrize("symmetric", [True, False])
def test_gpu_variable_patches_vs_oracle(patch_ctx, N, symmetric):
    """Five-point stencil (FiniteVolumeSolver.cpp:64-184), block-tridiagonal path: the variable beta
    (symmetric matrix, packed inverse pivot blocks) and a variable alpha (full blocks); N in
    {8, 16, 32} on the wave kernels, other sides on the workgroup-per-patch kernel."""
    from ellipticforest_amd import _lib
    say(f"patches five-point variable {'beta' if symmetric else 'alpha'} N={N}, 4 patches, 2 rhs")
    rng = np.random.default_rng(300 + N + (0 if symmetric else 1000))
    bounds = random_bounds(rng, 4)
    cell = variable_cells(bounds, N) if symmetric else alpha_cells(bounds, N)
    pv = _lib.Patches(patch_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
    assert pv.variable and pv.packed == (symmetric and N in (8, 16, 32))
    check_against_oracle(pv, bounds, N, rng, coeffs=cell)


@pytest.mark.gpu
def test_gpu_variable_beta_leaf_set_properties(patch_ctx):
    """The variable-beta leaf set ([-1,1]^2, uniform level 6, N = 32: 4,096 patches) with its
    coefficients: oracle on a sample, linearity, and mapD2N = T g + particular data."""
    from ellipticforest_amd import _lib
    N, lv = 32, 6
    w = 2.0 / (1 << lv)
    bounds = np.array([(-1 + i * w, -1 + (i + 1) * w, -1 + j * w, -1 + (j + 1) * w)
                       for i in range(1 << lv) for j in range(1 << lv)])
    say(f"variable-beta leaf set: {len(bounds)} patches, N={N}, variable beta, lambda=-1")
    cell = variable_cells(bounds, N)
    pv = _lib.Patches(patch_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)
    assert pv.variable and pv.packed
    rng = np.random.default_rng(12)
    P = len(bounds)
    g1, g2 = rng.standard_normal((2, 1, P, 4 * N))
    f1, f2 = rng.standard_normal((2, 1, P, N * N))
    u1, u2 = pv.solve(g1, f1), pv.solve(g2, f2)
    assert rel(pv.solve(g1 - 3.0 * g2, f1 - 3.0 * f2), u1 - 3.0 * u2) < 1e-11
    for p in rng.choice(P, 6, replace=False):
        assert rel(u1[0, p], oracle.leaf_solve(N, bounds[p], g1[0, p], f1[0, p], coeffs=cell[p])) < TOL, p
    sample = rng.choice(P, 4, replace=False)
    T = _lib.Patches(patch_ctx, bounds[sample], N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell[sample]).dtn()
    h, hp = pv.map_d2n(g1, f1), pv.map_d2n(None, f1)
    for i, p in enumerate(sample):
        assert rel(h[0, p], T[i] @ g1[0, p] + hp[0, p]) < 1e-11, p


@pytest.mark.parametrize("N", [8, 16])
def test_emu_shared_coefficient_tables(emu_ctx, N):
    """Patches of equal size share one coefficient entry and its 1/D table (the leaf-set case)."""
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(40 + N)
    bounds = np.array([(-1 + 0.5 * i, -0.5 + 0.5 * i, -1 + 0.5 * j, -0.5 + 0.5 * j) for i in range(4) for j in range(4)])
    pt = _lib.Patches(emu_ctx, bounds, N, lambda0=-0.5)
    check_against_oracle(pt, bounds, N, rng, nrhs=2, lam=-0.5)


@pytest.mark.parametrize("N", [40, 64])
def test_emu_large_patches(emu_ctx, N):
    """N > 32: fast diagonalisation as batched GEMMs (padding to a multiple of 16 for N = 40)."""
    from ellipticforest_amd import _lib
    rng = np.random.default_rng(50 + N)
    bounds = random_bounds(rng, 2)
    pt = _lib.Patches(emu_ctx, bounds, N, lambda0=-1.0)
    check_against_oracle(pt, bounds, N, rng, nrhs=1, lam=-1.0, dtn=False)  # DtN of large patches: GPU test
    with pytest.raises(_lib.EfError):  # coefficients varying within a patch stay at N <= 32
        cell = np.stack([oracle.sample_coeffs(N, b, lambda x, y: 1.0 + 0 * x, lambda x, y: 1.0 + x * y,
                                              lambda x, y: 0 * x) for b in bounds])
        _lib.Patches(emu_ctx, bounds, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=cell)


@pytest.mark.gpu
@pytest.mark.parametrize("N", [48, 64, 100, 128])
def test_gpu_large_patches_vs_oracle(patch_ctx, N):
    """Single grids beyond one wave's registers (the reference's solver test runs n up to 512):
    solve, mapD2N, particular data and (N <= 64) the DtN against the oracle."""
    from ellipticforest_amd import _lib
    say(f"large patches N={N}, 3 patches, 2 rhs, lambda=-1")
    rng = np.random.default_rng(500 + N)
    bounds = random_bounds(rng, 3)
    pt = _lib.Patches(patch_ctx, bounds, N, lambda0=-1.0)
    check_against_oracle(pt, bounds, N, rng, nrhs=2, lam=-1.0, dtn=(N <= 64))
