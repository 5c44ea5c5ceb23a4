"""The oracle (CPU restatement of the reference path) pinned against the reference's
own known answers, against an independent global 5-point solve, and against closed
forms that a dropped term, wrong sign, wrong index or transposed operand would break."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U = lambda x, y: np.sin(x) + np.sin(y)
F = lambda x, y: -(np.sin(x) + np.sin(y))


def known():
    with open(os.path.join(GOLD, "reference_known_answers.json")) as f:
        return json.load(f)


def test_dgesv_known_answer():
    k = known()["test_matrix.cpp:155-192 linear_solve"]
    A = np.array(k["A"], float)
    x = oracle.dgesv(A, np.array(k["b"], float))
    assert np.abs(x - k["x"]).max() <= k["tol"]
    X = oracle.dgesv(A, np.array(k["B"], float))
    assert np.abs(X - np.array(k["X"])).max() <= k["tol_matrix"]


def test_dgemm_known_answer():
    k = known()["test_matrix.cpp:194-268 math"]
    assert np.array_equal(oracle.dgemm(np.array(k["m6"], float), np.array(k["m7"], float)), np.array(k["m6m7"]))


def test_block_permute_known_answer():
    k = known()["test_matrix.cpp:270-300 block_permute"]
    out = oracle.block_permute(np.array(k["m1"], float), k["pi_rows"], k["pi_cols"], k["R"], k["C"])
    assert np.array_equal(out, np.array(k["expected"], float))


def test_grid_points_known_answer():
    k = known()["test_finite_volume_grid.cpp:30-49 points"]
    X, Y, dx, dy = oracle.cell_centres(4, (k["x"][0], k["x"][1], k["y"][0], k["y"][1]))
    xs = X.reshape(4, 4)[:, 0]
    ys = Y.reshape(4, 4)[0, :]
    assert np.abs(xs - k["xpoints"]).max() <= k["tol"]
    assert np.abs(ys - k["ypoints"]).max() <= k["tol"]


def test_leaf_solver_second_order():
    """test_finite_volume_solver.cpp:21-95: order > 1.95 between the last two grids."""
    k = known()["test_finite_volume_solver.cpp:21-95 solver_convergence_FISHPACK"]
    errs = []
    ns = [64, 128, 256]
    for n in ns:
        b = tuple(k["domain"])
        X, Y, dx, dy = oracle.cell_centres(n, b)
        xs = b[0] + dx / 2 + np.arange(n) * dx
        ys = b[2] + dy / 2 + np.arange(n) * dy
        g = np.concatenate([U(b[0], ys), U(b[1], ys), U(xs, b[2]), U(xs, b[3])])
        u = oracle.leaf_solve(n, b, g, F(X, Y))
        errs.append(np.abs(u - U(X, Y)).max())
    order = np.log(errs[-1] / errs[-2]) / np.log(ns[-2] / ns[-1])
    assert order > k["min_order_last_pair"]


def test_interpolation_operators():
    """L21 averages (rows sum to 1); L12 reproduces linear data exactly (SpecialMatrices.hpp:93-173)."""
    for nc in (3, 4, 8, 16):
        L21 = oracle.interp_fine2coarse(nc)
        assert np.allclose(L21.sum(1), 1.0) and L21.shape == (nc, 2 * nc)
        nf = 2 * nc
        L12 = oracle.interp_coarse2fine(nf)
        xc = (np.arange(nc) + 0.5) / nc
        xf = (np.arange(nf) + 0.5) / nf
        assert np.abs(L12 @ (3 * xc - 1) - (3 * xf - 1)).max() < 1e-14
        assert np.abs(L12 @ np.ones(nc) - 1).max() < 1e-14
        # quadratic end rows are exact on quadratics
        q = lambda x: x * x
        assert abs((L12 @ q(xc))[0] - q(xf[0])) < 1e-14 and abs((L12 @ q(xc))[-1] - q(xf[-1])) < 1e-14


def test_leaf_dtn_invariants():
    """T 1 = 0 for lambda = 0 (constant Dirichlet data gives a constant solution); T scales as 1/h;
    column j equals mapD2N(e_j)."""
    T = oracle.leaf_dtn(8, (0.0, 1.0, 0.0, 1.0))
    assert np.abs(T @ np.ones(32)).max() < 1e-10
    T2 = oracle.leaf_dtn(8, (0.0, 2.0, 0.0, 2.0))
    assert np.allclose(T, 2 * T2, rtol=1e-12, atol=1e-12)
    e = np.zeros(32)
    e[5] = 1
    assert np.allclose(T[:, 5], oracle.leaf_map_d2n(8, (0.0, 1.0, 0.0, 1.0), e, np.zeros(64)), atol=1e-13)
    # linear Dirichlet data: u = 2x + 3y exactly, so h = (du/dx, du/dy) = (2, 3) on every side
    X, Y, dx, dy = oracle.cell_centres(8, (0.0, 1.0, 0.0, 1.0))
    xs = dx / 2 + np.arange(8) * dx
    lin = lambda x, y: 2 * x + 3 * y
    g = np.concatenate([lin(0, xs), lin(1, xs), lin(xs, 0), lin(xs, 1)])
    h = T @ g
    assert np.allclose(h, np.concatenate([np.full(16, 2.0), np.full(16, 3.0)]), atol=1e-11)


@pytest.mark.parametrize("name", ["global_5pt_L2N8_pm1", "global_5pt_L2N8_pm10", "global_5pt_L3N16_pm10",
                                  "global_5pt_L2N8_helmholtz"])
def test_uniform_hps_equals_global_solve(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    level, N, bounds, lam = int(d["level"]), int(d["N"]), tuple(d["bounds"]), float(d["lam"])
    Fo = oracle.Forest(mesh.uniform(level), N, bounds, solver=oracle.FISHPACK, lambda0=lam)
    Fo.build()
    Fo.upwards(Fo.sample_rhs(F))
    r, a, b = Fo.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    u = Fo.solve(r, a, b)
    assert np.abs(u - d["u"]).max() / np.abs(d["u"]).max() < 1e-12


def test_adaptive_linear_exactness():
    """u = a x + b y + c with f = 0 is reproduced exactly on an adaptive mesh (5-point, L21, L12 exact on linears)."""
    leaves = mesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, 4)
    Fo = oracle.Forest(leaves, 8, (-10, 10, -10, 10), solver=oracle.FISHPACK)
    Fo.build()
    Fo.upwards(np.zeros((len(leaves), 64)))
    lin = lambda x, y: 0.3 * x - 0.7 * y + 2.0
    r, a, b = Fo.boundary_data(lambda s, x, y: (lin(x, y), 1.0, 0.0))
    u = Fo.solve(r, a, b)
    for k, bd in enumerate(Fo.leaf_bounds):
        X, Y, _, _ = oracle.cell_centres(8, bd)
        assert np.abs(u[k] - lin(X, Y)).max() < 1e-11


def test_five_point_constant_matches_fishpack():
    """FivePointStencil with alpha = beta = 1 and constant lambda is the FISHPACK system."""
    leaves = mesh.uniform(1)
    one = lambda x, y: np.ones_like(x)
    A = oracle.Forest(leaves, 8, (0, 1, 0, 1), solver=oracle.FIVE_POINT, coeff_fns=(one, one, lambda x, y: -np.ones_like(x)))
    B = oracle.Forest(leaves, 8, (0, 1, 0, 1), solver=oracle.FISHPACK, lambda0=-1.0)
    for Fo in (A, B):
        Fo.build()
        Fo.upwards(Fo.sample_rhs(F))
    r, a, b = A.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    assert np.abs(A.solve(r, a, b) - B.solve(r, a, b)).max() < 1e-12


def test_linearity_and_homogeneous_split():
    """u(f, g) = u(f, 0) + u(0, g): the upwards/solve pair is linear in the load and the data."""
    leaves = mesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, 3)
    Fo = oracle.Forest(leaves, 8, (-10, 10, -10, 10), solver=oracle.FISHPACK)
    Fo.build()
    f = Fo.sample_rhs(F)
    r, a, b = Fo.boundary_data(lambda s, x, y: (U(x, y), 1.0, 0.0))
    Fo.upwards(f)
    u_full = Fo.solve(r, a, b)
    u_g = Fo.solve(r * 0, a, b)
    Fo.upwards(0 * f)
    u_0 = Fo.solve(r, a, b)
    assert np.abs(u_full - (u_g + u_0)).max() < 1e-10 * np.abs(u_full).max()


def test_robin_root_condition_dirichlet_limit():
    """a u + b du/dn = r with b = 0 reproduces the Dirichlet solve; with b != 0 the root DtN closes the system."""
    leaves = mesh.uniform(2)
    Fo = oracle.Forest(leaves, 8, (0, 1, 0, 1), solver=oracle.FISHPACK)
    Fo.build()
    Fo.upwards(Fo.sample_rhs(lambda x, y: np.zeros_like(x)))
    # u = 1 + x + 2y is harmonic; its derivative data h = T g + h_p holds exactly for linear u
    lin = lambda x, y: 1 + x + 2 * y
    r, a, b = Fo.boundary_data(lambda s, x, y: (lin(x, y), 1.0, 0.0))
    u_d = Fo.solve(r, a, b)
    M = Fo.root_size()
    dudn = np.concatenate([np.full(M, 1.0), np.full(M, 1.0), np.full(M, 2.0), np.full(M, 2.0)])
    rr = r + 0.5 * dudn
    u_r = Fo.solve(rr, a, 0.5 * np.ones_like(a))
    assert np.abs(u_r - u_d).max() < 1e-10


def test_mesh_adaptive_poisson_leaf_counts():
    """refineByFunction(|f| > 1.2, 0..7) with corner balance: 262/1568/5920 leaves on levels 5/6/7."""
    leaves = mesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.2, 16, 16, -10, 10, -10, 10, 0, 7)
    levels = np.bincount([l for l, _, _ in leaves])
    assert len(leaves) == 7750 and list(levels[5:]) == [262, 1568, 5920]
    # 2:1 corner balance holds
    S = set(leaves)
    for (l, x, y) in leaves:
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                c = mesh.covering_leaf(S, l, x + dx, y + dy) if 0 <= x + dx < 2**l and 0 <= y + dy < 2**l else None
                assert c is None or c[0] >= l - 1


def test_variable_coefficient_second_order():
    """The variable-beta operator (beta = 1 + 0.5 sin 2pi x cos 2pi y, lambda = -1) with the manufactured
    u = sin(3x) cosh(y): the five-point HPS solution converges at second order."""
    import test_gpu_parity as gp
    errs = []
    for level in (1, 2, 3):
        Fo = oracle.Forest(mesh.uniform(level), 8, (-1.0, 1.0, -1.0, 1.0), solver=oracle.FIVE_POINT,
                           coeff_fns=(gp.ALPHA, gp.BETA, gp.LAMBDA))
        Fo.build()
        Fo.upwards(Fo.sample_rhs(gp.F3))
        r, a, b = Fo.boundary_data(lambda s, x, y: (gp.U3(x, y), 1.0, 0.0))
        u = Fo.solve(r, a, b)
        errs.append(max(np.abs(u[k] - gp.U3(*oracle.cell_centres(8, bd)[:2])).max()
                        for k, bd in enumerate(Fo.leaf_bounds)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert orders.min() > 1.8, (errs, orders)


@pytest.mark.parametrize("name", ["global_5pt_varbeta_L2N8", "global_5pt_varbeta_L3N16"])
def test_variable_hps_equals_global_solve(name):
    """Variable beta (sampled at faces), alpha = 1, lambda = -1: the uniform-mesh HPS solution is
    the global 5-point solution with face-sampled beta (scipy.sparse, independent of the oracle).
    The same solve with beta(y, x) (a transposed coefficient) must miss the fixture by far."""
    import test_gpu_parity as gp
    d = np.load(os.path.join(GOLD, name + ".npz"))
    level, N, bounds = int(d["level"]), int(d["N"]), tuple(d["bounds"])

    def run(beta):
        Fo = oracle.Forest(mesh.uniform(level), N, bounds, solver=oracle.FIVE_POINT,
                           coeff_fns=(gp.ALPHA, beta, gp.LAMBDA))
        Fo.build()
        Fo.upwards(Fo.sample_rhs(gp.F3))
        r, a, b = Fo.boundary_data(lambda s, x, y: (gp.U3(x, y), 1.0, 0.0))
        return Fo.solve(r, a, b)

    scale = np.abs(d["u"]).max()
    assert np.abs(run(gp.BETA) - d["u"]).max() / scale < 1e-12
    if level == 2:
        assert np.abs(run(lambda x, y: gp.BETA(y, x)) - d["u"]).max() / scale > 1e-4
