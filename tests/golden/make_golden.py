"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``; ``... make_golden.py varbeta`` writes only the
variable-beta global solutions).

Three kinds of data, none produced by running the reference (it cannot be built here:
no MPI, p4est, PETSc, FISHPACK90 or LAPACK):

1. reference_known_answers.json -- the known answers the reference's own gtest suite
   asserts, transcribed with their file:line (values only).
2. global_5pt_*.npz -- the uniform-mesh HPS solution must equal the global 5-point
   finite-volume solution (ghost = 2g - u on the physical boundary; at an interior
   face flux continuity gives g = (u_L + u_R)/2).  These fixtures are that global
   solution, computed here with scipy.sparse, independently of the oracle and of the
   GPU library.
3. oracle_*.npz -- small outputs of the oracle restatement (leaf DtN, one merge's
   X/S/T/H, an adaptive-mesh solution) that only a faithful restatement pins; the
   tests check the GPU library against them and the oracle against the closed forms
   that constrain them.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

KNOWN = {
    "test_matrix.cpp:155-192 linear_solve": {
        "A": [[1, 2, 3], [4, 5, 6], [7, 8, 10]], "b": [1, 1, 1], "x": [-1, 1, 0], "tol": 1e-15,
        "B": [[1, 2, 3, 4]] * 3, "X": [[-1, -2, -3, -4], [1, 2, 3, 4], [0, 0, 0, 0]], "tol_matrix": 1e-14,
    },
    "test_matrix.cpp:194-268 math": {
        "m6": [[1, 2, 3], [4, 5, 6]], "m7": [[1, 2], [2, 1], [2, 3]], "m6m7": [[11, 13], [26, 31]],
        "v1": [2, 2, 2], "m6v1": [12, 30],
    },
    "test_matrix.cpp:270-300 block_permute": {
        "m1": [[1, 1, 2, 2, 3, 3], [1, 1, 2, 2, 3, 3], [4, 4, 5, 5, 6, 6], [7, 7, 8, 8, 9, 9]],
        "pi_rows": [2, 1, 0], "pi_cols": [0, 2, 1], "R": [2, 1, 1], "C": [2, 2, 2],
        "expected": [[7, 7, 9, 9, 8, 8], [4, 4, 6, 6, 5, 5], [1, 1, 3, 3, 2, 2], [1, 1, 3, 3, 2, 2]],
    },
    "test_finite_volume_grid.cpp:30-49 points": {
        "nx": 4, "x": [0, 1], "ny": 4, "y": [1, 3],
        "xpoints": [0.125, 0.375, 0.625, 0.875], "ypoints": [1.25, 1.75, 2.25, 2.75], "tol": 1e-16,
    },
    "test_finite_volume_solver.cpp:21-95 solver_convergence_FISHPACK": {
        "domain": [-1, 1, -1, 1], "ns": [8, 16, 32, 64, 128, 256, 512],
        "u": "sin(x)+sin(y)", "f": "-(sin(x)+sin(y))", "min_order_last_pair": 1.95,
    },
}


def global_5pt(M, bounds, f, g, lam=0.0):
    """Cell-centred 5-point solution on an M x M grid (FISHPACK semantics, alpha=beta=1)."""
    xl, xu, yl, yu = bounds
    dx, dy = (xu - xl) / M, (yu - yl) / M
    xs = (xl + dx / 2) + np.arange(M) * dx
    ys = (yl + dy / 2) + np.arange(M) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    n = M * M
    idx = np.arange(n).reshape(M, M)  # idx[i, j] = j + i*M
    cx, cy = 1.0 / dx**2, 1.0 / dy**2
    diag = np.full((M, M), -2 * cx - 2 * cy + lam)
    rhs = f(X, Y).copy()
    rows, cols, vals = [], [], []
    # x neighbours
    rows.append(idx[1:, :].ravel()); cols.append(idx[:-1, :].ravel()); vals.append(np.full((M - 1) * M, cx))
    rows.append(idx[:-1, :].ravel()); cols.append(idx[1:, :].ravel()); vals.append(np.full((M - 1) * M, cx))
    rows.append(idx[:, 1:].ravel()); cols.append(idx[:, :-1].ravel()); vals.append(np.full((M - 1) * M, cy))
    rows.append(idx[:, :-1].ravel()); cols.append(idx[:, 1:].ravel()); vals.append(np.full((M - 1) * M, cy))
    # boundary: ghost = 2g - u
    diag[0, :] -= cx; rhs[0, :] -= 2 * cx * g(np.full(M, xl), ys)
    diag[-1, :] -= cx; rhs[-1, :] -= 2 * cx * g(np.full(M, xu), ys)
    diag[:, 0] -= cy; rhs[:, 0] -= 2 * cy * g(xs, np.full(M, yl))
    diag[:, -1] -= cy; rhs[:, -1] -= 2 * cy * g(xs, np.full(M, yu))
    rows.append(idx.ravel()); cols.append(idx.ravel()); vals.append(diag.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    u = spla.spsolve(A.tocsc(), rhs.ravel())
    return u.reshape(M, M)


def global_5pt_var(M, bounds, alpha, beta, lam, f, g):
    """Cell-centred 5-point solution of alpha div(beta grad u) + lambda u = f on an M x M grid:
    alpha and lambda at cell centres, beta at each face centre (sampled once per face and
    shared by the two cells on it), ghost = 2g - u on the physical boundary.  At an interior
    face the HPS interface condition (equal and opposite normal derivatives of the two leaf
    solutions, h_L + h_R = 0) gives g = (u_L + u_R)/2, so the uniform-mesh HPS solution of
    the variable-beta operator is this global solution."""
    xl, xu, yl, yu = bounds
    dx, dy = (xu - xl) / M, (yu - yl) / M
    xs = (xl + dx / 2) + np.arange(M) * dx
    ys = (yl + dy / 2) + np.arange(M) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    xf = xl + np.arange(M + 1) * dx  # x faces
    yf = yl + np.arange(M + 1) * dy
    bx = beta(*np.meshgrid(xf, ys, indexing="ij"))  # (M+1) x M: face between cells i-1 and i
    by = beta(*np.meshgrid(xs, yf, indexing="ij"))  # M x (M+1)
    a = alpha(X, Y)
    n = M * M
    idx = np.arange(n).reshape(M, M)
    cW, cE = a * bx[:-1, :] / dx**2, a * bx[1:, :] / dx**2
    cS, cN = a * by[:, :-1] / dy**2, a * by[:, 1:] / dy**2
    diag = -(cW + cE + cS + cN) + lam(X, Y)
    rhs = f(X, Y).copy()
    rows, cols, vals = [], [], []
    rows.append(idx[1:, :].ravel()); cols.append(idx[:-1, :].ravel()); vals.append(cW[1:, :].ravel())
    rows.append(idx[:-1, :].ravel()); cols.append(idx[1:, :].ravel()); vals.append(cE[:-1, :].ravel())
    rows.append(idx[:, 1:].ravel()); cols.append(idx[:, :-1].ravel()); vals.append(cS[:, 1:].ravel())
    rows.append(idx[:, :-1].ravel()); cols.append(idx[:, 1:].ravel()); vals.append(cN[:, :-1].ravel())
    diag[0, :] -= cW[0, :]; rhs[0, :] -= 2 * cW[0, :] * g(np.full(M, xl), ys)
    diag[-1, :] -= cE[-1, :]; rhs[-1, :] -= 2 * cE[-1, :] * g(np.full(M, xu), ys)
    diag[:, 0] -= cS[:, 0]; rhs[:, 0] -= 2 * cS[:, 0] * g(xs, np.full(M, yl))
    diag[:, -1] -= cN[:, -1]; rhs[:, -1] -= 2 * cN[:, -1] * g(xs, np.full(M, yu))
    rows.append(idx.ravel()); cols.append(idx.ravel()); vals.append(diag.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    u = spla.spsolve(A.tocsc(), rhs.ravel())
    return u.reshape(M, M)


def leaf_blocks(U, level, N):
    """Global grid -> per-leaf cell vectors in z-order (index j + i*N)."""
    from oracle import mesh
    out = []
    for (l, ix, iy) in mesh.uniform(level):
        out.append(U[ix * N:(ix + 1) * N, iy * N:(iy + 1) * N].ravel())
    return np.array(out)


def write_varbeta():
    """global_5pt_varbeta_*.npz: the variable-beta operator of the GPU parity tests (alpha = 1,
    beta = 1 + 0.5 sin(2 pi x) cos(2 pi y), lambda = -1, Dirichlet data and f of the manufactured
    u = sin(3x) cosh(y)) solved globally on uniform meshes."""
    sys.path.insert(0, os.path.dirname(HERE))
    import test_gpu_parity as gp
    for name, level, N in [("global_5pt_varbeta_L2N8", 2, 8), ("global_5pt_varbeta_L3N16", 3, 16)]:
        bounds = (-1.0, 1.0, -1.0, 1.0)
        M = N << level
        U = global_5pt_var(M, bounds, gp.ALPHA, gp.BETA, gp.LAMBDA, gp.F3, gp.U3)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), level=level, N=N, bounds=np.array(bounds),
                            u=leaf_blocks(U, level, N))


def main():
    if sys.argv[1:] == ["varbeta"]:
        write_varbeta()
        return
    with open(os.path.join(HERE, "reference_known_answers.json"), "w") as f:
        json.dump(KNOWN, f, indent=1)
    u_ex = lambda x, y: np.sin(x) + np.sin(y)
    f_ex = lambda x, y: -(np.sin(x) + np.sin(y))
    for name, level, N, bounds, lam in [("global_5pt_L2N8_pm1", 2, 8, (-1.0, 1.0, -1.0, 1.0), 0.0),
                                        ("global_5pt_L2N8_pm10", 2, 8, (-10.0, 10.0, -10.0, 10.0), 0.0),
                                        ("global_5pt_L3N16_pm10", 3, 16, (-10.0, 10.0, -10.0, 10.0), 0.0),
                                        ("global_5pt_L2N8_helmholtz", 2, 8, (-1.0, 1.0, -1.0, 1.0), -1.0)]:
        M = N << level
        U = global_5pt(M, bounds, f_ex, u_ex, lam)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), level=level, N=N, bounds=np.array(bounds), lam=lam,
                            u=leaf_blocks(U, level, N))
    # oracle-only fixtures
    import oracle
    from oracle import mesh
    T4 = oracle.leaf_dtn(4, (0.0, 1.0, 0.0, 1.0))
    T8 = oracle.leaf_dtn(8, (-1.0, 1.0, -1.0, 1.0), lambda0=-1.0)
    F = oracle.Forest(mesh.uniform(1), 4, (0.0, 1.0, 0.0, 1.0), solver=oracle.FISHPACK)
    F.build()
    merge = {w: F.get(0, w) for w in ("X", "S", "T", "H")}
    leaves = mesh.refine_by_function(lambda x, y: np.abs(f_ex(x, y)) > 1.2, 8, 8, -10, 10, -10, 10, 0, 3)
    G = oracle.Forest(leaves, 8, (-10, 10, -10, 10), solver=oracle.FISHPACK)
    G.build()
    fr = G.sample_rhs(f_ex)
    G.upwards(fr)
    r, a, b = G.boundary_data(lambda s, x, y: (u_ex(x, y), 1.0, 0.0))
    u = G.solve(r, a, b)
    np.savez_compressed(os.path.join(HERE, "oracle_fixtures.npz"), T4=T4, T8_helmholtz=T8,
                        merge_X=merge["X"], merge_S=merge["S"], merge_T=merge["T"], merge_H=merge["H"],
                        adaptive_leaves=np.array(leaves), adaptive_u=u)
    write_varbeta()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
