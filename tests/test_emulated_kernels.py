"""The product's HIP kernels and host planner, compiled for the CPU against the HIP shim
of tests/cpu_emu (guarded device buffers, emulated workgroups and f64 MFMA), checked
against the oracle.  This checks the kernels' indexing and arithmetic on the host; it
is not a fallback of the product (the product library refuses to run without a GPU)."""
import numpy as np
import pytest

import oracle
from oracle import mesh as omesh
import test_gpu_parity as gp


def test_emu_uniform_L1_all_operators(emu_ctx):
    O, D, uo, ug = gp.solve_both(emu_ctx, omesh.uniform(1), 8, (-1, 1, -1, 1), check_ops=True)
    assert gp.rel(ug, uo) < gp.TOL_U


def test_emu_golden_global_solve(emu_ctx):
    gp.test_uniform_equals_global_solve(emu_ctx, "global_5pt_L2N8_helmholtz")


def test_emu_adaptive_with_coarsening(emu_ctx):
    """Uneven siblings: coarsen_ (T), coarsenUpwards_ (h) and uncoarsen_ (g) all run."""
    leaves = omesh.refine_by_function(lambda x, y: (x < 0.2) & (y < 0.2), 4, 4, 0, 1, 0, 1, 0, 3)
    O, D, uo, ug = gp.solve_both(emu_ctx, leaves, 4, (0, 1, 0, 1), check_ops=True)
    assert max(D.node_info[:, 5]) >= 1  # some node was coarsened
    assert gp.rel(ug, uo) < gp.TOL_U


@pytest.mark.skipif(not __import__("os").environ.get("EF_SLOW_TESTS"), reason="several minutes on the emulator")
def test_emu_double_coarsening(emu_ctx):
    """A mesh with nodes coarsened twice (multi-step coarsen_/uncoarsen_)."""
    F = lambda x, y: -(np.sin(x) + np.sin(y))
    leaves = omesh.refine_by_function(lambda x, y: np.abs(F(x, y)) > 1.8, 4, 4, -10, 10, -10, 10, 0, 5)
    O, D, uo, ug = gp.solve_both(emu_ctx, leaves, 4, (-10, 10, -10, 10), check_ops=True, nodes=range(0, 40))
    assert max(D.node_info[:, 5]) >= 2
    assert gp.rel(ug, uo) < gp.TOL_U


def test_emu_robin_root(emu_ctx):
    gp.test_robin_root_condition(emu_ctx)


def test_emu_homogeneous_rhs(emu_ctx):
    gp.test_homogeneous_rhs_option(emu_ctx)


def test_emu_variable_coefficients(emu_ctx):
    """Block-tridiagonal leaves (bt_factor / bt_dtn / bt_solve) with cfg3's beta and lambda."""
    uo, ug, err = gp.variable_both(emu_ctx, omesh.uniform(1), 8, check_ops=True)
    assert gp.rel(ug, uo) < gp.TOL_U
