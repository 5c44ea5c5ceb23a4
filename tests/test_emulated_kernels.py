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


@pytest.mark.skipif(not __import__("os").environ.get("EF_SLOW_TESTS"), reason="a minute on the emulator; the GPU runs it")
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
    gp.test_robin_root_condition(emu_ctx, level=1)


def test_emu_homogeneous_rhs(emu_ctx):
    gp.test_homogeneous_rhs_option(emu_ctx, level=1)


def test_emu_variable_coefficients(emu_ctx):
    """Block-tridiagonal leaves (bt_factor / bt_dtn / bt_solve) with beta = 1 + 0.5 sin 2pi x cos 2pi y, lambda = -1."""
    uo, ug, err = gp.variable_both(emu_ctx, omesh.uniform(1), 8, check_ops=True)
    assert gp.rel(ug, uo) < gp.TOL_U


def test_emu_device_memory_and_stage_timers(emu_ctx):
    """ef_device_alloc / ef_memcpy round trip, the device-pointer stages against the host
    ones (bitwise: same kernels, same order), and the stage timers of ef_mesh_stage_ms."""
    from ellipticforest_amd import _lib
    leaves, N, bounds = omesh.uniform(1), 8, (-1, 1, -1, 1)
    D = _lib.DeviceMesh(emu_ctx, leaves, N, bounds)
    assert all(v == -1.0 for v in D.stage_ms().values())  # nothing run yet
    D.build()
    rng = np.random.default_rng(3)
    f = rng.standard_normal((2, len(leaves), N * N))
    r = rng.standard_normal((2, 4 * D.M))
    a = np.ones(4 * D.M)
    buf = _lib.DeviceArray(emu_ctx, (7,))
    buf.upload(np.arange(3.0), offset=2)
    assert np.array_equal(buf.download()[2:5], np.arange(3.0))
    with pytest.raises(ValueError):
        buf.upload(np.zeros(6), offset=2)
    other = _lib.DeviceArray(emu_ctx, (7,))
    _lib.check(emu_ctx.L.ef_memcpy(emu_ctx.h, other.ptr, buf.ptr, 7 * 8, 2), emu_ctx.L)  # copy kernel
    assert np.array_equal(other.download()[2:5], np.arange(3.0))
    D.upwards(f)
    u_host = D.solve(r, a)
    fd, rd, ad = _lib.DeviceArray(emu_ctx, f.shape), _lib.DeviceArray(emu_ctx, r.shape), _lib.DeviceArray(emu_ctx, a.shape)
    ud = _lib.DeviceArray(emu_ctx, f.shape)
    fd.upload(f); rd.upload(r); ad.upload(a)
    D.upwards_device(fd.ptr, 2)
    D.solve_device(fd.ptr, rd.ptr, ad.ptr, None, ud.ptr, 2)
    D.sync()
    assert np.array_equal(ud.download().reshape(u_host.shape), u_host)
    st = D.stage_ms()
    assert set(st) == {"build-stage", "upwards-stage", "solve-stage"}
    assert all(v >= 0.0 for v in st.values())


def test_emu_graph_replay(emu_ctx):
    """Captured sweeps (emulated capture: kernels recorded, copies and fills refused while a
    capture is open) replay bitwise like the eager launches."""
    leaves = omesh.refine_by_function(lambda x, y: (x < 0.2) & (y < 0.2), 4, 4, 0, 1, 0, 1, 0, 3)
    gp.graph_replay_case(emu_ctx, leaves, 4, (0, 1, 0, 1), k=2)


@pytest.mark.parametrize("m,B,integer", [(16, 3, False), (48, 1, True), (96, 1, False)])
def test_emu_batched_inverse_pivots(emu_ctx, m, B, integer):
    gp.inverse_case(emu_ctx, m, B, seed=m + B, integer=integer)


def test_emu_batched_inverse_pivot_ties(emu_ctx):
    gp.inverse_case(emu_ctx, 48, 1, seed=5, sparse=True)


def test_emu_mixed_precision(emu_ctx):
    """FP32 matrix-core products (emulated v_mfma_f32_16x16x4 layout) and the refinement loop."""
    gp.mixed_case(emu_ctx, 1, 8, variable=True)
    gp.test_mixed_precision_refuses_adaptive(emu_ctx)


def test_emu_multi_rhs_wide_batch(emu_ctx):
    """17 right-hand sides: the two-column-tile GEMM variants, bitwise against one at a time."""
    gp.test_multi_rhs_batch_equals_single(emu_ctx, k=17, maxl=2, cols=(0, 16))


def test_emu_sweep_tiles(emu_ctx):
    """Every sweep tile shape (forced and timed) gives the fixed rule's bits."""
    gp.tile_case(emu_ctx, 17, 1, 8)


def test_emu_variable_adaptive(emu_ctx):
    """Variable beta with coarsened children (five-point leaves of several sizes)."""
    gp.test_variable_coefficients_adaptive(emu_ctx, maxl=2, N=8)


def test_emu_variable_graph_widths(emu_ctx):
    """A captured variable-coefficient sweep is not replayed after its work buffer grew."""
    gp.variable_graph_widths_case(emu_ctx)
