"""The allocation-free variant of ef_patches_apply_device (ef_patches_apply_device_noalloc),
called through the C ABI: its output against ef_patches_apply_device's on the same inputs,
bitwise (the same kernels in the same order; only the work space differs: the library's fixed
scratch array instead of the batch's own buffers), on two calls in a row, the second with a
smaller input.  Covers the three kinds of batch: packed variable-coefficient Neumann maps and
large patches (both use the scratch array) and constant-coefficient patches (no work space),
and the refusal of a batch that needs more work space than the array holds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def beta_cells(bounds, N):
    beta = lambda x, y: 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    out = np.empty((len(bounds), 6, N * N))
    for p, (xl, xu, yl, yu) in enumerate(bounds):
        dx, dy = (xu - xl) / N, (yu - yl) / N
        X, Y = np.meshgrid(xl + (np.arange(N) + 0.5) * dx, yl + (np.arange(N) + 0.5) * dy, indexing="ij")
        X, Y = X.ravel(), Y.ravel()
        out[p] = [np.ones_like(X), beta(X - dx / 2, Y), beta(X + dx / 2, Y), beta(X, Y - dy / 2),
                  beta(X, Y + dy / 2), -np.ones_like(X)]
    return out


def grid_bounds(n):
    w = 2.0 / n
    return np.array([(-1 + i * w, -1 + (i + 1) * w, -1 + j * w, -1 + (j + 1) * w) for i in range(n) for j in range(n)])


def both(ctx, pt, op, N, k, seed, with_g=True):
    """(apply_device output, noalloc output) for random g and f."""
    from ellipticforest_amd import _lib
    import ellipticforest_amd as ef
    rng = np.random.default_rng(seed)
    P = pt.P
    g_h = rng.standard_normal((k, P, 4 * N))
    f_h = rng.standard_normal((k, P, N * N))
    out_n = k * P * (N * N if op == _lib.EF_PATCH_SOLVE else 4 * N)
    g, f = ef.DeviceArray(ctx, g_h.shape), ef.DeviceArray(ctx, f_h.shape)
    g.upload(g_h)
    f.upload(f_h)
    o1, o2 = ef.DeviceArray(ctx, (out_n,)), ef.DeviceArray(ctx, (out_n,))
    gp = g.ptr if with_g else None
    pt.apply_device(op, gp, f.ptr, o1.ptr, k)
    _lib.check(ctx.L.ef_patches_apply_device_noalloc(pt.h, op, gp, f.ptr, o2.ptr, k), ctx.L)
    pt.sync()
    return o1.download(), o2.download()


def test_noalloc_packed_variable_neumann_maps(ctx):
    """Packed symmetric variable-coefficient batch, particularNeumannData: the forward rows go
    through the scratch array.  Second call: a smaller batch."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    N = 32
    for n, k, seed in ((8, 2, 0), (4, 1, 1)):
        say(f"noalloc: packed variable-beta Neumann maps, {n * n} patches N={N}, {k} right-hand sides")
        b = grid_bounds(n)
        pt = ef.Patches(ctx, b, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=beta_cells(b, N))
        assert pt.packed
        a, c = both(ctx, pt, _lib.EF_PATCH_MAP_D2N, N, k, seed, with_g=False)
        assert np.array_equal(a, c)
        assert np.isfinite(a).all() and np.abs(a).max() > 0


def test_noalloc_large_patches(ctx):
    """Patches with N > 32 (batched-GEMM path): both GEMM work buffers in the scratch array.
    Second call: fewer patches and right-hand sides."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    N = 64
    for n, k, seed in ((3, 3, 2), (1, 1, 3)):
        say(f"noalloc: large patches, {n * n} patches N={N}, {k} right-hand sides")
        pt = ef.Patches(ctx, grid_bounds(n), N, lambda0=-1.0)
        for op in (_lib.EF_PATCH_SOLVE, _lib.EF_PATCH_MAP_D2N):
            a, c = both(ctx, pt, op, N, k, seed)
            assert np.array_equal(a, c)


def test_noalloc_constant_patches(ctx):
    """Constant-coefficient patches need no work space; the variant gives the same output."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    N = 16
    for n, k, seed in ((10, 4, 4), (2, 1, 5)):
        say(f"noalloc: constant patches, {n * n} patches N={N}, {k} right-hand sides")
        pt = ef.Patches(ctx, grid_bounds(n), N)
        for op in (_lib.EF_PATCH_SOLVE, _lib.EF_PATCH_MAP_D2N):
            a, c = both(ctx, pt, op, N, k, seed)
            assert np.array_equal(a, c)


def test_noalloc_refuses_oversized_batch(ctx):
    """A packed Neumann-map batch whose work space exceeds the scratch array (65^2 = 4,225
    patches x 32^2 doubles > 2^22) is refused before any launch, and the call after it still
    works."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    N = 32
    say("noalloc: oversized batch refused")
    b = grid_bounds(65)
    pt = ef.Patches(ctx, b, N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=beta_cells(b, N))
    assert pt.packed
    f = ef.DeviceArray(ctx, (pt.P, N * N))
    h = ef.DeviceArray(ctx, (pt.P, 4 * N))
    rc = ctx.L.ef_patches_apply_device_noalloc(pt.h, _lib.EF_PATCH_MAP_D2N, None, f.ptr, h.ptr, 1)
    assert rc == _lib.EF_ERR_ARG
    assert b"work space" in ctx.L.ef_last_error()
    small = ef.Patches(ctx, grid_bounds(2), N, solver_type=_lib.EF_SOLVER_FIVE_POINT, cell=beta_cells(grid_bounds(2), N))
    a, c = both(ctx, small, _lib.EF_PATCH_MAP_D2N, N, 1, 6, with_g=False)
    assert np.array_equal(a, c)
