"""C ABI of libefhps.so without a GPU: every symbol include/ef_hps.h declares is exported,
host-only entry points (mesh refinement, footprint planning) work, and the library
refuses to run without a device instead of falling back to the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import ellipticforest_amd as ef
from ellipticforest_amd import _lib
from oracle import mesh as omesh
from conftest import gpu_present

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F = lambda x, y: -(np.sin(x) + np.sin(y))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "ef_hps.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ef_[a-z_]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


@pytest.mark.parametrize("max_level", [3, 4, 5, 6, 7])
def test_product_mesh_builder_matches_oracle_restatement(max_level):
    fn = lambda x, y: np.abs(F(x, y)) > 1.2
    a = omesh.refine_by_function(fn, 16, 16, -10, 10, -10, 10, 0, max_level)
    b = ef.refine_by_function(fn, 16, 16, (-10, 10, -10, 10), 0, max_level)
    assert a == b


def test_mesh_builder_min_level_and_rectangle():
    fn = lambda x, y: (x - 0.3) ** 2 + (y - 1.1) ** 2 < 0.05
    a = omesh.refine_by_function(fn, 8, 8, 0.0, 1.0, 0.0, 2.0, 2, 6)
    b = ef.refine_by_function(fn, 8, 8, (0.0, 1.0, 0.0, 2.0), 2, 6)
    assert a == b and min(l for l, _, _ in b) >= 2


def test_mesh_builder_rejects_bad_levels():
    with pytest.raises(ef.EfError):
        ef.refine_by_function(lambda x, y: x > 0, 8, 8, (0, 1, 0, 1), 3, 2)


def test_footprint_scales_with_rhs():
    leaves = omesh.uniform(3)
    one = _lib.DeviceMesh.footprint(leaves, 16, (-10, 10, -10, 10), max_rhs=1)
    many = _lib.DeviceMesh.footprint(leaves, 16, (-10, 10, -10, 10), max_rhs=32)
    assert many > one > 0


def test_footprint_rejects_broken_leaf_lists():
    leaves = omesh.uniform(2)[:-1]  # a quadrant is missing
    with pytest.raises(ef.EfError):
        _lib.DeviceMesh.footprint(leaves, 8, (0, 1, 0, 1))
    with pytest.raises(ef.EfError):
        _lib.DeviceMesh.footprint(omesh.uniform(2), 7, (0, 1, 0, 1))  # odd N


@pytest.mark.skipif(gpu_present(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    with pytest.raises(ef.EfError):
        ef.Context(0)


def test_mesh_rejects_leaf_patches_above_32():
    """The mesh path's leaf kernels pad patches to 16 or 32 cells: N = 48 or 64 is refused at
    planning (single grids of any size run on ef_patches_*)."""
    from ellipticforest_amd import _lib
    for N in (48, 64):
        with pytest.raises(_lib.EfError):
            _lib.DeviceMesh.footprint([(1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1)], N, (0, 1, 0, 1))
    _lib.DeviceMesh.footprint([(1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1)], 32, (0, 1, 0, 1))
