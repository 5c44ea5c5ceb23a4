import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def gpu_present():
    """True when a HIP device is usable; decided without initialising torch."""
    return os.path.exists("/dev/kfd") and any(
        os.path.exists(f"/dev/dri/renderD{n}") for n in range(128, 192))


EMULATOR = os.environ.get("EF_TEST_EMULATOR") == "1"

def emulator_library():
    """The CPU-emulation build of the HIP sources (built on first use)."""
    from cpu_emu import build_emu
    from ellipticforest_amd import _lib
    if not os.path.exists(build_emu.LIB) or build_emu.stale():
        build_emu.build()
    return _lib.load_library(build_emu.LIB)


@pytest.fixture(autouse=True)
def _emulated_fixed_tiles(request, monkeypatch):
    """On the emulator, timing the sweep tile shapes measures nothing and costs minutes: keep the
    fixed rule there, except in the tests of the tile choice itself."""
    if ("emu_ctx" in request.fixturenames or EMULATOR) and "sweep_tiles" not in request.node.name:
        monkeypatch.setenv("EF_AUTOTUNE", "0")


@pytest.fixture(scope="session")
def emu_ctx():
    import ellipticforest_amd as ef
    return ef.Context(0, library=emulator_library())


@pytest.fixture(scope="session")
def ctx():
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    if EMULATOR:
        print("[gpu] context on the CPU emulator", flush=True)
        return ef.Context(0, library=emulator_library())
    if not gpu_present():
        pytest.fail("GPU test selected but no GPU device is visible")
    print(f"[gpu] creating context on device 0 ({_lib.LIB_PATH})", flush=True)
    return ef.Context(0)


@pytest.fixture(scope="session")
def patch_ctx():
    """Context for the batched patch entry points (ef_patches_*)."""
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    if EMULATOR:
        print("[gpu] patch context on the CPU emulator", flush=True)
        return ef.Context(0, library=emulator_library())
    if not gpu_present():
        pytest.fail("GPU test selected but no GPU device is visible")
    print(f"[gpu] creating context on device 0 ({_lib.LIB_PATH})", flush=True)
    return ef.Context(0)
