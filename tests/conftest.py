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


@pytest.fixture(scope="session")
def ctx():
    import ellipticforest_amd as ef
    from ellipticforest_amd import _lib
    if EMULATOR:
        # CPU emulation of the same HIP sources (tests/cpu_emu), for checking kernel
        # indexing on the host; never used outside the test suite
        from cpu_emu import build_emu
        _lib.LIB_PATH = build_emu.LIB
        assert _lib._lib is None or _lib._lib._name == build_emu.LIB, "real library already loaded"
    elif not gpu_present():
        pytest.fail("GPU test selected but no GPU device is visible")
    print(f"[gpu] creating context on device 0 ({_lib.LIB_PATH})", flush=True)
    return ef.Context(0)
