"""MI355X-native HPS factor-and-solve path of EllipticForest (single GPU).

The numerical core is ``libefhps.so`` (HIP, gfx950) behind the C ABI of
``include/ef_hps.h``; ``hps`` mirrors the reference's HPSAlgorithm / FiniteVolume*
interface on top of it.
"""
from ._lib import Context, DeviceArray, DeviceMesh, EfError, Patches, Timer, refine_by_function  # noqa: F401
