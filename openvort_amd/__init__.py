"""MI355X-native vortex-filament velocity / RK4 path of OpenVort.

The compute runs in ``libopenvort_gpu.so`` (HIP, gfx950) behind the C ABI of
``include/openvort_gpu.h``; ``openvort_amd.gpu`` binds it, ``openvort_amd.tangle``
holds the tangle layout and the seeded workload generators.
"""
from .tangle import Tangle, TangleBuilder, default_params  # noqa: F401

__all__ = ["Tangle", "TangleBuilder", "default_params"]
