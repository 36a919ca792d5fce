"""MI355X-native (gfx950 HIP) AWP-ODC time-stepping path.

The kernels and their C ABI live in libawp_hip.so (sources in csrc/,
declarations in include/awp_hip.h).  This package holds the host-side
setup (grid layout, medium, synthetic workloads) and a driver that calls the
C ABI in the reference's per-step order.
"""
from .grid import Dims, ALIGN, NGSL, NGSL2  # noqa: F401

__all__ = ["Dims", "ALIGN", "NGSL", "NGSL2"]
