"""MI355X-native dynamical hot path of DiskDyn: tiled gravity, grain forces and RK4.

The compute path is libddyn_hip.so (HIP kernels for gfx950 behind the C ABI in
include/ddyn.h); this package holds the host-side mirror of the reference's
state layout, the seeded synthetic systems and the ctypes binding.
"""
from .system import FP32_FORCE, FP64, FP64_RSQ, DiskState, Params, Workload, make_state  # noqa: F401
from .units import GRAV, LIGHT_AU  # noqa: F401
