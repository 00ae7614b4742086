"""MI355X-native transport hot path of bevy_lux (Split Radiance Cascades).

The compute path is liblux_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/lux.h); :mod:`bevy_penumbra_amd.lux` binds it with ctypes.
"""
__version__ = "0.1.0"
