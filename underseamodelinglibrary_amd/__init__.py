"""MI355X-native WaveQ3D wavefront-propagation hot path.

The compute path is libwq3d.so (HIP kernels for gfx950 behind the C ABI in
include/wq3d.h); usml/ mirrors the reference's usml::types, usml::ocean and
usml::waveq3d interfaces on top of it.
"""
__version__ = "0.1.0"
