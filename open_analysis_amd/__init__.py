"""MI355X-native oa-fdtd time-stepping engine (HIP kernels for gfx950 behind a C ABI)."""

__version__ = "0.1.0"
