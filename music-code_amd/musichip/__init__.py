"""musichip -- host side of the MI355X-native Music GCMC energy hot path.

The compute path is libmusic_hip.so (HIP kernels for gfx950 behind a C-ABI,
declared in include/music_hip.h); this package reads the reference's input
files, drives the library through ctypes and mirrors the reference's
mapmaker / GCMC driver interfaces.
"""
from . import constants  # noqa: F401
