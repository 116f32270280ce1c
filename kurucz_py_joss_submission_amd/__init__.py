"""MI355X-native synthe_py hot path: line opacity, continuum, JOSH RT, emergent flux.

The compute runs in hand-written HIP kernels for gfx950 (libsynthe_hip.so,
reached through a C ABI with ctypes); this package holds the host-side mirror
of the reference's engine/physics call signatures.
"""

__version__ = "0.1.0"
