"""taco_amd: an MI355X (gfx950, wave64) backend for taco's sparse hot path.

The native library libtaco_hip.so (taco_amd/csrc, C ABI in
include/taco_hip.h) holds the HIP kernels, the runtime and the code
generator; this package is its Python face.
"""
from . import _lib  # noqa: F401
