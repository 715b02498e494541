"""MI355X-native morph engine of atomorph (gfx950).

The engine is the C++ class am::morph (atomorph_amd/host) over the C ABI in
include/atomorph_hip.h (libatomorph_hip.so, hand-written HIP kernels).  This
Python package only builds the libraries, binds the C ABI with ctypes for
tests and benchmarks, and generates synthetic key frames.
"""
__all__ = ["_capi", "synth"]
