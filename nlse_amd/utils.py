"""Backend detection (reference: NLSE/utils.py:1-27).

The reference starts from "GPU" and falls back to "CPU" when CuPy or
PyOpenCL is missing.  This package implements only the GPU backend, on
libnlse_hip.so; its default is "GPU" and nothing falls back: when the
library or the GPU is absent, creating a solver raises ImportError.
"""

from . import _lib

__BACKEND__ = "GPU"
__HIP_AVAILABLE__ = _lib.available()
# Names of the reference's availability flags, kept for code that reads
# them (reference tests gate their GPU branches on __CUPY_AVAILABLE__).
__CUPY_AVAILABLE__ = False
__PYOPENCL_AVAILABLE__ = False
