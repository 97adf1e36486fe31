"""nlse_amd: MI355X-native split-step Fourier engine for the NLSE family.

Drop-in for the ``backend="GPU"`` path of Quantum-Optics-LKB/NLSE v2.3.0:
the classes keep the reference constructors, ``out_field`` signatures,
callbacks and leading-dimension broadcasting, and run on hand-written HIP
kernels for gfx950 in ``libnlse_hip.so`` (C ABI: include/nlse_hip.h).
"""

__version__ = "0.2.0"

from . import callbacks, utils  # noqa: F401
from .cnlse import CNLSE  # noqa: F401
from .cnlse_1d import CNLSE_1d  # noqa: F401
from .ddgpe import DDGPE  # noqa: F401
from .gpe import GPE  # noqa: F401
from .nlse import NLSE  # noqa: F401
from .nlse_1d import NLSE_1d  # noqa: F401
from .nlse_3d import NLSE_3d  # noqa: F401
from .batch import NLSEBatch  # noqa: F401,E402
