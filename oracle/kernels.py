"""Elementwise operators of the reference CPU path, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference compiles these with ``numba.njit(parallel=True,
fastmath=True)`` (NLSE/kernels_cpu.py).  numba promotes the float32 storage
to FP64 wherever a Python float or a complex literal enters the expression,
and stores the complex128 result back into the complex64 array.  Each
function below reproduces that: FP64 (complex128) intermediates, one
rounding at the store.  In the complex128 oracle mode the store is a no-op.

All functions mutate their first argument(s) in place and return None, like
the reference.  Parameters may be Python scalars or numpy arrays
broadcastable against the field (the reference CPU kernels only take
scalars; numpy broadcasting is used here so the oracle also covers the
GPU-only broadcasting feature, README.md:147-159).
"""

from __future__ import annotations

import numpy as np


def _f64(x):
    """Promote storage-precision data to FP64 the way numba does."""
    if isinstance(x, np.ndarray):
        if np.iscomplexobj(x):
            return x.astype(np.complex128, copy=False)
        return x.astype(np.float64, copy=False)
    return x


def square_mod(A: np.ndarray, A_sq: np.ndarray) -> None:
    """A_sq = re^2 + im^2 in the storage precision (kernels_cpu.py:186-200)."""
    A_sq[...] = A.real * A.real + A.imag * A.imag


def _store(A: np.ndarray, value: np.ndarray) -> None:
    A[...] = value.astype(A.dtype, copy=False)


def nl_prop(A, A_sq, dz, alpha, V, g, Isat, loss_saturates: bool = False) -> None:
    """Real-space step with potential (kernels_cpu.py:5-34).

    ``A *= exp(dz * (-alpha + 1j*g*A_sq*sat + 1j*V))``, sat = 1/(1+A_sq/Isat).
    ``loss_saturates=True`` gives the reference GPU semantics
    (kernels_gpu.py:4-35: loss term ``-alpha*sat``).
    """
    I = _f64(A_sq)
    sat = 1.0 / (1.0 + I / Isat)
    loss = -alpha * sat if loss_saturates else -alpha
    arg = loss + 1j * g * I * sat + 1j * _f64(V)
    _store(A, _f64(A) * np.exp(dz * arg))


def nl_prop_without_V(A, A_sq, dz, alpha, g, Isat, loss_saturates: bool = False) -> None:
    """Real-space step without potential (kernels_cpu.py:37-63; GPU form
    kernels_gpu.py:38-65 when ``loss_saturates``)."""
    I = _f64(A_sq)
    sat = 1.0 / (1.0 + I / Isat)
    loss = -alpha * sat if loss_saturates else -alpha
    arg = loss + 1j * g * I * sat
    _store(A, _f64(A) * np.exp(dz * arg))


def nl_prop_c(A1, A_sq_1, A_sq_2, dz, alpha, V, g11, g12, Isat1, Isat2) -> None:
    """Two-component real-space step with potential (kernels_cpu.py:66-104).

    sat = 1/(1 + I1/Isat1 + I2/Isat2); loss is saturated on both backends.
    """
    I1 = _f64(A_sq_1)
    I2 = _f64(A_sq_2)
    sat = 1.0 / (1.0 + I1 * 1 / Isat1 + I2 * 1 / Isat2)
    arg = -alpha * sat
    arg = arg + 1j * (g11 * I1 * sat + g12 * I2 * sat)
    arg = arg + 1j * _f64(V)
    _store(A1, _f64(A1) * np.exp(dz * arg))


def nl_prop_without_V_c(A1, A_sq_1, A_sq_2, dz, alpha, g11, g12, Isat1, Isat2) -> None:
    """Two-component real-space step without potential (kernels_cpu.py:107-141)."""
    I1 = _f64(A_sq_1)
    I2 = _f64(A_sq_2)
    sat = 1.0 / (1.0 + I1 * 1 / Isat1 + I2 * 1 / Isat2)
    arg = -alpha * sat
    arg = arg + 1j * (g11 * I1 * sat + g12 * I2 * sat)
    _store(A1, _f64(A1) * np.exp(dz * arg))


def rabi_coupling(A1, A2, dz, omega) -> None:
    """Rabi rotation between the two components (kernels_cpu.py:144-161).

    A1' = cos(w dz) A1 - i sin(w dz) A2 ;  A2' = cos(w dz) A2 - i sin(w dz) A1.
    """
    c = np.cos(omega * dz)
    s = np.sin(omega * dz)
    a1 = _f64(A1).copy()
    a2 = _f64(A2)
    new1 = c * a1 - 1j * s * a2
    new2 = c * a2 - 1j * s * a1
    _store(A1, new1)
    _store(A2, new2)
