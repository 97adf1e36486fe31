"""Hamiltonian of the single-component 2-D model, restated in numpy FP64.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The split step of nlse.py:372-487 integrates
    dA/dz = (i/2k) lap A + i (g I/(1 + I/Is) + V') A,   I = |A|^2,
with the linear factor exp(-i k_perp^2 dz/2k) (nlse.py:188-197) and the
nonlinear phase rate of kernels_cpu.nl_prop (g = k/2 n2 c eps0,
Is = 2 I_sat/(eps0 c), V' = k/2 Re V).  Without loss it conserves
    H = T - U,  T = dx dy sum (1/2k)|grad A|^2,
                U = dx dy sum [G(I) + V' I],  G(I) = g Is (I - Is ln(1 + I/Is))
(G = g I^2/2 for Is = inf), the reference has no such quantity: it is the
drift measure of the learned corrector's fallback.

Pinning (tests/test_oracle_pins.py): a plane wave gives the closed form
T = |k0|^2 a^2 area/2k, U = g a^4 area/2; G equals the quadrature of the
phase rate on both sides of the series switch; H is conserved (to the
splitting order) by the oracle's own lossless propagation while its kinetic
and potential parts each change by two orders of magnitude more.
"""

from __future__ import annotations

import numpy as np
from scipy.constants import c, epsilon_0


def G(I: np.ndarray, g: float, Is: float) -> np.ndarray:
    """Antiderivative of g s/(1 + s/Is) from 0 to I."""
    I = np.asarray(I, dtype=np.float64)
    if not np.isfinite(Is):
        return 0.5 * g * I * I
    x = I / Is
    # x - log1p(x), by its series where the difference cancels
    d = np.where(x < 1e-4, x * x * (0.5 - x * (1.0 / 3.0 - 0.25 * x)), x - np.log1p(np.maximum(x, 1e-4)))
    return g * Is * Is * d


def hamiltonian_terms(sim, A: np.ndarray):
    """(T, U) over the last two axes of A (complex, [..., NY, NX])."""
    A = np.asarray(A).astype(np.complex128)
    Ah = np.fft.fft2(A, axes=(-2, -1))
    k2 = np.asarray(sim.Kx, np.float64)[None, :] ** 2 + np.asarray(sim.Ky, np.float64)[:, None] ** 2
    npts = A.shape[-1] * A.shape[-2]
    cell = sim.delta_X * sim.delta_Y
    T = cell * (k2 * np.abs(Ah) ** 2).sum(axis=(-2, -1)) / npts / (2 * sim.k)
    g = sim.k / 2 * float(np.asarray(sim.n2).reshape(-1)[0]) * c * epsilon_0
    Is = 2 * float(np.asarray(sim.I_sat).reshape(-1)[0]) / (epsilon_0 * c)
    I = np.abs(A) ** 2
    u = G(I, g, Is)
    if sim.V is not None:
        u = u + sim.k / 2 * np.real(np.asarray(sim.V, dtype=np.complex128)) * I
    U = cell * u.sum(axis=(-2, -1))
    return T, U


def hamiltonian(sim, A: np.ndarray):
    T, U = hamiltonian_terms(sim, A)
    return T - U
