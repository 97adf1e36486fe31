"""Generic callbacks (reference NLSE/callbacks.py:1-112) for device fields.

Callbacks are called as ``cb(simu, A, z, i, *args)`` after every step, with
``A`` the device field (torch tensor).  Reductions run in FP64 on the device
(libnlse_hip.so ``nlse_sumsq``).
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.constants import c, epsilon_0

from . import kernels_hip


def sample(simu, A, z: float, i: int, save_every: int, E_samples) -> None:
    """Save a copy of the field every save_every steps (callbacks.py:7-29)."""
    if i % save_every == 0:
        if isinstance(E_samples, torch.Tensor):
            E_samples[i // save_every] = A
        else:
            E_samples[i // save_every] = A.detach().cpu().numpy()


def norm(simu, A, z: float, i: int, save_every: int, norms) -> None:
    """Save sum |A|^2 over all axes every save_every steps (callbacks.py:32-54;
    no dx*dy factor, like the reference)."""
    if i % save_every == 0:
        s = kernels_hip.sumsq(A.reshape(-1, A.shape[-1]), 1).sum()
        norms[i // save_every] = float(s)


def evaluate_delta_n(simu, A, z: float, i: int, save_every: int, delta_n) -> None:
    """Nonlinear index change c eps0/2 n2 |A|^2 / (1 + |A|^2/I_sat)
    (callbacks.py:57-82); stored on the host."""
    if i % save_every == 0:
        A_sq = (A.real * A.real + A.imag * A.imag).detach().cpu().numpy()
        n2 = np.asarray(simu.n2.detach().cpu().numpy() if isinstance(simu.n2, torch.Tensor) else simu.n2)
        isat = np.asarray(simu.I_sat.detach().cpu().numpy() if isinstance(simu.I_sat, torch.Tensor)
                          else simu.I_sat)
        delta_n[i // save_every] = c * epsilon_0 / 2 * n2 * A_sq / (1 + A_sq / isat)


def adapt_delta_z(simu, A, z: float, i: int, update_every: int, delta_z: list) -> None:
    """Step-size update (callbacks.py:85-112), including the reference's
    precedence (only im^2 is scaled by c eps0/2, callbacks.py:109).  The
    cached propagator is not rebuilt, as in the reference."""
    delta_z.append(simu.delta_z)
    if i % update_every == 0:
        re = A.real.detach().cpu().numpy().astype(np.float64)
        im = A.imag.detach().cpu().numpy().astype(np.float64)
        A_sq = re * re + im * im * c * epsilon_0 / 2
        n2 = np.asarray(simu.n2.detach().cpu().numpy() if isinstance(simu.n2, torch.Tensor) else simu.n2)
        isat = np.asarray(simu.I_sat.detach().cpu().numpy() if isinstance(simu.I_sat, torch.Tensor)
                          else simu.I_sat)
        delta_n = np.abs(n2) * A_sq / (1 + A_sq / isat)
        z_nl = float(1 / (simu.k * delta_n.max()))
        simu.delta_z = np.abs(z_nl) / 12
