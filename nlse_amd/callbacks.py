"""Generic callbacks (reference NLSE/callbacks.py:1-112) for device fields.

Callbacks are called as ``cb(simu, A, z, i, *args)`` after every step, with
``A`` the device field (torch tensor).  Power sums run in FP64 on the device
(libnlse_hip.so ``nlse_sumsq``); the nonlinear index change and its maximum
in FP32 on the device (``nlse_delta_n``, ``nlse_delta_n_max``).
"""

from __future__ import annotations

import numpy as np
import torch
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
        s = kernels_hip.sumsq(A.reshape(-1, A.shape[-1]), 1).cpu().numpy().sum()
        norms[i // save_every] = float(s)


def evaluate_delta_n(simu, A, z: float, i: int, save_every: int, delta_n) -> None:
    """Nonlinear index change c eps0/2 n2 |A|^2 / (1 + |A|^2/I_sat)
    (callbacks.py:57-82), computed on the device in FP32 (nlse_delta_n);
    stored on the device when delta_n is a device tensor, else on the host."""
    if i % save_every == 0:
        dn = kernels_hip.delta_n(A, simu.n2, simu.I_sat, tdims=len(simu._last_axes))
        if isinstance(delta_n, torch.Tensor):
            delta_n[i // save_every] = dn
        else:
            delta_n[i // save_every] = dn.cpu().numpy()


def adapt_update(simu, A) -> None:
    """The update of callbacks.py:107-112: delta_z = |1/(k max delta_n)|/12
    with delta_n = |n2| q/(1 + q/I_sat) and q = re^2 + im^2 c eps0/2 (the
    reference's operator precedence).  The maximum is reduced on the device
    (nlse_delta_n_max); one float is read back.  The cached propagator is
    not rebuilt, as in the reference, unless simu.adapt_rebuilds_propagator
    is set (no reference counterpart)."""
    m = kernels_hip.delta_n_max(A, simu.n2, simu.I_sat, tdims=len(simu._last_axes))
    with np.errstate(divide="ignore"):
        z_nl = float(1 / (simu.k * np.float64(m)))
    simu.delta_z = np.abs(z_nl) / 12
    if getattr(simu, "adapt_rebuilds_propagator", False):
        simu._own_propagator()


def adapt_delta_z(simu, A, z: float, i: int, update_every: int, delta_z: list) -> None:
    """Step-size update (callbacks.py:85-112): records the step size of every
    step and updates it every update_every steps.  When it is the only
    callback of NLSE.out_field the loop runs the steps between two updates
    as one engine call (same step sizes, same z bookkeeping)."""
    delta_z.append(simu.delta_z)
    if i % update_every == 0:
        adapt_update(simu, A)
