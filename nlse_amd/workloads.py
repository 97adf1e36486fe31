"""Synthetic inputs of the benchmark workloads (frozen recipes).

Each builder takes the solver class (``nlse_amd.NLSE`` or an oracle class)
so the engine and the CPU oracle are set up from the same numbers.

  readme_nlse            2-D NLSE, README example form: 512^2, alpha=20,
                         P=1.05 W, window 4 w (w = 2.23 mm), n2=-1.6e-9,
                         Isat=1e5, L=10 mm, dz=1e-5 (1001 steps by the
                         reference loop rule), Gaussian beam times (1 + 0.01
                         complex Gaussian noise, seed 0), potential
                         V = -1e-4 exp(-r^2/(70 um)^2).
  speckle_nonlocal_nlse  the same physics at 2048^2, dz=5e-6 (2001 steps),
                         nonlocal length 20 um (optional), Gaussian beam
                         plus a random-phase speckle low-passed to 1/8 of
                         the band (seed 1).
  uneven_batch           64 uneven instances (256^2, 512^2, 1024^2) with
                         random n2, power and Isat, 500 steps each.
  gvd_3d                 3-D NLSE 512x512x64 with GVD (D0 = -1e-26), sech
                         pulse.
"""

from __future__ import annotations

import numpy as np

WAIST = 2.23e-3
WINDOW = 4 * WAIST
L = 10e-3


def _common(cls, N, **kw):
    args = dict(alpha=20, power=1.05, window=WINDOW, n2=-1.6e-9, V=None, L=L, NX=N, NY=N, Isat=1e5, wvl=780e-9)
    args.update(kw)
    return cls(**args)


def _potential(sim):
    return (-1e-4 * np.exp(-(sim.XX**2 + sim.YY**2) / (70e-6) ** 2)).astype(np.float32)


def readme_nlse(cls, N: int = 512, **kw):
    sim = _common(cls, N, **kw)
    sim.delta_z = 1e-5
    sim.V = _potential(sim)
    return sim


def gaussian(sim):
    return np.exp(-(sim.XX.astype(np.float64) ** 2 + sim.YY.astype(np.float64) ** 2) / WAIST**2)


def readme_field(sim, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    G = gaussian(sim)
    nr = rng.standard_normal(G.shape)
    ni = rng.standard_normal(G.shape)
    return (G * (1 + 0.01 * (nr + 1j * ni))).astype(np.complex64)


def speckle_nonlocal_nlse(cls, N: int = 2048, nonlocal_length: float = 20e-6, **kw):
    sim = _common(cls, N, nl_length=nonlocal_length, **kw)
    sim.delta_z = 5e-6
    sim.V = _potential(sim)
    return sim


def speckle_field(sim, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    ny, nx = sim.XX.shape
    phi = rng.uniform(0.0, 2 * np.pi, (ny, nx))
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    # |k| <= Kmax/8 on both axes (Kmax = pi/dx <-> |fftfreq| = 1/2)
    mask = (np.abs(kx) <= 0.5 / 8) & (np.abs(ky) <= 0.5 / 8)
    S = np.fft.ifft2(np.fft.fft2(np.exp(1j * phi)) * mask)
    S /= np.sqrt(np.mean(np.abs(S) ** 2))
    return (gaussian(sim) * (1 + S)).astype(np.complex64)


def uneven_batch(cls, count: int = 64, seed: int = 2, sizes=(256, 512, 1024), **kw):
    """Uneven batch: N_i uniform on `sizes`, n2_i = -10^U(-10,-8),
    P_i ~ U(0.1, 2) W, Isat_i ~ U(1e4, 1e6) W/m^2 (rng(seed), drawn in that
    order per instance), README-example material otherwise, V = None,
    dz = 1e-5; E0_i = Gaussian x (1 + 0.01 complex noise from rng(1000+i))."""
    rng = np.random.default_rng(seed)
    sims, fields = [], []
    for i in range(count):
        N = int(sizes[rng.integers(len(sizes))])
        n2 = -(10 ** rng.uniform(-10, -8))
        power = rng.uniform(0.1, 2.0)
        isat = rng.uniform(1e4, 1e6)
        sim = _common(cls, N, n2=n2, power=power, Isat=isat, **kw)
        sim.delta_z = 1e-5
        r = np.random.default_rng(1000 + i)
        G = gaussian(sim)
        noise = r.standard_normal(G.shape) + 1j * r.standard_normal(G.shape)
        sims.append(sim)
        fields.append((G * (1 + 0.01 * noise)).astype(np.complex64))
    return sims, fields


def gvd_3d(cls, NXY: int = 512, NT: int = 64, **kw):
    """3-D (x, y, t): NLSE_3d, window [4e-3, 4e-3, 2e-12], D0 = -1e-26,
    n2 = -1e-9, alpha = 0, Isat = inf, dz = 1e-5; the energy is chosen so
    the peak power is 1 W for the sech pulse of sech_pulse_field."""
    T0 = 200e-15
    t, dT = np.linspace(-1e-12, 1e-12, NT, retstep=True)
    energy = float(np.sum(1.0 / np.cosh(t / T0) ** 2) * dT)
    sim = cls(alpha=0, energy=energy, window=[4e-3, 4e-3, 2e-12], n2=-1e-9, D0=-1e-26, vg=1.0, V=None, L=L,
              NX=NXY, NY=NXY, NZ=NT, Isat=np.inf, **kw)
    sim.delta_z = 1e-5
    return sim


def sech_pulse_field(sim) -> np.ndarray:
    """Gaussian in space (waist 1 mm) times sech(t / 200 fs), layout (NY, NX, NT)."""
    XX = sim.XX.astype(np.float64)
    YY = sim.YY.astype(np.float64)
    return (np.exp(-(XX**2 + YY**2) / (1e-3) ** 2) / np.cosh(sim.TT / 200e-15)).astype(np.complex64)
