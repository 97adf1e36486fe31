"""Split-step solvers of the reference CPU path, restated in numpy/scipy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Nothing in the product
package ``nlse_amd`` imports this module.

Every class follows the reference file it names, method by method:

  NLSE      NLSE/nlse.py:47-155 (init), 188-197 (propagator), 257-307
            (output array / normalisation), 372-487 (split_step),
            489-619 (out_field loop)
  NLSE_1d   NLSE/nlse_1d.py:18-105
  NLSE_3d   NLSE/nlse_3d.py:20-150
  GPE       NLSE/gpe.py:18-119
  CNLSE     NLSE/cnlse.py:22-148, 179-423
  CNLSE_1d  NLSE/cnlse_1d.py:16-131
  DDGPE     NLSE/ddgpe.py:17-94, 127-156, 204-457, 519-565

The FFTs of pyFFTW (forward unnormalised, inverse scaled by 1/N,
nlse.py:440-442) are replaced by scipy.fft (pocketfft), which keeps single
precision for complex64 data like FFTW's fftwf plans do.

``mode="c64"`` reproduces the reference CPU numerics; ``mode="c128"`` is
the FP64 oracle (same scheme in complex128).  ``loss_saturates=True``
switches the single-component loss to the reference GPU semantics
(kernels_gpu.py:30,62).

Deliberate deviations (the reference result is undefined or an exception):
  * batched fields with a 2-D V / nonlocal kernel are broadcast correctly
    (the numba kernel reads V out of bounds, kernels_cpu.py:26-34, and
    oaconvolve rejects unequal ndim);
  * GPE with normalize=False copies E_in (the reference leaves A
    uninitialised, gpe.py:102-119).
"""

from __future__ import annotations

import os
from typing import Callable, Union

import numpy as np
import scipy.fft as sfft
from scipy import signal, special
from scipy.constants import atomic_mass, c, epsilon_0, hbar

from . import kernels


def count_steps(z: float, delta_z: float) -> int:
    """Number of loop iterations of ``while abs(z_prop) < z`` (nlse.py:566-585)
    for a fixed ``delta_z``: the float accumulation decides, e.g. 1001 steps
    for z=10e-3, dz=1e-5."""
    z_prop = 0
    n = 0
    while abs(z_prop) < z:
        z_prop += delta_z
        n += 1
    return n


def _workers() -> int:
    return os.cpu_count() or 1


def _np126_times_scalar(arr: np.ndarray, s: float) -> np.ndarray:
    """array * numpy-float64-scalar under numpy 1.26 value-based casting: a
    scalar of the same kind does not upcast a float32 array."""
    return arr * arr.dtype.type(s)


class NLSE:
    """2-D NLSE oracle (nlse.py:41-619)."""

    def __init__(
        self,
        alpha,
        power,
        window,
        n2,
        V,
        L,
        NX: int = 1024,
        NY: int = 1024,
        Isat=np.inf,
        nl_length: float = 0,
        wvl: float = 780e-9,
        mode: str = "c64",
        loss_saturates: bool = False,
        workers: int | None = None,
        fft: str = "ideal",
        threads: int = 1,
    ) -> None:
        if mode not in ("c64", "c128"):
            raise ValueError("mode must be 'c64' or 'c128'")
        if fft not in ("ideal", "pocketfft32"):
            raise ValueError("fft must be 'ideal' or 'pocketfft32'")
        self.mode = mode
        self.fft_kind = fft
        self.threads = max(1, int(threads))
        self.loss_saturates = loss_saturates
        self.workers = workers or _workers()
        # nlse.py:90-114
        self.n2 = n2
        self.V = V
        self.wl = wvl
        self.k = 2 * np.pi / self.wl
        self.L = L
        self.alpha = alpha
        self.power = power
        self.I_sat = Isat
        self.NX = NX
        self.NY = NY
        if isinstance(window, (float, int)):
            self.window = [window, window]
        else:
            self.window = window
        Dn = self.n2 * self.power / min(self.window) ** 2
        z_nl = 1 / (self.k * abs(Dn))
        if isinstance(z_nl, np.ndarray):
            z_nl = float(z_nl.min())
        self.delta_z = 5e-3 * z_nl
        # nlse.py:116-139: float32 coordinates, float64 step and K
        self.X, self.delta_X = np.linspace(
            -self.window[0] / 2, self.window[0] / 2, num=NX, endpoint=False,
            retstep=True, dtype=np.float32,
        )
        self.Y, self.delta_Y = np.linspace(
            -self.window[1] / 2, self.window[1] / 2, num=NY, endpoint=False,
            retstep=True, dtype=np.float32,
        )
        self._last_axes = (-2, -1)
        self.XX, self.YY = np.meshgrid(self.X, self.Y)
        self.Kx = 2 * np.pi * np.fft.fftfreq(self.NX, d=self.delta_X)
        self.Ky = 2 * np.pi * np.fft.fftfreq(self.NY, d=self.delta_Y)
        self.Kxx, self.Kyy = np.meshgrid(self.Kx, self.Ky)
        self.propagator = None
        self.nl_length = nl_length
        # nlse.py:143-155: K0 Bessel kernel on (6d+1)^2 pixels
        if self.nl_length > 0:
            self.nl_profile = nonlocal_profile(self.nl_length, self.delta_X)
        else:
            self.nl_profile = np.ones((self.NY, self.NX), dtype=np.float32)

    # ---------------------------------------------------------------- tables
    def _prop_cast(self, p: np.ndarray) -> np.ndarray:
        return p.astype(np.complex64) if self.mode == "c64" else p

    def _build_propagator(self) -> np.ndarray:
        """nlse.py:188-197 (complex64 in the reference)."""
        return self._prop_cast(
            np.exp(-1j * 0.5 * (self.Kxx**2 + self.Kyy**2) / self.k * self.delta_z)
        )

    # ---------------------------------------------------------------- arrays
    def _field_dtype(self, E_in: np.ndarray):
        return np.complex128 if self.mode == "c128" else E_in.dtype

    def _normalise(self, E: np.ndarray, target, with_ceps: bool, cell: float):
        """Normalisation of nlse.py:296-304 with numpy-1.26 dtypes."""
        rho = E.real * E.real + E.imag * E.imag
        integral = _np126_times_scalar(rho, cell).sum(axis=self._last_axes)
        if np.ndim(integral) == 0:
            # 0-d result: scalar * Python float promotes to float64 in 1.26
            integral = float(integral)
            if with_ceps:
                integral = integral * c * epsilon_0 / 2
            E_00 = (target / integral) ** 0.5
            return (E_00 * E).astype(E.dtype, copy=False)
        if with_ceps:
            integral = integral * c * epsilon_0 / 2
        target = np.asarray(target)
        if target.ndim > integral.ndim:  # parameter arrays shaped (B,1,1)
            target = target.reshape(integral.shape)
        E_00 = (target / integral) ** 0.5
        return (E_00.T * E.T).T.astype(E.dtype, copy=False)

    def _cell(self) -> float:
        return self.delta_X * self.delta_Y

    def _prepare_output_array(self, E_in: np.ndarray, normalize: bool):
        """nlse.py:257-307 (CPU branch)."""
        E = np.asarray(E_in).astype(self._field_dtype(E_in), copy=True)
        A_sq = np.zeros(E.shape, dtype=E.real.dtype)
        if normalize:
            A = self._normalise(E, self.power, True, self._cell())
        else:
            A = E
        return np.ascontiguousarray(A), A_sq

    # ----------------------------------------------------------------- FFTs
    def _fft(self, A: np.ndarray) -> None:
        A[...] = sfft.fftn(self._fft_in(A), axes=self._last_axes, workers=self.workers)

    def _ifft(self, A: np.ndarray) -> None:
        A[...] = sfft.ifftn(self._fft_in(A), axes=self._last_axes, workers=self.workers)

    def _fft_in(self, A: np.ndarray) -> np.ndarray:
        """"ideal": the transform is computed in FP64 and rounded once to the
        storage precision on assignment -- a correctly rounded complex64 FFT,
        which is what FFTW's single-precision plans approximate.
        "pocketfft32": scipy's own single-precision transform (its error
        carries a gain bias of about -1e-7 per forward+inverse pair, which
        accumulates linearly over a propagation)."""
        if self.fft_kind == "ideal" and A.dtype == np.complex64:
            return A.astype(np.complex128)
        return A

    def _convolution(self, A_sq: np.ndarray) -> np.ndarray:
        """oaconvolve(A_sq, nl_profile, mode='same', axes=last_axes)
        (nlse.py:186,406-409); FP64 result like scipy's float32 x float64.
        nonlocal_mode = "lorentzian" (no reference counterpart): the
        periodic convolution with spectrum 1/(1 + |k|^2 l^2), FP64 FFTs."""
        if getattr(self, "nonlocal_mode", "k0") == "lorentzian":
            Kx = 2 * np.pi * np.fft.fftfreq(self.NX, d=self.delta_X)
            Ky = 2 * np.pi * np.fft.fftfreq(self.NY, d=self.delta_Y)
            spec = 1.0 / (1.0 + (Kx[None, :] ** 2 + Ky[:, None] ** 2) * self.nl_length**2)
            return np.real(np.fft.ifft2(np.fft.fft2(np.asarray(A_sq, np.float64), axes=(-2, -1)) * spec,
                                        axes=(-2, -1)))
        kern = self.nl_profile
        while kern.ndim < A_sq.ndim:
            kern = kern[np.newaxis]
        return signal.oaconvolve(A_sq, kern, mode="same", axes=self._last_axes)

    # ------------------------------------------------------------ operators
    def _Vprime(self, V):
        """``self.k / 2 * V`` (nlse.py:425,464,484): float32 (or complex64)
        product for float32 V in the reference; FP64 in the c128 oracle."""
        if V is None:
            return None
        if self.mode == "c128":
            V = np.asarray(V).astype(
                np.complex128 if np.iscomplexobj(V) else np.float64)
        return self.k / 2 * V

    def _rows_parallel(self, fn, A, A_sq, V) -> None:
        """Apply an elementwise operator over chunks of rows on a thread pool
        (numpy releases the GIL), like numba's prange over the points."""
        ny = A.shape[-2] if A.ndim >= 2 else 1
        if self.threads <= 1 or A.ndim < 2 or ny < 2 * self.threads:
            fn(A, A_sq, V)
            return
        from concurrent.futures import ThreadPoolExecutor

        bounds = np.linspace(0, ny, self.threads + 1).astype(int)

        def part(i):
            sl = (Ellipsis, slice(bounds[i], bounds[i + 1]), slice(None))
            Vs = None if V is None or np.ndim(V) < 2 else V[sl]
            fn(A[sl], A_sq[sl], V if Vs is None and V is not None and np.ndim(V) < 2 else Vs)

        with ThreadPoolExecutor(self.threads) as ex:
            list(ex.map(part, range(self.threads)))

    def _nl(self, A, A_sq, V, dz) -> None:
        self._rows_parallel(lambda a, s, v: self._nl_serial(a, s, v, dz), A, A_sq, V)

    def _nl_serial(self, A, A_sq, V, dz) -> None:
        g = self.k / 2 * self.n2 * c * epsilon_0
        Isat = 2 * self.I_sat / (epsilon_0 * c)
        if V is None:
            kernels.nl_prop_without_V(A, A_sq, dz, self.alpha / 2, g, Isat,
                                      self.loss_saturates)
        else:
            kernels.nl_prop(A, A_sq, dz, self.alpha / 2, self._Vprime(V), g, Isat,
                            self.loss_saturates)

    def _intensity(self, A, A_sq):
        self._rows_parallel(lambda a, s, v: kernels.square_mod(a, s), A, A_sq, None)
        if self.nl_length > 0:
            A_sq[...] = self._convolution(A_sq)

    def split_step(self, A, A_sq, V, propagator, precision: str = "single") -> None:
        """nlse.py:372-487."""
        if precision == "double":
            self._intensity(A, A_sq)
            self._nl(A, A_sq, V, self.delta_z / 2)
        self._fft(A)
        np.multiply(A, propagator, out=A, casting="same_kind")
        self._ifft(A)
        self._intensity(A, A_sq)
        if precision == "double":
            self._nl(A, A_sq, V, self.delta_z / 2)
        else:
            self._nl(A, A_sq, V, self.delta_z)

    # ------------------------------------------------------------------ loop
    def out_field(
        self,
        E_in: np.ndarray,
        z: float,
        precision: str = "single",
        normalize: bool = True,
        callback: Union[list, Callable, None] = None,
        callback_args=(),
    ) -> np.ndarray:
        """nlse.py:489-619, CPU branch, without plotting / progress bar."""
        assert E_in.shape[self._last_axes[0]:] == self.XX.shape[self._last_axes[0]:], \
            "Shape mismatch"
        assert E_in.dtype in [np.complex64, np.complex128], \
            "Type mismatch, E_in should be complex64 or complex128"
        if self.propagator is None:
            self.propagator = self._build_propagator()
        V = None if self.V is None else np.array(self.V, copy=True)
        A, A_sq = self._prepare_output_array(E_in, normalize)
        n2_old = self.n2
        z_prop = 0
        i = 0
        while abs(z_prop) < z:
            if z > self.L:
                self.n2 = 0
            self.split_step(A, A_sq, V, self.propagator, precision)
            if callback is not None:
                if isinstance(callback, Callable):
                    callback(self, A, z, i, *callback_args)
                elif isinstance(callback, list) and isinstance(callback[0], Callable):
                    for cb, ca in zip(callback, callback_args):
                        cb(self, A, z, i, *ca)
                else:
                    raise ValueError(
                        "callbacks should be a callable or a list of callables")
            z_prop += self.delta_z
            i += 1
        self.n2 = n2_old
        return A


def nonlocal_profile(nl_length: float, delta_X: float) -> np.ndarray:
    """K0 Bessel nonlocal kernel (nlse.py:143-153): d = nl_length // dX whole
    pixels, support (6d+1)^2, centre set to the largest finite value,
    normalised to unit sum."""
    d = nl_length // delta_X
    x = np.arange(-3 * d, 3 * d + 1)
    y = np.arange(-3 * d, 3 * d + 1)
    XX, YY = np.meshgrid(x, y)
    R = np.hypot(XX, YY)
    prof = special.kn(0, R / d)
    prof[prof.shape[0] // 2, prof.shape[1] // 2] = np.nanmax(
        prof[np.logical_not(np.isinf(prof))])
    prof /= prof.sum()
    return prof


class NLSE_1d(NLSE):
    """1-D NLSE oracle (nlse_1d.py:15-105)."""

    def __init__(self, alpha, power, window, n2, V, L, NX: int = 1024, Isat=np.inf,
                 nl_length: float = 0, wvl: float = 780e-9, **kw) -> None:
        super().__init__(alpha=alpha, power=power, window=window, n2=n2, V=V, L=L,
                         NX=NX, Isat=Isat, nl_length=nl_length, wvl=wvl, **kw)
        self._last_axes = (-1,)
        # nlse_1d.py:65-66: edge row of the 2-D kernel, renormalised
        self.nl_profile = self.nl_profile[0]
        self.nl_profile /= self.nl_profile.sum()

    def _cell(self) -> float:
        return self.delta_X**2  # nlse_1d.py:89 (dX^2, quirk kept)

    def _prepare_output_array(self, E_in, normalize):
        """nlse_1d.py:68-96: the 1-D integral is an array op with ``*=``."""
        E = np.asarray(E_in).astype(self._field_dtype(E_in), copy=True)
        A_sq = np.zeros(E.shape, dtype=E.real.dtype)
        if normalize:
            A = self._normalise(E, self.power, True, self._cell())
        else:
            A = E
        return A, A_sq

    def _build_propagator(self) -> np.ndarray:
        """nlse_1d.py:98-105: complex128, no cast."""
        return np.exp(-1j * 0.5 * (self.Kx**2) / self.k * self.delta_z)


class NLSE_3d(NLSE):
    """3-D (x, y, t) NLSE oracle (nlse_3d.py:15-150); layout (NY, NX, NT)."""

    def __init__(self, alpha, energy, window, n2, D0, vg, V, L, NX: int = 1024,
                 NY: int = 1024, NZ: int = 1024, Isat=np.inf, nl_length: float = 0,
                 wvl: float = 780e-9, **kw) -> None:
        if len(window) == 2:
            window = [window[0], window[0], window[-1]]
        super().__init__(alpha=alpha, power=energy, window=window[0:2], n2=n2, V=V,
                         L=L, NX=NX, NY=NY, Isat=Isat, nl_length=nl_length, wvl=wvl,
                         **kw)
        self.energy = self.power
        self.NZ = NZ
        self.window_t = window[-1]
        self.power = self.energy / self.window_t
        Dn = self.n2 * self.power / min(self.window[0:2]) ** 2
        z_nl = 1 / (self.k * abs(Dn))
        if isinstance(z_nl, np.ndarray):
            z_nl = z_nl.min()
        self.delta_z = 0.5e-2 * z_nl
        # nlse_3d.py:95-98: endpoint-inclusive time grid (quirk Q7)
        self.T, self.delta_T = np.linspace(-self.window_t / 2, self.window_t / 2,
                                           self.NZ, retstep=True)
        self.omega = 2 * np.pi * np.fft.fftfreq(self.NZ, self.delta_T)
        self.D0 = D0
        self.vg = vg
        self.XX, self.YY, self.TT = np.meshgrid(self.X, self.Y, self.T)
        self.Kxx, self.Kyy, self.Omega = np.meshgrid(self.Kx, self.Ky, self.omega)
        self._last_axes = (-3, -2, -1)

    def _build_propagator(self) -> np.ndarray:
        """nlse_3d.py:105-114: complex64 2-D part times complex128 temporal
        part, which has no dz factor (quirk Q6)."""
        prop_2d = super()._build_propagator()
        prop_t = np.exp(-1j * self.D0 / 2 * self.Omega**2)
        return prop_2d * prop_t

    def _cell(self) -> float:
        return self.delta_X * self.delta_Y * self.delta_T

    def _prepare_output_array(self, E_in, normalize):
        """nlse_3d.py:116-150 (energy normalisation)."""
        E = np.asarray(E_in).astype(self._field_dtype(E_in), copy=True)
        A_sq = np.zeros(E.shape, dtype=E.real.dtype)
        if normalize:
            A = self._normalise(E, self.energy, True, self._cell())
        else:
            A = E
        return A, A_sq


class GPE(NLSE):
    """GPE oracle (gpe.py:15-119)."""

    def __init__(self, gamma, N, window, g, V, m=87 * atomic_mass, NX: int = 1024,
                 NY: int = 1024, sat=np.inf, nl_length: float = 0, **kw) -> None:
        super().__init__(alpha=gamma, power=N, window=window, n2=g, V=V, L=0, NX=NX,
                         NY=NY, Isat=sat, nl_length=nl_length, wvl=2 * np.pi / m, **kw)
        self.g = g
        self.V = V
        self.gamma = self.alpha
        self.N = self.power
        self.m = self.k
        self.sat = sat
        self.delta_t = self.delta_z
        self.I_sat *= epsilon_0 * c / 2

    def _build_propagator(self) -> np.ndarray:
        """gpe.py:81-90: uses delta_t frozen at construction (quirk Q11)."""
        return self._prop_cast(np.exp(
            -1j * 0.5 * hbar * (self.Kxx**2 + self.Kyy**2) / self.m * self.delta_t))

    def _prepare_output_array(self, E_in, normalize):
        """gpe.py:92-119: N normalisation without c*eps0/2."""
        E = np.asarray(E_in).astype(self._field_dtype(E_in), copy=True)
        A_sq = np.zeros(E.shape, dtype=E.real.dtype)
        if normalize:
            A = self._normalise(E, self.N, False, self._cell())
        else:
            A = E
        return A, A_sq


class CNLSE(NLSE):
    """Two-component NLSE oracle (cnlse.py:19-423); layout (..., 2, NY, NX)."""

    def __init__(self, alpha, power, window, n2, n12, V, L, NX: int = 1024,
                 NY: int = 1024, Isat=np.inf, nl_length: float = 0,
                 wvl: float = 780e-9, omega=None, **kw) -> None:
        super().__init__(alpha=alpha, power=power, window=window, n2=n2, V=V, L=L,
                         NX=NX, NY=NY, Isat=Isat, nl_length=nl_length, wvl=wvl, **kw)
        self.I_sat2 = self.I_sat
        self.n12 = n12
        self.n22 = self.n2
        self.omega = omega
        self.alpha2 = self.alpha
        self.k2 = self.k
        self.power2 = self.power

    def _prepare_output_array(self, E, normalize):
        """cnlse.py:99-148: per-component powers, complex power array."""
        E = np.asarray(E).astype(self._field_dtype(E), copy=True)
        A_sq = np.zeros(E.shape, dtype=E.real.dtype)
        if not normalize:
            return E, A_sq
        puiss = np.array([self.power, self.power2], dtype=E.dtype)
        rho = E.real * E.real + E.imag * E.imag
        integral = _np126_times_scalar(rho, self._cell()).sum(axis=self._last_axes)
        integral = integral * c * epsilon_0 / 2
        E_00 = (puiss / integral) ** 0.5
        A = np.empty_like(E)
        nd = len(self._last_axes)
        sl = (slice(None),) * nd
        na = (np.newaxis,) * nd
        A[(Ellipsis, 0) + sl] = E_00[(Ellipsis, 0) + na] * E[(Ellipsis, 0) + sl]
        A[(Ellipsis, 1) + sl] = E_00[(Ellipsis, 1) + na] * E[(Ellipsis, 1) + sl]
        return A, A_sq

    def _build_propagator(self) -> np.ndarray:
        """cnlse.py:179-190."""
        p1 = super()._build_propagator()
        p2 = self._prop_cast(np.exp(
            -1j * 0.5 * (self.Kxx**2 + self.Kyy**2) / self.k2 * self.delta_z))
        return np.array([p1, p2])

    def _take_components(self, A):
        """cnlse.py:192-202."""
        return A[..., 0, :, :], A[..., 1, :, :]

    def _conv_components(self, A_sq):
        I1, I2 = self._take_components(A_sq)
        if self.nl_length > 0:
            # cnlse.py:243-249: rebinding, so the FP64 result is not rounded
            I1 = self._convolution(I1)
            I2 = self._convolution(I2)
        return I1, I2

    def _nl_c(self, A1, A2, I1, I2, V, dz) -> None:
        """The two nl_prop[_without_V]_c calls of cnlse.py:251-298/374-421."""
        g11 = self.k / 2 * self.n2 * c * epsilon_0
        g12a = self.k / 2 * self.n12 * c * epsilon_0
        g22 = self.k2 / 2 * self.n22 * c * epsilon_0
        g12b = self.k2 / 2 * self.n12 * c * epsilon_0
        Is1 = 2 * self.I_sat / (epsilon_0 * c)
        Is2 = 2 * self.I_sat2 / (epsilon_0 * c)
        if V is None:
            kernels.nl_prop_without_V_c(A1, I1, I2, dz, self.alpha / 2, g11, g12a, Is1, Is2)
            kernels.nl_prop_without_V_c(A2, I2, I1, dz, self.alpha2 / 2, g22, g12b, Is2, Is1)
        else:
            kernels.nl_prop_c(A1, I1, I2, dz, self.alpha / 2, self._Vprime(V), g11, g12a,
                              Is1, Is2)
            V2 = self.k2 / 2 * (V if self.mode == "c64" else np.asarray(V, np.float64))
            kernels.nl_prop_c(A2, I2, I1, dz, self.alpha2 / 2, V2, g22, g12b, Is2, Is1)

    def split_step(self, A, A_sq, V, propagator, precision: str = "single") -> None:
        """cnlse.py:204-423."""
        A1, A2 = self._take_components(A)
        if precision == "double":
            kernels.square_mod(A, A_sq)
            I1, I2 = self._conv_components(A_sq)
            self._nl_c(A1, A2, I1, I2, V, self.delta_z / 2)
        self._fft(A)
        np.multiply(A, propagator, out=A, casting="same_kind")
        self._ifft(A)
        kernels.square_mod(A, A_sq)
        I1, I2 = self._conv_components(A_sq)
        if precision == "double":
            self._nl_c(A1, A2, I1, I2, V, self.delta_z / 2)
        else:
            self._nl_c(A1, A2, I1, I2, V, self.delta_z)
            if self.omega is not None:  # single mode only (quirk Q14)
                kernels.rabi_coupling(A1, A2, self.delta_z, self.omega / 2)


class CNLSE_1d(CNLSE):
    """1-D two-component NLSE oracle (cnlse_1d.py:13-131); layout (..., 2, NX)."""

    def __init__(self, alpha, power, window, n2, n12, V, L, NX: int = 1024,
                 Isat=np.inf, nl_length: float = 0, wvl: float = 780e-9, omega=None,
                 **kw) -> None:
        super().__init__(alpha=alpha, power=power, window=window, n2=n2, n12=n12, V=V,
                         L=L, NX=NX, Isat=Isat, nl_length=nl_length, wvl=wvl,
                         omega=omega, **kw)
        self._last_axes = (-1,)
        self.nl_profile = self.nl_profile[0]
        self.nl_profile /= self.nl_profile.sum()

    def _cell(self) -> float:
        return self.delta_X**2

    def _take_components(self, A):
        return A[..., 0, :], A[..., 1, :]

    def _build_propagator(self) -> np.ndarray:
        """cnlse_1d.py:119-131: complex128."""
        p1 = np.exp(-1j * 0.5 * (self.Kx**2) / self.k * self.delta_z)
        p2 = np.exp(-1j * 0.5 * (self.Kx**2) / self.k2 * self.delta_z)
        return np.array([p1, p2])


class DDGPE(CNLSE):
    """Driven-dissipative GPE oracle (ddgpe.py:14-565); layout (..., 2, NY, NX)
    with component 0 the exciton and 1 the cavity field."""

    def __init__(self, gamma, power, window, g, omega, T, omega_exc, omega_cav,
                 detuning, k_z, V=None, g12=0, NX: int = 1024, NY: int = 1024,
                 Isat=np.inf, nl_length: float = 0, **kw) -> None:
        super().__init__(alpha=gamma, power=power, window=window, n2=-g, n12=g12, V=V,
                         L=T, NX=NX, NY=NY, Isat=Isat, nl_length=nl_length, wvl=1e-30,
                         omega=omega, **kw)
        self.g = self.n2
        self.g12 = self.n12
        self.g2 = 0
        self.k_z = k_z
        self.gamma = gamma
        self.gamma2 = self.gamma
        self.omega_exc = omega_exc
        self.omega_cav = omega_cav
        self.detuning = detuning
        omega_lp = (omega_exc + omega_cav) / 2 - 0.5 * np.sqrt(
            (omega_exc - omega_cav) ** 2 + (omega) ** 2)
        self.omega_pump = omega_lp + detuning

    @staticmethod
    def laser_excitation(simu, A, t, i, F_pump_r, F_pump_t, F_probe_r, F_probe_t):
        """ddgpe.py:127-156: cavity drive, two separate subtractions."""
        A[..., 1, :, :] -= F_pump_r * F_pump_t[i] * simu.delta_z * 1j
        A[..., 1, :, :] -= F_probe_r * F_probe_t[i] * simu.delta_z * 1j

    def _build_propagator(self) -> np.ndarray:
        """ddgpe.py:204-223: exciton (flat) and cavity (polariton dispersion)."""
        p1 = self._prop_cast(np.exp(
            -1j * (self.omega_exc * (1 + 0 * self.Kxx**2) - self.omega_pump)
            * self.delta_z))
        p2 = self._prop_cast(np.exp(
            -1j * (self.omega_cav * np.sqrt(1 + (self.Kxx**2 + self.Kyy**2) / self.k_z**2)
                   - self.omega_pump) * self.delta_z))
        return np.array([p1, p2])

    def _prepare_output_array(self, E_in, normalize):
        """ddgpe.py:226-245: never normalises."""
        E = np.asarray(E_in).astype(self._field_dtype(E_in), copy=True)
        return E, np.zeros(E.shape, dtype=E.real.dtype)

    def _V(self, V):
        if V is None or self.mode == "c64":
            return V
        return np.asarray(V).astype(np.complex128 if np.iscomplexobj(V) else np.float64)

    def split_step(self, A, A_sq, V, propagator, precision: str = "single") -> None:
        """ddgpe.py:247-457, including the inconsistent double mode (Q15)."""
        A1, A2 = self._take_components(A)
        g, g12, Is1, Is2 = self.g, self.g12, self.I_sat, self.I_sat2
        dz = self.delta_z
        Vr = self._V(V)
        if precision == "double":
            kernels.square_mod(A, A_sq)
            I1, I2 = self._conv_components(A_sq)
            if V is None:
                kernels.nl_prop_without_V_c(A1, I1, I2, dz / 2, self.gamma / 2, g, g12, Is1, Is2)
                kernels.nl_prop_without_V_c(A2, I2, I1, dz / 2, self.gamma2 / 2, g, g12, Is1, Is2)
            else:
                kernels.nl_prop_c(A1, I1, I2, dz, self.gamma / 2, Vr, g, g12, Is1, Is2)
                kernels.nl_prop_c(A2, I2, I1, dz, self.gamma2 / 2, Vr, self.g2, g12, Is1, Is2)
        self._fft(A)
        np.multiply(A, propagator, out=A, casting="same_kind")
        self._ifft(A)
        kernels.square_mod(A, A_sq)
        I1, I2 = self._conv_components(A_sq)
        if precision == "double":
            if V is None:
                kernels.nl_prop_without_V_c(A1, I1, I2, dz / 2, self.gamma / 2, g, g12, Is1, Is2)
                kernels.nl_prop_without_V_c(A2, I2, I1, dz / 2, self.gamma2 / 2, g, g12, Is1, Is2)
            else:
                kernels.nl_prop_c(A1, I1, I2, dz, self.gamma / 2, self.k / 2 * Vr, g, g12,
                                  Is1, Is2)
                kernels.nl_prop_c(A2, I2, I1, dz, self.gamma2 / 2, Vr, self.g2, g12, Is1, Is2)
        else:
            if V is None:
                kernels.nl_prop_without_V_c(A1, I1, I2, dz, self.alpha / 2, g, g12, Is1, Is2)
                kernels.nl_prop_without_V_c(A2, I2, I1, dz, self.gamma2 / 2, self.g2, g12,
                                            Is1, Is2)
            else:
                kernels.nl_prop_c(A1, I1, I2, dz, self.gamma / 2, Vr, g, g12, Is1, Is2)
                kernels.nl_prop_c(A2, I2, I1, dz, self.gamma2 / 2, Vr, self.g2, g12, Is1, Is2)
            if self.omega is not None:
                kernels.rabi_coupling(A1, A2, dz, self.omega / 2)

    def out_field(self, E_in, t, laser_excitation, precision: str = "single",
                  callback=None, callback_args=None):
        """ddgpe.py:519-565: drive inserted as callback 0, normalize=False,
        returns None (quirk Q16); the caller's callback list is mutated."""
        if laser_excitation is None:
            callback.insert(0, self.laser_excitation)
        else:
            callback.insert(0, laser_excitation)
        self._last_field = NLSE.out_field(self, E_in, t, precision=precision,
                                          normalize=False, callback=callback,
                                          callback_args=callback_args)
        return None
