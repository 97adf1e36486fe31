"""Pins of the CPU oracle (oracle/), CPU only.

The reference holds no golden vectors for this path (its tests check plan
types, a dz=0 identity and norm conservation), so each oracle function is
checked against what the mathematics fixes: closed forms, exact special
cases, independent library routines (numpy.fft, a DFT matrix,
scipy.signal.convolve) and invariants.  Each check is chosen so that a
plausible mistake -- a dropped term, a wrong sign, a wrong axis, a missing
saturation -- fails it.  The reference's own tests are re-expressed where
they pin something (propagator closed forms, normalisation, norm
conservation, the dz=0 identity).
"""

import numpy as np
import pytest
from scipy import signal, special
from scipy.constants import atomic_mass, c, epsilon_0, hbar

import oracle
import oracle.kernels as K

waist = 2.23e-3
window = 4 * waist


def small_nlse(N=64, **kw):
    args = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, V=None, L=10e-3, NX=N, NY=N, Isat=1e5)
    args.update(kw)
    return oracle.NLSE(**args)


# ----------------------------------------------------------- step counting
@pytest.mark.parametrize("z,dz,n", [(10e-3, 1e-5, 1001), (10e-3, 5e-6, 2001), (10e-3, 8e-5, 125),
                                    (2e-3, 1e-4, 20)])
def test_step_count_rule(z, dz, n):
    """The loop `while abs(z_prop) < z: z_prop += dz` (nlse.py:570,584): the
    float accumulation decides the count."""
    assert oracle.count_steps(z, dz) == n


# ------------------------------------------------------------ propagators
def test_propagator_2d_closed_form():
    """test_nlse.py:32-50, and complex64 storage."""
    s = small_nlse(128)
    p = s._build_propagator()
    assert p.dtype == np.complex64
    assert np.allclose(p, np.exp(-1j * 0.5 * (s.Kxx**2 + s.Kyy**2) / s.k * s.delta_z))
    # fftfreq ordering: the zero frequency is at [0, 0], the Nyquist row at N/2
    assert p[0, 0] == 1
    assert np.isclose(s.Kx[s.NX // 2], -np.pi / s.delta_X)


def test_propagator_1d_and_3d():
    """test_nlse_1d.py:28-37 (complex128) and test_nlse3d.py:34-59, which pins
    the temporal factor without dz (quirk kept by the reference)."""
    s1 = oracle.NLSE_1d(20, 1.05, window, -1.6e-9, None, 1e-3, NX=256, Isat=1e5)
    p1 = s1._build_propagator()
    assert p1.dtype == np.complex128
    assert np.allclose(p1, np.exp(-1j * 0.5 * (s1.Kx**2) / s1.k * s1.delta_z))
    s3 = oracle.NLSE_3d(alpha=20, energy=1.05 * 2e-6, window=np.array([4 * waist, 16e-6]), n2=-1.6e-9, D0=1e-16,
                        vg=0.1 * c, V=None, L=1e-2, NX=32, NY=16, NZ=8, Isat=1e5)
    p3 = s3._build_propagator()
    th = np.exp(-1j * 0.5 * (s3.Kxx**2 + s3.Kyy**2) / s3.k * s3.delta_z) * np.exp(-1j * s3.D0 / 2 * s3.Omega**2)
    assert p3.shape == (16, 32, 8) and p3.dtype == np.complex128
    assert np.allclose(p3, th)
    # the endpoint-inclusive time grid
    assert np.isclose(s3.delta_T, 16e-6 / (8 - 1))


def test_propagator_gpe_ddgpe_cnlse1d():
    g = oracle.GPE(gamma=0, N=1e6, window=1e-3, g=1e3 / (1e6 / 1e-3**2), V=None, m=87 * atomic_mass, NX=64, NY=64)
    assert np.allclose(g._build_propagator(),
                       np.exp(-1j * 0.5 * hbar * (g.Kxx**2 + g.Kyy**2) / g.m * g.delta_t))
    c1 = oracle.CNLSE_1d(20, 1.05, window, -1.6e-9, -1e-10, None, 1e-3, NX=128, Isat=1e5)
    p = c1._build_propagator()
    assert np.allclose(p, np.array([np.exp(-1j * 0.5 * c1.Kx**2 / c1.k * c1.delta_z),
                                    np.exp(-1j * 0.5 * c1.Kx**2 / c1.k2 * c1.delta_z)]))
    h = 0.654
    d = oracle.DDGPE(0, 0.17 / h / (1e-2 / h), 256, 1e-2 / h, 5.07 / h, 1, 1484.44 / h, 1482.76 / h, 0.17 / h, 27,
                     NX=32, NY=32)
    pd = d._build_propagator()
    assert np.allclose(pd[0], np.exp(-1j * (d.omega_exc - d.omega_pump) * d.delta_z))
    cav = np.exp(-1j * (d.omega_cav * np.sqrt(1 + (d.Kxx**2 + d.Kyy**2) / d.k_z**2) - d.omega_pump) * d.delta_z)
    assert np.allclose(pd[1], cav)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("fft", ["ideal", "pocketfft32"])
def test_fft_is_the_dft(fft):
    """Oracle transforms equal an explicit DFT matrix (tiny grid) with the
    reference's normalisation (forward unnormalised, inverse 1/N)."""
    s = small_nlse(8, fft=fft)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))).astype(np.complex64)
    n = np.arange(8)
    Fm = np.exp(-2j * np.pi * np.outer(n, n) / 8)
    ref = Fm @ x.astype(np.complex128) @ Fm.T
    a = x.copy()
    s._fft(a)
    assert np.allclose(a, ref, rtol=1e-6, atol=1e-5)
    s._ifft(a)
    assert np.allclose(a, x, atol=1e-6)


# ------------------------------------------------------------ operators
def test_nl_prop_phase_and_loss_closed_form():
    """For alpha = 0 the modulus is kept and the phase advances by
    dz*(g*I/(1+I/Isat) + V) (saturable Kerr); with alpha > 0 and I = 0 the
    amplitude decays by exp(-alpha dz) (CPU semantics: unsaturated loss)."""
    rng = np.random.default_rng(1)
    A = (1e4 * (rng.standard_normal(256) + 1j * rng.standard_normal(256))).astype(np.complex64)
    I = (A.real * A.real + A.imag * A.imag).astype(np.float32)
    V = (50 * rng.standard_normal(256)).astype(np.float32)
    dz, g, Isat = 1e-5, -2e-5, 1e8
    B = A.copy()
    K.nl_prop(B, I, dz, 0.0, V, g, Isat)
    # modulus in FP64 (a float32 np.abs adds its own 1-ulp rounding)
    assert np.allclose(np.abs(B.astype(np.complex128)), np.abs(A.astype(np.complex128)), rtol=2e-7)
    dphi = np.angle(B.astype(np.complex128) / A)
    expect = dz * (g * I.astype(np.float64) / (1 + I.astype(np.float64) / Isat) + V)
    assert np.allclose(dphi, expect, atol=1e-6)
    Z = np.ones(8, np.complex64)
    K.nl_prop_without_V(Z, np.zeros(8, np.float32), dz, 30.0, g, Isat)
    assert np.allclose(Z, np.exp(-30.0 * dz), rtol=1e-7)
    # GPU semantics multiply the loss by the saturation factor
    Zs = np.ones(8, np.complex64) * 1e4
    K.nl_prop_without_V(Zs, np.full(8, 1e8, np.float32), dz, 30.0, 0.0, 1e8, loss_saturates=True)
    assert np.allclose(np.abs(Zs), 1e4 * np.exp(-30.0 * dz * 0.5), rtol=1e-7)


def test_nl_prop_special_cases():
    rng = np.random.default_rng(2)
    A = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    I = np.abs(A).astype(np.float32) ** 2
    a1, a2 = A.copy(), A.copy()
    K.nl_prop(a1, I, 1e-3, 3.0, np.zeros(64, np.float32), 0.5, 10.0)
    K.nl_prop_without_V(a2, I, 1e-3, 3.0, 0.5, 10.0)
    assert np.array_equal(a1, a2)  # V = 0 == V = None
    b1, b2 = A.copy(), A.copy()
    K.nl_prop_without_V(b1, I, 1e-3, 0.0, 0.5, np.inf)
    assert np.allclose(np.angle(b1 / A), 1e-3 * 0.5 * I, atol=1e-6)  # Isat = inf -> sat = 1


def test_nl_prop_c_reduces_to_single_component():
    """With g12 = 0 and I2 = 0 the two-component step of component 1 is the
    single-component step with saturated loss."""
    rng = np.random.default_rng(3)
    A = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    I = np.abs(A).astype(np.float32) ** 2
    V = rng.standard_normal(64).astype(np.float32)
    a1, a2 = A.copy(), A.copy()
    K.nl_prop_c(a1, I, np.zeros(64, np.float32), 1e-3, 3.0, V, 0.7, 0.0, 5.0, 1.0)
    K.nl_prop(a2, I, 1e-3, 3.0, V, 0.7, 5.0, loss_saturates=True)
    assert np.array_equal(a1, a2)


def test_rabi_rotation():
    """omega dz = pi/2 swaps the components up to -i; two half steps equal
    one full step; |A1|^2 + |A2|^2 is conserved."""
    rng = np.random.default_rng(4)
    A1 = (rng.standard_normal(32) + 1j * rng.standard_normal(32)).astype(np.complex128)
    A2 = (rng.standard_normal(32) + 1j * rng.standard_normal(32)).astype(np.complex128)
    x1, x2 = A1.copy(), A2.copy()
    K.rabi_coupling(x1, x2, 1.0, np.pi / 2)
    assert np.allclose(x1, -1j * A2) and np.allclose(x2, -1j * A1)
    h1, h2 = A1.copy(), A2.copy()
    K.rabi_coupling(h1, h2, 0.5, 0.3)
    K.rabi_coupling(h1, h2, 0.5, 0.3)
    f1, f2 = A1.copy(), A2.copy()
    K.rabi_coupling(f1, f2, 1.0, 0.3)
    assert np.allclose(h1, f1) and np.allclose(h2, f2)
    assert np.allclose(np.abs(f1) ** 2 + np.abs(f2) ** 2, np.abs(A1) ** 2 + np.abs(A2) ** 2)


# -------------------------------------------------------------- nonlocal
def test_nonlocal_profile_and_convolution():
    """K0 kernel of nlse.py:143-153 and the zero-padded 'same' convolution,
    against scipy.special / scipy.signal.convolve."""
    s = small_nlse(64, nl_length=4 * (4 * waist / 64) + 1e-9)
    d = s.nl_length // s.delta_X
    assert d == 4 and s.nl_profile.shape == (25, 25)
    assert np.isclose(s.nl_profile.sum(), 1.0)
    assert np.allclose(s.nl_profile, s.nl_profile.T) and np.allclose(s.nl_profile, s.nl_profile[::-1])
    r = np.hypot(1, 0) / d
    assert np.isclose(s.nl_profile[12, 13] / s.nl_profile[12, 14], special.kn(0, r) / special.kn(0, 2 / d))
    rng = np.random.default_rng(5)
    I = rng.random((64, 64)).astype(np.float32)
    assert np.allclose(s._convolution(I), signal.convolve(I.astype(np.float64), s.nl_profile, mode="same"))
    s1 = oracle.NLSE_1d(20, 1.05, window, -1.6e-9, None, 1e-3, NX=64, nl_length=s.nl_length)
    assert s1.nl_profile.shape == (25,) and np.isclose(s1.nl_profile.sum(), 1)  # edge row, renormalised


# ------------------------------------------------------------ whole loop
def test_split_step_identity():
    """test_nlse.py:243-282: delta_z = 0, field of ones, 'double' step."""
    s = small_nlse(64)
    s.delta_z = 0
    s.propagator = s._build_propagator()
    A, A_sq = s._prepare_output_array(np.ones((64, 64), np.complex64), False)
    s.split_step(A, A_sq, None, s.propagator, "double")
    assert np.allclose(A, 1)


@pytest.mark.parametrize("mode", ["c64", "c128"])
def test_uniform_plane_wave_recursion(mode):
    """A uniform field only has the k = 0 mode, where the propagator is 1:
    the whole loop reduces to a scalar recursion of the nonlinear step."""
    N = 32
    s = small_nlse(N, alpha=8.0, mode=mode)
    s.delta_z = 2e-5
    s.V = None
    a0 = 3e4
    E = np.full((N, N), a0, np.complex64)
    out = s.out_field(E, 1e-3, normalize=False)
    g = s.k / 2 * s.n2 * c * epsilon_0
    isat = 2 * s.I_sat / (epsilon_0 * c)
    a = complex(a0)
    for _ in range(oracle.count_steps(1e-3, s.delta_z)):
        I = abs(a) ** 2
        a *= np.exp(s.delta_z * (-s.alpha / 2 + 1j * g * I / (1 + I / isat)))
    tol = 2e-6 if mode == "c64" else 1e-12
    assert np.allclose(out, a, rtol=tol)


def test_linear_diffraction_exact():
    """n2 = 0, alpha = 0, V = None: n steps equal IFFT(P^n FFT(A0))."""
    N = 64
    s = small_nlse(N, alpha=0, mode="c128")
    s.n2 = 0.0  # set after construction: the constructor divides by n2 (nlse.py:108-109)
    s.delta_z = 1e-4
    rng = np.random.default_rng(6)
    E0 = (np.exp(-(s.XX**2 + s.YY**2) / (1e-3) ** 2) * (1 + 0.1 * rng.standard_normal((N, N)))).astype(
        np.complex64)
    n = oracle.count_steps(5e-3, s.delta_z)
    out = s.out_field(E0, 5e-3, normalize=False)
    P = np.exp(-1j * 0.5 * (s.Kxx**2 + s.Kyy**2) / s.k * s.delta_z)
    ref = np.fft.ifft2(P**n * np.fft.fft2(E0.astype(np.complex128)))
    assert np.allclose(out, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


def test_linear_loss_decay():
    """Linear regime with loss: the power decays as exp(-alpha z) exactly
    (CPU semantics, loss -alpha/2 on the amplitude)."""
    s = small_nlse(64, alpha=20.0, mode="c128")
    s.n2 = 0.0
    s.delta_z = 1e-4
    E0 = np.exp(-(s.XX**2 + s.YY**2) / (1.5e-3) ** 2).astype(np.complex64)
    out = s.out_field(E0, 5e-3)
    n = oracle.count_steps(5e-3, s.delta_z)
    P = (np.abs(out) ** 2).sum() * s.delta_X * s.delta_Y * c * epsilon_0 / 2
    assert np.isclose(P, 1.05 * np.exp(-20.0 * s.delta_z * n), rtol=1e-10)


def test_norm_conservation():
    """test_nlse.py:287-311: alpha = 0 conserves the power (rtol 1e-4)."""
    s = small_nlse(128, alpha=0)
    out = s.out_field(np.ones((128, 128), np.complex64), 1e-3)
    norm = np.sum(np.abs(out.astype(np.complex128)) ** 2 * s.delta_X * s.delta_Y) * c * epsilon_0 / 2
    assert np.isclose(norm, s.power, rtol=1e-4)


def test_splitting_orders():
    """Global error of Lie ('single') is O(dz), of Strang ('double') O(dz^2),
    measured against a fine-step FP64 run on a tiny grid.  z is a power of two
    so that every z/n accumulates exactly (no extra step from the loop
    rule)."""
    N, z = 32, 2.0**-9
    rng = np.random.default_rng(7)

    def run(dz, precision):
        s = small_nlse(N, alpha=0, n2=-4e-8, mode="c128")
        s.delta_z = dz
        E0 = np.exp(-(s.XX**2 + s.YY**2) / (2.5e-3) ** 2) * (1 + 0.2 * np.cos(s.XX / 4e-4))
        return s.out_field(E0.astype(np.complex64), z, precision=precision)

    ref = run(z / 2048, "double")
    e_s = [np.linalg.norm(run(z / n, "single") - ref) for n in (64, 128)]
    e_d = [np.linalg.norm(run(z / n, "double") - ref) for n in (64, 128)]
    assert 1.7 < e_s[0] / e_s[1] < 2.3
    assert 3.4 < e_d[0] / e_d[1] < 4.6


def test_normalisation_variants():
    """_prepare_output_array of each class integrates to its power/energy/N
    (nlse.py:296-304, nlse_1d.py:86-93, nlse_3d.py:269-279, gpe.py:110-118,
    cnlse.py:124-145)."""
    rng = np.random.default_rng(8)
    s = small_nlse(32)
    E = (rng.random((32, 32)) + 1j * rng.random((32, 32))).astype(np.complex64)
    A, _ = s._prepare_output_array(E, True)
    assert np.isclose((np.abs(A) ** 2).sum() * s.delta_X * s.delta_Y * c * epsilon_0 / 2, s.power, rtol=1e-6)
    s1 = oracle.NLSE_1d(0, 2.0, window, -1e-9, None, 1e-3, NX=64)
    A1, _ = s1._prepare_output_array(E[0], True)
    assert np.isclose((np.abs(A1) ** 2).sum() * s1.delta_X**2 * c * epsilon_0 / 2, 2.0, rtol=1e-6)
    g = oracle.GPE(gamma=0, N=1e6, window=1e-3, g=1e-6, V=None, NX=32, NY=32)
    Ag, _ = g._prepare_output_array(E, True)
    assert np.isclose((np.abs(Ag) ** 2).sum() * g.delta_X * g.delta_Y, 1e6, rtol=1e-6)
    cn = oracle.CNLSE(0, 1.05, window, -1e-9, -1e-10, None, 1e-3, NX=32, NY=32)
    cn.power2 = 0.3
    Ac, _ = cn._prepare_output_array(np.stack([E, E]), True)
    pw = (np.abs(Ac) ** 2).sum(axis=(-2, -1)) * cn.delta_X * cn.delta_Y * c * epsilon_0 / 2
    assert np.allclose(pw, [1.05, 0.3], rtol=1e-6)


def test_z_beyond_L_zeroes_n2():
    """nlse.py:571-572: z > L switches the nonlinearity off for the whole run
    (and n2 is restored afterwards)."""
    s = small_nlse(32, L=1e-4)
    s.delta_z = 1e-5
    E0 = np.exp(-(s.XX**2 + s.YY**2) / (2e-3) ** 2).astype(np.complex64)
    out = s.out_field(E0, 2e-4)
    assert s.n2 == -1.6e-9
    lin = small_nlse(32)
    lin.n2 = 0.0
    lin.delta_z = 1e-5
    assert np.allclose(out, lin.out_field(E0, 2e-4), rtol=1e-6, atol=1e-6 * np.abs(out).max())


def test_lorentzian_response_closed_form():
    """The oracle's Fourier-space Lorentzian response maps 1 + cos(k0 x) to
    1 + cos(k0 x) / (1 + k0^2 l^2) (a single mode is an eigenfunction)."""
    import numpy as np

    import oracle

    o = oracle.NLSE(0, 1.0, 4e-3, -1e-9, None, 1e-3, NX=64, NY=32, nl_length=30e-6)
    o.nonlocal_mode = "lorentzian"
    k0 = 2 * np.pi * 5 / (64 * o.delta_X)
    I = 1 + np.cos(k0 * np.arange(64) * o.delta_X)[None, :] * np.ones((32, 1))
    out = o._convolution(I)
    expect = 1 + np.cos(k0 * np.arange(64) * o.delta_X)[None, :] / (1 + k0**2 * 30e-6**2)
    assert np.abs(out - expect).max() < 1e-12


def test_delta_n_closed_form_and_precedence():
    """oracle.callbacks: a constant field gives |n2| q/(1 + q/I_sat) with
    q = re^2 + im^2 c eps0/2 (adapt) or re^2 + im^2 (evaluate), and the
    adaptive update dz = 1/(12 k delta_n); dropping the precedence quirk
    changes the value by orders of magnitude (so the pin separates them)."""
    import oracle.callbacks as CB

    a, b, n2, isat = 3e3, 4e3, -1.6e-9, 1e5
    A = np.full((8, 8), a + 1j * b, np.complex64)
    q_adapt = a * a + b * b * c * epsilon_0 / 2
    q_eval = a * a + b * b
    want_adapt = abs(n2) * q_adapt / (1 + q_adapt / isat)
    want_eval = c * epsilon_0 / 2 * n2 * q_eval / (1 + q_eval / isat)
    assert np.allclose(CB.delta_n_adapt(A, n2, isat), want_adapt, rtol=2e-6)
    assert np.allclose(CB.delta_n(A, n2, isat), want_eval, rtol=2e-6)
    assert abs(want_adapt - abs(n2) * q_eval / (1 + q_eval / isat)) > 1e-3 * want_adapt
    assert np.allclose(CB.delta_n_adapt(A, n2, np.inf), abs(n2) * q_adapt, rtol=2e-6)

    class S:
        pass

    s = S()
    s.n2, s.I_sat, s.k, s.delta_z = n2, isat, 2 * np.pi / 780e-9, 1e-5
    hist = []
    for i in range(3):
        CB.adapt_delta_z(s, A, 1.0, i, 2, hist)
    assert hist[0] == 1e-5 and hist[1] == hist[2]
    assert np.isclose(s.delta_z, 1 / (s.k * want_adapt) / 12, rtol=2e-6)


def test_vortex_closed_form_and_winding():
    """oracle.callbacks.vortex: charge 1 is atan2(y, x); along a loop around
    the core the phase winds by 2 pi * charge."""
    import oracle.callbacks as CB

    jj, ii = np.meshgrid(np.arange(64), np.arange(48))
    im = np.zeros((48, 64))
    CB.vortex(im, 20.0, 30.0, ii, jj, 1)
    assert np.allclose(im, np.arctan2(jj - 30.0, ii - 20.0), atol=1e-12)
    for ll in (-2, 3):
        im = np.zeros((48, 64))
        CB.vortex(im, 20.0, 30.0, ii, jj, ll)
        loop = [(20 + 5, 30 + t) for t in range(-5, 5)] + [(20 - t, 35) for t in range(-5, 5)] + \
               [(15, 30 - t) for t in range(-5, 5)] + [(20 + t, 25) for t in range(-5, 5)]
        ph = np.array([im[p] for p in loop + loop[:1]])
        d = np.angle(np.exp(1j * np.diff(ph)))
        assert round(d.sum() / (2 * np.pi)) == ll


def test_philox_and_box_muller():
    """oracle.callbacks.philox_raw is numpy.random.Philox; an independent
    big-integer restatement of Philox4x64-10 (Salmon et al. 2011) agrees on
    counters with carries; Box-Muller normals have unit variance."""
    import oracle.callbacks as CB

    M = 2**64

    def block(ctr, k, b):
        v = (sum(x << (64 * i) for i, x in enumerate(ctr)) + b + 1) % 2**256
        x = [(v >> (64 * i)) % M for i in range(4)]
        k0, k1 = k
        for r in range(10):
            if r:
                k0, k1 = (k0 + 0x9E3779B97F4A7C15) % M, (k1 + 0xBB67AE8584CAA73B) % M
            p0, p1 = 0xD2E7470EE14C6C93 * x[0], 0xCA5A826395121157 * x[2]
            x = [(p1 >> 64) ^ x[1] ^ k0, p1 % M, (p0 >> 64) ^ x[3] ^ k1, p0 % M]
        return x

    for ctr, key in [((0, 0, 0, 0), (0, 0)), ((M - 2, M - 1, 3, 0), (M - 1, 1))]:
        raw = CB.philox_raw(3, ctr, key)
        assert [int(w) for w in raw] == [w for b in range(3) for w in block(ctr, key, b)]
    w = CB.philox_raw(50000, (1, 2, 0, 0), (7, 9)).reshape(-1, 4)
    z0, z1 = CB.box_muller(w[:, 0], w[:, 1])
    z = np.concatenate([z0, z1])
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.02 and abs(np.mean(z0 * z1)) < 0.01


# ------------------------------------------------------------ Hamiltonian
def test_hamiltonian_plane_wave_closed_form():
    """A plane wave a exp(i kx0 x) (grid wavenumber), Kerr, no V:
    T = kx0^2 a^2 area / 2k, U = g a^4 area / 2."""
    from oracle import energy as OE

    s = oracle.NLSE(0, 1.05, window, -1.6e-9, None, 1e-3, NX=64, NY=32)
    a, kx0 = 3.0, s.Kx[5]
    A = (a * np.exp(1j * kx0 * s.XX)).astype(np.complex128)
    T, U = OE.hamiltonian_terms(s, A)
    area = s.NX * s.NY * s.delta_X * s.delta_Y
    g = s.k / 2 * s.n2 * c * epsilon_0
    assert np.isclose(T, kx0**2 * a**2 * area / (2 * s.k), rtol=1e-12)
    assert np.isclose(U, g * a**4 * area / 2, rtol=1e-12)


def test_hamiltonian_potential_is_quadrature():
    """G(I) equals the integral of the phase rate g s/(1 + s/Is) from 0 to
    I, on both sides of the series switch (I/Is = 1e-4)."""
    from scipy.integrate import quad

    from oracle import energy as OE

    g, Is = -2.5, 7.0
    for I in (7e-7, 6e-5 * 7, 1e-3 * 7, 0.5 * 7, 30.0):
        want = quad(lambda u: g * u / (1 + u / Is), 0, I, epsabs=0, epsrel=1e-13)[0]
        assert np.isclose(OE.G(I, g, Is), want, rtol=1e-11, atol=0), I
    assert np.isclose(OE.G(2.0, g, np.inf), g * 2.0, rtol=1e-15)


def test_hamiltonian_conserved_by_lossless_propagation():
    """alpha = 0, saturable, real V: the oracle's own split steps keep H
    within 1e-4 of |T| + |U| while the kinetic part and the V part each
    move by more than 1e-3 of it (a wrong sign or factor in any term
    breaks the balance)."""
    from oracle import energy as OE

    N = 128
    s = oracle.NLSE(alpha=0, power=1.05, window=window, n2=-1.6e-9, V=None, L=1e-2, NX=N, NY=N, Isat=1e5,
                    mode="c128")
    s.V = (-3e-4 * np.exp(-(s.XX**2 + s.YY**2) / (400e-6) ** 2)).astype(np.float32)
    s.delta_z = 2e-5
    rng = np.random.default_rng(0)
    E0 = (np.exp(-(s.XX**2 + s.YY**2) / waist**2) * (1 + 0.05 * rng.standard_normal(s.XX.shape))).astype(
        np.complex64)
    A0, _ = s._prepare_output_array(E0, True)
    out = s.out_field(E0, 5e-3)
    T0, U0 = OE.hamiltonian_terms(s, A0)
    T1, U1 = OE.hamiltonian_terms(s, out)
    scale = abs(T0) + abs(U0)
    cell = s.delta_X * s.delta_Y
    uv0 = cell * (s.k / 2 * s.V * np.abs(A0) ** 2).sum()
    uv1 = cell * (s.k / 2 * s.V * np.abs(out) ** 2).sum()
    assert abs((T1 - U1) - (T0 - U0)) / scale < 1e-4
    assert abs(T1 - T0) / scale > 1e-3 and abs(uv1 - uv0) / scale > 1e-3
