"""NLSE_1d, NLSE_3d, GPE, CNLSE, CNLSE_1d and DDGPE on the HIP engine against
the CPU oracle classes of the same names, plus the reference tests of those
classes re-expressed for the GPU backend."""

import numpy as np
import pytest
import torch
from scipy.constants import atomic_mass, c, epsilon_0

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

PARITY = 1e-5
waist = 2.23e-3
window = 4 * waist


def _noise(shape, seed, amp=0.01):
    rng = np.random.default_rng(seed)
    return 1 + amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


# ------------------------------------------------------------------ 1-D
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("N,z", [(256, 10e-3), (2048, 5e-4)])
def test_nlse_1d_matches_oracle(gpu, precision, N, z):
    from nlse_amd import NLSE_1d

    kw = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, V=None, L=10e-3, NX=N, Isat=1e5)
    sims = [NLSE_1d(**kw), oracle.NLSE_1d(**kw)]
    for s in sims:
        s.delta_z = 1e-5
        s.V = (-1e-4 * np.exp(-(s.X**2) / (70e-6) ** 2)).astype(np.float32)
    E0 = (np.exp(-(sims[0].X.astype(np.float64) ** 2) / waist**2) * _noise(N, 1)).astype(np.complex64)
    out = sims[0].out_field(E0, z, precision=precision, verbose=False)
    ref = sims[1].out_field(E0, z, precision=precision)
    err = rel_l2(out, ref)
    print(f"NLSE_1d N={N} steps={oracle.count_steps(z, 1e-5)} {precision}: {err:.2e}")
    assert out.shape == (N,) and err < PARITY


def test_nlse_1d_batch_and_norm(gpu):
    """Batched lines with per-line n2 equal independent runs; alpha = 0
    conserves the dX^2-normalised power (test_nlse_1d.py:123-134)."""
    from nlse_amd import NLSE_1d

    N, B = 512, 3
    n2 = np.array([-1.6e-9, -8e-10, -1e-10]).reshape(B, 1)
    sim = NLSE_1d(0, 1.05, window, n2, None, 1e-3, NX=N, Isat=1e5)
    sim.delta_z = 1e-5
    E0 = np.exp(-(sim.X.astype(np.float64) ** 2) / waist**2).astype(np.complex64)
    out = sim.out_field(np.stack([E0] * B), 1e-3, verbose=False)
    for b in range(B):
        o = oracle.NLSE_1d(0, 1.05, window, float(n2[b, 0]), None, 1e-3, NX=N, Isat=1e5)
        o.delta_z = 1e-5
        assert rel_l2(out[b], o.out_field(E0, 1e-3)) < PARITY
        norm = (np.abs(out[b].astype(np.complex128)) ** 2 * sim.delta_X**2).sum() * c * epsilon_0 / 2
        assert np.isclose(norm, 1.05, rtol=1e-4)


# ------------------------------------------------------------------ 3-D
def _nlse3d_pair(NX=64, NY=32, NT=16, **extra):
    kw = dict(alpha=0, energy=1.0 * 2e-12, window=[4e-3, 4e-3, 2e-12], n2=-1e-9, D0=-1e-26, vg=1.0, V=None,
              L=10e-3, NX=NX, NY=NY, NZ=NT, Isat=np.inf)
    kw.update(extra)
    from nlse_amd import NLSE_3d

    return NLSE_3d(**kw), oracle.NLSE_3d(**kw)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_nlse_3d_matches_oracle(gpu, precision):
    g, o = _nlse3d_pair(NX=64, NY=32, NT=16, alpha=20, Isat=1e5)
    for s in (g, o):
        s.delta_z = 1e-5
    pulse = (np.exp(-(g.XX.astype(np.float64) ** 2 + g.YY.astype(np.float64) ** 2) / (1e-3) ** 2)
             / np.cosh(g.TT / 200e-15))
    E0 = (pulse * _noise(pulse.shape, 3)).astype(np.complex64)
    z = 2e-4
    out = g.out_field(E0, z, precision=precision, verbose=False)
    ref = o.out_field(E0, z, precision=precision)
    err = rel_l2(out, ref)
    print(f"NLSE_3d {E0.shape} {precision}: {err:.2e}")
    assert out.shape == E0.shape and err < PARITY


def test_nlse_3d_identity_and_norm(gpu):
    """test_nlse3d.py:265-301 (delta_z = 0 double step on ones, asserted on
    the stepped field) and :306-336 (energy conservation at alpha = 0)."""
    g, _ = _nlse3d_pair(NX=32, NY=32, NT=16, alpha=20, Isat=1e5)
    g.delta_z = 0
    g.propagator = g._build_propagator()
    A, A_sq = g._prepare_output_array(np.ones((32, 32, 16), np.complex64), normalize=False)
    g.plans = g._build_fft_plan(A)
    g._send_arrays_to_gpu()
    g.split_step(A, A_sq, g.V, g.propagator, g.plans, precision="double")
    assert torch.allclose(A, torch.ones_like(A))
    g2, _ = _nlse3d_pair(NX=64, NY=64, NT=32)
    E = g2.out_field(np.ones((64, 64, 32), np.complex64), 20 * g2.delta_z, verbose=False)
    norm = np.sum(np.abs(E.astype(np.complex128)) ** 2 * g2.delta_X * g2.delta_Y * g2.delta_T) * c * epsilon_0 / 2
    assert np.isclose(norm, g2.energy, rtol=1e-4)


# ------------------------------------------------------------------ GPE
def test_gpe_matches_oracle_and_norm(gpu):
    """test_gpe.py:107-126 physics at 256^2 (delta_t frozen at init, loop in
    delta_z), against the oracle and with N conserved."""
    from nlse_amd import GPE

    N_at = 1e6
    kw = dict(gamma=0, N=N_at, window=1e-3, g=1e3 / (N_at / 1e-3**2), V=None, m=87 * atomic_mass, NX=256, NY=256)
    g, o = GPE(**kw), oracle.GPE(**kw)
    for s in (g, o):
        s.delta_t = 1e-8
        s.delta_z = 1e-8
    psi0 = (np.exp(-(g.XX**2 + g.YY**2) / (1e-4) ** 2) * _noise(g.XX.shape, 4)).astype(np.complex64)
    out = g.out_field(psi0, 1e-6, verbose=False)
    ref = o.out_field(psi0, 1e-6)
    assert rel_l2(out, ref) < PARITY
    norm = np.sum(np.abs(out.astype(np.complex128)) ** 2 * g.delta_X * g.delta_Y)
    assert np.isclose(norm, N_at, rtol=1e-4)


# ---------------------------------------------------------------- CNLSE
def _cnlse_pair(N=128, omega=None, V=True):
    from nlse_amd import CNLSE

    kw = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, n12=-1e-10, V=None, L=10e-3, NX=N, NY=N,
              Isat=1e5, omega=omega)
    out = []
    for cls in (CNLSE, oracle.CNLSE):
        s = cls(**kw)
        s.delta_z = 1e-5
        s.power2 = 0.3
        s.n22 = -4e-10
        s.k2 = 2 * np.pi / 795e-9
        s.I_sat2 = 2e5
        s.alpha2 = 5.0
        if V:
            s.V = (-1e-4 * np.exp(-(s.XX**2 + s.YY**2) / (70e-6) ** 2)).astype(np.float32)
        out.append(s)
    return out


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("omega", [None, 2e3])
def test_cnlse_matches_oracle(gpu, precision, omega):
    g, o = _cnlse_pair(128, omega)
    G = np.exp(-(g.XX.astype(np.float64) ** 2 + g.YY.astype(np.float64) ** 2) / waist**2)
    E0 = np.stack([G * _noise(G.shape, 5), G * _noise(G.shape, 6) * 0.5]).astype(np.complex64)
    z = 5e-4
    out = g.out_field(E0, z, precision=precision, verbose=False)
    ref = o.out_field(E0, z, precision=precision)
    err = rel_l2(out, ref)
    print(f"CNLSE omega={omega} {precision}: {err:.2e}")
    assert out.shape == E0.shape and err < PARITY


def test_cnlse_norm_and_components(gpu):
    """test_cnlse.py:188-217 (_take_components of a (3,2,N,N) batch) and
    :264-287 (per-component power conservation at alpha = 0)."""
    from nlse_amd import CNLSE

    N = 256
    simu = CNLSE(0, 1.05, window, -1.6e-9, -1e-10, None, 1e-3, NX=N, NY=N, Isat=1e5)
    A1, A2 = simu._take_components(np.ones((3, 2, N, N), np.complex64))
    assert A1.shape == A2.shape == (3, N, N)
    E = simu.out_field(np.ones((2, N, N), np.complex64), 1e-3, verbose=False)
    norm = np.sum(np.abs(E.astype(np.complex128)) ** 2 * simu.delta_X * simu.delta_Y * c * epsilon_0 / 2,
                  axis=(-2, -1))
    assert np.allclose(norm, [simu.power, simu.power2], rtol=1e-4)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_cnlse_1d_matches_oracle(gpu, precision):
    from nlse_amd import CNLSE_1d

    kw = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, n12=-1e-10, V=None, L=1e-3, NX=1024, Isat=1e5,
              omega=3e3)
    sims = [CNLSE_1d(**kw), oracle.CNLSE_1d(**kw)]
    for s in sims:
        s.delta_z = 1e-5
        s.power2 = 10e-3
        s.n22 = 1e-10
        s.k2 = 2 * np.pi / 795e-9
    X = sims[0].X.astype(np.float64)
    E0 = np.stack([np.exp(-X**2 / waist**2), np.exp(-X**2 / (70e-6) ** 2)]).astype(np.complex64)
    out = sims[0].out_field(E0, 1e-3, precision=precision, verbose=False)
    ref = sims[1].out_field(E0, 1e-3, precision=precision)
    assert rel_l2(out, ref) < PARITY


# ---------------------------------------------------------------- DDGPE
def _ddgpe_pair(N=128):
    from nlse_amd import DDGPE

    h = 0.654
    args = (0.07 / h, 0.17 / h / (1e-2 / h), 256, 1e-2 / h, 5.07 / h, 1, 1484.44 / h, 1482.76 / h, 0.17 / h, 27)
    out = []
    for cls in (DDGPE, oracle.DDGPE):
        s = cls(*args, NX=N, NY=N)
        s.delta_z = 0.1 / 32
        out.append(s)
    return out


def _ddgpe_drive(s, nsteps):
    time = np.arange(nsteps + 1, dtype=np.float32) * np.float32(s.delta_z)
    Fr = (0.3 * np.exp(-((s.XX**2 + s.YY**2) / 50**2))).astype(np.complex64)
    Ft = np.ones(time.shape, np.complex64)
    Ft[time < 2] = np.exp(-1 * (time[time < 2] - 2) ** 2 / 1.0**2)
    Fq = (0.05 * np.exp(-(((s.XX - 20) ** 2 + s.YY**2) / 10**2))).astype(np.complex64)
    Fqt = np.full(time.shape, 0.5 + 0.1j, np.complex64)
    return [Fr, Ft, Fq, Fqt]


@pytest.mark.parametrize("with_sampler", [False, True])
def test_ddgpe_matches_oracle(gpu, with_sampler):
    """Driven-dissipative run with pump and probe (the reference test runs
    with F = 0, test_ddgpe.py:313-377; here the drive is non-zero).  Without
    other callbacks the drive is fused into the row sweep; with a sampler
    callback each step returns to the host."""
    g, o = _ddgpe_pair(128)
    T = 1.0
    nsteps = oracle.count_steps(T, g.delta_z)
    drive = _ddgpe_drive(g, nsteps)
    E0 = np.zeros((2, 128, 128), np.complex64)
    samples = np.zeros(nsteps + 1, np.float64)

    def sampler(simu, A, z, i):
        samples[i] = float(torch.view_as_real(A[..., 1, :, :]).double().pow(2).sum())

    cb_g = [sampler] if with_sampler else []
    args_g = [drive, ()] if with_sampler else [drive]
    assert g.out_field(E0, T, g.laser_excitation, verbose=False, callback=cb_g, callback_args=args_g) is None
    assert cb_g[0] is type(g).laser_excitation  # the caller's list is mutated (Q16)
    o.out_field(E0, T, o.laser_excitation, callback=[], callback_args=[drive])
    out = g.field
    ref = o._last_field
    err = rel_l2(out, ref)
    print(f"DDGPE steps={nsteps} sampler={with_sampler}: {err:.2e}")
    assert np.abs(ref[1]).max() > 0 and err < PARITY
    if with_sampler:
        assert samples[nsteps - 1] > 0


@pytest.mark.parametrize("name,shape", [("NLSE_1d", (256,)), ("CNLSE", (2, 64, 64)), ("CNLSE_1d", (2, 256))])
def test_one_step_identity(gpu, name, shape):
    """test_nlse_1d.py:96-121, test_cnlse.py:220-259, test_cnlse_1d.py:116-150:
    delta_z = 0, a field of ones, one 'double' split step returns ones
    (asserted on the stepped field; the reference's CNLSE variants assert on
    the input)."""
    import nlse_amd

    kw = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, V=None, L=10e-3, Isat=1e5)
    if name.startswith("CNLSE"):
        kw["n12"] = -1e-10
    if name.endswith("_1d"):
        kw["NX"] = shape[-1]
    else:
        kw.update(NX=shape[-1], NY=shape[-2])
    simu = getattr(nlse_amd, name)(**kw)
    simu.delta_z = 0
    simu.propagator = simu._build_propagator()
    A, A_sq = simu._prepare_output_array(np.ones(shape, np.complex64), normalize=False)
    simu.plans = simu._build_fft_plan(A)
    simu._send_arrays_to_gpu()
    simu.split_step(A, A_sq, simu.V, simu.propagator, simu.plans, precision="double")
    simu._retrieve_arrays_from_gpu()
    assert torch.allclose(A, torch.ones_like(A))
