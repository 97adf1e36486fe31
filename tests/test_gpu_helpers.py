"""Adaptive step size and the initial-state / drive helpers on the GPU:
the nonlinear index change and its maximum (bit-exact against the oracle's
float32 restatement of callbacks.py:57-112), the adaptive loop (fused chunks
and step-by-step) against the oracle loop with the oracle callback, the
vortex imprint (kernels_gpu.py:159-176), Philox4x64-10 against
numpy.random.Philox and DDGPE.add_noise (ddgpe.py:96-125)."""

import numpy as np
import pytest
import torch

import oracle
import oracle.callbacks as OC
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _field(shape, seed, scale=1e4):
    rng = np.random.default_rng(seed)
    return (scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)


@pytest.mark.parametrize("isat", [1e5, np.inf])
def test_delta_n_bit_exact(gpu, isat):
    """Both forms of delta_n, per point and the maximum, equal the oracle's
    float32 arithmetic bit for bit; per-realisation n2 and I_sat."""
    from nlse_amd import kernels_hip as KH

    E = _field((3, 64, 96), 1)
    A = torch.from_numpy(E).cuda()
    n2 = -1.6e-9
    got = KH.delta_n(A, n2, isat).cpu().numpy()
    assert np.array_equal(got, OC.delta_n(E, n2, isat))
    m = KH.delta_n_max(A, n2, isat)
    assert m == OC.delta_n_adapt(E, n2, isat).max()
    n2s = np.array([-1e-9, -2e-9, 3e-9]).reshape(3, 1, 1)
    isats = np.array([1e5, 2e5, 5e4]).reshape(3, 1, 1)
    got = KH.delta_n(A, n2s, isats).cpu().numpy()
    for b in range(3):
        assert np.array_equal(got[b], OC.delta_n(E[b], n2s[b, 0, 0], isats[b, 0, 0]))
    want = max(OC.delta_n_adapt(E[b], n2s[b, 0, 0], isats[b, 0, 0]).max() for b in range(3))
    assert KH.delta_n_max(A, n2s, isats) == want


def _pair(N=128, precision="single"):
    from nlse_amd import NLSE

    kw = dict(alpha=20.0, power=1.05, window=4 * 2.23e-3, n2=-1.6e-9, V=None, L=10e-3, NX=N, NY=N, Isat=1e5)
    g = NLSE(**kw)
    o = oracle.NLSE(**kw)
    for s in (g, o):
        s.delta_z = 1e-5
    w = 2.23e-3
    G = np.exp(-(o.XX.astype(np.float64) ** 2 + o.YY.astype(np.float64) ** 2) / w**2)
    rng = np.random.default_rng(5)
    E0 = (G * (1 + 0.01 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)))).astype(np.complex64)
    return g, o, E0


@pytest.mark.parametrize("precision,every", [("single", 5), ("double", 3), ("single", 1), ("single", 40)])
def test_adaptive_loop_matches_oracle(gpu, precision, every):
    """adapt_delta_z as the only callback (fused chunks, graph replay for
    long chunks): same number of steps and step sizes as the oracle loop
    (rtol 1e-5) and the field within 1e-5."""
    from nlse_amd import callbacks

    g, o, E0 = _pair()
    z = 4e-3
    hg, ho = [], []
    out = g.out_field(E0, z, precision=precision, verbose=False, callback=callbacks.adapt_delta_z,
                      callback_args=(every, hg))
    ref = o.out_field(E0, z, precision=precision, callback=OC.adapt_delta_z, callback_args=(every, ho))
    print(f"steps {len(hg)} / {len(ho)}, last dz {hg[-1]:.4e} / {ho[-1]:.4e}, field {rel_l2(out, ref):.2e}")
    assert len(hg) == len(ho) and len(hg) > 10
    assert np.allclose(hg, ho, rtol=1e-5)
    assert g.delta_z == pytest.approx(o.delta_z, rel=1e-5)
    assert rel_l2(out, ref) < 1e-5


def test_adaptive_step_by_step_and_rebuild(gpu):
    """With a second callback the loop returns to the host every step and
    must agree with the fused chunks; adapt_rebuilds_propagator rebuilds the
    propagator for the new step size (checked against the oracle loop with
    a callback that does the same)."""
    from nlse_amd import callbacks

    g, o, E0 = _pair()
    z = 3e-3
    h1, h2, norms = [], [], np.zeros(10000)
    fused = g.out_field(E0, z, verbose=False, callback=callbacks.adapt_delta_z, callback_args=(4, h1))
    g2, _, _ = _pair()
    steps = g2.out_field(E0, z, verbose=False, callback=[callbacks.adapt_delta_z, callbacks.norm],
                         callback_args=[(4, h2), (1, norms)])
    assert len(h1) == len(h2) and np.allclose(h1, h2, rtol=1e-6)
    assert rel_l2(fused, steps) < 2e-6

    g3, o3, _ = _pair()
    g3.adapt_rebuilds_propagator = True
    h3, h4 = [], []

    def rebuild(simu, A, zz, i, every, hist):
        OC.adapt_delta_z(simu, A, zz, i, every, hist)
        if i % every == 0:
            simu.propagator = simu._build_propagator()

    out = g3.out_field(E0, z, verbose=False, callback=callbacks.adapt_delta_z, callback_args=(4, h3))
    ref = o3.out_field(E0, z, callback=rebuild, callback_args=(4, h4))
    assert len(h3) == len(h4) and np.allclose(h3, h4, rtol=1e-5)
    assert rel_l2(out, ref) < 1e-5
    assert rel_l2(out, fused) > 1e-4  # the rebuilt propagator changes the result


def test_vortex_matches_oracle(gpu):
    """Phase imprint on float32 and float64 images with integer and float
    coordinates: equal to the oracle modulo 2 pi (the core excluded, where
    the reference's phase is undefined for negative charges)."""
    from nlse_amd import kernels_hip as KH

    jj, ii = np.meshgrid(np.arange(96), np.arange(80))
    base = np.random.default_rng(2).uniform(-1, 1, (80, 96))
    for dt, ct, ll in [(np.float32, np.int64, 1), (np.float64, np.float64, 3), (np.float32, np.float32, -2)]:
        im = base.astype(dt)
        ref = im.copy()
        OC.vortex(ref, 40.0, 50.0, ii.astype(ct), jj.astype(ct), ll)
        t = torch.from_numpy(im.copy()).cuda()
        KH.vortex_cp(t, 40, 50, torch.from_numpy(ii.astype(ct)).cuda(), torch.from_numpy(jj.astype(ct)).cuda(), ll)
        got = t.cpu().numpy()
        d = np.angle(np.exp(1j * (got.astype(np.float64) - ref.astype(np.float64))))
        d[40, 50] = 0.0
        tol = 1e-6 if dt is np.float32 else 1e-12
        assert np.abs(d).max() < tol, (dt, ct, ll)


def test_philox_bit_exact_vs_numpy(gpu):
    from nlse_amd import kernels_hip as KH

    M = 2**64
    for ctr, key, n in [((0, 0, 0, 0), (0, 0), 1000), ((M - 3, 5, 0, 0), (12345, 678), 4096),
                        ((7, M - 1, M - 1, 2), (M - 1, M - 1), 17)]:
        got = KH.philox4x64(n, ctr, key).cpu().numpy().view(np.uint64)
        assert np.array_equal(got, OC.philox_raw(n, ctr, key)), (ctr, key)


def test_add_noise_matches_oracle_and_statistics(gpu):
    """DDGPE.add_noise: the device draw equals the oracle's Philox/Box-Muller
    restatement (1e-6 relative, FP64 transcendental ulps), is shared by
    every realisation, advances with the call counter, and has the
    reference's variance (noise^2 gamma/(4 dX dY) dz^2 per quadrature)."""
    from nlse_amd import DDGPE
    from nlse_amd import kernels_hip as KH

    E = _field((2, 2, 128, 128), 3, scale=1.0)
    key, amp0, amp1, dz = (99, 7), 3.0, 5.0, 0.25
    A = torch.from_numpy(E.copy()).cuda()
    KH.add_noise(A, dz, amp0, amp1, (0, 4, 0, 0), key)
    ref = E.copy()
    OC.add_noise(ref, dz, amp0, amp1, (0, 4, 0, 0), key)
    got = A.cpu().numpy()
    assert rel_l2(got - E, ref - E) < 1e-6
    d = (got - E).astype(np.complex128)
    assert np.allclose(d[0], d[1], atol=1e-6)
    assert abs(np.std(d[0, 0].real) / (amp0 * dz) - 1) < 0.02
    assert abs(np.std(d[0, 1].imag) / (amp1 * dz) - 1) < 0.02
    assert abs(np.corrcoef(d[0, 0].real.ravel(), d[0, 1].real.ravel())[0, 1]) < 0.02

    # the class method: counter advances per call, amplitudes from gamma
    sim = DDGPE(gamma=0.1, power=1.0, window=50e-6, g=1e-30, omega=1.0, T=1.0, omega_exc=1.0, omega_cav=1.0,
                detuning=0.0, k_z=1e7, NX=128, NY=128)
    sim.delta_z = 0.5
    B = torch.zeros((2, 128, 128), dtype=torch.complex64, device="cuda")
    DDGPE.add_noise(sim, B, 0.0, 0, noise=2.0)
    first = B.clone()
    DDGPE.add_noise(sim, B, 0.0, 1, noise=2.0)
    second = (B - first).cpu().numpy()
    first = first.cpu().numpy()
    amp = 2.0 * np.sqrt(sim.gamma / (4 * sim.delta_X * sim.delta_Y))
    assert abs(np.std(first[0].real) / (amp * sim.delta_z) - 1) < 0.03
    assert rel_l2(second, first) > 1.0  # fresh draw, not a repeat


@pytest.mark.parametrize("ny,nx,dz", [(2048, 2048, 5e-6), (128, 256, 1e-4), (64, 64, 3.3e-3)])
def test_propagator_table_equals_numpy(gpu, ny, nx, dz):
    """NLSE._build_propagator builds the dense table on the device: equal,
    value for value, to the reference's numpy expression (nlse.py:188-197)."""
    from nlse_amd import NLSE

    s = NLSE(alpha=20.0, power=1.05, window=[nx * 3.76e-6, ny * 3.76e-6], n2=-1.6e-9, V=None, L=1e-3, NX=nx,
             NY=ny)
    s.delta_z = dz
    got = s._build_propagator()
    want = np.exp(-1j * 0.5 * (s.Kxx**2 + s.Kyy**2) / s.k * s.delta_z).astype(np.complex64)
    assert got.dtype == np.complex64 and got.shape == (ny, nx)
    assert np.array_equal(got, want), np.count_nonzero(got != want)


def test_sumsq_diff_matches_fp64(gpu):
    """nlse_sumsq_diff: per-item FP64 sum of |A - B|^2 (the corrector's error
    and drift measure) against numpy in FP64; zero for equal fields."""
    from nlse_amd import kernels_hip as KH

    a = _field((3, 50, 70), 11, scale=1.0)
    b = (a + 1e-3 * _field((3, 50, 70), 12, scale=1.0)).astype(np.complex64)
    got = KH.sumsq_diff(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), 2).cpu().numpy()
    want = (np.abs(a.astype(np.complex128) - b.astype(np.complex128)) ** 2).sum(axis=(1, 2))
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    A = torch.from_numpy(a).cuda()
    assert np.all(KH.sumsq_diff(A, A.clone(), 2).cpu().numpy() == 0)


@pytest.mark.parametrize("isat,vkind,amp", [(np.inf, None, 1.0), (1e5, "real", 1.0), (1e5, "complex", 1.0),
                                            (1e5, "real", 1e-3)])
def test_hamiltonian_terms_match_oracle(gpu, isat, vkind, amp):
    """The corrector's Hamiltonian terms (FP32 transform + FP64 sums on the
    device, nlse_sumsq_weighted / nlse_nl_energy) against the oracle's FP64
    restatement; amp = 1e-3 puts every point on the small-I series branch."""
    from nlse_amd import NLSE
    from nlse_amd import corrector as CR
    from oracle import energy as OE

    N = 128
    kw = dict(alpha=20.0, power=1.05, window=4 * 2.23e-3, n2=-1.6e-9, V=None, L=1e-3, NX=N, NY=N, Isat=isat)
    g, o = NLSE(**kw), oracle.NLSE(**kw)
    if vkind is not None:
        rng = np.random.default_rng(7)
        V = -1e-4 * np.exp(-(o.XX**2 + o.YY**2) / (300e-6) ** 2)
        V = V.astype(np.float32) if vkind == "real" else (V + 1j * 1e-5 * rng.random(V.shape)).astype(np.complex64)
        g.V, o.V = V, V
    w = 2.23e-3
    rng = np.random.default_rng(8)
    E = (np.exp(-(o.XX**2 + o.YY**2) / w**2) * (1 + 0.05 * rng.standard_normal(o.XX.shape))).astype(np.complex64)
    A, _ = g._prepare_output_array(E, True)
    if amp != 1.0:
        A = (A * amp).contiguous()
    T, U = CR.hamiltonian_terms(g, A)
    To, Uo = OE.hamiltonian_terms(o, A.cpu().numpy())
    print(f"isat {isat} V {vkind} amp {amp}: T {T:.6e}/{To:.6e} U {U:.6e}/{Uo:.6e}")
    assert np.isclose(T, To, rtol=2e-6) and np.isclose(U, Uo, rtol=1e-9)
