"""2-D NLSE on the HIP engine against the CPU oracle (reference CPU path
restated), plus the reference's own NLSE tests re-expressed for the GPU
backend (tests/test_nlse.py of the reference)."""

import numpy as np
import pytest
import torch
from scipy.constants import c, epsilon_0

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# Parity bound: relative L2 <= 1e-5 against the complex64 reference-CPU
# oracle.
PARITY = 1e-5

waist = 2.23e-3
window = 4 * waist


def readme_nlse(cls, N, mode=None, **extra):
    kw = dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, V=None, L=10e-3, NX=N, NY=N, Isat=1e5,
              wvl=780e-9)
    kw.update(extra)
    if mode is not None:
        s = cls(mode=mode, **kw) if mode != "pocket" else cls(mode="c64", fft="pocketfft32", **kw)
    else:
        s = cls(**kw)
    s.delta_z = 1e-5
    s.V = (-1e-4 * np.exp(-(s.XX**2 + s.YY**2) / (70e-6) ** 2)).astype(np.float32)
    return s


def readme_field(s, seed=0):
    rng = np.random.default_rng(seed)
    G = np.exp(-(s.XX.astype(np.float64) ** 2 + s.YY.astype(np.float64) ** 2) / waist**2)
    nr = rng.standard_normal(G.shape)
    ni = rng.standard_normal(G.shape)
    return (G * (1 + 0.01 * (nr + 1j * ni))).astype(np.complex64)


@pytest.mark.parametrize("N,z", [(128, 5e-4), (256, 3e-4)])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_out_field_matches_oracle(gpu, N, z, precision):
    from nlse_amd import NLSE

    sim = readme_nlse(NLSE, N)
    E0 = readme_field(sim)
    out = sim.out_field(E0, z, precision=precision, verbose=False)
    ref = readme_nlse(oracle.NLSE, N, mode="c64").out_field(E0, z, precision=precision)
    err = rel_l2(out, ref)
    err64 = rel_l2(out, readme_nlse(oracle.NLSE, N, mode="c128").out_field(E0, z, precision=precision))
    print(f"N={N} steps={oracle.count_steps(z, 1e-5)} {precision}: vs c64 oracle {err:.2e}, vs FP64 {err64:.2e}")
    assert out.shape == E0.shape and out.dtype == np.complex64
    assert err < PARITY


def test_loss_semantics_switch(gpu):
    """loss_saturates=True follows the reference GPU kernels (kernels_gpu.py:30)."""
    from nlse_amd import NLSE

    N, z = 128, 2e-4
    sim = readme_nlse(NLSE, N)
    sim.loss_saturates = True
    E0 = readme_field(sim)
    out = sim.out_field(E0, z, verbose=False)
    ref_o = readme_nlse(oracle.NLSE, N, mode="c64")
    ref_o.loss_saturates = True
    ref = ref_o.out_field(E0, z)
    assert rel_l2(out, ref) < PARITY
    cpu_sem = readme_nlse(oracle.NLSE, N, mode="c64").out_field(E0, z)
    assert rel_l2(out, cpu_sem) > 10 * PARITY  # the two semantics do differ


def test_broadcasting_parameters(gpu):
    """(B,1,1) parameters over a (B,NY,NX) field (README.md:147-159,
    tests/test_broadcasting.py) equal B independent runs."""
    from nlse_amd import NLSE

    N, z, B = 64, 1e-4, 3
    n2 = np.linspace(-1.6e-9, -1e-10, B).reshape(B, 1, 1)
    alpha = np.array([20.0, 10.0, 0.0]).reshape(B, 1, 1)
    sim = NLSE(alpha, 1.05, window, n2, None, 10e-3, NX=N, NY=N, Isat=1e5)
    sim.delta_z = 1e-5
    E0 = readme_field(sim)
    Eb = np.stack([E0] * B)
    out = sim.out_field(Eb, z, verbose=False)
    assert out.shape == Eb.shape
    for b in range(B):
        o = oracle.NLSE(float(alpha[b, 0, 0]), 1.05, window, float(n2[b, 0, 0]), None, 10e-3, NX=N, NY=N,
                        Isat=1e5)
        o.delta_z = 1e-5
        assert rel_l2(out[b], o.out_field(E0, z)) < PARITY


def test_split_step_identity(gpu):
    """Reference test_nlse.py:243-282: delta_z = 0, field of ones, one
    'double' step returns ones."""
    from nlse_amd import NLSE

    N = 256
    simu = NLSE(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    simu.delta_z = 0
    simu.propagator = simu._build_propagator()
    E = np.ones((N, N), dtype=np.complex64)
    A, A_sq = simu._prepare_output_array(E, normalize=False)
    simu.plans = simu._build_fft_plan(A)
    simu._send_arrays_to_gpu()
    simu.split_step(A, A_sq, simu.V, simu.propagator, simu.plans, precision="double")
    assert torch.allclose(A, torch.ones((N, N), dtype=torch.complex64, device=A.device))


def test_prepare_output_array(gpu):
    """Reference test_nlse.py:102-175: normalisation to the power, shape kept."""
    from nlse_amd import NLSE

    N = 512
    simu = NLSE(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    rng = np.random.default_rng(11)
    A = (rng.random((N, N)) + 1j * rng.random((N, N))).astype(np.complex64)
    out, out_sq = simu._prepare_output_array(A, normalize=True)
    assert out.is_contiguous() and out_sq.is_contiguous()
    o = out.cpu().numpy().astype(np.complex128)
    integral = (np.abs(o) ** 2 * simu.delta_X * simu.delta_Y).sum() * c * epsilon_0 / 2
    assert np.allclose(integral, simu.power)
    assert out.shape == (N, N)
    o /= np.abs(o).max()
    assert np.allclose(o, A / np.abs(A).max())


def test_build_propagator(gpu):
    """Reference test_nlse.py:32-50 and the separable factors used by the
    fused column pass (complex64 of the FP64 product)."""
    from nlse_amd import NLSE

    N = 2048
    simu = NLSE(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    prop = simu._build_propagator()
    assert np.allclose(prop, np.exp(-1j * 0.5 * (simu.Kxx**2 + simu.Kyy**2) / simu.k * simu.delta_z))
    px, py = simu._separable_factors()
    sep = (py[0][:, None] * px[0][None, :]).astype(np.complex64)
    mism = np.count_nonzero(sep != prop)
    assert mism <= prop.size * 1e-6, mism


def test_out_field_norm_conservation(gpu):
    """Reference test_nlse.py:287-311: alpha = 0 conserves the power (rtol
    1e-4) over L = 10 mm at the default step (341 steps), N = 2048."""
    from nlse_amd import NLSE

    N = 2048
    E = np.ones((N, N), dtype=np.complex64)
    simu = NLSE(0, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    E = simu.out_field(E, 10e-3, verbose=False, plot=False, precision="single")
    norm = np.sum(np.abs(E.astype(np.complex128)) ** 2 * simu.delta_X * simu.delta_Y) * c * epsilon_0 / 2
    assert E.shape == (N, N)
    assert np.allclose(norm, simu.power, rtol=1e-4)


def test_callbacks_and_step_rule(gpu):
    """Callbacks get (simu, A, z, i, *args) once per step; the step count
    follows the float accumulation rule (1001 steps for 10 mm at 1e-5)."""
    from nlse_amd import NLSE, callbacks

    N = 64
    simu = NLSE(0, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    simu.delta_z = 1e-5
    seen = []
    norms = np.zeros(1001)

    def cb(s, A, z, i):
        seen.append((z, i))

    E0 = np.ones((N, N), np.complex64)
    out = simu.out_field(E0, 10e-3, verbose=False, callback=[cb, callbacks.norm],
                         callback_args=[(), (1, norms)])
    assert len(seen) == 1001 and seen[-1] == (10e-3, 1000)
    assert np.allclose(norms, norms[0], rtol=1e-5)
    ref = simu.out_field(E0, 10e-3, verbose=False)
    assert rel_l2(out, ref) < 1e-6


@pytest.mark.parametrize("precision", ["single", "double"])
def test_full_length_parity(gpu, precision):
    """README-example physics over the full L = 10 mm (1001 steps) at 128^2: the
    accumulated error stays within the parity bound against the complex64
    oracle with a correctly rounded FFT; the distance to the FP64 oracle and
    to scipy's own single-precision FFT (whose gain bias accumulates) is
    reported."""
    from nlse_amd import NLSE

    N, z = 128, 10e-3
    sim = readme_nlse(NLSE, N)
    E0 = readme_field(sim)
    out = sim.out_field(E0, z, precision=precision, verbose=False)
    ref = readme_nlse(oracle.NLSE, N, mode="c64").out_field(E0, z, precision=precision)
    ref64 = readme_nlse(oracle.NLSE, N, mode="c128").out_field(E0, z, precision=precision)
    pocket = readme_nlse(oracle.NLSE, N, mode="pocket").out_field(E0, z, precision=precision)
    err, err64, errp = rel_l2(out, ref), rel_l2(out, ref64), rel_l2(out, pocket)
    print(f"1001 steps {precision}: vs c64 oracle {err:.2e}, vs FP64 oracle {err64:.2e} "
          f"(c64 oracle vs FP64 {rel_l2(ref, ref64):.2e}), vs pocketfft32 c64 {errp:.2e}")
    assert err < PARITY


def test_full_length_parity_1024(gpu):
    """README-example physics over the full L = 10 mm (1001 steps) at 1024^2, where
    the row sweep runs with 8 points per thread: the accumulated error stays
    within the parity bound against the complex64 oracle."""
    from nlse_amd import NLSE

    N, z = 1024, 10e-3
    sim = readme_nlse(NLSE, N)
    E0 = readme_field(sim)
    out = sim.out_field(E0, z, verbose=False)
    ref = readme_nlse(oracle.NLSE, N, mode="c64").out_field(E0, z)
    err = rel_l2(out, ref)
    print(f"1001 steps at 1024^2: vs c64 oracle {err:.2e}")
    assert err < PARITY
