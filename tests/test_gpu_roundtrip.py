"""Host-side API of the reference tests: parameter host<->device round
trips (test_nlse.py:178-240, test_cnlse.py:91-185, test_ddgpe.py:85-200):
after _send_arrays_to_gpu the propagator, V and the per-realisation array
parameters are device tensors, after _retrieve_arrays_from_gpu numpy arrays
again, with unchanged values; an out_field call leaves the instance's own
propagator as a host array (the engine applies its separable factors and
never moves the dense table); FFT plan type and shape (test_nlse.py:53-99,
test_nlse3d.py:62-103); _prepare_output_array of every class against the
oracle class."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

N = 128
window = 4 * 2.23e-3


def _dev(x) -> bool:
    return isinstance(x, torch.Tensor) and x.is_cuda


def _per_item(v, extra_axes: int):
    return np.repeat(v, 2).reshape((2,) + (1,) * extra_axes)


def _check_round_trip(simu, names):
    simu.propagator = simu._build_propagator()
    before = {n: np.array(getattr(simu, n)) for n in names + ["propagator", "V"]}
    simu._send_arrays_to_gpu()
    for n in names + ["propagator", "V"]:
        assert _dev(getattr(simu, n)), f"{n} is not a device tensor after _send_arrays_to_gpu"
    simu._retrieve_arrays_from_gpu()
    for n in names + ["propagator", "V"]:
        v = getattr(simu, n)
        assert isinstance(v, np.ndarray), f"{n} is not a numpy array after _retrieve_arrays_from_gpu"
        assert np.array_equal(v, before[n]), n


def test_nlse_round_trip(gpu):
    from nlse_amd import NLSE

    rng = np.random.default_rng(1)
    V = rng.random((N, N)) + 1j * rng.random((N, N))
    simu = NLSE(_per_item(20, 2), 1.05, window, _per_item(-1.6e-9, 2), V, 10e-3, NX=N, NY=N,
                Isat=_per_item(10e4, 2))
    _check_round_trip(simu, ["alpha", "n2", "I_sat"])


def test_cnlse_round_trip(gpu):
    from nlse_amd import CNLSE

    rng = np.random.default_rng(2)
    V = rng.random((N, N)) + 1j * rng.random((N, N))
    n2 = _per_item(-1.6e-9, 3)
    # the reference test builds n12 by repeating n2 (test_cnlse.py:102-103)
    n12 = np.repeat(n2, 2)[..., np.newaxis, np.newaxis, np.newaxis]
    simu = CNLSE(_per_item(20, 3), 1.05, window, n2, n12, V, 10e-3, NX=N, NY=N, Isat=_per_item(10e4, 3))
    _check_round_trip(simu, ["alpha", "n2", "n12", "I_sat"])


def test_ddgpe_round_trip(gpu):
    from nlse_amd import DDGPE

    h = 0.654
    rng = np.random.default_rng(3)
    V = rng.random((N, N)) + 1j * rng.random((N, N))
    g = _per_item(1e-2 / h, 3)
    simu = DDGPE(_per_item(0.0, 3), 0.17 / h / (1e-2 / h), 256, g, _per_item(5.07 / h, 3), 1,
                 _per_item(1484.44 / h, 3), _per_item(1482.76 / h, 3), 0.17 / h, 27, V=V, NX=N, NY=N)
    _check_round_trip(simu, ["gamma", "g", "omega", "omega_cav", "omega_exc"])


def test_out_field_keeps_own_propagator_on_host(gpu):
    """The instance's own propagator (built by the first out_field) is a host
    array equal to the reference's expression after every call, and stays
    one object across calls; a propagator assigned from outside still goes
    to the device for the call and comes back."""
    from nlse_amd import NLSE

    simu = NLSE(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    simu.delta_z = 1e-4
    E0 = np.exp(-(simu.XX**2 + simu.YY**2) / 2.23e-3**2).astype(np.complex64)
    out1 = simu.out_field(E0, 5e-4, verbose=False)
    p = simu.propagator
    want = np.exp(-1j * 0.5 * (simu.Kxx**2 + simu.Kyy**2) / simu.k * simu.delta_z).astype(np.complex64)
    assert isinstance(p, np.ndarray) and np.array_equal(p, want)
    assert isinstance(simu.nl_profile, np.ndarray)
    out2 = simu.out_field(E0, 5e-4, verbose=False)
    assert simu.propagator is p
    assert np.array_equal(out1, out2)
    # assigned from outside: the dense table is applied and round-trips
    simu.propagator = want.copy()
    out3 = simu.out_field(E0, 5e-4, verbose=False)
    assert isinstance(simu.propagator, np.ndarray) and np.array_equal(simu.propagator, want)
    assert np.abs(out3 - out1).max() <= 1e-6 * np.abs(out1).max()


def test_build_fft_plan(gpu):
    """test_nlse.py:53-99 and test_nlse3d.py:62-103 (GPU branch): one plan
    over the field's shape, here a HipFFTPlan (the VkFFTApp stand-in), and it
    transforms: forward then inverse returns the field."""
    from nlse_amd import NLSE, NLSE_3d, NLSE_1d
    from nlse_amd.fft import HipFFTPlan

    rng = np.random.default_rng(4)
    s2 = NLSE(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, NY=N, Isat=1e5)
    s3 = NLSE_3d(alpha=20, energy=1e-6, window=[window, window, 2e-12], n2=-1.6e-9, D0=1e-26, vg=1.0, V=None,
                 L=1e-3, NX=32, NY=32, NZ=16, Isat=1e5)
    s1 = NLSE_1d(20, 1.05, window, -1.6e-9, None, 10e-3, NX=N, Isat=1e5)
    for simu, shape in ((s2, (N, N)), (s3, (32, 32, 16)), (s1, (3, N))):
        A = torch.from_numpy((rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)).cuda()
        plans = simu._build_fft_plan(A)
        assert len(plans) == 1 and isinstance(plans[0], HipFFTPlan)
        assert plans[0].shape0 == shape
        B = A.clone()
        plans[0].fft(B, B)
        plans[0].ifft(B, B)
        assert torch.allclose(B, A, atol=1e-5)


def _classes():
    from scipy.constants import atomic_mass

    h = 0.654
    return [
        ("NLSE_1d", dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, V=None, L=1e-3, NX=256, Isat=1e5), (3, 256)),
        ("NLSE_3d", dict(alpha=20, energy=1e-6, window=[window, window, 2e-12], n2=-1.6e-9, D0=1e-26, vg=1.0, V=None,
                         L=1e-3, NX=32, NY=32, NZ=16, Isat=1e5), (32, 32, 16)),
        ("GPE", dict(gamma=0, N=1e6, window=1e-3, g=1e-6, V=None, m=87 * atomic_mass, NX=64, NY=64), (2, 64, 64)),
        ("CNLSE", dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, n12=-1e-10, V=None, L=1e-3, NX=64, NY=64,
                       Isat=1e5), (2, 64, 64)),
        ("CNLSE_1d", dict(alpha=20, power=1.05, window=window, n2=-1.6e-9, n12=-1e-10, V=None, L=1e-3, NX=256,
                          Isat=1e5), (3, 2, 256)),
        ("DDGPE", dict(gamma=0.0, power=17.0, window=256, g=1e-2 / h, omega=5.07 / h, T=1, omega_exc=1484.44 / h,
                       omega_cav=1482.76 / h, detuning=0.17 / h, k_z=27, NX=64, NY=64), (2, 64, 64)),
    ]


@pytest.mark.parametrize("name,kw,shape", _classes(), ids=[c[0] for c in _classes()])
@pytest.mark.parametrize("normalize", [True, False])
def test_prepare_output_array_all_classes(gpu, name, kw, shape, normalize):
    """test_nlse_1d.py:39-94, test_nlse3d.py:106-173, test_gpe.py:52-104,
    test_cnlse.py:31-88, test_cnlse_1d.py:51-114, test_ddgpe.py:32-82:
    contiguous outputs of the input's shape, normalised as the class
    normalises (equal to the oracle class, whose normalisations are pinned
    to their integrals)."""
    import nlse_amd
    import oracle

    rng = np.random.default_rng(5)
    E = (rng.random(shape) + 1j * rng.random(shape)).astype(np.complex64)
    g = getattr(nlse_amd, name)(**kw)
    o = getattr(oracle, name)(**kw)
    out, out_sq = g._prepare_output_array(E, normalize)
    ref, _ = o._prepare_output_array(E, normalize)
    assert out.is_contiguous() and out_sq.is_contiguous()
    assert tuple(out.shape) == shape and tuple(out_sq.shape) == shape
    got = out.cpu().numpy().astype(np.complex128)
    ref = np.asarray(ref).astype(np.complex128)
    assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)
