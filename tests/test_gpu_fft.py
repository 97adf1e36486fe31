"""HIP FFT (replacement of VkFFTApp, nlse.py:207-218) against numpy.fft in
FP64, plus the bias check that long split-step runs depend on."""

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

# fp32 FFT: ~0.5 ulp per stage of <= 4 stages; rms relative error well
# below 5e-7 for every size handled by the engine
FFT_TOL = 5e-7


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_fft_1d(gpu, n):
    from nlse_amd.fft import HipFFTPlan

    x = _rand((6, n), n)
    plan = HipFFTPlan(x.shape, np.complex64, ndim=1)
    t = torch.from_numpy(x).to(gpu)
    plan.fft(t, t)
    ref = np.fft.fft(x.astype(np.complex128), axis=-1)
    assert rel_l2(t.cpu().numpy(), ref) < FFT_TOL
    plan.ifft(t, t)
    assert rel_l2(t.cpu().numpy(), x) < 2 * FFT_TOL


@pytest.mark.parametrize("shape", [(2, 8, 8), (64, 64), (1, 128, 256), (3, 256, 128), (512, 512),
                                   (2048, 2048), (16, 4096), (4096, 16)])
def test_fft_2d(gpu, shape):
    from nlse_amd.fft import HipFFTPlan

    x = _rand(shape, sum(shape))
    plan = HipFFTPlan(x.shape, np.complex64, ndim=2)
    assert plan.shape0 == tuple(shape)
    t = torch.from_numpy(x).to(gpu)
    plan.fft(t, t)
    ref = np.fft.fft2(x.astype(np.complex128))
    assert rel_l2(t.cpu().numpy(), ref) < FFT_TOL
    plan.ifft(t, t)
    assert rel_l2(t.cpu().numpy(), x) < 2 * FFT_TOL


@pytest.mark.parametrize("shape", [(32, 16, 8), (2, 64, 32, 16), (128, 128, 64)])
def test_fft_3d(gpu, shape):
    from nlse_amd.fft import HipFFTPlan

    x = _rand(shape, 7)
    plan = HipFFTPlan(x.shape, np.complex64, ndim=3)
    t = torch.from_numpy(x).to(gpu)
    plan.fft(t, t)
    ref = np.fft.fftn(x.astype(np.complex128), axes=(-3, -2, -1))
    assert rel_l2(t.cpu().numpy(), ref) < FFT_TOL
    plan.ifft(t, t)
    assert rel_l2(t.cpu().numpy(), x) < 2 * FFT_TOL


def test_fft_known_answers(gpu):
    """delta <-> constant, plane wave <-> delta (exact in fp32)."""
    from nlse_amd.fft import HipFFTPlan

    n = 256
    plan = HipFFTPlan((n, n), np.complex64, ndim=2)
    d = torch.zeros((n, n), dtype=torch.complex64, device=gpu)
    d[0, 0] = 1
    plan.fft(d, d)
    assert torch.all(d == 1)
    ky, kx = 5, 17
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    w = np.exp(2j * np.pi * (ky * yy + kx * xx) / n).astype(np.complex64)
    t = torch.from_numpy(w).to(gpu)
    plan.fft(t, t)
    out = t.cpu().numpy()
    assert abs(out[ky, kx] - n * n) / (n * n) < 1e-6
    out[ky, kx] = 0
    assert np.abs(out).max() / (n * n) < 1e-5


def test_fft_roundtrip_unbiased(gpu):
    """The gain of a forward+inverse pair must be unbiased: a split-step run
    applies it once per step, so a bias b becomes n*b after n steps (scipy's
    single precision transform has b ~ -1e-7)."""
    from nlse_amd.fft import HipFFTPlan

    for n in (512, 2048):
        x = _rand((n, n), 3)
        plan = HipFFTPlan(x.shape, np.complex64, ndim=2)
        t = torch.from_numpy(x).to(gpu)
        plan.fft(t, t)
        plan.ifft(t, t)
        y = t.cpu().numpy().astype(np.complex128)
        x64 = x.astype(np.complex128)
        e = y - x64
        bias = np.real(np.vdot(x64, e)) / np.real(np.vdot(x64, x64))
        print(f"n={n} round-trip rel err {np.linalg.norm(e) / np.linalg.norm(x64):.3e} gain bias {bias:.3e}")
        assert abs(bias) < 1e-8
