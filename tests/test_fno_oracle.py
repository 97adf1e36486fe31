"""Pins of the CPU restatement of the corrector network (oracle/fno.py):
gradients against central finite differences in FP64, and the truncated
spectral transform against a direct DFT sum."""

import math

import numpy as np

from oracle import fno as F


def test_truncated_transform_is_a_direct_dft_sum():
    rng = np.random.default_rng(0)
    H, W = 40, 36
    h = rng.standard_normal((1, 1, H, W))
    Rc = np.zeros((32, 32, 1, 1), complex)
    Rc[3, 30, 0, 0] = 1.0  # keep one mode: ky = 3, kx = W - 2
    s, hh = F.spectral_fwd(h, Rc)
    ky, kx = F.mode_freqs(H)[3], F.mode_freqs(W)[30]
    assert kx == W - 2
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    ph = 2 * np.pi * (ky * yy / H + kx * xx / W)
    direct = (h[0, 0] * np.exp(-1j * ph)).sum() / math.sqrt(H * W)
    assert abs(hh[0, 0, 3, 30] - direct) < 1e-12
    # s = Re(o e^{+i phi}) / sqrt(HW) for the single kept mode
    assert np.allclose(s[0, 0], np.real(direct * np.exp(1j * ph)) / math.sqrt(H * W), atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    cin, width, layers, cout, B, H, W = 4, 3, 2, 2, 2, 32, 32
    lay, n = F.layout(cin, width, layers, cout)
    w = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((B, cin, H, W))
    tgt = rng.standard_normal((B, cout, H, W))

    def loss(wv):
        out, _ = F.forward(wv, x, cin, width, layers, cout)
        return F.mse_loss_grad(out, tgt)[0]

    out, cache = F.forward(w, x, cin, width, layers, cout)
    _, dout = F.mse_loss_grad(out, tgt)
    g = F.backward(w, cache, dout, cin, width, layers, cout)
    idx = []
    for name, (o, s) in lay.items():
        size = int(np.prod(s))
        idx += list(o + rng.choice(size, size=min(3, size), replace=False))
    eps = 1e-6
    for i in idx:
        e = np.zeros(n)
        e[i] = eps
        fd = (loss(w + e) - loss(w - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
