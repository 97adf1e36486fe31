"""CPU restatement (FP64, numpy) of the learned corrector network of
nlse_amd.fno -- TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The network has no counterpart in the reference; this module states its
definition independently of the GPU formulation: the spectral layers use
numpy FFTs (the GPU multiplies by truncated cos/sin DFT matrices), and the
backward pass is derived by hand from the FFT form:
  forward   h_hat = fft2(h)[modes] / sqrt(HW)
            o_hat = R h_hat (per mode, complex C x C)
            s     = Re(ifft2(scatter(o_hat))) * sqrt(HW)
  adjoint   d o_hat = fft2(ds)[modes] / sqrt(HW)   (adjoint of Re o ifft2)
            dR      = sum_b d o_hat conj(h_hat)^T,  d h_hat = R^H d o_hat
            dh      = Re(ifft2(scatter(d h_hat))) * sqrt(HW)
Pinned by finite differences (tests/test_fno_oracle.py) and by a direct
DFT sum for the truncated transform.
"""

from __future__ import annotations

import math

import numpy as np

MODES = 32


def mode_freqs(n: int) -> np.ndarray:
    half = MODES // 2
    return np.concatenate([np.arange(half), np.arange(n - half, n)])


def layout(cin: int, width: int, layers: int, cout: int):
    shapes = [("P", (cin, width)), ("p", (width,))]
    for l in range(layers):
        shapes += [(f"R{l}", (MODES * MODES, width, width, 2)), (f"W{l}", (width, width)), (f"b{l}", (width,))]
    shapes += [("Q", (width, cout)), ("q", (cout,))]
    out, off = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (off, shp)
        off += n
    return out, off


def unpack(w: np.ndarray, lay: dict) -> dict:
    return {k: np.asarray(w[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in lay.items()}


def gelu(x):
    k0, k1 = math.sqrt(2 / math.pi), 0.044715
    return 0.5 * x * (1 + np.tanh(k0 * (x + k1 * x**3)))


def gelu_grad(x):
    k0, k1 = math.sqrt(2 / math.pi), 0.044715
    t = np.tanh(k0 * (x + k1 * x**3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * k0 * (1 + 3 * k1 * x * x)


def _modes(H, W):
    return mode_freqs(H), mode_freqs(W)


def spectral_fwd(h, Rc):
    """h (B, C, H, W) real, Rc (32, 32, Cout, Cin) complex -> (s, h_hat)."""
    B, C, H, W = h.shape
    ky, kx = _modes(H, W)
    hh = np.fft.fft2(h)[:, :, ky][:, :, :, kx] / math.sqrt(H * W)  # (B, C, 32, 32)
    oh = np.einsum("yxoi,biyx->boyx", Rc, hh)
    full = np.zeros((B, Rc.shape[2], H, W), complex)
    full[:, :, ky[:, None], kx[None, :]] = oh
    s = np.real(np.fft.ifft2(full)) * math.sqrt(H * W)
    return s, hh


def spectral_bwd(ds, hh, Rc):
    B, C, H, W = ds.shape
    ky, kx = _modes(H, W)
    doh = np.fft.fft2(ds)[:, :, ky][:, :, :, kx] / math.sqrt(H * W)
    dR = np.einsum("boyx,biyx->yxoi", doh, np.conj(hh))
    dhh = np.einsum("yxoi,boyx->biyx", np.conj(Rc), doh)
    full = np.zeros((B, Rc.shape[3], H, W), complex)
    full[:, :, ky[:, None], kx[None, :]] = dhh
    dh = np.real(np.fft.ifft2(full)) * math.sqrt(H * W)
    return dh, dR


def _rc(R):
    """[1024][Cout][Cin][2] -> complex (32, 32, Cout, Cin)."""
    return (R[..., 0] + 1j * R[..., 1]).reshape(MODES, MODES, R.shape[1], R.shape[2])


def forward(w, x_in, cin, width, layers, cout):
    """x_in (B, cin, H, W) -> out (B, cout, H, W), cache."""
    lay, _ = layout(cin, width, layers, cout)
    p = unpack(w, lay)
    x = np.asarray(x_in, np.float64)
    h = np.einsum("bchw,co->bohw", x, p["P"]) + p["p"][None, :, None, None]
    cache = {"x": x, "h": [h], "z": [], "hh": []}
    for l in range(layers):
        s, hh = spectral_fwd(h, _rc(p[f"R{l}"]))
        z = s + np.einsum("bchw,co->bohw", h, p[f"W{l}"]) + p[f"b{l}"][None, :, None, None]
        h = gelu(z) if l < layers - 1 else z
        cache["z"].append(z)
        cache["hh"].append(hh)
        cache["h"].append(h)
    out = np.einsum("bchw,co->bohw", h, p["Q"]) + p["q"][None, :, None, None]
    return out, cache


def backward(w, cache, dout, cin, width, layers, cout) -> np.ndarray:
    lay, n = layout(cin, width, layers, cout)
    p = unpack(w, lay)
    g = {}
    hL = cache["h"][-1]
    g["Q"] = np.einsum("bchw,bohw->co", hL, dout)
    g["q"] = dout.sum(axis=(0, 2, 3))
    dh = np.einsum("bohw,co->bchw", dout, p["Q"])
    for l in range(layers - 1, -1, -1):
        z = cache["z"][l]
        h = cache["h"][l]
        dz = dh * gelu_grad(z) if l < layers - 1 else dh
        g[f"W{l}"] = np.einsum("bchw,bohw->co", h, dz)
        g[f"b{l}"] = dz.sum(axis=(0, 2, 3))
        dh_pw = np.einsum("bohw,co->bchw", dz, p[f"W{l}"])
        dh_sp, dR = spectral_bwd(dz, cache["hh"][l], _rc(p[f"R{l}"]))
        g[f"R{l}"] = np.stack([dR.real, dR.imag], axis=-1).reshape(MODES * MODES, width, width, 2)
        dh = dh_pw + dh_sp
    g["P"] = np.einsum("bchw,bohw->co", cache["x"], dh)
    g["p"] = dh.sum(axis=(0, 2, 3))
    flat = np.zeros(n)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = g[k].reshape(-1)
    return flat


def mse_loss_grad(out, tgt):
    e = out - tgt
    return float(np.mean(e * e)), 2 * e / e.size
