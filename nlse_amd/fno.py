"""Learned corrector for coarse split steps: a Fourier neural operator (FNO)
trained and run on the engine's own BF16 matrix-core kernels.

No counterpart in the reference.  A coarse split step of dz_c = R * dz
(R = 8) is cheap but carries splitting error; the FNO maps the coarse-step
field Y (and the field X before the step) to the residual F - Y, where F is
the result of R fine steps of the plain solver.  The corrected field
Y + s * r * FNO(Y/s, X/s) (s = rms(Y), r the RMS relative residual of the
training pairs) replaces the R fine steps, with the
plain fine steps as fallback when the corrected step changes the power by
more than a threshold relative to the coarse step.

Network (per sample, grid H x W):
  lift        [Re Y, Im Y, Re X, Im X] -> C channels (1x1, bias)
  L layers    h <- gelu(K_l(h) + h W_l + b_l)  (no gelu after the last)
              K_l: truncated 2-D DFT to 32 x 32 modes (|k| < 16 on each
              axis, both signs), per-mode complex C x C mixing, inverse
  projection  C -> 2 (1x1, bias): the residual / (s r)

Every product is nlse_fno_gemm (BF16 operands, FP32 accumulation); the
DFTs are products with fixed cos/sin matrices (orthonormal scaling), so
forward and backward passes are sequences of the same GEMM with other
strides.  Master weights and activations are FP32.  The CPU restatement in
oracle/fno.py (FP64) is the reference for the outputs and gradients.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _device as D
from . import _lib

MODES = 32  # retained modes per axis


def mode_freqs(n: int) -> np.ndarray:
    """Integer frequencies of the 32 retained modes of an axis of n points:
    0..15 and -16..-1 (as n-16..n-1)."""
    half = MODES // 2
    return np.concatenate([np.arange(half), np.arange(n - half, n)])


def dft_mats(H: int, W: int):
    """Fx (W x 64), Fy (64 x H) of the truncated forward DFT and their
    transposes for the inverse (orthonormal scaling)."""
    kx = mode_freqs(W)
    ky = mode_freqs(H)
    x = np.arange(W)
    y = np.arange(H)
    thx = 2 * np.pi * np.outer(x, kx) / W  # (W, 32)
    thy = 2 * np.pi * np.outer(ky, y) / H  # (32, H)
    Fx = np.concatenate([np.cos(thx), -np.sin(thx)], axis=1) / math.sqrt(W)
    Fy = np.concatenate([np.cos(thy), np.sin(thy)], axis=0) / math.sqrt(H)
    return Fx, Fy


def param_layout(cin: int, width: int, layers: int, cout: int):
    """Names, shapes and offsets of the flat parameter vector."""
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


def init_params(cin: int, width: int, layers: int, cout: int, seed: int = 0) -> np.ndarray:
    """Uniform fan-in initialisation of the 1x1 maps, small spectral
    weights (1/(width^2) scale), zero biases."""
    lay, n = param_layout(cin, width, layers, cout)
    rng = np.random.default_rng(seed)
    w = np.zeros(n, np.float32)
    for name, (off, shp) in lay.items():
        size = int(np.prod(shp))
        if name.startswith("R"):
            vals = rng.uniform(-1, 1, size) / (width * width)
        elif name in ("P", "Q") or name.startswith("W"):
            vals = rng.uniform(-1, 1, size) / math.sqrt(shp[0])
        else:
            vals = np.zeros(size)
        w[off:off + size] = vals
    return w


def _f(t) -> int:
    return 0 if t is None else t.data_ptr()


class FNO:
    """The corrector network on the GPU for batches of B samples on an
    H x W grid (H, W >= 32)."""

    def __init__(self, H: int, W: int, width: int = 32, layers: int = 4, cin: int = 4, cout: int = 2,
                 seed: int = 0, params: np.ndarray | None = None):
        if H < MODES or W < MODES:
            raise ValueError(f"grid {H}x{W} smaller than the {MODES} retained modes")
        self.H, self.W, self.C, self.L, self.cin, self.cout = H, W, width, layers, cin, cout
        self.layout, self.nparams = param_layout(cin, width, layers, cout)
        w0 = init_params(cin, width, layers, cout, seed) if params is None else np.asarray(params, np.float32)
        if w0.size != self.nparams:
            raise ValueError("parameter vector has the wrong size")
        dev = torch.device("cuda")
        self.w = torch.from_numpy(w0.copy()).to(dev)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.t = 0
        Fx, Fy = dft_mats(H, W)
        self.Fx = torch.from_numpy(Fx.astype(np.float32)).to(dev).contiguous()  # (W, 64)
        self.Fy = torch.from_numpy(Fy.astype(np.float32)).to(dev).contiguous()  # (64, H)
        self._bufs = {}
        self.lib = _lib.load()
        # the residual F - Y is a small fraction of the field: the network
        # predicts (F - Y) / (s * res_scale), an O(1) quantity
        self.res_scale = 1.0

    # ------------------------------------------------------------ helpers
    def view(self, name: str, which: torch.Tensor | None = None) -> torch.Tensor:
        off, shp = self.layout[name]
        t = self.w if which is None else which
        return t[off:off + int(np.prod(shp))].view(*shp)

    def params(self) -> np.ndarray:
        return self.w.cpu().numpy()

    def _buf(self, name: str, shape) -> torch.Tensor:
        b = self._bufs.get(name)
        n = int(np.prod(shape))
        if b is None or b.numel() < n:
            b = torch.empty(n, dtype=torch.float32, device="cuda")
            self._bufs[name] = b
        return b[:n].view(*shape)

    def _gemm(self, M, N, K, batch, A, sam, sak, sa, B, sbk, sbn, sb, C, scm, scn, sc, beta=0.0, bias=None,
              act=None):
        _lib.check(self.lib.nlse_fno_gemm(int(M), int(N), int(K), int(batch), A.data_ptr(), sam, sak, sa,
                                          B.data_ptr(), sbk, sbn, sb, C.data_ptr(), scm, scn, sc, 1.0,
                                          float(beta), _f(bias), _f(act), D.stream_ptr()), "fno_gemm")

    # (cin, cout) pairs with a dedicated pointwise kernel (nlse_fno_pointwise)
    _PW = {(4, 32), (32, 32), (32, 2), (4, 16), (16, 16), (16, 2), (4, 8), (8, 8), (8, 2)}

    def _pointwise(self, X, Wt, bias, Y, B, cin, cout, beta=0.0, act=None):
        """Y[b][o][p] = sum_c X[b][c][p] Wt[c][o] (+ bias[o]) for all b."""
        hw = self.H * self.W
        if (cin, cout) in self._PW:
            _lib.check(self.lib.nlse_fno_pointwise(X.data_ptr(), Wt.data_ptr(), _f(bias), Y.data_ptr(), _f(act), B,
                                                   cin, cout, hw, float(beta), D.stream_ptr()), "fno_pointwise")
            return
        self._gemm(hw, cout, cin, B, X, 1, hw, cin * hw, Wt, cout, 1, 0, Y, 1, hw, cout * hw, beta, bias, act)

    def _weight_grad(self, X, dY, B, cin, cout, out):
        """out[c][o] = sum_b sum_p X[b][c][p] dY[b][o][p] (split over p)."""
        hw = self.H * self.W
        S = max(1, min(64, hw // 1024))
        if hw % S:
            S = 1
        part = self._buf("wg_part", (B, S, cin, cout))
        chunk = hw // S
        for b in range(B):
            Xb = X[b].reshape(-1)
            dYb = dY[b].reshape(-1)
            self._gemm(cin, cout, chunk, S, Xb, hw, 1, chunk, dYb, 1, hw, chunk, part[b], cout, 1, cin * cout)
        _lib.check(self.lib.nlse_fno_sum_parts(part.data_ptr(), B * S, cin * cout, out.data_ptr(),
                                               D.stream_ptr()), "sum_parts")

    def _bias_grad(self, dY, B, cout, out):
        hw = self.H * self.W
        part = self._buf("bg_part", (B, cout))
        for b in range(B):
            _lib.check(self.lib.nlse_fno_colsum(dY[b].data_ptr(), hw, cout, 1, hw, part[b].data_ptr(),
                                                D.stream_ptr()), "colsum")
        _lib.check(self.lib.nlse_fno_sum_parts(part.data_ptr(), B, cout, out.data_ptr(), D.stream_ptr()),
                   "sum_parts")

    # ------------------------------------------------------------ forward
    def forward(self, x_in: torch.Tensor, keep: bool = False) -> torch.Tensor:
        """x_in [B][cin][H][W] float32 -> [B][cout][H][W]."""
        B = x_in.shape[0]
        H, W, C, hw = self.H, self.W, self.C, self.H * self.W
        h = self._buf("h0", (B, C, H, W))
        self._pointwise(x_in, self.view("P"), self.view("p"), h, B, self.cin, C)
        saved = [(x_in, None, None, h)]
        for l in range(self.L):
            last = l == self.L - 1
            T = self._buf("T", (B * C * H, 64))
            self._gemm(B * C * H, 64, W, 1, h, W, 1, 0, self.Fx, 64, 1, 0, T, 64, 1, 0)
            P1 = self._buf(f"P1_{l}" if keep else "P1", (B * C, 64, 64))
            self._gemm(64, 64, H, B * C, self.Fy, H, 1, 0, T, 64, 1, H * 64, P1, 64, 1, 4096)
            Q = self._buf("Q", (B * C, 64, 64))
            _lib.check(self.lib.nlse_fno_mix(P1.data_ptr(), Q.data_ptr(), self.view(f"R{l}").data_ptr(), B, C, C,
                                             D.stream_ptr()), "fno_mix")
            U = self._buf("U", (B * C * H, 64))
            self._gemm(H, 64, 64, B * C, self.Fy, 1, H, 0, Q, 64, 1, 4096, U, 64, 1, H * 64)
            # inference keeps no pre-activation: the layer's GELU is applied
            # in place, alternating two buffers (h is the other one)
            inplace = not keep and not last
            zname = f"z_{l}" if keep else (("hA" if l % 2 == 0 else "hB") if inplace else "z")
            z = self._buf(zname, (B, C, H, W))
            self._gemm(B * C * H, W, 64, 1, U, 64, 1, 0, self.Fx, 1, 64, 0, z, W, 1, 0)
            hn = z if inplace else self._buf(f"h{l + 1}" if keep else ("hA" if l % 2 == 0 else "hB"), (B, C, H, W))
            self._pointwise(h, self.view(f"W{l}"), self.view(f"b{l}"), z, B, C, C, beta=1.0,
                            act=None if last else hn)
            if last:
                hn = z
            saved.append((h, P1, z, hn))
            h = hn
        out = self._buf("out", (B, self.cout, H, W))
        self._pointwise(h, self.view("Q"), self.view("q"), out, B, C, self.cout)
        if keep:
            self._saved = saved
        return out

    # ----------------------------------------------------------- backward
    def backward(self, dout: torch.Tensor) -> None:
        """Gradients of all parameters into self.g, given dL/dout
        (forward must have run with keep=True)."""
        saved = self._saved
        B = dout.shape[0]
        H, W, C, hw = self.H, self.W, self.C, self.H * self.W
        g = self.g
        hL = saved[-1][3]
        self._weight_grad(hL, dout, B, C, self.cout, self.view("Q", g))
        self._bias_grad(dout, B, self.cout, self.view("q", g))
        dh = self._buf("dh", (B, C, H, W))
        Qt = self.view("Q")
        self._gemm(hw, C, self.cout, B, dout, 1, hw, self.cout * hw, Qt, 1, self.cout, 0, dh, 1, hw, C * hw)
        for l in range(self.L - 1, -1, -1):
            h, P1, z, _ = saved[l + 1]
            if l < self.L - 1:
                dz = self._buf("dz", (B, C, H, W))
                _lib.check(self.lib.nlse_fno_gelu_bwd(z.data_ptr(), dh.data_ptr(), dz.data_ptr(), dz.numel(),
                                                      D.stream_ptr()), "gelu_bwd")
            else:
                dz = self._buf("dz", (B, C, H, W))
                dz.copy_(dh)
            self._weight_grad(h, dz, B, C, C, self.view(f"W{l}", g))
            self._bias_grad(dz, B, C, self.view(f"b{l}", g))
            # pointwise path: dh = dz W^T
            self._gemm(hw, C, C, B, dz, 1, hw, C * hw, self.view(f"W{l}"), 1, C, 0, dh, 1, hw, C * hw)
            # spectral path
            dU = self._buf("U", (B * C * H, 64))
            self._gemm(B * C * H, 64, W, 1, dz, W, 1, 0, self.Fx, 64, 1, 0, dU, 64, 1, 0)
            dQ = self._buf("Q", (B * C, 64, 64))
            self._gemm(64, 64, H, B * C, self.Fy, H, 1, 0, dU, 64, 1, H * 64, dQ, 64, 1, 4096)
            dP1 = self._buf("dP1", (B * C, 64, 64))
            _lib.check(self.lib.nlse_fno_mix_bwd(P1.data_ptr(), dQ.data_ptr(), self.view(f"R{l}").data_ptr(),
                                                 dP1.data_ptr(), self.view(f"R{l}", g).data_ptr(), B, C, C,
                                                 D.stream_ptr()), "fno_mix_bwd")
            dT = self._buf("T", (B * C * H, 64))
            self._gemm(H, 64, 64, B * C, self.Fy, 1, H, 0, dP1, 64, 1, 4096, dT, 64, 1, H * 64)
            self._gemm(B * C * H, W, 64, 1, dT, 64, 1, 0, self.Fx, 1, 64, 0, dh, W, 1, 0, beta=1.0)
        x_in = saved[0][0]
        self._weight_grad(x_in, dh, B, self.cin, C, self.view("P", g))
        self._bias_grad(dh, B, C, self.view("p", g))

    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        _lib.check(self.lib.nlse_fno_adam(self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                          self.v.data_ptr(), self.nparams, lr, b1, b2, eps, self.t,
                                          D.stream_ptr()), "fno_adam")

    # ---------------------------------------------------- field interface
    def _scales(self, Y: torch.Tensor) -> np.ndarray:
        """Per-sample rms s = sqrt(sum |Y|^2 / (H W)) (FP64 sum on the
        device, 1 where Y is zero), as host float32."""
        from . import kernels_hip

        ss = kernels_hip.sumsq(Y, 2).reshape(Y.shape[0]).cpu().numpy()
        s = np.sqrt(ss / (self.H * self.W)).astype(np.float32)
        return np.where(s > 0, s, np.float32(1.0)).astype(np.float32)

    @staticmethod
    def _dev(v: np.ndarray) -> torch.Tensor:
        return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).to(D.device())

    def features(self, Y: torch.Tensor, X: torch.Tensor):
        """Network input and per-sample scales (host) for complex64 fields
        Y, X [B][H][W]."""
        B = Y.shape[0]
        s = self._scales(Y)
        inv = self._dev(np.float32(1.0) / s)
        x_in = self._buf("x_in", (B, 4, self.H, self.W))
        _lib.check(self.lib.nlse_fno_pack(Y.data_ptr(), X.data_ptr(), inv.data_ptr(), x_in.data_ptr(), B,
                                          self.H * self.W, D.stream_ptr()), "fno_pack")
        return x_in, s

    def correct(self, Y: torch.Tensor, X: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Corrected fields Y + s * FNO(Y/s, X/s)."""
        x_in, s = self.features(Y, X)
        o = self.forward(x_in)
        A = torch.empty_like(Y) if out is None else out
        s2 = self._dev(s * np.float32(self.res_scale))
        _lib.check(self.lib.nlse_fno_apply(Y.data_ptr(), o.data_ptr(), s2.data_ptr(), A.data_ptr(), Y.shape[0],
                                           self.H * self.W, D.stream_ptr()), "fno_apply")
        return A

    def train_step(self, Y: torch.Tensor, X: torch.Tensor, F: torch.Tensor, lr: float) -> float:
        """One Adam step on the batch (Y, X) -> F; returns the loss (MSE of
        the scaled residual)."""
        B = Y.shape[0]
        x_in, s = self.features(Y, X)
        tgt = self._buf("tgt", (B, 2, self.H, self.W))
        inv2 = self._dev(np.float32(1.0) / s / np.float32(self.res_scale))
        _lib.check(self.lib.nlse_fno_residual(F.data_ptr(), Y.data_ptr(), inv2.data_ptr(), tgt.data_ptr(), B,
                                              self.H * self.W, D.stream_ptr()), "fno_residual")
        out = self.forward(x_in, keep=True)
        dout = self._buf("dout", (B, 2, self.H, self.W))
        nparts = 256
        part = torch.empty(nparts, dtype=torch.float64, device="cuda")
        _lib.check(self.lib.nlse_fno_mse(out.data_ptr(), tgt.data_ptr(), dout.data_ptr(), out.numel(),
                                         part.data_ptr(), nparts, D.stream_ptr()), "fno_mse")
        self.backward(dout)
        self.adam(lr)
        return float(part.cpu().numpy().sum()) / out.numel()
