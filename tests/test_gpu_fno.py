"""The learned corrector on the GPU: the BF16 matrix-core GEMM against
numpy, the network's forward pass and every parameter gradient against the
FP64 oracle (oracle/fno.py), training on pairs from the plain solver with a
held-out gain over the uncorrected coarse step, and the drift fallback
(power and Hamiltonian)."""

import numpy as np
import pytest
import torch

from oracle import fno as F

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K,batch,ta,tb,tc", [(100, 70, 300, 1, False, False, False),
                                                  (64, 64, 64, 3, True, False, False),
                                                  (33, 130, 1000, 2, False, True, False),
                                                  (4096, 32, 32, 1, True, True, False),
                                                  (4096, 32, 32, 2, True, False, True),
                                                  (1000, 70, 45, 1, False, False, True)])
def test_gemm_bf16_matches_numpy(gpu, M, N, K, batch, ta, tb, tc):
    """C = A B + 0.5 C + bias with BF16-rounded operands and FP32
    accumulation, for row- and column-major operands and results (a
    column-major C takes the transposed-tile kernel), batched, with the
    GELU side output."""
    from nlse_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((batch, M, K)).astype(np.float32)
    B = rng.standard_normal((batch, K, N)).astype(np.float32)
    C0 = rng.standard_normal((batch, M, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    Ad = torch.from_numpy(np.ascontiguousarray(A.transpose(0, 2, 1)) if ta else A).cuda()
    Bd = torch.from_numpy(np.ascontiguousarray(B.transpose(0, 2, 1)) if tb else B).cuda()
    Cd = torch.from_numpy(np.ascontiguousarray(C0.transpose(0, 2, 1)) if tc else C0).cuda()
    act = torch.empty_like(Cd)
    bd = torch.from_numpy(bias).cuda()
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    scm, scn = (1, M) if tc else (N, 1)
    rc = lib.nlse_fno_gemm(M, N, K, batch, Ad.data_ptr(), sam, sak, M * K, Bd.data_ptr(), sbk, sbn, K * N,
                           Cd.data_ptr(), scm, scn, M * N, 1.0, 0.5, bd.data_ptr(), act.data_ptr(), 0)
    assert rc == 0, lib.nlse_last_error()
    ref = np.einsum("bmk,bkn->bmn", _bf16(A), _bf16(B)) + 0.5 * C0 + bias
    got = Cd.cpu().numpy()
    ga = act.cpu().numpy()
    if tc:
        got, ga = got.transpose(0, 2, 1), ga.transpose(0, 2, 1)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5
    assert np.abs(ga - F.gelu(got.astype(np.float64))).max() < 1e-5


def _net_and_params(H, W, C, L, seed=3):
    from nlse_amd import fno as G

    rng = np.random.default_rng(seed)
    w = G.init_params(4, C, L, 2, seed=seed)
    lay, _ = G.param_layout(4, C, L, 2)
    for name, (o, s) in lay.items():
        n = int(np.prod(s))
        if name.startswith("R"):
            w[o:o + n] = (rng.uniform(-1, 1, n) / C).astype(np.float32)
        if name.startswith("b") or name in ("p", "q"):
            w[o:o + n] = rng.uniform(-0.1, 0.1, n).astype(np.float32)
    return G.FNO(H, W, width=C, layers=L, params=w), w, lay


def test_fno_forward_and_gradients_match_oracle(gpu):
    """BF16 operands through four spectral layers: outputs within 1e-2 and
    every parameter group's gradient within 3e-2 (relative L2) of FP64."""
    H, W, C, L, B = 64, 48, 8, 4, 2
    net, w, lay = _net_and_params(H, W, C, L)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((B, 4, H, W)).astype(np.float32)
    tgt = rng.standard_normal((B, 2, H, W))
    out = net.forward(torch.from_numpy(x).cuda(), keep=True).cpu().numpy()
    ref, cache = F.forward(w.astype(np.float64), x, 4, C, L, 2)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print(f"forward rel L2 {err:.2e}")
    assert err < 1e-2
    _, dref = F.mse_loss_grad(ref, tgt)
    gref = F.backward(w.astype(np.float64), cache, dref, 4, C, L, 2)
    net.backward(torch.from_numpy(dref.astype(np.float32)).cuda().contiguous())
    g = net.g.cpu().numpy()
    for name, (o, s) in lay.items():
        n = int(np.prod(s))
        e = np.linalg.norm(g[o:o + n] - gref[o:o + n]) / (np.linalg.norm(gref[o:o + n]) + 1e-30)
        print(f"grad {name}: {e:.2e}")
        assert e < 3e-2, name
    # inference (no saved activations, GELU in place) gives the same output
    out_inf = net.forward(torch.from_numpy(x).cuda(), keep=False).cpu().numpy()
    assert np.array_equal(out_inf, out)


def _loop_errors(net, seed, nfine=64):
    """Relative errors of the plain coarse loop and of the corrected loop
    against the fine-step solver, for one held-out beam."""
    from nlse_amd import NLSE
    from nlse_amd import corrector as CR

    rng = np.random.default_rng(seed)
    mat = CR.material(rng)

    def sim():
        s = NLSE(mat["alpha"], mat["power"], CR.WINDOW, mat["n2"], None, 10e-3, NX=128, NY=128, Isat=mat["Isat"])
        s.delta_z = 1e-5
        return s

    E0 = CR._speckle_field(sim(), rng)
    z = (nfine - 0.5) * 1e-5
    fine = sim().out_field(E0, z, verbose=False)
    coarse_sim = sim()
    coarse_sim.delta_z = 8e-5
    coarse = coarse_sim.out_field(E0, (nfine // 8 - 0.5) * 8e-5, verbose=False)
    out, stats = CR.CorrectedSolver(sim(), net, drift_tol=1e-3).out_field(E0, z)
    assert stats["coarse_steps"] == nfine // 8
    print(f"seed {seed}: fallbacks {stats['fallbacks']} (power {stats['fallbacks_power']}, H {stats['fallbacks_h']}); "
          f"max drift power {stats['max_power_drift']:.2e}, H {stats['max_h_drift']:.2e}")
    e_c = np.linalg.norm(coarse - fine) / np.linalg.norm(fine)
    e_n = np.linalg.norm(out - fine) / np.linalg.norm(fine)
    return e_c, e_n, (sim, E0, z, fine)


def test_training_gain_and_fallback(gpu):
    """Trained on 1600 pairs of the plain solver (README-example material x
    U(0.5,1.5), speckle inputs), the corrected coarse step is closer to the
    8 fine steps than the plain coarse step on held-out trajectories, and
    the corrected loop beats the coarse loop against the fine-step solver
    on average over four held-out beams; a corrupted network trips the
    drift fallback on every step and reproduces the fine solver."""
    from nlse_amd import corrector as CR
    from nlse_amd.fno import FNO

    pairs = CR.make_pairs(N=128, trajectories=80, coarse_steps=20, seed=1)
    held = CR.make_pairs(N=128, trajectories=4, coarse_steps=20, seed=99)
    net = FNO(128, 128, width=32, layers=4, seed=0)
    hist = CR.train(net, pairs, steps=800, batch=8, lr=2e-3)
    ev = CR.evaluate(net, held)
    print(f"loss {hist[0]:.3e} -> {np.mean(hist[-50:]):.3e}; held out {ev}")
    assert np.mean(hist[-50:]) < 0.2 * hist[0]
    assert ev["gain"] > 1.5

    errs = [_loop_errors(net, sd) for sd in (123, 124, 125, 126)]
    e_c = np.mean([e[0] for e in errs])
    e_n = np.mean([e[1] for e in errs])
    print(f"loop over 64 fine-step equivalents, 4 beams: coarse {e_c:.2e} corrected {e_n:.2e}")
    assert e_n < e_c

    # a corrupted network: every corrected step changes the power -> fallback
    sim, E0, z, fine = errs[0][2]
    net.view("q")[:] = 50.0
    out, stats = CR.CorrectedSolver(sim(), net, drift_tol=1e-4).out_field(E0, z)
    assert stats["fallbacks"] == stats["coarse_steps"] == 8
    # eight runs of 8 fine steps vs one run of 64: equal up to the extra
    # x-transform round trips at the run boundaries
    assert np.linalg.norm(out - fine) / np.linalg.norm(fine) < 1e-5
    # a grossly corrupted network with the power test off: the Hamiltonian
    # test alone catches it
    net.view("q")[:] = 5000.0
    out, stats = CR.CorrectedSolver(sim(), net, drift_tol=np.inf, h_tol=1e-3).out_field(E0, z)
    print(f"corrupted, power test off: {stats}")
    assert stats["fallbacks_h"] == stats["fallbacks"] == 8 and stats["fallbacks_power"] == 0
    assert np.linalg.norm(out - fine) / np.linalg.norm(fine) < 1e-5


@pytest.mark.parametrize("cin,cout,beta,with_act", [(4, 32, 0.0, False), (32, 32, 1.0, True), (32, 2, 0.0, False),
                                                    (16, 16, 1.0, True)])
def test_pointwise_matches_gemm_and_numpy(gpu, cin, cout, beta, with_act):
    """nlse_fno_pointwise (the 1x1 layers) against numpy with BF16-rounded
    operands and against nlse_fno_gemm on the same product."""
    from nlse_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(cin * 100 + cout)
    B, hw = 2, 3000
    X = rng.standard_normal((B, cin, hw)).astype(np.float32)
    W = rng.standard_normal((cin, cout)).astype(np.float32)
    Y0 = rng.standard_normal((B, cout, hw)).astype(np.float32)
    bias = rng.standard_normal(cout).astype(np.float32)
    Xd, Wd, bd = (torch.from_numpy(a).cuda() for a in (X, W, bias))
    outs = []
    for kind in ("pointwise", "gemm"):
        Yd = torch.from_numpy(Y0.copy()).cuda()
        act = torch.empty_like(Yd) if with_act else None
        ap = act.data_ptr() if with_act else None
        if kind == "pointwise":
            rc = lib.nlse_fno_pointwise(Xd.data_ptr(), Wd.data_ptr(), bd.data_ptr(), Yd.data_ptr(), ap, B, cin, cout,
                                        hw, beta, 0)
        else:
            rc = lib.nlse_fno_gemm(hw, cout, cin, B, Xd.data_ptr(), 1, hw, cin * hw, Wd.data_ptr(), cout, 1, 0,
                                   Yd.data_ptr(), 1, hw, cout * hw, 1.0, beta, bd.data_ptr(), ap, 0)
        assert rc == 0, lib.nlse_last_error()
        outs.append((Yd.cpu().numpy(), act.cpu().numpy() if with_act else None))
    ref = np.einsum("bcp,co->bop", _bf16(X), _bf16(W)) + beta * Y0 + bias[None, :, None]
    for y, a in outs:
        assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-5
        if with_act:
            assert np.abs(a - F.gelu(y.astype(np.float64))).max() < 1e-5
    assert np.abs(outs[0][0] - outs[1][0]).max() / np.abs(ref).max() < 1e-6
