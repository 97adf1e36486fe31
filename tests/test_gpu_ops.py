"""Operator kernels (kernels_hip, the kernels_gpu.py API) against the oracle's
restatement of the reference numba kernels (kernels_cpu.py)."""

import numpy as np
import pytest
import torch
from scipy.constants import c, epsilon_0

import oracle.kernels as OK
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# complex64 storage: one rounding of the result plus fp32 evaluation of a
# phase of order 1e-2 rad; per element relative error ~1e-7
OP_TOL = 1e-6


def _field(shape, seed, amp=1e4):
    rng = np.random.default_rng(seed)
    return (amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)


def test_square_mod(gpu):
    from nlse_amd import kernels_hip as K

    A = _field((3, 64, 32), 0)
    S = np.zeros(A.shape, np.float32)
    OK.square_mod(A, S)
    tA = torch.from_numpy(A).to(gpu)
    tS = torch.zeros(A.shape, dtype=torch.float32, device=gpu)
    K.square_mod(tA, tS)
    np.testing.assert_allclose(tS.cpu().numpy(), S, rtol=2e-7)


@pytest.mark.parametrize("vkind", ["none", "real", "complex"])
@pytest.mark.parametrize("loss_sat", [False, True])
def test_nl_prop(gpu, vkind, loss_sat):
    from nlse_amd import kernels_hip as K

    A = _field((2, 64, 64), 1)
    S = np.zeros(A.shape, np.float32)
    OK.square_mod(A, S)
    rng = np.random.default_rng(2)
    V = None
    if vkind == "real":
        V = (4e2 * rng.standard_normal((64, 64))).astype(np.float32)
    elif vkind == "complex":
        V = (4e2 * (rng.standard_normal((64, 64)) + 0.1j * rng.standard_normal((64, 64)))).astype(np.complex64)
    k = 2 * np.pi / 780e-9
    dz, alpha = 1e-5, 10.0
    g = np.array([k / 2 * -1.6e-9 * c * epsilon_0, k / 2 * -0.4e-9 * c * epsilon_0]).reshape(2, 1, 1)
    Isat = 2 * 1e5 / (epsilon_0 * c)
    ref = A.copy()
    if V is None:
        OK.nl_prop_without_V(ref, S, dz, alpha, g, Isat, loss_saturates=loss_sat)
    else:
        OK.nl_prop(ref, S, dz, alpha, V, g, Isat, loss_saturates=loss_sat)
    tA = torch.from_numpy(A).to(gpu)
    K.nl_prop(tA, torch.from_numpy(S).to(gpu), dz, alpha, None if V is None else torch.from_numpy(V).to(gpu),
              g, Isat, loss_saturates=loss_sat)
    assert rel_l2(tA.cpu().numpy(), ref) < OP_TOL


def test_nl_prop_special_cases(gpu):
    """Isat = inf -> sat = 1; alpha = 0 preserves |A|; V = None == V = 0."""
    from nlse_amd import kernels_hip as K

    A = _field((64, 64), 3)
    S = (A.real * A.real + A.imag * A.imag).astype(np.float32)
    tS = torch.from_numpy(S).to(gpu)
    a0 = torch.from_numpy(A).to(gpu)
    K.nl_prop_without_V(a0, tS, 1e-5, 0.0, -1e-5, np.inf)
    out = a0.cpu().numpy()
    # one rounding per component of the updated value: |A| within a few ulp
    np.testing.assert_allclose(np.abs(out), np.abs(A), rtol=5e-7)
    # closed form with sat = 1: A exp(i dz g |A|^2)
    exact = A.astype(np.complex128) * np.exp(1j * 1e-5 * -1e-5 * S.astype(np.float64))
    assert rel_l2(out, exact) < 2e-7
    a1 = torch.from_numpy(A).to(gpu)
    K.nl_prop(a1, tS, 1e-5, 3.0, torch.zeros((64, 64), dtype=torch.float32, device=gpu), -1e-5, 1e8)
    a2 = torch.from_numpy(A).to(gpu)
    K.nl_prop_without_V(a2, tS, 1e-5, 3.0, -1e-5, 1e8)
    assert torch.equal(a1, a2)


def test_nl_prop_c_and_rabi(gpu):
    from nlse_amd import kernels_hip as K

    A = _field((3, 2, 32, 32), 4)
    S = (A.real * A.real + A.imag * A.imag).astype(np.float32)
    V = (1e2 * np.random.default_rng(5).standard_normal((32, 32))).astype(np.float32)
    args = dict(dz=1e-5, alpha=10.0, g11=-3e-5, g12=-1e-6, Isat1=1e9, Isat2=5e8)
    ref = A.copy()
    OK.nl_prop_c(ref[:, 0], S[:, 0], S[:, 1], args["dz"], args["alpha"], V, args["g11"], args["g12"],
                 args["Isat1"], args["Isat2"])
    OK.nl_prop_without_V_c(ref[:, 1], S[:, 1], S[:, 0], args["dz"], args["alpha"], args["g11"], args["g12"],
                           args["Isat2"], args["Isat1"])
    OK.rabi_coupling(ref[:, 0], ref[:, 1], 1e-5, 3e3)
    tA = torch.from_numpy(A).to(gpu)
    tS = torch.from_numpy(S).to(gpu)
    K.nl_prop_c(tA[:, 0], tS[:, 0], tS[:, 1], args["dz"], args["alpha"], torch.from_numpy(V).to(gpu),
                args["g11"], args["g12"], args["Isat1"], args["Isat2"])
    K.nl_prop_without_V_c(tA[:, 1], tS[:, 1], tS[:, 0], args["dz"], args["alpha"], args["g11"], args["g12"],
                          args["Isat2"], args["Isat1"])
    K.rabi_coupling(tA[:, 0], tA[:, 1], 1e-5, 3e3)
    assert rel_l2(tA.cpu().numpy(), ref) < OP_TOL


def test_rabi_quarter_period_swaps(gpu):
    """omega dz = pi/2 maps (A1, A2) -> (-i A2, -i A1)."""
    from nlse_amd import kernels_hip as K

    A = _field((2, 16, 16), 6, amp=1.0)
    t = torch.from_numpy(A).to(gpu)
    K.rabi_coupling(t[0], t[1], 1.0, np.pi / 2)
    out = t.cpu().numpy()
    np.testing.assert_allclose(out[0], -1j * A[1], atol=1e-6)
    np.testing.assert_allclose(out[1], -1j * A[0], atol=1e-6)


def test_mul_propagator(gpu):
    from nlse_amd import kernels_hip as K

    A = _field((2, 32, 64), 7, amp=1.0)
    P64 = np.exp(-1j * np.random.default_rng(8).uniform(0, 6, (32, 64)))
    for P in (P64.astype(np.complex64), P64):
        t = torch.from_numpy(A).to(gpu)
        K.mul_propagator(t, torch.from_numpy(P).to(gpu))
        ref = (A.astype(np.complex128) * P).astype(np.complex64) if P.dtype == np.complex128 else A * P
        assert rel_l2(t.cpu().numpy(), ref) < 2e-7


def test_sumsq_fp64(gpu):
    from nlse_amd import kernels_hip as K

    A = _field((3, 128, 256), 9, amp=1.0)
    t = torch.from_numpy(A).to(gpu)
    s = K.sumsq(t, 2).cpu().numpy()
    ref = (np.abs(A.astype(np.complex128)) ** 2).sum(axis=(-2, -1))
    np.testing.assert_allclose(s, ref, rtol=1e-12)
