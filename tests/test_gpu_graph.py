"""hipGraph replay of the steady steps (nlse_run_graph) against direct
launches (nlse_run): the same kernels with the same arguments, so the
results must be bit-identical.  Each case replays the captured graph
several times on the same field buffer with other device work and
synchronisations in between, and checks every replay; a table change
between runs must trigger a re-capture."""

import numpy as np
import pytest
import torch

import oracle  # noqa: F401  (keeps the test layout of the other GPU modules)

pytestmark = pytest.mark.gpu

waist = 2.23e-3
window = 4 * waist


def _noise(shape, seed, amp=0.01):
    rng = np.random.default_rng(seed)
    return 1 + amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _setup(sim, E0, precision):
    sim._own_propagator()
    sim._send_arrays_to_gpu()
    A, _ = sim._prepare_output_array(E0, True)
    eng = sim._engine_for(A)
    sim._configure(eng, A, sim.V, sim.propagator, precision)
    return A, eng


def _compare(eng, A0, nsteps, precision, chunk=4, reps=3):
    buf_g = torch.empty_like(A0)
    buf_e = torch.empty_like(A0)
    for rep in range(reps):
        buf_g.copy_(A0)
        eng.graph_chunk = chunk
        eng.run(buf_g, nsteps, precision)
        # unrelated device work and a synchronisation between replays
        junk = torch.view_as_real(buf_g).abs().sum()
        torch.cuda.synchronize()
        buf_e.copy_(A0)
        eng.graph_chunk = 0
        eng.run(buf_e, nsteps, precision)
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        same = torch.equal(torch.view_as_real(buf_g), torch.view_as_real(buf_e))
        print(f"replay {rep}: bit-identical={same}")
        assert same
    return buf_g


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nl_px", [0.0, 3.5])
def test_graph_matches_direct_2d(gpu, precision, nl_px):
    from nlse_amd import NLSE

    N = 128
    sim = NLSE(20, 1.05, window, -1.6e-9, None, 1e-2, NX=N, NY=N, Isat=1e5, nl_length=nl_px * window / N)
    sim.delta_z = 1e-5
    sim.V = (-1e-4 * np.exp(-(sim.XX**2 + sim.YY**2) / (70e-6) ** 2)).astype(np.float32)
    G = np.exp(-(sim.XX.astype(np.float64) ** 2 + sim.YY.astype(np.float64) ** 2) / waist**2)
    E0 = (G * _noise(G.shape, 31)).astype(np.complex64)
    A, eng = _setup(sim, E0, "double" if precision else "single")
    _compare(eng, A.clone(), 23, precision)
    # a new potential re-captures: graph result still equals direct launches
    sim.V = (2e-4 * np.exp(-(sim.XX**2 + sim.YY**2) / (50e-6) ** 2)).astype(np.float32)
    sim._send_arrays_to_gpu()
    sim._configure(eng, A, sim.V, sim.propagator, "double" if precision else "single")
    _compare(eng, A.clone(), 23, precision, reps=2)


def test_graph_matches_direct_cnlse_and_3d(gpu):
    from nlse_amd import CNLSE, NLSE_3d

    N = 128
    sim = CNLSE(20, 1.05, window, -1.6e-9, -1e-10, None, 1e-2, NX=N, NY=N, Isat=1e5, omega=2e3)
    sim.delta_z = 1e-5
    G = np.exp(-(sim.XX.astype(np.float64) ** 2 + sim.YY.astype(np.float64) ** 2) / waist**2)
    E0 = np.stack([G * _noise(G.shape, 32), G * _noise(G.shape, 33) * 0.5]).astype(np.complex64)
    A, eng = _setup(sim, E0, "single")
    _compare(eng, A.clone(), 17, 0, chunk=5)

    s3 = NLSE_3d(0, 2e-12, [4e-3, 4e-3, 2e-12], -1e-9, -1e-26, 1.0, None, 1e-2, NX=64, NY=32, NZ=16)
    s3.delta_z = 1e-5
    pulse = (np.exp(-(s3.XX.astype(np.float64) ** 2 + s3.YY.astype(np.float64) ** 2) / (1e-3) ** 2)
             / np.cosh(s3.TT / 200e-15))
    A, eng = _setup(s3, (pulse * _noise(pulse.shape, 34)).astype(np.complex64), "single")
    _compare(eng, A.clone(), 11, 0, chunk=3)
