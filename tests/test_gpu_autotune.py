"""Autotuner on the GPU: tuning a plan and a batch writes the cache, the
engines then use the tuned choices, and the results stay bit-identical
(every candidate runs the same arithmetic)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

waist = 2.23e-3
window = 4 * waist


def _field(sim, seed):
    rng = np.random.default_rng(seed)
    G = np.exp(-(sim.XX.astype(np.float64) ** 2 + sim.YY.astype(np.float64) ** 2) / waist**2)
    return (G * (1 + 0.01 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)))).astype(
        np.complex64)


def test_tune_plan_and_use(gpu, tmp_path, monkeypatch):
    from nlse_amd import NLSE, autotune

    monkeypatch.setenv("NLSE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sim = NLSE(20, 1.05, window, -1.6e-9, None, 1e-2, NX=256, NY=256, Isat=1e5, nl_length=3.5 * window / 256)
    sim.delta_z = 1e-5
    E0 = _field(sim, 41)
    ref = sim.out_field(E0, 2e-4, verbose=False)  # engine defaults
    res = sim.tune(E0)
    print("tuned:", res["col_width"], res["graph_chunk"], {k: round(v, 1) for k, v in res["col_width_us"].items()})
    assert res["col_width"] in (1, 2, 4, 8, 16, 32) and res["graph_chunk"] in autotune.GRAPH_CHUNKS
    assert len(autotune.load()) == 1
    out = sim.out_field(E0, 2e-4, verbose=False)
    assert sim._engine.config()[1] == res["col_width"]
    assert np.array_equal(out, ref)


def test_tune_batch_and_use(gpu, tmp_path, monkeypatch):
    from nlse_amd import NLSE, NLSEBatch, autotune

    monkeypatch.setenv("NLSE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sims = [NLSE(20, 1.0, window, -1e-9, None, 1e-2, NX=n, NY=n, Isat=1e5) for n in (128, 256, 512, 256)]
    for s in sims:
        s.delta_z = 1e-5
    fields = [_field(s, 50 + i) for i, s in enumerate(sims)]
    b = NLSEBatch(sims)
    ref = b.out_field(fields, 30.5e-5)
    res = b.tune(fields)
    print("tuned batch:", res)
    assert len(autotune.load()) == 1
    out = b.out_field(fields, 30.5e-5)
    for o, r in zip(out, ref):
        assert np.array_equal(o, r)
