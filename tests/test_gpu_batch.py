"""Uneven batches (nlse_batch_*): fields of different grids and parameters
advanced together must equal their separate out_field runs (same kernels,
same arithmetic) and the oracle; graph replay equals direct launches."""

import numpy as np
import pytest
import torch

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

waist = 2.23e-3
window = 4 * waist


def _instances(cls, shapes, seed=2):
    rng = np.random.default_rng(seed)
    sims, fields = [], []
    for i, (ny, nx) in enumerate(shapes):
        n2 = -10 ** rng.uniform(-10, -8)
        power = rng.uniform(0.1, 2)
        isat = rng.uniform(1e4, 1e6)
        s = cls(20, power, window, n2, None, 10e-3, NX=nx, NY=ny, Isat=isat)
        s.delta_z = 1e-5
        if i % 3 == 1:
            s.V = (-1e-4 * np.exp(-(s.XX**2 + s.YY**2) / (70e-6) ** 2)).astype(np.float32)
        r2 = np.random.default_rng(1000 + i)
        G = np.exp(-(s.XX.astype(np.float64) ** 2 + s.YY.astype(np.float64) ** 2) / waist**2)
        noise = r2.standard_normal(G.shape) + 1j * r2.standard_normal(G.shape)
        sims.append(s)
        fields.append((G * (1 + 0.01 * noise)).astype(np.complex64))
    return sims, fields


SHAPES = [(256, 256), (128, 128), (512, 512), (256, 128), (128, 512), (512, 256), (256, 256)]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_batch_equals_separate_runs(gpu, precision):
    from nlse_amd import NLSE, NLSEBatch

    sims, fields = _instances(NLSE, SHAPES)
    z = 40.5e-5  # 41 steps
    outs = NLSEBatch(sims).out_field(fields, z, precision=precision)
    for i, (s, E, o) in enumerate(zip(sims, fields, outs)):
        ref = s.out_field(E, z, precision=precision, verbose=False)
        err = rel_l2(o, ref)
        print(f"field {i} {s.NY}x{s.NX}: vs separate run {err:.2e} identical={np.array_equal(o, ref)}")
        assert o.shape == E.shape and err < 1e-6


def test_batch_matches_oracle_and_graph(gpu):
    from nlse_amd import NLSE, NLSEBatch

    sims, fields = _instances(NLSE, SHAPES[:4], seed=5)
    z = 99.5e-5  # 100 steps
    b = NLSEBatch(sims)
    b.graph_chunk = 8
    outs = b.out_field(fields, z)
    b.graph_chunk = 0
    outs0 = b.out_field(fields, z)
    for i, (s, E, o, o0) in enumerate(zip(sims, fields, outs, outs0)):
        assert np.array_equal(o, o0), f"field {i}: graph replay differs from direct launches"
        kw = dict(alpha=20, power=s.power, window=window, n2=s.n2, V=s.V, L=10e-3, NX=s.NX, NY=s.NY, Isat=s.I_sat)
        o_ref = oracle.NLSE(**kw)
        o_ref.delta_z = 1e-5
        err = rel_l2(o, o_ref.out_field(E, z))
        print(f"field {i} {s.NY}x{s.NX} vs oracle: {err:.2e}")
        assert err < 1e-5


def test_batch_config_balanced_units(gpu):
    """Every row unit holds 4096 points and every column unit 8192, whatever
    the grid, so the persistent grids stride over equal-cost units."""
    from nlse_amd import NLSE
    from nlse_amd.batch import NLSEBatch

    sims, fields = _instances(NLSE, SHAPES)
    eng, buf, n = NLSEBatch(sims).prepare(fields, 1e-5)
    cfg = eng.config()
    P = sum(ny * nx for ny, nx in SHAPES)
    assert cfg["row_units"] * 4096 == P and cfg["col_units"] * 8192 == P
    assert 0 < cfg["row_grid"] <= cfg["row_units"] and 0 < cfg["col_grid"] <= cfg["col_units"]
