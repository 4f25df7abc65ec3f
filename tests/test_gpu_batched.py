"""Batched independent systems: one context, one queue, one graph per step.
Graph replay (static and learned order) must equal eager stepping bit for
bit, and each system must match its own oracle result."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

from oracle import oracle as orc  # noqa: E402
from barneshut import costmodel, inputs  # noqa: E402
from barneshut.batched import BatchedSystems  # noqa: E402


def _setup(k=12, seed=3):
    sizes, grids = inputs.batched_sizes(k, 1024, 30000, seed=seed)
    data = inputs.batched_uniform(sizes, seed=seed)
    return sizes, grids, data


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replay_equals_eager(precision):
    sizes, grids, data = _setup()
    law = (1e-6, 1e-8, 0.0)
    a = BatchedSystems(sizes, grids, data, law, precision, 1e-3, graph=False)
    b = BatchedSystems(sizes, grids, data, law, precision, 1e-3, graph=True)
    for _ in range(4):
        a.step()
        b.step()
        ha, hb = a.eng.fetch(), b.eng.fetch()
        for k in ("px", "py", "vx", "vy", "pid"):
            np.testing.assert_array_equal(ha[k], hb[k])
    a.close()
    b.close()


def test_learned_graph_step_equals_static():
    sizes, grids, data = _setup(k=16, seed=4)
    law = (1e-6, 1e-8, 0.0)
    model = costmodel.CostModel(costmodel.init_params(2), device="cuda")
    a = BatchedSystems(sizes, grids, data, law, "fp32", 1e-3, graph=True)
    b = BatchedSystems(sizes, grids, data, law, "fp32", 1e-3, graph=True, model=model)
    for _ in range(3):
        a.step()
        b.step()
    ha, hb = a.eng.fetch(), b.eng.fetch()
    for k in ("px", "py", "vx", "vy", "pid"):
        np.testing.assert_array_equal(ha[k], hb[k])
    a.close()
    b.close()


def test_each_system_matches_oracle_fp32():
    sizes, grids, data = _setup(k=6, seed=5)
    b = BatchedSystems(sizes, grids, data, (1.0, 1e-8, 0.0), "fp32", 0.0, graph=True)
    e = b.eng
    e.create_tree()
    e.summarize()
    e.evaluate(1.0, 1e-8, 0.0, "fp32")
    g = e.fetch()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, sl in enumerate(b.system_slices()):
        G = int(grids[s])
        ks = e.keys()[sl] - int((grids[:s].astype(np.int64) ** 2).sum())
        _, o = orc.counts_offsets(ks, G * G)
        com = orc.summarize(o, g["px"][sl], g["py"][sl], g["mass"][sl])
        rx, ry = orc.forces(G, o, ks, g["px"][sl], g["py"][sl], g["mass"][sl], com, 1.0, 1e-8, 0.0)
        num = np.hypot(g["ax"][sl] - rx, g["ay"][sl] - ry)
        den = np.hypot(rx, ry)
        assert np.sqrt((num ** 2).sum() / (den ** 2).sum()) < 1e-5, s
    assert off[-1] == e.N
    b.close()
