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


def test_autotuner_picks_admissible_shape_and_caches(tmp_path):
    """The autotuner times every candidate on the loaded state, admits only
    FP32 shapes within max_rel_err of an FP64 pass, sets the fastest, and
    answers the same request from its cache."""
    from barneshut import autotune
    from barneshut.engine import Engine
    sizes, grids = inputs.batched_sizes(24, 1024, 65536, seed=4)
    data = inputs.batched_uniform(sizes, seed=4)
    e = Engine(sizes, grids)
    e.load(data)
    e.create_tree()
    e.summarize()
    cand = dict(p=(1, 2), yspan=(64, 512), waves_per_cu=(16,), tail_split=(1, 8))
    cache = str(tmp_path / "tune.json")
    law = (1e-6, 1e-12, 0.0)
    r = autotune.tune(e, law, "fp32", candidates=cand, max_rel_err=1e-5, cache=cache)
    assert len(r["table"]) == 8
    assert all(row["rel_err"] <= 1e-5 for row in r["table"])
    best = min(r["table"], key=lambda row: row["ms"])
    assert r["tuning"] == best["tuning"]
    assert e.ctx.get_tuning() == r["tuning"]
    print(f"[autotune] best {r['tuning']} {r['ms']:.3f} ms", flush=True)
    hit = autotune.tune(e, law, "fp32", candidates=cand, cache=cache)
    assert hit["tuning"] == r["tuning"] and "table" not in hit
    with pytest.raises(RuntimeError):     # an impossible error bound admits nothing
        autotune.tune(e, law, "fp32", candidates=cand, max_rel_err=1e-15, use_cache=False,
                      cache=None)
    e.close()


def test_batched_systems_autotuned_step_matches_eager():
    """BatchedSystems(tuning="auto") runs with the tuner's shape, and its graph
    steps equal eager steps under that shape."""
    sizes, grids = inputs.batched_sizes(16, 1024, 32768, seed=5)
    data = inputs.batched_uniform(sizes, seed=5)
    law = (1e-6, 1e-8, 0.0)
    cand = dict(p=(1, 2), yspan=(512,), waves_per_cu=(16,), tail_split=(1,))
    b = BatchedSystems(sizes, grids, data, law, "fp32", 1e-4, tuning="auto",
                       autotune_candidates=cand, graph=True)
    tun = b.eng.ctx.get_tuning()
    assert tun == b.tune_result["tuning"]
    e = BatchedSystems(sizes, grids, data, law, "fp32", 1e-4, tuning=tun, graph=False)
    for _ in range(2):
        b.step()
        e.step()
    g1, g2 = b.eng.fetch(), e.eng.fetch()
    for k in ("px", "py", "vx", "vy", "pid"):
        np.testing.assert_array_equal(g1[k], g2[k])
    b.close()
    e.close()
