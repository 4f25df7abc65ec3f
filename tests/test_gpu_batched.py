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
    hit = autotune.tune(e, law, "fp32", candidates=cand, max_rel_err=1e-5, cache=cache)
    assert hit["tuning"] == r["tuning"] and "table" not in hit and hit["cached"]
    assert hit["rel_err"] <= 1e-5      # admitted again on this state
    # the key covers every system, the law and the threshold: a batch that
    # differs only in its last system, another law or another threshold misses
    import types
    k0 = autotune._key(e, "fp32", law, 1e-5)
    n2 = sizes.copy()
    n2[-1] += 1
    other = types.SimpleNamespace(n=n2, grid_dim=grids, dev=e.dev)
    assert autotune._key(other, "fp32", law, 1e-5) != k0
    assert autotune._key(e, "fp32", (1e-6, 0.0, 0.0), 1e-5) != k0
    assert autotune._key(e, "fp32", law, 1e-6) != k0
    assert autotune._key(types.SimpleNamespace(n=sizes.copy(), grid_dim=grids.copy(), dev=e.dev),
                         "fp32", law, 1e-5) == k0
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


def test_full_size_512_systems_sampled_against_oracle():
    """The 512-system batch at full size (27.5M particles, one graph step of
    the bench's line): for the smallest, a middle and the largest system,
    placement keys and counts bit-exact against the oracle, rows sorted in
    the reference's order, COMs and sampled FP32 accelerations within
    tolerance."""
    from barneshut.batched import BATCHED_TUNING
    sizes, grids = inputs.batched_sizes(512, 1024, 262144, seed=0)
    data = inputs.batched_uniform(sizes, seed=0)
    law = (1e-6, 1e-12, 0.0)
    b = BatchedSystems(sizes, grids, data, law, "fp32", 0.0, tuning=BATCHED_TUNING, graph=True)
    e = b.eng
    e.create_tree()
    e.summarize()
    e.evaluate(*law, "fp32")
    e.check()
    g = e.fetch()
    keys = e.keys()
    keys_in = e.keys_in()
    coff = np.concatenate([[0], np.cumsum(grids.astype(np.int64) ** 2)])
    slices = b.system_slices()
    order = np.argsort(sizes, kind="stable")
    rng = np.random.default_rng(8)
    for s in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        sl, G = slices[s], int(grids[s])
        n = sl.stop - sl.start
        px, py = data["px"][sl], data["py"][sl]
        bb = orc.bbox(px, py, G)
        _, _, key = orc.place(px, py, bb[0], bb[1], bb[4], G)
        np.testing.assert_array_equal(keys_in[sl] - coff[s], key)
        ks = keys[sl] - coff[s]
        assert (np.diff(ks) >= 0).all()
        cnt, off = orc.counts_offsets(ks, G * G)
        np.testing.assert_array_equal(cnt, np.bincount(key, minlength=G * G))
        x, y, m = g["px"][sl], g["py"][sl], g["mass"][sl]
        com = orc.summarize(off, x, y, m)
        np.testing.assert_allclose(e.com().reshape(-1, 3)[coff[s]:coff[s + 1]], com,
                                   rtol=1e-12, atol=1e-15)
        idx = np.sort(rng.choice(n, min(n, 512), replace=False))
        rx, ry = orc.forces(G, off, ks, x, y, m, com, *law, targets=idx)
        num = np.hypot(g["ax"][sl][idx] - rx, g["ay"][sl][idx] - ry)
        norm = float(np.sqrt((num ** 2).sum() / (np.hypot(rx, ry) ** 2).sum()))
        print(f"[512 systems] system {s}: n={n} G={G} sampled normwise {norm:.3e}", flush=True)
        assert norm <= 1e-5
    b.close()
