"""Cost-model kernels (MLP forward/backward/Adam, cost ordering, item
features) against the FP64 numpy oracle, and the learned schedule's effect:
identical physics, a trained model that ranks item costs."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

from oracle import mlp as omlp  # noqa: E402
from barneshut import costmodel, inputs  # noqa: E402
from barneshut.kernels import hiplib  # noqa: E402


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32 if dtype is None else dtype),
                           device="cuda")


def test_mlp_forward_matches_oracle():
    import torch
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, omlp.NPARAM).astype(np.float32)
    X = rng.normal(size=(3000, 8)).astype(np.float32)
    y = torch.empty(3000, dtype=torch.float32, device="cuda")
    hiplib.mlp_forward(_t(p), _t(X), y)
    ref = omlp.forward(p.astype(np.float64), X.astype(np.float64))
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=2e-5, atol=2e-5)


def test_mlp_backward_matches_oracle():
    import torch
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.3, omlp.NPARAM).astype(np.float32)
    X = rng.normal(size=(5000, 8)).astype(np.float32)
    t = rng.normal(size=5000).astype(np.float32)
    g = torch.empty(omlp.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    hiplib.mlp_backward(_t(p), _t(X), _t(t), g, loss)
    pd = p.astype(np.float64)
    ref = omlp.backward(pd, X.astype(np.float64), t.astype(np.float64))
    gg = g.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(gg - ref).max() <= 1e-4 * scale
    assert abs(float(loss.item()) - omlp.loss(pd, X, t)) <= 1e-4 * omlp.loss(pd, X, t)


def test_mlp_adam_matches_oracle():
    rng = np.random.default_rng(3)
    p = rng.normal(size=omlp.NPARAM).astype(np.float32)
    g = rng.normal(size=omlp.NPARAM).astype(np.float32)
    m = rng.normal(size=omlp.NPARAM).astype(np.float32) * 0.1
    v = np.abs(rng.normal(size=omlp.NPARAM)).astype(np.float32) * 0.1
    dp, dm, dv = _t(p), _t(m), _t(v)
    hiplib.mlp_adam(dp, _t(g), dm, dv, 5, 1e-2)
    rp, rm, rv = omlp.adam(p.astype(np.float64), g.astype(np.float64), m.astype(np.float64),
                           v.astype(np.float64), 5, 1e-2)
    np.testing.assert_allclose(dp.cpu().numpy(), rp, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dm.cpu().numpy(), rm, rtol=1e-5, atol=1e-7)


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(4)
    X = rng.normal(size=(20000, 8))
    cost = np.exp(1.0 + 0.8 * X[:, 0] - 0.5 * np.maximum(X[:, 1], 0) + 0.3 * X[:, 2] * X[:, 3])
    model, hist = costmodel.train(X, cost, epochs=40, batch=4096, lr=3e-3)
    assert hist[-1][1] < 0.1 * hist[0][1]
    pred = model.predict(_t(X)).cpu().numpy()
    assert costmodel.rank_quality(pred, cost) > 0.95
    # the folded device parameters equal the oracle forward on standardised inputs
    ref = omlp.forward(model.raw.astype(np.float64), (X - model.mu) / model.sd)
    np.testing.assert_allclose(pred, ref, rtol=1e-3, atol=1e-3)


def _engine(d, G, tuning=None):
    from barneshut.engine import Engine
    e = Engine([len(d["px"])], [G], tuning=tuning)
    e.load(d)
    e.create_tree()
    e.summarize()
    e.ctx.build_items()
    return e


def test_order_by_cost_is_descending_permutation():
    import torch
    d = inputs.gaussian_mixture(50000, 8, (0.01, 0.05), seed=5)
    e = _engine(d, 64, dict(p=1, yspan=8, waves_per_cu=8))
    n = int(e.ctx.stats()[hiplib.ST_NITEMS])
    rng = np.random.default_rng(6)
    cap = e.ctx.query(4)   # buffers are sized for the work-item capacity
    cost = np.zeros(cap, np.float32)
    cost[:n] = rng.normal(size=n)
    order = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    e.ctx.order_by_cost(_t(cost), order)
    o = order[:n].cpu().numpy()
    assert sorted(o.tolist()) == list(range(n))
    c = cost[o]
    width = (cost[:n].max() - cost[:n].min()) / 4095
    assert (np.diff(c) <= width * 1.001).all()
    # a non-finite prediction falls back to the static order
    cost[n // 2] = np.nan
    e.ctx.order_by_cost(_t(cost), order)
    np.testing.assert_array_equal(order[:n].cpu().numpy(), np.arange(n))
    e.close()


def test_item_features_match_items():
    import torch
    d = inputs.gaussian_mixture(40000, 8, (0.01, 0.05), seed=7)
    e = _engine(d, 48, dict(p=2, yspan=16, waves_per_cu=8))
    items = e.items()
    n = len(items)
    feat = torch.empty(e.ctx.query(4) * 8, dtype=torch.float32, device="cuda")
    e.ctx.item_features(feat)
    f = feat[: n * 8].view(n, 8).cpu().numpy()
    np.testing.assert_allclose(f[:, 0], np.log1p(items[:, 1]), rtol=1e-6)
    np.testing.assert_allclose(f[:, 1], items[:, 1] / 128.0, rtol=1e-6)
    K = int((e.counts() > 0).sum())
    np.testing.assert_allclose(f[:, 4], np.log1p(K), rtol=1e-6)
    np.testing.assert_allclose(f[:, 5], np.log1p(items[:, 5] - items[:, 4] + 1), rtol=1e-6)
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_learned_order_leaves_physics_unchanged(precision):
    import torch
    d = inputs.gaussian_mixture(60000, 16, (0.005, 0.05), seed=8)
    e = _engine(d, 96)
    e.ctx.force_pass(e.parts, 1.0, 1e-8, 0.0, precision)
    ref = e.fetch()
    model = costmodel.CostModel(costmodel.init_params(1), device="cuda")
    order = torch.empty(e.ctx.query(4), dtype=torch.int32, device="cuda")
    e.ctx.learned_order(model.params, order)
    n = int(e.ctx.stats()[hiplib.ST_NITEMS])
    assert sorted(order[:n].cpu().tolist()) == list(range(n))
    e.parts["ax"].zero_()
    e.parts["ay"].zero_()
    e.ctx.force_pass(e.parts, 1.0, 1e-8, 0.0, precision, order)
    got = e.fetch()
    np.testing.assert_array_equal(got["ax"], ref["ax"])
    np.testing.assert_array_equal(got["ay"], ref["ay"])
    e.close()


def test_collect_train_rank_on_logged_costs():
    from barneshut.engine import Engine
    d = inputs.gaussian_mixture(200000, 32, (0.002, 0.05), seed=9, mass="lognormal")
    e = Engine([len(d["px"])], [256], tuning=dict(p=2, yspan=32, waves_per_cu=8))
    e.load(d)
    X, c = costmodel.collect(e, 4, (1e-6, 1e-10, 0.0), "fp32", dt=1e-3)
    assert len(X) == len(c) and (c > 0).all()
    model, hist = costmodel.train(X, c, epochs=30, batch=8192, lr=3e-3)
    assert hist[-1][1] < hist[0][1]
    pred = model.predict(_t(X)).cpu().numpy()
    assert np.isfinite(pred).all()
    print(f"rank quality on logged items: {costmodel.rank_quality(pred, c):.3f}", flush=True)
    e.close()


def _batched_engine(sizes, seed):
    from barneshut.batched import BATCHED_TUNING
    from barneshut.engine import Engine
    d = inputs.batched_uniform(sizes, seed=seed)
    grids = np.ceil(np.sqrt(np.asarray(sizes) / 16.0)).astype(np.int32)
    e = Engine(sizes, grids, tuning=BATCHED_TUNING)
    e.load(d)
    return e


def _batched_model():
    import os
    import torch
    from conftest import PKG
    return costmodel.CostModel.load(os.path.join(PKG, "costmodels", "batched512.npz"),
                                    device=torch.device("cuda"))


def test_batched_model_ranks_logged_item_costs():
    """The shipped batched model against item durations the force kernel logs
    on a batch it was not trained on (64 uneven systems, another seed): the
    rank correlation of predicted and measured costs must clear 0.4 (a model
    that predicts noise scores ~0; measured 0.59 on MI355X).  Items of one system share their
    far-field length K, so the signal is K across systems plus the near field
    within them."""
    sizes, _ = inputs.batched_sizes(64, 1024, 262144, seed=77)
    e = _batched_engine(sizes, 77)
    X, c = costmodel.collect(e, 2, (1e-6, 1e-12, 0.0), "fp32", dt=1e-5)
    model = _batched_model()
    pred = model.predict(_t(X)).cpu().numpy()
    rq = costmodel.rank_quality(pred, c)
    print(f"[rank] {len(c)} logged items, rank correlation {rq:.3f}", flush=True)
    assert rq >= 0.4
    e.close()


def test_trained_model_gain_on_small_batches():
    """Where the queue is short against the resident waves -- batches of 32
    systems of uneven size, a few milliseconds of force work -- the order of
    the items decides the tail, and the shipped batched model's
    longest-first order pays off: measured 1.003, 1.047 and 1.072 (geometric
    mean 1.041) on MI355X, 1.012, 1.031 and 1.108 on another box.  The
    geometric mean must reach 1.01.  (Elsewhere the order gains nothing
    measurable: with 8 waves per SIMD the last long items run faster as the
    other waves drain, so even a batch sorted smallest-first loses < 1 % in
    the static order.)"""
    import os
    import torch
    from conftest import PKG
    from barneshut.batched import BATCHED_TUNING
    from barneshut.engine import Engine
    model = costmodel.CostModel.load(os.path.join(PKG, "costmodels", "batched512.npz"),
                                     device=torch.device("cuda"))
    gains = []
    for seed in (0, 1, 2):
        sizes, grids = inputs.batched_sizes(32, 1024, 262144, seed=seed)
        d = inputs.batched_uniform(sizes, seed=seed)
        e = Engine(sizes, grids, tuning=BATCHED_TUNING)
        e.load(d)
        e.create_tree()
        e.summarize()
        e.ctx.build_items()
        order = torch.empty(e.ctx.query(4), dtype=torch.int32, device="cuda")
        e.ctx.learned_order(model.params, order)
        t = {"static": [], "learned": []}
        for _ in range(5):
            for name, o in (("static", None), ("learned", order)):
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                e.ctx.force_pass(e.parts, 1e-6, 1e-12, 0.0, "fp32", o)
                b.record()
                b.synchronize()
                t[name].append(a.elapsed_time(b))
        s, l = min(t["static"]), min(t["learned"])
        gains.append(s / l)
        print(f"[small batch seed={seed} N={int(sizes.sum())}] static {s:.3f} ms, "
              f"learned {l:.3f} ms, gain {s / l:.3f}", flush=True)
        e.close()
    gm = float(np.exp(np.mean(np.log(gains))))
    print(f"[small batches] geometric-mean gain {gm:.3f}", flush=True)
    assert gm >= 1.01
