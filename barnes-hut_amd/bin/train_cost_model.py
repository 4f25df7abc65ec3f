#!/usr/bin/env python3
"""Trains the learned per-item cost model on logged steps and measures what
it gains over the static queue order.

    train_cost_model.py --workload {clusters256|batched512} --steps 200 --out model.npz

clusters256: N = 4,194,304, 256 Gaussian clusters (sigma log-uniform
             [0.001, 0.03]), masses lognormal(0, 0.5), 1024 x 1024 grid.
batched512:  512 independent uniform systems, N log-uniform [1024, 262144],
             grid side ceil(sqrt(N/16)), masses U[1,2].
Both: FP32 cell-relative forces, G = 1e-6, dt = 1e-3, eps2 = 1e-12.

Every logged step runs with the static order and records each work item's
duration (bh_set_cost_log); the MLP is trained on all but the last 10 % of
the steps with the library's GPU kernels and scored on the rest (rank
correlation).  The A/B then alternates static and learned order on the same
states for --ab-steps steps (force kernel alone, HIP events), and for the
batched workload also whole graph-replayed steps.  One JSON report line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))

from barneshut import costmodel, inputs  # noqa: E402
from barneshut.engine import Engine  # noqa: E402

LAW = (1e-6, 1e-12, 0.0)
DT = 1e-3


def log(m):
    print(f"[train] {m}", file=sys.stderr, flush=True)


def workload(name, seed):
    if name == "clusters256":
        d = inputs.gaussian_mixture(4194304, 256, (0.001, 0.03), seed=seed, mass="lognormal")
        return [len(d["px"])], [1024], d
    if name == "batched512":
        n, g = inputs.batched_sizes(512, 1024, 262144, seed=seed)
        return n, g, inputs.batched_uniform(n, seed=seed)
    raise SystemExit(f"unknown workload {name}")


def ab_force(eng, model, steps):
    """Alternating static / learned force passes on identical states."""
    import torch
    order = torch.empty(eng.ctx.query(4), dtype=torch.int32, device=eng.dev)
    t_static, t_learned, t_order = [], [], []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for s in range(steps):
        eng.create_tree()
        eng.summarize()
        eng.ctx.build_items()
        runs = ("static", "learned") if s % 2 == 0 else ("learned", "static")
        for r in runs:
            eng.parts["ax"].zero_()
            eng.parts["ay"].zero_()
            if r == "learned":
                ev[2].record()
                eng.ctx.learned_order(model.params, order)
                ev[3].record()
            ev[0].record()
            eng.ctx.force_pass(eng.parts, *LAW, "fp32", order if r == "learned" else None)
            ev[1].record()
            torch.cuda.synchronize()
            (t_learned if r == "learned" else t_static).append(ev[0].elapsed_time(ev[1]))
            if r == "learned":
                t_order.append(ev[2].elapsed_time(ev[3]))
        eng.timestep(DT)
    eng.check()
    return dict(static_ms=float(np.mean(t_static)), learned_ms=float(np.mean(t_learned)),
                order_ms=float(np.mean(t_order)),
                gain=float(np.mean(t_static) / (np.mean(t_learned) + np.mean(t_order))))


def ab_graph(sizes, grids, d, model, steps):
    from barneshut.batched import BatchedSystems
    res = {}
    for name, m in (("static", None), ("learned", model)):
        b = BatchedSystems(sizes, grids, d, LAW, "fp32", DT, graph=True, model=m)
        res[name + "_ms_per_step"] = b.run(steps, warmup=1) * 1e3
        b.close()
    res["gain"] = res["static_ms_per_step"] / res["learned_ms_per_step"]
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="clusters256")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--ab-steps", type=int, default=6)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    sizes, grids, d = workload(a.workload, a.seed)
    from barneshut.batched import BATCHED_TUNING
    # log and compare with the launch shape each workload runs with
    eng = Engine(sizes, grids, tuning=BATCHED_TUNING if len(sizes) > 1 else None)
    eng.load(d)
    t0 = time.perf_counter()
    X, c = costmodel.collect(eng, a.steps, LAW, "fp32", dt=DT,
                             log=lambda m: None)
    t_collect = time.perf_counter() - t0
    per_step = len(X) // a.steps
    n_train = per_step * int(a.steps * 0.9)
    log(f"logged {len(X)} items over {a.steps} steps in {t_collect:.1f}s")
    t0 = time.perf_counter()
    model, hist = costmodel.train(X[:n_train], c[:n_train], epochs=a.epochs, log=log)
    t_train = time.perf_counter() - t0
    import torch
    pred = model.predict(torch.as_tensor(X[n_train:].astype(np.float32), device=eng.dev))
    rq = costmodel.rank_quality(pred.cpu().numpy(), c[n_train:])
    model.meta.update(workload=a.workload, steps=a.steps)
    model.save(a.out)
    log(f"trained: mse {hist[-1][1]:.4f}, held-out rank correlation {rq:.3f}")
    ab = ab_force(eng, model, a.ab_steps)
    log(f"A/B force kernel: static {ab['static_ms']:.2f} ms, learned {ab['learned_ms']:.2f} ms "
        f"+ order {ab['order_ms']:.3f} ms")
    eng.close()
    rep = dict(workload=a.workload, logged_steps=a.steps, logged_items=int(len(X)),
               items_per_step=int(per_step), collect_s=t_collect, train_s=t_train,
               epochs=a.epochs, final_mse=hist[-1][1], first_mse=hist[0][1],
               heldout_rank_corr=rq, duration_ticks_median=float(np.median(c)),
               duration_ticks_p99=float(np.quantile(c, 0.99)), ab_force=ab, model=a.out)
    if len(sizes) > 1:
        rep["ab_graph"] = ab_graph(sizes, grids, d, model, a.ab_steps)
        log(f"A/B graph steps: {rep['ab_graph']}")
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
