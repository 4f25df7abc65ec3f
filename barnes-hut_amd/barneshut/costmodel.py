"""Learned per-item cost model for the force kernel's work queue.

The model is an MLP 8 -> 32 -> 32 -> 1 (FP32, ReLU) over the item features
of bh_item_features (particles, lane fill, near-field rows, masked COM
entries, far-field length K, row span, neighbour mass, cells).  It is trained
on the GPU with the library's own forward/backward/Adam kernels
(bh_mlp_forward / bh_mlp_backward / bh_mlp_adam) to predict the log of each
item's measured duration, logged by the force kernel itself
(bh_set_cost_log).  The queue is then ordered longest-predicted-first
(bh_learned_order); a non-finite prediction falls back to the static order.
The order changes only which wave runs which item, not any particle's sum.

Training standardises features and targets; the saved parameters have the
feature standardisation folded into the first layer, so the device-side
inference (bh_learned_order) runs on raw features.
"""
from __future__ import annotations

import numpy as np

from .kernels import hiplib

NF, NH = hiplib.MLP_NFEAT, 32
OFF_W1, OFF_B1 = 0, NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1
assert NPARAM == hiplib.MLP_NPARAM


def init_params(seed=0):
    """He-style initialisation (numpy float32 vector of NPARAM)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / NF), NH * NF)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    p[OFF_B1:OFF_W2] = 0.01
    p[OFF_B2:OFF_W3] = 0.01
    return p


def fold_standardisation(params, mu, sd):
    """Parameters for raw features x given parameters trained on (x - mu) / sd."""
    q = np.array(params, dtype=np.float64)
    W1 = q[OFF_W1:OFF_B1].reshape(NH, NF)
    b1 = q[OFF_B1:OFF_W2]
    W1f = W1 / sd[None, :]
    b1f = b1 - W1f @ mu
    q[OFF_W1:OFF_B1] = W1f.ravel()
    q[OFF_B1:OFF_W2] = b1f
    return q.astype(np.float32)


class CostModel:
    """Device parameters (folded) plus the training statistics."""

    def __init__(self, params, mu=None, sd=None, device=None, meta=None):
        import torch
        self.raw = np.asarray(params, np.float32)
        self.mu = np.zeros(NF) if mu is None else np.asarray(mu, np.float64)
        self.sd = np.ones(NF) if sd is None else np.asarray(sd, np.float64)
        self.meta = meta or {}
        folded = fold_standardisation(self.raw, self.mu, self.sd)
        self.params = torch.as_tensor(folded, device=device)

    def save(self, path):
        np.savez(path, params=self.raw, mu=self.mu, sd=self.sd,
                 meta=np.array(repr(self.meta)))

    @staticmethod
    def load(path, device=None):
        z = np.load(path, allow_pickle=False)
        return CostModel(z["params"], z["mu"], z["sd"], device=device)

    def predict(self, feat):
        """feat: device float32 [n, 8] raw features -> device float32 [n]."""
        import torch
        y = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
        hiplib.mlp_forward(self.params, feat.contiguous(), y)
        return y


def collect(eng, steps, law, precision="fp32", dt=0.0, order_fn=None, log=None):
    """Run `steps` full steps on `eng` with per-item duration logging.
    Returns (features [m, 8] float32, durations [m] int64) on the host."""
    import torch
    feats, costs = [], []
    cap = eng.ctx.query(4)
    saved = eng.ctx.get_tuning()
    eng.ctx.set_tuning(**dict(saved, tail_split=1, tail_items=0))   # durations of whole items
    fbuf = torch.empty(cap * NF, dtype=torch.float32, device=eng.dev)
    for s in range(steps):
        eng.create_tree()
        eng.summarize()
        eng.ctx.build_items()
        n = int(eng.ctx.stats()[hiplib.ST_NITEMS])
        eng.ctx.item_features(fbuf)
        order = order_fn() if order_fn else None
        eng.ctx.set_cost_log(True)
        eng.ctx.force_pass(eng.parts, *law, precision, order)
        eng.ctx.set_cost_log(False)
        c = eng.ctx.export_host(hiplib.EXPORT_COST, n, np.int64)
        feats.append(fbuf[: n * NF].view(n, NF).cpu().numpy().copy())
        costs.append(c)
        if dt:
            eng.timestep(dt)
        if log:
            log(f"collect step {s}: {n} items, median {np.median(c)} ticks")
    eng.ctx.set_tuning(**saved)
    return np.concatenate(feats), np.concatenate(costs)


def train(X, durations, epochs=60, batch=65536, lr=3e-3, seed=0, device="cuda", log=None):
    """Fit log(duration) with the library's GPU kernels (MSE, Adam).
    Returns (CostModel, history of (epoch, loss))."""
    import torch
    X = np.asarray(X, np.float64)
    y = np.log(np.maximum(np.asarray(durations, np.float64), 1.0))
    mu = X.mean(0)
    sd = X.std(0)
    sd[sd < 1e-6] = 1.0
    ymu, ysd = y.mean(), max(y.std(), 1e-6)
    Xs = torch.as_tensor(((X - mu) / sd).astype(np.float32), device=device)
    ys = torch.as_tensor(((y - ymu) / ysd).astype(np.float32), device=device)
    params = torch.as_tensor(init_params(seed), device=device)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    grad = torch.zeros_like(params)
    loss = torch.zeros(1, dtype=torch.float32, device=device)
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    n = Xs.shape[0]
    t = 0
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(device)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            xb = Xs[idx].contiguous()
            yb = ys[idx].contiguous()
            hiplib.mlp_backward(params, xb, yb, grad, loss)
            t += 1
            hiplib.mlp_adam(params, grad, m, v, t, lr)
            tot += float(loss.item()) * len(idx)
        hist.append((ep, tot / n))
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log(f"train epoch {ep}: mse {tot / n:.4f}")
    model = CostModel(params.cpu().numpy(), mu, sd, device=device,
                      meta=dict(ymu=float(ymu), ysd=float(ysd), rows=int(n), epochs=epochs))
    return model, hist


def rank_quality(pred, durations):
    """Spearman-like check: correlation of predicted and measured log costs."""
    a = np.asarray(pred, np.float64)
    b = np.log(np.maximum(np.asarray(durations, np.float64), 1.0))
    ra = np.argsort(np.argsort(a))
    rb = np.argsort(np.argsort(b))
    return float(np.corrcoef(ra, rb)[0, 1])
