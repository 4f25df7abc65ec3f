"""The allocation-free step (bh_step_noalloc, all scratch in the library's
static per-GPU array) against bh_step on a context with the same sizes,
tuning and recompute threshold: the same kernels, so the sorted rows,
accelerations, ticked positions and velocities must agree bit for bit.
Several calls in a row share the static array; each later call is smaller
than the one before (or changes precision and law), so state left over from
an earlier call would show."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

from barneshut import inputs  # noqa: E402

TUNING = dict(p=2, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0)
FIELDS = ("px", "py", "mass", "vx", "vy", "ax", "ay", "pid")


def _with_ctx(d, G, law, prec, dt, kappa):
    from barneshut.engine import Engine
    n = len(d["px"])
    e = Engine([n], [G], tuning=TUNING)
    e.ctx.set_recompute(kappa)
    e.load(d)
    e.step(*law, prec, dt)
    st = e.ctx.stats()
    h = e.fetch()
    e.close()
    return h, st


def _noalloc(d, G, law, prec, dt, kappa):
    import torch
    from barneshut.engine import alloc_particles
    from barneshut.kernels import hiplib
    n = len(d["px"])
    dev = torch.device("cuda:0")
    a, b = alloc_particles(n, dev), alloc_particles(n, dev)
    for k in FIELDS:
        src = d.get(k, np.zeros(n))
        a[k].copy_(torch.as_tensor(np.ascontiguousarray(
            src, dtype=np.int32 if k == "pid" else np.float64)))
    b["ax"].fill_(123.0)   # garbage in the output buffer: the step must not read it
    stats = torch.full((16,), -7, dtype=torch.int32, device=dev)
    hiplib.step_noalloc(a, b, G, *law, prec, dt, True, None, kappa, stats)
    torch.cuda.synchronize()
    return {k: b[k].cpu().numpy() for k in FIELDS}, stats.cpu().numpy()


CALLS = [
    # n, G, clusters (0: uniform), law (G, eps2, r2_cut), precision, dt, kappa
    (200000, 112, 24, (1e-6, 1e-10, 0.0), "fp32", 1e-5, 2.0),
    (30000, 40, 6, (1.0, 1e-6, 1e-5), "fp32", 1e-4, 2.0),     # smaller, cutoff law
    (9000, 24, 0, (1.0, 1e-6, 0.0), "fp64", 1e-4, 2.0),
    (4096, 16, 3, (2.5, 1e-6, 0.0), "fp32_kahan_term", 1e-4, 0.0),
]


def test_noalloc_matches_bh_step_on_calls_in_a_row():
    for i, (n, G, nc, law, prec, dt, kappa) in enumerate(CALLS):
        print(f"[noalloc] call {i}: n={n} G={G} {prec} law={law} kappa={kappa}", flush=True)
        d = (inputs.gaussian_mixture(n, nc, (0.01, 0.1), seed=500 + i) if nc
             else inputs.uniform(n, seed=500 + i))
        ref, st_ref = _with_ctx(d, G, law, prec, dt, kappa)
        got, st = _noalloc(d, G, law, prec, dt, kappa)
        for k in FIELDS:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"call {i} field {k}")
        assert st[0] == 0 and st_ref[0] == 0
        assert st[2] == st_ref[2]                        # work items
        assert st[6] == st_ref[6]                        # particles recomputed in FP64
        print(f"[noalloc] call {i}: bit-identical, {st[2]} items, {st[6]} recomputed", flush=True)


def test_noalloc_at_its_limits_and_refusals():
    """The largest system the static array holds (n = 262144 particles on a
    256 x 256 grid) matches bh_step; one particle or one grid row more, or a
    split tuning, is refused with an error before any launch."""
    import torch
    from barneshut.engine import alloc_particles
    from barneshut.kernels import hiplib
    from barneshut.kernels.hiplib import BhError
    n, G = hiplib.NOALLOC_MAX_N, hiplib.NOALLOC_MAX_GRID
    d = inputs.gaussian_mixture(n, 16, (0.005, 0.05), seed=520)
    law = (1e-6, 1e-10, 0.0)
    ref, _ = _with_ctx(d, G, law, "fp32", 1e-5, 2.0)
    got, st = _noalloc(d, G, law, "fp32", 1e-5, 2.0)
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert st[0] == 0
    dev = torch.device("cuda:0")
    big = alloc_particles(n + 1, dev), alloc_particles(n + 1, dev)
    with pytest.raises(BhError, match="static scratch"):
        hiplib.step_noalloc(big[0], big[1], G, *law, "fp32", 1e-5)
    small = alloc_particles(1000, dev), alloc_particles(1000, dev)
    with pytest.raises(BhError, match="static scratch"):
        hiplib.step_noalloc(small[0], small[1], G + 1, *law, "fp32", 1e-5)
    with pytest.raises(BhError, match="tail_split"):
        hiplib.step_noalloc(small[0], small[1], 8, *law, "fp32", 1e-5,
                            tuning=dict(TUNING, tail_split=4))
