"""The drop-in: python-bh's BarnesHut(ini) / run() / run_2d.py interface
driving the HIP implementation, the --check harness, pid-ordered results,
and the device error path."""
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

from oracle import oracle as orc  # noqa: E402
from barneshut import inputs  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ini(tmp_path, impl="singlegpu", precision="fp64", grid_dim=0, eps2=0.0, rounds=1,
         skip=False, G=1.0, tick=0.005, extra=""):
    f = tmp_path / "t.ini"
    f.write_text(f"""[general]
implementation = {impl}
sample_check_size = 10
skip_timestep = {skip}
evaluation_rounds = 1
rounds = {rounds}
[grid]
max_particles_per_box = 1000
[bh]
grav_constant = {G}
tick_seconds = {tick}
[hip]
precision = {precision}
eps2 = {eps2}
r2_cut = 0
grid_dim = {grid_dim}
{extra}
""")
    return str(f)


def _input(tmp_path, n=3000, seed=1):
    d = inputs.reference_normal(n, seed)
    f = tmp_path / "in.txt"
    inputs.write_particles_file(str(f), d)
    return str(f), d


def test_run_2d_cli_with_check_equals_direct_sum_on_2x2_grid(tmp_path):
    """grid_dim = 2: every cell is a neighbour, so --check compares the
    implementation with the exact direct sum."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "run_2d", os.path.join(ROOT, "barnes-hut_amd", "bin", "run_2d.py"))
    run_2d = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(run_2d)
    fin, d = _input(tmp_path)
    ini = _ini(tmp_path, grid_dim=2, eps2=1e-6)
    bh = run_2d.main([fin, "1", "2", ini, "--check"])
    chk = bh.bh.last_check
    assert max(chk["vector_rel"]) < 1e-10
    bh.bh.close()


def _load_bin(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "barnes-hut_amd", "bin", f"{name}.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _oracle_steps(d, G, steps, Gc=1.0, dt=0.005, eps2=1e-6):
    n = len(d["px"])
    st = dict(d, ax=np.zeros(n), ay=np.zeros(n))
    for _ in range(steps):
        o = orc.step(st["px"], st["py"], st["mass"], st["vx"], st["vy"], st["pid"], G, Gc, dt,
                     eps2, 0.0, st["ax"], st["ay"])
        inv = np.argsort(o["pid"])
        st = dict(px=o["px_next"][inv], py=o["py_next"][inv], mass=o["m"][inv],
                  vx=o["vx_next"][inv], vy=o["vy_next"][inv], pid=o["pid"][inv],
                  ax=np.zeros(n), ay=np.zeros(n))
    return st


def test_run_2d_concurrent_cli_copies_equal_single_system_and_oracle(tmp_path, capsys):
    """run_2d_concurrent.py input_file nwarmups nruns concurrent configfile
    --check (the reference's CLI, bin/run_2d_concurrent.py:29-90): four copies
    of one input stepped as four systems of one context.  Every copy equals
    system 0 bit for bit after every run; system 0 equals the single-system
    drop-in (run_2d.py with the same INI and runs) bit for bit and tracks the
    oracle over the 6 steps taken (1 warm-up + 2 runs of 2 rounds); on a 2x2
    grid the check against the direct sum is exact."""
    run_c = _load_bin("run_2d_concurrent")
    run_1 = _load_bin("run_2d")
    fin, d = _input(tmp_path, 3000, seed=7)
    ini = _ini(tmp_path, rounds=2, eps2=1e-6, grid_dim=12)
    bhc = run_c.main([fin, "1", "2", "4", ini, "--check"])
    out = capsys.readouterr().out
    assert out.count("4 concurrent systems identical: True") == 3
    assert "warmup done" in out and "end-to-end" in out and "evaluation" in out
    assert bhc.bh.systems_identical()
    got = bhc.bh.get_particles()
    bhc.bh.close()
    bh1 = run_1.main([fin, "1", "2", ini, "--check"])
    one = bh1.bh.get_particles()
    bh1.bh.close()
    np.testing.assert_array_equal(got, one)
    st = _oracle_steps(d, 12, 6)
    np.testing.assert_array_equal(got[:, 9], np.arange(3000))
    np.testing.assert_allclose(got[:, 0], st["px"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(got[:, 3], st["vx"], rtol=1e-7, atol=1e-9)
    ini2 = _ini(tmp_path, rounds=1, eps2=1e-6, grid_dim=2)
    bh2 = run_c.main([fin, "0", "1", "3", ini2, "--check"])
    assert max(bh2.bh.last_check["vector_rel"]) < 1e-10
    bh2.bh.close()


def test_rounds_track_oracle_steps(tmp_path):
    fin, d = _input(tmp_path, 4000, seed=2)
    G = orc.grid_dim(4000, 1000)
    ini = _ini(tmp_path, rounds=3, eps2=1e-6)
    from barneshut import BarnesHut
    bh = BarnesHut(ini)
    bh.read_input_file(fin)
    bh.run()
    got = bh.bh.get_particles()
    st = dict(d)
    st["ax"] = np.zeros(4000)
    st["ay"] = np.zeros(4000)
    for _ in range(3):
        o = orc.step(st["px"], st["py"], st["mass"], st["vx"], st["vy"], st["pid"], G, 1.0, 0.005,
                     1e-6, 0.0, st["ax"], st["ay"])
        inv = np.argsort(o["pid"])
        st = dict(px=o["px_next"][inv], py=o["py_next"][inv], mass=o["m"][inv],
                  vx=o["vx_next"][inv], vy=o["vy_next"][inv], pid=o["pid"][inv],
                  ax=np.zeros(4000), ay=np.zeros(4000))
    np.testing.assert_array_equal(got[:, 9], np.arange(4000))
    np.testing.assert_allclose(got[:, 0], st["px"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(got[:, 3], st["vx"], rtol=1e-7, atol=1e-9)
    bh.bh.close()


def test_grid_columns_and_pid_order(tmp_path):
    fin, d = _input(tmp_path, 2000, seed=3)
    ini = _ini(tmp_path, skip=True)
    from barneshut import BarnesHut
    bh = BarnesHut(ini)
    bh.read_input_file(fin)
    bh.run()
    a = bh.bh.get_particles()
    G = bh.bh.grid_dim_
    bb = orc.bbox(d["px"], d["py"], G)
    gx, gy, _ = orc.place(d["px"], d["py"], bb[0], bb[1], bb[4], G)
    np.testing.assert_array_equal(a[:, 7], gx)
    np.testing.assert_array_equal(a[:, 8], gy)
    np.testing.assert_array_equal(a[:, 0], d["px"])      # skip_timestep: positions unchanged
    assert np.abs(a[:, 5]).max() > 0                      # accelerations kept
    bh.bh.close()


def test_fp32_precision_through_ini(tmp_path):
    fin, d = _input(tmp_path, 20000, seed=4)
    from barneshut import BarnesHut
    r = {}
    shapes = {}
    for prec in ("fp64", "fp32", "fp64_ieee"):
        bh = BarnesHut(_ini(tmp_path, precision=prec, skip=True, grid_dim=64))
        bh.read_input_file(fin)
        bh.run()
        r[prec] = bh.bh.get_particles()[:, 5:7].copy()
        shapes[prec] = bh.bh.engine.ctx.get_tuning()
        bh.bh.close()
    num = np.linalg.norm(r["fp32"] - r["fp64"])
    assert num / np.linalg.norm(r["fp64"]) < 1e-5
    assert np.linalg.norm(r["fp64_ieee"] - r["fp64"]) <= 1e-14 * np.linalg.norm(r["fp64"])
    # the FP64 precisions run their own launch shape ([hip] p64, yspan64,
    # waves_per_cu64), never split into slices; FP32 the [hip] p ... shape
    fp64_shape = dict(p=2, yspan=512, waves_per_cu=32, tail_split=1, tail_items=0)
    assert shapes["fp64"] == fp64_shape and shapes["fp64_ieee"] == fp64_shape
    assert shapes["fp32"]["tail_split"] == 16


def test_nonfinite_coordinate_raises(tmp_path):
    """A NaN coordinate is flagged by the bounding-box kernel; every kernel
    still completes (keys are clamped, the row order stays total) and the
    implementation raises."""
    d = inputs.uniform(3000, seed=5)
    d["px"][17] = np.nan
    f = tmp_path / "nan.txt"
    inputs.write_particles_file(str(f), d)
    from barneshut import BarnesHut
    from barneshut.kernels.hiplib import BhError
    bh = BarnesHut(_ini(tmp_path, grid_dim=16))
    bh.read_input_file(str(f))
    with pytest.raises(BhError, match="non-finite"):
        bh.run()
    bh.bh.close()


def test_invalid_arguments_raise():
    from barneshut.engine import Engine
    from barneshut.kernels.hiplib import BhError
    with pytest.raises(BhError):
        Engine([100], [0])
    with pytest.raises(BhError, match="32767"):
        Engine([100], [32768])     # beyond the FP32 COM list's 16-bit cell packing
    e = Engine([1000], [8])
    with pytest.raises(BhError):
        e.ctx.set_tuning(p=3)
    with pytest.raises(BhError):
        e.ctx.evaluate(e.parts, 1.0, -1.0, 0.0, "fp64")
    with pytest.raises(BhError):
        e.ctx.sort(e.parts, e.parts)
    with pytest.raises(ValueError, match="elements"):   # a short particle buffer
        e.ctx.place({k: v[:500] for k, v in e.parts.items()})
    e.close()


def test_refused_reconfiguration_keeps_the_context_usable():
    """bh_set_systems validates every system before it releases the old
    buffers: a refused call (grid side beyond the cap, an empty system) leaves
    the configured context as it was, and its next step equals the step of a
    context that never saw the refused call."""
    import torch
    from barneshut.engine import Engine
    from barneshut.kernels.hiplib import BhError
    d = inputs.gaussian_mixture(3000, 4, (0.02, 0.1), seed=41)
    out = []
    for refuse in (False, True):
        e = Engine([3000], [16])
        e.load(d)
        if refuse:
            with pytest.raises(BhError, match="32767"):
                e.ctx.set_systems([3000], [32768])
            with pytest.raises(BhError, match="n < 1"):
                e.ctx.set_systems([3000, 0], [16, 4])
            assert e.ctx.query(0) == 3000 and e.ctx.query(1) == 256
        e.step(1.0, 1e-6, 0.0, "fp64", 1e-4)
        e.check()
        torch.cuda.synchronize()
        out.append(e.fetch())
        e.close()
    for k in ("px", "py", "vx", "vy", "pid"):
        np.testing.assert_array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("n", [1, 64])
def test_degenerate_box_raises(n):
    """All particles at one point (or a single particle): the bounding box has
    zero extent, the cell size is 0; keys stay inside the grid and the
    context reports the condition instead of producing garbage cells."""
    from barneshut.engine import Engine
    from barneshut.kernels.hiplib import BhError
    d = inputs.uniform(n, seed=9)
    d["px"][:] = 0.3
    d["py"][:] = 0.7
    e = Engine([n], [4])
    e.load(d)
    e.create_tree()
    with pytest.raises(BhError, match="zero-extent"):
        e.check()
    assert (e.keys_in() >= 0).all() and (e.keys_in() < 16).all()
    e.close()
