"""Pins the CPU oracle (oracle/bh_oracle.c, oracle/mlp.py) before it is
trusted as the checker of the HIP path.

Each phase is checked against something that does not restate its formula:
  * grid sizing      -> numpy replay of base.py + the reference log's 120x120 grid
  * placement        -> numpy's own structured sort call of sequential.py:37,
                        and the placement rows the reference printed in its logs
  * counts/offsets   -> np.bincount / np.cumsum
  * COM              -> exact rational arithmetic
  * forces           -> O(N^2) direct sum where the method must equal it
                        (grid <= 2x2; single-particle cells), two-body closed
                        forms, momentum conservation, the cutoff boundary
  * tick             -> Python float arithmetic (IEEE, no FMA)
  * interactions     -> brute-force pair enumeration
  * MLP              -> finite differences, hand-computed cases
  * energy           -> two-body closed form; its gradient against the direct
                        force of the same law (finite differences)
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import mlp as omlp
from oracle import oracle as orc
from barneshut import inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _aos(d):
    n = len(d["px"])
    a = np.zeros((n, 10))
    for c, k in enumerate(("px", "py", "mass", "vx", "vy")):
        a[:, c] = d[k]
    a[:, 9] = d["pid"]
    return a


def _ties(n, seed):
    d = inputs.gaussian_mixture(n, 4, sigma=(0.05, 0.4), seed=seed)
    rng = np.random.default_rng(seed)
    i = rng.choice(n, n // 5, replace=False)
    d["px"][i[: n // 10]] = d["px"][i[n // 10: 2 * (n // 10)]]
    j = rng.choice(n, 10, replace=False)
    d["px"][j[:5]] = d["px"][j[5:]]
    d["py"][j[:5]] = d["py"][j[5:]]
    d["px"][j[0]] = -0.0 if d["px"][j[0]] == 0.0 else d["px"][j[0]]
    return d


@pytest.mark.parametrize("n,ppb", [(16384, 23), (16384, 24), (20000, 1000), (10_000_000, 1000),
                                   (4194304, 4), (1000, 1), (5, 2), (123457, 77)])
def test_grid_dim_matches_numpy_replay(n, ppb):
    assert orc.grid_dim(n, ppb) == orc.grid_dim_numpy(n, ppb)


def test_grid_dim_reference_log():
    # "grid size 14400" for N=10M, ppb=1000 (data/frontera/nbody_out_singlegpu.dat:11)
    assert orc.grid_dim(10_000_000, 1000) ** 2 == 14400
    assert orc.grid_dim(123, 1) == 1   # ppb == 1 -> one leaf (base.py:223-224)


@pytest.mark.parametrize("case", ["uniform", "normal", "ties", "one_cell"])
def test_placement_matches_numpy_structured_sort(case):
    if case == "uniform":
        d, G = inputs.uniform(5000, seed=1), 16
    elif case == "normal":
        d, G = inputs.reference_normal(6000, seed=2), 7
    elif case == "ties":
        d, G = _ties(6000, seed=3), 12
    else:
        d, G = inputs.uniform(800, seed=4), 1
    srt, (mnx, mny, step) = orc.place_numpy(_aos(d), G)
    bb = orc.bbox(d["px"], d["py"], G)
    assert (bb[0], bb[1], bb[4]) == (mnx, mny, step)
    gx, gy, key = orc.place(d["px"], d["py"], bb[0], bb[1], bb[4], G)
    z = np.zeros(len(key))
    perm = orc.sort_perm(key, (d["px"], d["py"], d["mass"], d["vx"], d["vy"], z, z,
                               d["pid"].astype(float)))
    np.testing.assert_array_equal(d["pid"][perm], srt[:, 9].astype(np.int32))
    np.testing.assert_array_equal(gx[perm], srt[:, 7].astype(np.int32))
    np.testing.assert_array_equal(gy[perm], srt[:, 8].astype(np.int32))
    np.testing.assert_array_equal(d["px"][perm], srt[:, 0])


def test_placement_reference_log_rows():
    """Rows the reference printed after placement (first round of its
    Parla logs): N=10M uniform in [0,499)^2, 120x120 grid."""
    with open(os.path.join(GOLDEN, "ref_log_placement.json")) as f:
        fx = json.load(f)
    r = np.array(fx["rows"])
    G = fx["grid_dim"]
    step = (fx["domain"][1] - fx["domain"][0]) / G
    gx, gy, key = orc.place(r[:, 0], r[:, 1], 0.0, 0.0, step, G)
    np.testing.assert_array_equal(gx, r[:, 7].astype(np.int32))
    np.testing.assert_array_equal(gy, r[:, 8].astype(np.int32))
    np.testing.assert_array_equal(key, (r[:, 7] * G + r[:, 8]).astype(np.int32))


def test_counts_offsets_match_bincount():
    rng = np.random.default_rng(5)
    key = np.sort(rng.integers(0, 400, 10000)).astype(np.int32)
    c, o = orc.counts_offsets(key, 400)
    np.testing.assert_array_equal(c, np.bincount(key, minlength=400))
    np.testing.assert_array_equal(o, np.concatenate([[0], np.cumsum(np.bincount(key, minlength=400))]))


def test_com_exact_rational():
    rng = np.random.default_rng(6)
    n = 200
    x = rng.random(n)
    y = rng.random(n)
    m = rng.uniform(1, 2, n)
    off = np.array([0, 1, 3, 3, 50, 200], np.int64)   # one particle, two, empty, ...
    com = orc.summarize(off, x, y, m)
    for c in range(len(off) - 1):
        a, b = off[c], off[c + 1]
        if a == b:
            assert tuple(com[c]) == (0.0, 0.0, 0.0)
            continue
        M = sum(Fraction(v) for v in m[a:b])
        X = sum(Fraction(u) * Fraction(v) for u, v in zip(x[a:b], m[a:b])) / M
        Y = sum(Fraction(u) * Fraction(v) for u, v in zip(y[a:b], m[a:b])) / M
        assert abs(com[c, 2] - float(M)) <= 2 * np.spacing(float(M))
        assert abs(com[c, 0] - float(X)) <= 4 * np.spacing(float(X))
        assert abs(com[c, 1] - float(Y)) <= 4 * np.spacing(float(Y))
    assert com[0, 2] == m[0]


@pytest.mark.parametrize("eps2", [0.0, 1e-3])
def test_direct_two_body_closed_form(eps2):
    x = np.array([0.1, 0.4])
    y = np.array([0.2, -0.2])
    m = np.array([2.0, 3.0])
    G = 1.7
    ax, ay = orc.direct(x, y, m, G, eps2)
    r2 = 0.3 ** 2 + 0.4 ** 2
    assert np.isclose(ax[0], G * 3.0 * 0.3 / (r2 + eps2), rtol=1e-14)
    assert np.isclose(ay[0], G * 3.0 * -0.4 / (r2 + eps2), rtol=1e-14)
    assert np.isclose(ax[1], -G * 2.0 * 0.3 / (r2 + eps2), rtol=1e-14)
    # |a| = G m / r for the unsoftened 2-D law
    if eps2 == 0.0:
        assert np.isclose(np.hypot(ax[0], ay[0]), G * 3.0 / np.sqrt(r2), rtol=1e-14)


def test_direct_momentum_conservation():
    d = inputs.uniform(600, seed=7)
    ax, ay = orc.direct(d["px"], d["py"], d["mass"], 1.0, 1e-6)
    px = (d["mass"] * ax).sum()
    scale = (d["mass"] * np.abs(ax)).sum()
    assert abs(px) <= 1e-12 * scale
    assert abs((d["mass"] * ay).sum()) <= 1e-12 * (d["mass"] * np.abs(ay)).sum()


def _method(d, G, Gc=1.0, eps2=0.0, r2=0.0):
    out = orc.step(d["px"], d["py"], d["mass"], d["vx"], d["vy"], d["pid"], G, Gc, 0.0, eps2,
                   r2, do_tick=False)
    return out


@pytest.mark.parametrize("G", [1, 2])
@pytest.mark.parametrize("law", [(1.0, 1e-6, 0.0), (2.0, 0.0, 1e-5)])
def test_method_equals_direct_on_tiny_grids(G, law):
    d = inputs.uniform(1500, seed=8 + G)
    o = _method(d, G, *law)
    rx, ry = orc.direct(o["px"], o["py"], o["m"], *law)
    err = np.hypot(o["ax"] - rx, o["ay"] - ry) / np.hypot(rx, ry)
    assert err.max() < 1e-11


def test_method_single_particle_cells_equals_direct():
    G = 40
    rng = np.random.default_rng(9)
    cells = rng.choice(np.arange(1, G * G - 1), 300, replace=False)
    x = (cells // G + rng.uniform(0.2, 0.8, 300)) / G
    y = (cells % G + rng.uniform(0.2, 0.8, 300)) / G
    x[:2] = [0.0, 1.0]
    y[:2] = [0.0, 1.0]
    d = inputs._pack(x, y, rng.uniform(1, 2, 300))
    o = _method(d, G)
    assert o["counts"].max() == 1
    rx, ry = orc.direct(o["px"], o["py"], o["m"], 1.0)
    err = np.hypot(o["ax"] - rx, o["ay"] - ry) / np.hypot(rx, ry)
    assert err.max() < 1e-12


def test_method_differs_from_direct_when_far_cells_hold_several():
    """Guard against a far field that silently sums particles instead of COMs."""
    d = inputs.uniform(3000, seed=10)
    o = _method(d, 16)
    rx, ry = orc.direct(o["px"], o["py"], o["m"], 1.0)
    err = np.hypot(o["ax"] - rx, o["ay"] - ry) / np.hypot(rx, ry)
    assert 1e-6 < np.median(err) < 0.05


def test_cutoff_boundary():
    cut = 1e-5
    for r2, nonzero in ((cut * 0.999, False), (cut, False), (cut * 1.001, True)):
        x = np.array([0.0, np.sqrt(r2)])
        y = np.zeros(2)
        ax, _ = orc.direct(x, y, np.ones(2), 1.0, 0.0, cut)
        assert (ax[0] != 0.0) == nonzero


def test_tick_matches_python_floats():
    rng = np.random.default_rng(11)
    n = 64
    vals = [rng.normal(size=n) for _ in range(6)]
    x, y, vx, vy, ax, ay = [v.copy() for v in vals]
    dt = 0.005
    orc.tick(x, y, vx, vy, ax, ay, dt)
    for i in range(n):
        pvx = float(vals[2][i]) + float(vals[4][i]) * dt
        pvy = float(vals[3][i]) + float(vals[5][i]) * dt
        assert vx[i] == pvx and vy[i] == pvy
        assert x[i] == float(vals[0][i]) + pvx * dt
        assert y[i] == float(vals[1][i]) + pvy * dt
        assert ax[i] == 0.0 and ay[i] == 0.0


def test_interactions_bruteforce():
    d = inputs.gaussian_mixture(700, 3, (0.05, 0.2), seed=12)
    G = 9
    o = _method(d, G)
    gx, gy = o["key"] // G, o["key"] % G
    near = 0
    for i in range(len(gx)):
        nb = (np.abs(gx - gx[i]) <= 1) & (np.abs(gy - gy[i]) <= 1)
        near += int(nb.sum()) - 1
    ne = np.nonzero(o["counts"])[0]
    ex, ey = ne // G, ne % G
    far = sum(int(((np.abs(ex - gx[i]) > 1) | (np.abs(ey - gy[i]) > 1)).sum())
              for i in range(len(gx)))
    assert orc.interactions(G, o["counts"]) == (near, far)


# ---------------------------------------------------------------- MLP oracle
def test_mlp_backward_finite_differences():
    rng = np.random.default_rng(13)
    p = rng.normal(0, 0.4, omlp.NPARAM)
    X = rng.normal(size=(37, omlp.NF))
    t = rng.normal(size=37)
    g = omlp.backward(p, X, t)
    idx = rng.choice(omlp.NPARAM, 80, replace=False)
    idx = np.concatenate([idx, [omlp.OFF_B3, omlp.OFF_B1, omlp.OFF_W3]])
    h = 1e-6
    for k in idx:
        e = np.zeros(omlp.NPARAM)
        e[k] = h
        fd = (omlp.loss(p + e, X, t) - omlp.loss(p - e, X, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_mlp_forward_hand_case():
    p = np.zeros(omlp.NPARAM)
    W1 = np.zeros((32, 8))
    W1[0, 0] = 1.0
    W1[1, 1] = -1.0
    p[omlp.OFF_W1:omlp.OFF_B1] = W1.ravel()
    W2 = np.eye(32)
    p[omlp.OFF_W2:omlp.OFF_B2] = W2.ravel()
    p[omlp.OFF_W3] = 2.0
    p[omlp.OFF_W3 + 1] = 3.0
    p[omlp.OFF_B3] = 0.5
    X = np.zeros((3, 8))
    X[0, 0], X[1, 1], X[2, 0] = 1.5, -2.0, -1.0
    # y = 2*relu(x0) + 3*relu(-x1) + 0.5
    np.testing.assert_allclose(omlp.forward(p, X), [3.5, 6.5, 0.5])


def test_adam_first_step():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 0.0])
    q, m, v = omlp.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.1)
    # first step: m_hat = g, v_hat = g^2 -> update -lr * sign(g)
    np.testing.assert_allclose(q, [0.9, -1.9, 3.0], atol=1e-7)


def test_energy_two_body_closed_form():
    x, y = np.array([0.1, 0.4]), np.array([0.2, 0.6])
    m, vx, vy = np.array([1.5, 2.0]), np.array([0.3, -0.1]), np.array([0.0, 0.2])
    U, K = orc.energy(x, y, m, vx, vy, Gc=0.7, eps2=0.0)
    r = np.hypot(0.3, 0.4)
    assert U == pytest.approx(0.7 * 1.5 * 2.0 * np.log(r), rel=1e-14)
    assert K == pytest.approx(0.5 * (1.5 * 0.09 + 2.0 * (0.01 + 0.04)), rel=1e-14)


def test_energy_gradient_is_the_force_law():
    """-dU/dx_i = m_i a_i with a from the direct sum of the same softened law
    (central differences): pins the potential against the force."""
    rng = np.random.default_rng(7)
    n = 7
    x, y, m = rng.random(n), rng.random(n), rng.uniform(1, 2, n)
    z = np.zeros(n)
    Gc, eps2, h = 1.3, 1e-3, 1e-6
    ax, ay = orc.direct(x, y, m, Gc, eps2, 0.0)
    for i in range(n):
        for arr, acc in ((x, ax), (y, ay)):
            arr[i] += h
            up = orc.energy(x, y, m, z, z, Gc, eps2)[0]
            arr[i] -= 2 * h
            dn = orc.energy(x, y, m, z, z, Gc, eps2)[0]
            arr[i] += h
            grad = (up - dn) / (2 * h)
            assert -grad == pytest.approx(m[i] * acc[i], rel=1e-6, abs=1e-8)


def _grid_energy(o, G, Gc, eps2):
    phi = orc.potential(G, o["offsets"], o["key"], o["px"], o["py"], o["m"], o["com"], eps2)
    return 0.25 * Gc * float(np.sum(o["m"] * phi)), phi


@pytest.mark.parametrize("G", [1, 2])
def test_grid_potential_equals_direct_energy_on_tiny_grids(G):
    """G <= 2: every cell is a neighbour, so the grid potential is the direct
    pair sum and U_grid equals the direct energy."""
    d = inputs.gaussian_mixture(2500, 3, (0.02, 0.2), seed=30 + G)
    o = _method(d, G, 1.0, 1e-6, 0.0)
    U, _ = _grid_energy(o, G, 1.7, 1e-6)
    Ud, _ = orc.energy(o["px"], o["py"], o["m"], o["vx"], o["vy"], 1.7, 1e-6)
    assert U == pytest.approx(Ud, rel=1e-13)


def test_grid_potential_single_particle_cells_equals_direct_energy():
    """Every non-empty cell holds one particle: each far COM is the particle
    itself, so the far field is exact and U_grid equals the direct energy.
    On a dense grid (several particles per far cell) it does not: the
    monopole far field is an approximation, guarding against a far field that
    sums particles."""
    G = 40
    rng = np.random.default_rng(31)
    cells = rng.choice(np.arange(1, G * G - 1), 300, replace=False)
    x = (cells // G + rng.uniform(0.2, 0.8, 300)) / G
    y = (cells % G + rng.uniform(0.2, 0.8, 300)) / G
    x[:2] = [0.0, 1.0]
    y[:2] = [0.0, 1.0]
    d = inputs._pack(x, y, rng.uniform(1, 2, 300))
    o = _method(d, G)
    assert o["counts"].max() == 1
    U, _ = _grid_energy(o, G, 1.0, 0.0)
    Ud, _ = orc.energy(o["px"], o["py"], o["m"], o["vx"], o["vy"], 1.0, 0.0)
    assert U == pytest.approx(Ud, rel=1e-13)
    d2 = inputs.uniform(3000, seed=32)
    o2 = _method(d2, 16)
    U2, _ = _grid_energy(o2, 16, 1.0, 0.0)
    Ud2, _ = orc.energy(o2["px"], o2["py"], o2["m"], o2["vx"], o2["vy"], 1.0, 0.0)
    assert 1e-9 < abs(U2 - Ud2) / abs(Ud2) < 1e-2


def test_grid_potential_two_body_closed_form():
    """Two particles far apart on a 16 x 16 grid (each alone in its cell, the
    other a far COM): phi_0 = m_1 log(r^2 + eps2), phi_1 = m_0 log(r^2 + eps2)."""
    x, y = np.array([0.05, 0.93]), np.array([0.1, 0.77])
    d = inputs._pack(x, y, np.array([1.5, 2.5]))
    o = _method(d, 16)
    _, phi = _grid_energy(o, 16, 1.0, 1e-4)
    r2 = (0.93 - 0.05) ** 2 + (0.77 - 0.1) ** 2
    m_of = dict(zip(o["pid"].tolist(), o["m"].tolist()))
    for k in range(2):
        other = m_of[1 - int(o["pid"][k])]
        assert phi[k] == pytest.approx(other * np.log(r2 + 1e-4), rel=1e-14)


def test_oracle_refuses_inconsistent_inputs():
    """The C restatement indexes per-cell arrays by key and offset: its
    wrappers refuse keys, offsets, shapes and targets that would run out of
    bounds (found with an AddressSanitizer build of the oracle)."""
    with pytest.raises(ValueError):
        orc.counts_offsets(np.array([0, 5], np.int32), 4)
    with pytest.raises(ValueError):
        orc.counts_offsets(np.array([-1, 0], np.int32), 4)
    x = np.array([0.1, 0.9]); m = np.ones(2)
    with pytest.raises(ValueError):
        orc.summarize(np.array([0, 1, 3]), x, x, m)          # offsets end past n
    key = np.array([0, 3], np.int32)
    cnt, off = orc.counts_offsets(key, 4)
    com = orc.summarize(off, x, x, m)
    with pytest.raises(ValueError):
        orc.forces(2, off, np.array([0, 4], np.int32), x, x, m, com)   # key past G*G
    with pytest.raises(ValueError):
        orc.forces(3, off, key, x, x, m, com)                 # offsets of another grid
    with pytest.raises(ValueError):
        orc.forces(2, off, key, x, x, m, com, targets=np.array([2]))
    ax, ay = orc.forces(2, off, key, x, x, m, com)
    assert np.isfinite(ax).all() and np.isfinite(ay).all()


# --------------------------------------------------------------------------
# cloud-to-cloud gravity (cpu_numba.py:8-92), pinned by closed forms, exact
# dyadic cutoff cases, the direct sum (itself pinned above) and linearity
# --------------------------------------------------------------------------
def _cloud_rows(rng, n, lo=0.0, hi=1.0):
    r = np.zeros((n, 10))
    r[:, 0:2] = rng.uniform(lo, hi, (n, 2))
    r[:, 2] = rng.uniform(1.0, 2.0, n)
    return r


def test_cloud_gravity_two_body_closed_form():
    A = np.array([[0.0, 0.0, 2.0]])
    B = np.array([[3.0, 4.0, 5.0]])
    # a_A = G m_B (x_B - x_A) / r^2 = 1.5 * 5 * (3, 4) / 25; a_B = 1.5 * 2 * (-3, -4) / 25
    accA, accB = orc.cloud_gravity(A, B, Gc=1.5, update_other=True)
    assert np.allclose(accA, [[0.9, 1.2]], rtol=1e-15, atol=0)
    assert np.allclose(accB, [[-0.36, -0.48]], rtol=1e-15, atol=0)
    accA1, none = orc.cloud_gravity(A, B, Gc=1.5, update_other=False)
    assert none is None and np.array_equal(accA1, accA)
    acc, _ = orc.cloud_gravity(np.vstack([A, B]), Gc=1.5)       # one cloud of both rows
    assert np.allclose(acc, [[0.9, 1.2], [-0.36, -0.48]], rtol=1e-15, atol=0)


def test_cloud_gravity_cutoff_on_exact_dyadic_distances():
    # r^2 = 2^-18 (3.8e-6) <= 1e-5: dropped; r^2 = 2^-16 (1.5e-5): kept, a = m / dx
    for dx, want in ((2.0 ** -9, 0.0), (2.0 ** -8, 3.0 * 256.0)):
        A = np.array([[0.25, 0.5, 1.0]])
        B = np.array([[0.25 + dx, 0.5, 3.0]])
        acc, _ = orc.cloud_gravity(A, B, Gc=1.0, r2_cut=1e-5)
        assert acc[0, 0] == want and acc[0, 1] == 0.0


def test_cloud_gravity_one_cloud_equals_direct_sum():
    rng = np.random.default_rng(5)
    A = _cloud_rows(rng, 300)
    A[7, :2] = A[3, :2] + 1e-4        # a pair inside the cutoff (r^2 = 2e-8)
    acc, _ = orc.cloud_gravity(A, Gc=0.7, r2_cut=1e-5)
    ax, ay = orc.direct(A[:, 0], A[:, 1], A[:, 2], Gc=0.7, r2_cut=1e-5)
    ref = np.stack([ax, ay], 1)
    assert np.linalg.norm(acc - ref) <= 1e-13 * np.linalg.norm(ref)


def test_cloud_gravity_other_cloud_is_union_minus_self():
    rng = np.random.default_rng(6)
    A = _cloud_rows(rng, 200, 0.0, 1.0)
    B = _cloud_rows(rng, 150, 0.5, 1.5)
    accA, accB = orc.cloud_gravity(A, B, Gc=2.0, update_other=True)
    U = np.vstack([A, B])
    ux, uy = orc.direct(U[:, 0], U[:, 1], U[:, 2], Gc=2.0, r2_cut=1e-5)
    for acc, rows, sl in ((accA, A, slice(0, 200)), (accB, B, slice(200, 350))):
        sx, sy = orc.direct(rows[:, 0], rows[:, 1], rows[:, 2], Gc=2.0, r2_cut=1e-5)
        ref = np.stack([ux[sl] - sx, uy[sl] - sy], 1)
        assert np.linalg.norm(acc - ref) <= 1e-12 * np.linalg.norm(ref)


def test_cloud_gravity_conserves_momentum():
    rng = np.random.default_rng(8)
    A = _cloud_rows(rng, 120)
    B = _cloud_rows(rng, 90, 0.2, 0.9)
    acc, _ = orc.cloud_gravity(A, Gc=1.0)
    scale = np.abs(A[:, 2:3] * acc).sum()
    assert np.abs((A[:, 2:3] * acc).sum(0)).max() <= 1e-13 * scale
    accA, accB = orc.cloud_gravity(A, B, Gc=1.0, update_other=True)
    tot = (A[:, 2:3] * accA).sum(0) + (B[:, 2:3] * accB).sum(0)
    scale = np.abs(A[:, 2:3] * accA).sum() + np.abs(B[:, 2:3] * accB).sum()
    assert np.abs(tot).max() <= 1e-13 * scale
