"""Pins of the oracle's plain CA loop (CellularAutomataManager_3D.cpp:440-669).

The reference ships no tests or golden outputs for this path, so the oracle's
time stepping and capture sweep are pinned by closed forms, symmetry and
invariants that a wrong formula, a wrong neighbour order, a wrong claim order or
a wrong deactivation rule would break.
"""
import numpy as np
import pytest


def velocity(p, t):
    T = p.T0 + p.Tdot * t
    dT = p.Tm - T
    v = 0.0
    if dT > 0:
        v = p.a1 * dT + p.a2 * dT * dT + p.a3 * dT * dT * dT
        v = max(v, 0.0)
    return v


def test_first_nucleation_and_dt_sequence(oracle):
    """A single site with critical undercooling dTc nucleates at the first step
    whose start time t has Tm - (T0 + Tdot t) > dTc; before it dt = maxDt (no
    active grain: maxV = 1e-10); afterwards dt = min(0.2 h / v(t), maxDt)."""
    p = oracle.plain_params(nx=5, ny=5, nz=5, h=2e-6, T0=-0.2, Tdot=-200.0, max_dt=1e-4,
                            max_steps=300)
    c = oracle.cell_index(p, 3, 3, 3)
    r = oracle.run_plain(p, given=([c], [1.7], [[0.1, 0.2, 0.3]]))
    s = r["steps"]
    t = 0.0
    nucleated = False
    for k in range(len(s)):
        assert s["time_before"][k] == t
        if not nucleated and (p.Tm - (p.T0 + p.Tdot * t)) > 1.7:
            assert s["n_nucleated"][k] == 1
            nucleated = True
        else:
            assert s["n_nucleated"][k] == 0
        maxv = max(1e-10, velocity(p, t)) if nucleated else 1e-10
        dt = min(p.tsf * p.h / maxv, p.max_dt)
        assert s["dt"][k] == dt
        t = t + dt
    assert nucleated


def test_two_cell_capture_step(oracle):
    """Two cells along x, identity orientation: the nucleus captures its
    neighbour (L1 distance h) in the first step whose grown length reaches h."""
    p = oracle.plain_params(nx=2, ny=1, nz=1, h=1e-6, T0=-5.0, Tdot=-200.0, max_dt=1e-4,
                            max_steps=200)
    c = oracle.cell_index(p, 1, 1, 1)
    r = oracle.run_plain(p, given=([c], [0.0], [[0.0, 0.0, 0.0]]))
    s = r["steps"]
    L = 0.0
    gap = (p.x0 + p.h * (2 - 0.5)) - (p.x0 + p.h * (1 - 0.5))  # centre distance
    for k in range(len(s)):
        v = velocity(p, s["time_before"][k])
        L = L + s["dt"][k] * v
        if L >= gap:
            assert s["n_captured"][k] == 1 and r["done"] == 1 and k == len(s) - 1
            break
        assert s["n_captured"][k] == 0
    assert r["ori"][oracle.cell_index(p, 2, 1, 1)] == 0
    assert r["gid"][oracle.cell_index(p, 2, 1, 1)] == 1


def test_single_cell_domain(oracle):
    p = oracle.plain_params(nx=1, ny=1, nz=1, nv=1e12, h=1e-3, dTv_mean=1.0, dTv_sd=0.1)
    r = oracle.run_plain(p)
    assert r["done"] == 1 and r["n_grains"] == 1 and r["n_active"] == 0


def test_no_sites_runs_to_termination(oracle):
    p = oracle.plain_params(nx=4, ny=4, nz=4, final_time=0.0105, max_dt=1e-3)
    r = oracle.run_plain(p)
    assert r["done"] == 0 and r["n_grains"] == 0
    assert r["n_steps"] == 11  # 0.001 added 11 times first exceeds 0.0105
    assert (oracle.interior(p, r["ori"]) == -1).all()


@pytest.mark.parametrize("steps", [15, 40, 80])
def test_identity_nucleus_has_cube_symmetry(oracle, steps):
    """One nucleus with identity orientation at the centre: the solid set is
    invariant under the cube group (three mirrors, three axis swaps)."""
    n = 21
    p = oracle.plain_params(nx=n, ny=n, nz=n, h=1e-6, T0=-5.0, Tdot=-200.0, max_dt=1e-4,
                            max_steps=steps)
    c = oracle.cell_index(p, 11, 11, 11)
    r = oracle.run_plain(p, given=([c], [0.5], [[0.0, 0.0, 0.0]]))
    solid = oracle.interior(p, r["ori"]) >= 0
    assert solid.sum() > 1
    for t in (solid[::-1], solid[:, ::-1], solid[:, :, ::-1], solid.transpose(1, 0, 2),
              solid.transpose(2, 1, 0), solid.transpose(0, 2, 1)):
        assert np.array_equal(solid, t)


def _check_invariants(oracle, p, r):
    ng = r["n_grains"]
    gid = r["gid"]
    ori = r["ori"]
    ioriv = oracle.interior(p, ori)
    igid = oracle.interior(p, gid)
    solid = ioriv >= 0
    # one grain per solid cell, ids 0..n-1, each grain's cell holds its id
    assert solid.sum() == ng
    assert np.array_equal(np.sort(igid[solid]), np.arange(ng))
    assert (gid[r["grain_cell"]] == np.arange(ng)).all()
    assert (ori[r["grain_cell"]] == r["grain_orient"]).all()
    assert (r["grain_orient"] < r["n_orient"]).all()
    # liquid cells carry -1 in both fields
    assert (igid[~solid] == -1).all()
    # every solid neighbour of a liquid cell is an active grain: a grain is only
    # deactivated when none of its neighbours is liquid, and nothing melts here
    active_cell = np.zeros(ori.shape, bool)
    active_cell[r["grain_cell"][r["grain_active"] == 1]] = True
    act = oracle.interior(p, active_cell)
    liquid = ~solid
    pad_l = np.pad(liquid, 1)
    near_liquid = np.zeros_like(liquid)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if dk == dj == di == 0:
                    continue
                near_liquid |= pad_l[1 + dk:pad_l.shape[0] - 1 + dk, 1 + dj:pad_l.shape[1] - 1 + dj,
                                     1 + di:pad_l.shape[2] - 1 + di]
    assert (act[solid & near_liquid]).all()
    # every captured grain has a same-orientation solid neighbour (its parent)
    s = r["steps"]
    assert int(s["n_nucleated"].sum()) == r["n_orient"]
    assert int(s["n_nucleated"].sum() + s["n_captured"].sum()) == ng


@pytest.mark.parametrize("seed", [0, 1, 7])
@pytest.mark.parametrize("steps", [30, 120])
def test_run_invariants(oracle, seed, steps):
    p = oracle.plain_params(nx=24, ny=20, nz=16, h=2e-6, nv=5e13, dTv_mean=3.0, dTv_sd=1.0,
                            ns=1e9, dTs_mean=1.0, dTs_sd=0.5, Tdot=-200.0, max_dt=1e-4,
                            seed=seed, max_steps=steps)
    r = oracle.run_plain(p)
    _check_invariants(oracle, p, r)


def test_site_counts_and_draws(oracle):
    """numSites = int(round(area * density * numCells) + 0.5); surface sites are
    drawn on the six-face list (edges twice, corners three times)."""
    p = oracle.plain_params(nx=60, ny=60, nz=60, h=5e-4, nv=1e7, dTv_mean=3.0, dTv_sd=1.0,
                            max_steps=0)
    r = oracle.run_plain(p, want_grains=False)
    assert r["n_site_draws"] == 270
    assert len(r["sites_cell"]) <= 270 and np.all(np.diff(r["sites_cell"]) > 0)
    n = 20
    p = oracle.plain_params(nx=n, ny=n, nz=n, h=1e-4, ns=1e7, dTs_mean=1.0, dTs_sd=0.5,
                            max_steps=0)
    r = oracle.run_plain(p, want_grains=False)
    face_entries = 6 * n * n
    assert r["n_site_draws"] == int(round(1e-8 * 1e7 * face_entries) + 0.5)
    i = (r["sites_cell"] % (n + 2))
    j = (r["sites_cell"] // (n + 2)) % (n + 2)
    k = r["sites_cell"] // ((n + 2) * (n + 2))
    on_face = (i == 1) | (i == n) | (j == 1) | (j == n) | (k == 1) | (k == n)
    assert on_face.all()
