"""Pins of the coupled CA-FE (remelting) oracle, and the host voxel mapping
against it.  CPU only.

Each check is a closed form or a hand case whose answer follows from the
rules of CellularAutomataManager_3DRemelting.cpp / MapVoxelManager.cpp /
LoadDataFromFile.cpp, not from re-running the oracle's own formulas."""
import numpy as np
import pytest

import oracle_lib as ol
import remelt_cases as rc

pytestmark = pytest.mark.usefixtures("oracle")


def _cell_centres(grid):
    nx, ny, nz, h = grid["nx"], grid["ny"], grid["nz"], grid["h"]
    nxG, nyG, nzG = nx + 2, ny + 2, nz + 2
    k, j, i = np.meshgrid(np.arange(nzG), np.arange(nyG), np.arange(nxG), indexing="ij")
    return np.stack([grid["x0"] + h * (i.ravel() - 0.5), grid["y0"] + h * (j.ravel() - 0.5),
                     grid["z0"] + h * (k.ravel() - 0.5)], axis=1)


def _hex8(par):
    xi, eta, zeta = par[:, 0], par[:, 1], par[:, 2]
    return 0.125 * np.stack([(1 - eta) * (1 - xi) * (1 - zeta), (1 - eta) * (1 + xi) * (1 - zeta),
                             (1 + eta) * (1 + xi) * (1 - zeta), (1 + eta) * (1 - xi) * (1 - zeta),
                             (1 - eta) * (1 - xi) * (1 + zeta), (1 - eta) * (1 + xi) * (1 + zeta),
                             (1 + eta) * (1 + xi) * (1 + zeta), (1 + eta) * (1 - xi) * (1 + zeta)],
                            axis=1)


def _element_nodes(shape, e):
    fnx, fny, _ = shape
    ncx, ncy = fnx - 1, fny - 1
    ei, ej, ek = e % ncx, (e // ncx) % ncy, e // (ncx * ncy)
    n0 = (ek * fny + ej) * fnx + ei
    n1 = n0 + 1
    n2 = n1 + fnx
    n3 = n2 - 1
    nxy = fnx * fny
    return np.stack([n0, n1, n2, n3, n0 + nxy, n1 + nxy, n2 + nxy, n3 + nxy], axis=1)


def _one_step(case, steps=1):
    c = dict(case)
    c["max_steps"] = steps
    return ol.run_remelt(c)


@pytest.fixture(scope="module")
def track():
    return rc.track_case(nx=30, ny=12, nz=10, n_frames=8, P=30.0)


def test_voxel_map_axis_aligned_closed_form(track):
    """On box elements xi = 2(x - x_lo)/dx - 1 (MapVoxelManager.cpp:522-690),
    and the tag is the element whose box holds the cell centre."""
    r = _one_step(track)
    g, fe = track["grid"], track["fe"]
    cen = _cell_centres(g)
    tag, par = r["tag"], r["par"]
    ok = tag >= 0
    assert ok.sum() == len(tag)  # the mesh covers the ghosted box
    ax = fe["axes"]
    idx = [np.searchsorted(ax[q], cen[:, q]) - 1 for q in range(3)]
    ncx, ncy = fe["shape"][0] - 1, fe["shape"][1] - 1
    expect_tag = (idx[2] * ncy + idx[1]) * ncx + idx[0]
    np.testing.assert_array_equal(tag, expect_tag)
    for q in range(3):
        lo, hi = ax[q][idx[q]], ax[q][idx[q] + 1]
        np.testing.assert_allclose(par[:, q], 2 * (cen[:, q] - lo) / (hi - lo) - 1, atol=1e-12)


def test_voxel_map_distorted_mesh_forward_maps_to_centre():
    """On a non-affine mesh the Newton result maps back to the cell centre."""
    case = rc.track_case(nx=20, ny=10, nz=8, n_frames=4, P=30.0)
    fe = case["fe"]
    rng = np.random.default_rng(5)
    h = case["grid"]["h"]
    xyz = fe["xyz"] + rng.uniform(-0.3, 0.3, size=fe["xyz"].shape) * h
    fe["xyz"] = xyz
    r = _one_step(case)
    cen = _cell_centres(case["grid"])
    tag, par = r["tag"], r["par"]
    ok = tag >= 0
    assert ok.mean() > 0.8
    nodes = _element_nodes(fe["shape"], tag[ok])
    sf = _hex8(par[ok])
    back = np.einsum("cm,cmq->cq", sf, xyz[nodes])
    np.testing.assert_allclose(back, cen[ok], atol=1e-9 * h, rtol=0)
    assert np.all(np.abs(par[ok]) <= 1.0)
    assert np.all(par[~ok] == -2.0)


@pytest.mark.parametrize("distort", [False, True])
def test_host_voxel_map_matches_oracle(amca, distort):
    """The library's host mapping (no GPU involved) equals the oracle's bit for bit."""
    from amca3d_amd import remelt
    case = rc.track_case(nx=26, ny=11, nz=9, n_frames=4, P=30.0)
    if distort:
        rng = np.random.default_rng(9)
        case["fe"]["xyz"] = case["fe"]["xyz"] + rng.uniform(-0.25, 0.25, case["fe"]["xyz"].shape) * 2.5e-6
    r = _one_step(case)
    tag, par = remelt.map_voxels(case["grid"], case["fe"], bin=case["bin"], tol=case["tol"],
                                 max_iter=case["max_iter"])
    np.testing.assert_array_equal(tag, r["tag"])
    np.testing.assert_array_equal(par.view(np.uint64), r["par"].view(np.uint64))


def test_temperature_exact_for_affine_field(track):
    """evaluate_temperature reproduces an affine nodal field (MapVoxelManager.cpp:746-799)."""
    case = dict(track)
    fe = dict(case["fe"])
    xyz = fe["xyz"]
    coef = np.array([3e5, 2e5, 1e5])
    th = 1000.0 + xyz @ coef
    fe["theta"] = np.stack([th] * len(fe["times"]))
    case["fe"] = fe
    r = _one_step(case)
    cen = _cell_centres(case["grid"])
    np.testing.assert_allclose(r["T"], 1000.0 + cen @ coef, rtol=1e-13)


def test_void_nodes_remove_material(track):
    """A cell whose element has a node below void + 1e-5 has no material: T = -1."""
    case = dict(track)
    fe = dict(case["fe"])
    th = np.full(len(fe["xyz"]), 1000.0)
    th[0] = case["void_temp"]  # first node: element 0 only
    fe["theta"] = np.stack([th] * len(fe["times"]))
    case["fe"] = fe
    r = _one_step(case)
    bad = r["tag"] == 0
    assert bad.sum() > 0
    assert np.all(r["T"][bad] == -1.0)
    np.testing.assert_allclose(r["T"][~bad], 1000.0, rtol=1e-14)


def test_nodal_time_interpolation_closed_form(track):
    """theta(t) between frames is r*(th0 - th1) + th1 (LoadDataFromFile.cpp:281-290):
    exact for fields affine in time; void nodes stay void; a node leaving the
    void jumps to its new frame value."""
    case = dict(track)
    fe = dict(case["fe"])
    n = len(fe["xyz"])
    rng = np.random.default_rng(2)
    A = rng.uniform(900.0, 1200.0, n)
    B = rng.uniform(-2e6, 2e6, n)
    t = fe["times"] - fe["times"][0]
    th = A[None, :] + B[None, :] * t[:, None]
    void = case["void_temp"]
    th[:, :5] = void - 1.0  # always void
    th[:3, 5:9] = void - 1.0  # void in frames 0..2, hot afterwards
    fe["theta"] = th
    case["fe"] = fe
    for steps in (1, 2, 4):
        r = _one_step(case, steps)
        assert not r["frames_exhausted"]
        tf = r["time_fe"]
        f = np.searchsorted(t, tf, side="right")  # theta_[0] is frame f
        assert 1 <= f < len(t)
        aff = slice(9, None)
        np.testing.assert_allclose(r["theta"][aff], A[aff] + B[aff] * tf, rtol=1e-12)
        assert np.all(r["theta"][:5] == void)
        if f <= 3:
            assert np.all(r["theta"][5:9] == (th[f, 5:9] if f == 3 else void))


def test_single_molten_cell_melts_flags_and_cools():
    """Step 1 melts exactly the hot cell; step 2 flags its 26 solid neighbours
    and marks the cells whose temperature fell as cooling (:338-404, :485-542)."""
    case = rc.single_cell_case()
    hc = case["hot_cell"]
    r1 = _one_step(case, 1)
    liquid = np.flatnonzero(r1["ori"] < 0)
    g = case["grid"]
    interior = (np.abs(_cell_centres(g) - np.array([g["x0"], g["y0"], g["z0"]])
                       - 0.5 * g["h"] * np.array([g["nx"], g["ny"], g["nz"]])) <
                0.5 * g["h"] * np.array([g["nx"], g["ny"], g["nz"]])).all(1)
    assert set(liquid[interior[liquid]]) == {hc}
    assert r1["melted"][hc] == 1 and r1["melted"].sum() == 1
    assert r1["T"][hc] == pytest.approx(1605.0, abs=1e-9)
    r2 = _one_step(case, 2)
    assert r2["steps"]["n_flagged"][1] == 26
    nxG, nyG = g["nx"] + 2, g["ny"] + 2
    faces = {hc + d for d in (-1, 1, -nxG, nxG, -nxG * nyG, nxG * nyG)}
    assert set(np.flatnonzero(r2["cooling"])) == faces | {hc}


@pytest.mark.parametrize("fresh", [False, True])
def test_single_molten_cell_goes_to_the_first_cell_in_x_major_order(fresh):
    """All six face neighbours reach the molten cell in the same step; the
    sweep runs i outermost (:552-554), so the -x neighbour captures it.  With
    no grain record on the molten cell a new record is appended (:831-845)."""
    case = rc.single_cell_case(fresh=fresh)
    hc = case["hot_cell"]
    st = case["state"]
    r = ol.run_remelt(case)
    assert r["ori"][hc] == st["ori"][hc - 1]
    others = np.ones(len(st["ori"]), bool)
    others[hc] = False
    np.testing.assert_array_equal(r["ori"][others], st["ori"][others])
    n0 = len(st["grain_cell"])
    if fresh:
        assert r["n_grains"] == n0 + 1
        assert r["gid"][hc] == n0 and r["grain_cell"][n0] == hc
    else:
        assert r["n_grains"] == n0 and r["gid"][hc] == st["gid"][hc]
    assert r["grain_orient"][r["gid"][hc]] == st["ori"][hc - 1]
    # closed-form time steps: dt = min(next frame - t_FE, 0.2 h / v(Tm - 5))
    ph, h = case["physics"], case["grid"]["h"]
    v = ph["a1"] * 5.0 + ph["a2"] * 25.0
    t = case["fe"]["times"]
    steps = r["steps"]
    ncap = np.flatnonzero(steps["n_captured"])
    assert len(ncap) == 1 and steps["n_captured"][ncap[0]] == 1
    assert steps["n_fresh"][ncap[0]] == (1 if fresh else 0)
    for s in range(1, ncap[0] + 1):
        t0 = steps["time_before"][s]
        nxt = t[np.searchsorted(t, t0, side="right")]
        assert steps["dt"][s] == pytest.approx(min(nxt - t0, 0.2 * h / v), rel=1e-12)
    # growth starts in step 2; the face neighbour captures once L = v*elapsed >= h
    grown = v * (steps["time_before"] + steps["dt"] - t[1])
    assert grown[ncap[0]] >= h and grown[ncap[0] - 1] < h


def test_nucleation_sites_match_the_plain_sampler():
    """The remelting oracle draws its sites with the same sampler as the plain
    path (CellularAutomataManager.cpp:153-213), whose draws are pinned by KATs."""
    nuc = dict(surface=(4e11, 2.0, 0.5), bulk=(2e16, 5.0, 1.0))
    case = rc.track_case(nx=24, ny=10, nz=8, n_frames=4, P=30.0, nucleation=nuc, seed=7)
    case["max_steps"] = 0
    r = ol.run_remelt(case)
    g = case["grid"]
    prm = ol.plain_params(nx=g["nx"], ny=g["ny"], nz=g["nz"], h=g["h"], ns=4e11, dTs_mean=2.0,
                          dTs_sd=0.5, nv=2e16, dTv_mean=5.0, dTv_sd=1.0, seed=7, final_time=0.0)
    p = ol.run_plain(prm, want_grains=False)
    assert len(r["site_cell"]) > 20
    np.testing.assert_array_equal(r["site_cell"], p["sites_cell"])
    np.testing.assert_array_equal(r["site_dt"], p["sites_dt"])


def test_track_bookkeeping_invariants():
    """Over a melt track: grain records = loaded + fresh captures + fresh nuclei,
    orientation rows = loaded + nuclei, and cells never melted keep their
    loaded orientation."""
    nuc = dict(surface=(0.0, 1.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=40, ny=14, nz=12, n_frames=14, P=30.0, nucleation=nuc)
    r = ol.run_remelt(case)
    st = case["state"]
    s = r["steps"]
    assert s["n_captured"].sum() > 100
    assert r["n_orient"] == len(st["euler"]) + s["n_nucleated"].sum()
    fresh_nuclei = r["n_grains"] - len(st["grain_cell"]) - s["n_fresh"].sum()
    assert 0 <= fresh_nuclei <= s["n_nucleated"].sum()
    never = (r["melted"] == 0) & (st["ori"] >= 0)
    np.testing.assert_array_equal(r["ori"][never], st["ori"][never])
    # every solid interior cell's orientation is its grain record's
    solid = (r["ori"] >= 0) & (r["gid"] >= 0)
    np.testing.assert_array_equal(r["grain_orient"][r["gid"][solid]], r["ori"][solid])


def test_remelting_velocity_pow_is_not_the_product():
    """The reference's remelting velocity (ProblemPhysics.cpp:142-158) takes
    pow(dT, b); the oracle calls glibc's pow.  glibc's pow(x, 2) is not
    correctly rounded: it is the neighbouring double of x*x for a small share of
    arguments, among them the undercooling (1600 K minus this cell temperature)
    that a grain of the 240 x 64 x 40 GPU track meets at step 393, so the
    velocity cannot be taken as products."""
    import math
    Tc = float.fromhex("0x1.86299cd5e950cp+10")
    x = 1600.0 - Tc
    assert math.pow(x, 1.0) == x
    assert (x * x).hex() == "0x1.831a0b997f4fdp+10"
    assert math.pow(x, 2.0).hex() == "0x1.831a0b997f4fep+10"
    rng = np.random.default_rng(7)
    xs = rng.uniform(0.0, 300.0, 200_000)
    share = np.mean([math.pow(v, 2.0) != v * v for v in xs])
    assert 1e-4 < share < 1e-2, share
