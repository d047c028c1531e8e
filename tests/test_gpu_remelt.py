"""Coupled CA-FE (remelting) path on the GPU, bit-exact against the oracle.

Every case runs the same inputs through the device domain (amca_remelt_* C
ABI) and the oracle, then compares the per-cell fields (orientation, grain id,
temperature, melted, cooling), the grain records, the nodal field, the step log
and the counters word for word."""
import numpy as np
import pytest

import oracle_lib as ol
import remelt_cases as rc

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle")]

IDENT = np.array([1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0])


def _desc(case):
    g, ph = case["grid"], case["physics"]
    d = {k: g[k] for k in ("nx", "ny", "nz", "h", "x0", "y0", "z0")}
    d.update({k: ph[k] for k in ("Tm", "Ts", "Tmin_nuc", "a1", "a2", "a3", "b1", "b2", "b3")})
    d.update(trunc=case["trunc"], time_step_factor=case["tsf"], cutoff_factor=case["cutoff_factor"],
             final_time=case["final_time"], void_temperature=case["void_temp"], bin=case["bin"],
             tolerance=case["tol"], max_iterations=case["max_iter"],
             initial_melted=case["initial_melted"])
    return d


def build_own_sampler(case, reserve=None):
    """A domain for `case` with the library's own host sampler and rand()
    stream (amca_remelt_draw_sites) instead of inputs taken from the oracle
    run; returns (domain, site count)."""
    from amca3d_amd.remelt import RemeltDomain
    st = case["state"]
    n0 = len(st["euler"])
    nuc = case["nucleation"]
    d = RemeltDomain(_desc(case), n0 + 200000)
    d.push_orientations(st["euler"])  # rotations computed by the library
    d.set_fe(case["fe"])
    d.set_state(st)
    n_sites = d.draw_sites(case["seed"], nuc["surface"], nuc["bulk"])
    if reserve is not None:
        d.set_row_reserve(reserve)
    d.enable_log(100000)
    return d, n_sites


def collect(d, c, n_sites=None):
    out = d.download()
    out.update(theta=d.theta(), log=d.read_log(), counters=c, tag_par=d.voxel_map(),
               euler=d.euler(), n_sites=n_sites)
    d.close()
    return out


def run_device_own_sampler(case, orc, reserve=None):
    """Run `case` with the library's own sampler (see build_own_sampler)."""
    d, n_sites = build_own_sampler(case, reserve)
    return collect(d, d.run(case["max_steps"]), n_sites)


def run_device(case, orc, chunks=None, window=True):
    """Run `case` on the device.  Nucleation sites and the orientation rows drawn
    after the loaded ones are the host-side inputs of the path; they come from
    the oracle run so both sides consume the same rand() stream."""
    from amca3d_amd.remelt import RemeltDomain
    st = case["state"]
    n0 = len(st["euler"])
    extra = orc["euler"][n0:]
    n_sites = len(orc["site_cell"]) if case.get("nucleation") else 0
    cap = n0 + len(extra) + 64 * max(n_sites, 1) + 16
    d = RemeltDomain(_desc(case), cap)
    d.push_orientations(st["euler"], ol.rotations(st["euler"]))
    rows = ol.rotations(extra) if len(extra) else np.zeros((0, 9))
    pos = [0]

    def source(n):
        take = rows[pos[0]:pos[0] + n]
        pos[0] += len(take)
        pad = n - len(take)
        eu = np.zeros((n, 3))
        r = np.concatenate([take, np.tile(IDENT, (pad, 1))]) if pad else take
        return eu, r

    d.set_orientation_source(source)
    d.set_fe(case["fe"])
    d.set_state(st)
    if n_sites:
        d.set_sites(orc["site_cell"], orc["site_dt"])
    if not window:
        d.set_window(False)
    d.enable_log(100000)
    if chunks:
        for k in chunks:
            d.run(k)
        c = d.counters()
    else:
        c = d.run(case["max_steps"])
    out = d.download()
    out.update(theta=d.theta(), log=d.read_log(), counters=c, tag_par=d.voxel_map())
    d.close()
    return out


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def compare(case, orc, dev):
    c = dev["counters"]
    assert c["step"] == orc["n_steps"]
    assert c["n_grains"] == orc["n_grains"]
    assert c["n_orient"] == orc["n_orient"]
    assert c["frames_exhausted"] == orc["frames_exhausted"]
    assert bits(np.float64(c["time"])) == bits(np.float64(orc["time"]))
    assert bits(np.float64(c["time_fe"])) == bits(np.float64(orc["time_fe"]))
    tag, par = dev["tag_par"]
    np.testing.assert_array_equal(tag, orc["tag"])
    np.testing.assert_array_equal(bits(par), bits(orc["par"]))
    np.testing.assert_array_equal(dev["ori"], orc["ori"])
    np.testing.assert_array_equal(dev["gid"], orc["gid"])
    np.testing.assert_array_equal(bits(dev["T"]), bits(orc["T"]))
    np.testing.assert_array_equal(dev["melted"], orc["melted"])
    np.testing.assert_array_equal(dev["cooling"], orc["cooling"])
    np.testing.assert_array_equal(bits(dev["theta"]), bits(orc["theta"]))
    n = orc["n_grains"]
    np.testing.assert_array_equal(dev["grain_cell"], orc["grain_cell"][:n])
    np.testing.assert_array_equal(dev["grain_orient"], orc["grain_orient"][:n])
    np.testing.assert_array_equal(bits(dev["grain_env"]), bits(orc["grain_env"][:n]))
    lg, so = dev["log"], orc["steps"]
    assert len(lg) == len(so)
    for k in ("n_nucleated", "n_captured", "n_fresh", "n_flagged"):
        np.testing.assert_array_equal(lg[k], so[k], err_msg=k)
    for k in ("time_before", "dt"):
        np.testing.assert_array_equal(bits(lg[k]), bits(so[k]), err_msg=k)


def _holes(case, frac, seed):
    """Cells that stay solid but lose their grain record id (gid = -1): when
    they melt and are recaptured, the capture appends a fresh record."""
    st = case["state"]
    rng = np.random.default_rng(seed)
    cells = st["grain_cell"]
    pick = cells[rng.uniform(size=len(cells)) < frac]
    st["gid"] = st["gid"].copy()
    st["gid"][pick] = -1
    return case


@pytest.mark.parametrize("fresh", [False, True])
def test_single_molten_cell(amca, fresh):
    case = rc.single_cell_case(fresh=fresh)
    orc = ol.run_remelt(case)
    dev = run_device(case, orc)
    compare(case, orc, dev)
    assert dev["ori"][case["hot_cell"]] == case["state"]["ori"][case["hot_cell"] - 1]


@pytest.fixture(scope="module")
def track_nuc():
    nuc = dict(surface=(0.0, 1.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=80, ny=24, nz=20, n_frames=26, P=30.0, nucleation=nuc)
    return case, ol.run_remelt(case)


def test_track_with_nucleation(amca, track_nuc):
    case, orc = track_nuc
    assert orc["steps"]["n_captured"].sum() > 1000
    assert orc["steps"]["n_nucleated"].sum() > 50
    compare(case, orc, run_device(case, orc))


def test_track_with_the_library_sampler(amca, track_nuc):
    """Sites and nucleus orientations drawn by the library's host rand() stream
    give the oracle's run, including every Euler angle."""
    case, orc = track_nuc
    dev = run_device_own_sampler(case, orc)
    assert dev["n_sites"] == len(orc["site_cell"])
    compare(case, orc, dev)
    np.testing.assert_array_equal(bits(dev["euler"]), bits(orc["euler"]))


@pytest.mark.parametrize("reserve", [0, 3])
def test_track_with_rows_drawn_on_demand(amca, track_nuc, reserve):
    """With (almost) no orientation rows drawn ahead, steps whose liquid sites
    could outrun the pushed rows stop and are repeated after the host draws
    more: the run is the oracle's."""
    case, orc = track_nuc
    dev = run_device_own_sampler(case, orc, reserve=reserve)
    compare(case, orc, dev)
    np.testing.assert_array_equal(bits(dev["euler"]), bits(orc["euler"]))


def test_run_many_equals_the_oracle(amca):
    """Three tracks of different size, power and seed run side by side
    (amca_remelt_run_many, one stream each): each equals its oracle run."""
    from amca3d_amd import remelt
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    cases = [rc.track_case(nx=48 + 16 * q, ny=18 + 2 * q, nz=14, n_frames=14 + 4 * q,
                           P=30.0 + 5.0 * q, nucleation=nuc, seed=21 + q) for q in range(3)]
    orcs = [ol.run_remelt(c) for c in cases]
    built = [build_own_sampler(c) for c in cases]
    counters = remelt.run_many([d for d, _ in built])
    for case, orc, (d, ns), c in zip(cases, orcs, built, counters):
        dev = collect(d, c, ns)
        assert dev["n_sites"] == len(orc["site_cell"])
        compare(case, orc, dev)
        np.testing.assert_array_equal(bits(dev["euler"]), bits(orc["euler"]))


@pytest.mark.parametrize("graph_steps", ["0", "5"])
def test_graph_length_changes_nothing(amca, track_nuc, graph_steps, monkeypatch):
    """Steps launched one by one (AMCA_REMELT_GRAPH_STEPS=0) or replayed from
    5-step graphs give the oracle's run."""
    monkeypatch.setenv("AMCA_REMELT_GRAPH_STEPS", graph_steps)
    case, orc = track_nuc
    compare(case, orc, run_device(case, orc))


def test_graph_length_is_checked(amca, track_nuc, monkeypatch):
    monkeypatch.setenv("AMCA_REMELT_GRAPH_STEPS", "-1")
    case, orc = track_nuc
    with pytest.raises(Exception, match="AMCA_REMELT_GRAPH_STEPS"):
        run_device(case, orc)


def test_track_in_chunks_equals_one_run(amca, track_nuc):
    case, orc = track_nuc
    dev = run_device(case, orc, chunks=[1, 2, 5, 13, 40, 1000])
    compare(case, orc, dev)


@pytest.mark.parametrize("chunks", [None, [1, 3, 7, 50, 1000]])
def test_active_box_equals_whole_grid(amca, chunks):
    """A track long and wide enough that each step's active box holds a small
    part of the grid: the boxed run, in one go or in chunks, equals the
    whole-grid run and the oracle bit for bit."""
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=120, ny=32, nz=24, n_frames=40, P=30.0, nucleation=nuc, seed=5)
    orc = ol.run_remelt(case)
    assert orc["steps"]["n_captured"].sum() > 1000
    boxed = run_device(case, orc, chunks=chunks)
    compare(case, orc, boxed)
    whole = run_device(case, orc, window=False)
    compare(case, orc, whole)
    ng = 122 * 34 * 26
    nb = boxed["log"]["n_box"]
    assert nb[0] == ng and (whole["log"]["n_box"] == ng).all()
    assert np.median(nb[1:]) < 0.25 * ng, np.median(nb[1:])


def test_wider_hotter_track(amca):
    """240 x 64 x 40 cells under 60 frames of a 60 W source, surface and bulk
    nucleation, seed 9 (about 630 steps, 1900 nuclei, 47 000 captures): the
    boxed run equals the oracle bit for bit (fields, temperatures, flags, grain
    ids and records, step log), and the box stays a small part of the grid.
    At step 393 one cell's undercooling is an argument where glibc's
    pow(dT, 2) is not dT*dT (test_remelting_velocity_pow_is_not_the_product
    pins it): the device's restated pow gives glibc's value."""
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=240, ny=64, nz=40, n_frames=60, P=60.0, nucleation=nuc, seed=9)
    orc = ol.run_remelt(case)
    assert orc["steps"]["n_nucleated"].sum() > 1000
    assert orc["steps"]["n_captured"].sum() > 10000
    dev = run_device(case, orc)
    compare(case, orc, dev)
    nb = dev["log"]["n_box"]
    assert np.median(nb[1:]) < 0.25 * (242 * 66 * 42), np.median(nb[1:])


def test_pbf_deck_kinetics(amca):
    """The growth coefficients of the reference's own PBF decks
    (PBF_x/Inputfile.i:73-79: Tm 1715.5, a1 = 0, a2 = 2.49e-3, a3 = 6.2e-4),
    whose cube term glibc evaluates as pow(dT, 3), on a 128 x 40 x 28 track with
    nucleation: bit-exact against the oracle over about 1700 steps whose dt
    each come from the velocity at a different minimum temperature.  The
    solidus is 20 K below Tm (the deck's is 117 K: at its 1000 m/s a 10 us
    frame would take ~10^4 steps, too long for a test)."""
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=128, ny=40, nz=28, n_frames=24, P=90.0, nucleation=nuc, seed=17)
    case["physics"].update(Tm=1715.5, Ts=1695.5, Tmin_nuc=1695.5, a1=0.0, a2=2.49e-3, a3=6.2e-4)
    orc = ol.run_remelt(case)
    st = orc["steps"]
    assert orc["n_steps"] > 1000 and len(np.unique(st["dt"])) > 1000
    assert st["n_nucleated"].sum() > 50
    assert st["n_captured"].sum() > 10000
    compare(case, orc, run_device(case, orc))


def test_device_pow_equals_glibc(amca, tmp_path):
    """The device velocity's pow (amca_pow_selftest: glibc's pow with the
    exact-product shortcut for the exponents 1, 2, 3) equals glibc's pow bit
    for bit: random undercoolings for the exponents 1, 2, 3, arguments whose
    exact square or cube lies within 1e-3 ulp of a rounding midpoint and within
    0.09 ulp (both sides of the shortcut's threshold), and the subnormal and
    rescaled branches."""
    import pow_check
    from amca3d_amd.remelt import pow_selftest
    _, args, glibc_pow = pow_check.build(tmp_path)
    sets = [(y, args(200_000, y, -45, 12, seed=3 + int(y))) for y in (1.0, 2.0, 3.0)]
    sets += [(y, args(50_000, y, -30, 11, seed=7 + int(y), hard=1, band=1e-3)) for y in (2.0, 3.0)]
    sets += [(y, args(100_000, y, -30, 11, seed=9 + int(y), hard=1, band=0.09)) for y in (2.0, 3.0)]
    sets += [(3.0, args(50_000, 3.0, -1074, -1023, seed=5)), (3.0, args(50_000, 3.0, -400, -100, seed=6)),
             (2.5, args(50_000, 2.5, -1074, 1023, seed=8))]
    for y, x in sets:
        got = pow_selftest(x, y)
        ref = glibc_pow(x, y)
        bad = np.flatnonzero(got.view(np.uint64) != ref.view(np.uint64))
        assert bad.size == 0, (y, x[bad[:3]], got[bad[:3]], ref[bad[:3]])


def test_track_with_fresh_captures(amca):
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = _holes(rc.track_case(nx=64, ny=20, nz=16, n_frames=22, P=30.0, nucleation=nuc,
                                seed=11), 0.3, 4)
    orc = ol.run_remelt(case)
    assert orc["steps"]["n_fresh"].sum() > 100
    compare(case, orc, run_device(case, orc))


def test_track_without_nucleation_stops_at_max_steps(amca):
    case = rc.track_case(nx=48, ny=16, nz=14, n_frames=16, P=30.0, max_steps=37)
    orc = ol.run_remelt(case)
    assert orc["n_steps"] == 37
    compare(case, orc, run_device(case, orc))


def test_track_runs_off_the_end_of_the_frames(amca):
    case = rc.track_case(nx=40, ny=14, nz=12, n_frames=6, P=30.0)
    orc = ol.run_remelt(case)
    assert orc["frames_exhausted"] == 1
    compare(case, orc, run_device(case, orc))


def test_termination_time(amca):
    case = rc.track_case(nx=40, ny=14, nz=12, n_frames=20, P=30.0, final_time=7.3e-5)
    orc = ol.run_remelt(case)
    assert orc["frames_exhausted"] == 0 and orc["n_steps"] > 5
    compare(case, orc, run_device(case, orc))


def test_distorted_mesh_and_hot_track(amca):
    """A non-affine FE mesh (Newton mapping) and a pool that melts through the
    top surface, with melted-from-the-start cells."""
    case = rc.track_case(nx=56, ny=20, nz=16, n_frames=14, P=80.0, seed=5)
    rng = np.random.default_rng(1)
    case["fe"]["xyz"] = case["fe"]["xyz"] + rng.uniform(-0.2, 0.2, case["fe"]["xyz"].shape) * 2.5e-6
    case["initial_melted"] = 1
    case["state"]["melted"] = np.ones_like(case["state"]["melted"])
    orc = ol.run_remelt(case)
    assert orc["steps"]["n_captured"].sum() > 500
    compare(case, orc, run_device(case, orc))


def test_state_validation(amca):
    from amca3d_amd._lib import AmcaError
    from amca3d_amd.remelt import RemeltDomain
    case = rc.single_cell_case()
    st = dict(case["state"])
    d = RemeltDomain(_desc(case), len(st["euler"]) + 8)
    d.push_orientations(st["euler"], ol.rotations(st["euler"]))
    d.set_fe(case["fe"])
    shared = st["gid"].copy()
    shared[st["grain_cell"][1]] = shared[st["grain_cell"][0]]
    with pytest.raises(AmcaError, match="shared"):
        d.set_state(dict(st, gid=shared))
    bad = st["grain_orient"].copy()
    bad[0] = len(st["euler"])
    with pytest.raises(AmcaError, match="orientation"):
        d.set_state(dict(st, grain_orient=bad))
    d.set_state(st)
    with pytest.raises(AmcaError, match="ascend"):
        d.set_sites([30, 20], [1.0, 1.0])
    d.close()


@pytest.mark.parametrize("distort", [False, True])
def test_device_voxel_map_equals_host(amca, distort):
    """tag_voxels on the GPU (one thread per cell) equals the host mapper bit for
    bit -- tags and parametric coordinates -- on box hexes and on distorted
    hexes, with bins small enough that cells search several of them."""
    from amca3d_amd import remelt
    case = rc.track_case(nx=41, ny=17, nz=13, n_frames=2, P=30.0)
    if distort:
        rng = np.random.default_rng(11)
        case["fe"]["xyz"] = case["fe"]["xyz"] + rng.uniform(-0.3, 0.3, case["fe"]["xyz"].shape) * 2.5e-6
    kw = dict(bin=(3, 2, 4), tol=case["tol"], max_iter=case["max_iter"])
    th, ph = remelt.map_voxels(case["grid"], case["fe"], **kw)
    td, pd = remelt.map_voxels(case["grid"], case["fe"], device=True, **kw)
    assert (th >= 0).sum() > 1000
    np.testing.assert_array_equal(td, th)
    np.testing.assert_array_equal(pd.view(np.uint64), ph.view(np.uint64))
