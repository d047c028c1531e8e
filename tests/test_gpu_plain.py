"""GPU parity of the plain CA step (realm type `cellular_automata`).

Every case runs the HIP path through the C ABI and the CPU oracle on the same
input deck (same seed, same host-drawn sites and orientations) and requires
bit-exact equality of the grain-ID fields (cellOrientationID_, cellGrainID_),
of the FP64 envelope (centre and half-diagonal) of every grain, of the live
nucleation sites and of the per-step log (time, dt, counts).  Full-size runs are
checked through size-independent invariants.
"""
import numpy as np
import pytest

from amca3d_amd import decks

pytestmark = pytest.mark.gpu

TRUNC = float(np.sqrt(np.float32(3.0)))


def case(nx, ny, nz, h, *, surface=(0.0, 2.0, 0.5), faces="all", bulk=(0.0, 3.0, 1.0),
         T0=-0.2, Tdot=-20.0, time_step=1e-3, termination_time=17.81, seed=0, lateral=None,
         max_half_diagonal_length=None):
    text = decks.plain_deck(nx, ny, nz, h, surface=surface, surface_faces=faces, bulk=bulk,
                            T0=T0, Tdot=Tdot, time_step=time_step,
                            termination_time=termination_time, random_seed=seed or None,
                            lateral=lateral, max_half_diagonal_length=max_half_diagonal_length)
    prm = dict(nx=nx, ny=ny, nz=nz, h=h, ns=surface[0], dTs_mean=surface[1], dTs_sd=surface[2],
               nv=bulk[0], dTv_mean=bulk[1], dTv_sd=bulk[2], surface_mode=1 if faces == "z_min" else 0,
               seed=seed, T0=T0, Tdot=Tdot, max_dt=time_step, final_time=termination_time)
    if max_half_diagonal_length is not None:
        # envelope.max_half_diagonal_length f: truncation sqrtf(2) * f * 0.5
        # (Octahedron.cpp:316-320), in double after the float square root
        prm["trunc"] = float(np.sqrt(np.float32(2.0))) * float(max_half_diagonal_length) * 0.5
    return text, prm


def run_gpu(amca, text, log_cap=1 << 16, max_steps=-1):
    s = amca.Simulation(text).initialize()
    d = s.domain
    d.enable_log(log_cap)
    c = s.run(max_steps)
    out = d.download()
    out["grains"] = d.download_grains()
    out["log"] = d.read_log()
    out["counters"] = c
    out["initial_sites"] = s.initial_sites()
    s.close()
    return out


def compare(oracle, prm_kw, g, r=None):
    p = oracle.plain_params(**prm_kw)
    if r is None:
        r = oracle.run_plain(p)
    assert g["counters"]["step"] == r["n_steps"]
    assert g["counters"]["time"] == r["time"]
    assert g["counters"]["n_grains"] == r["n_grains"]
    assert g["counters"]["n_orient"] == r["n_orient"]
    assert g["counters"]["n_active"] == r["n_active"]
    assert np.array_equal(g["ori"], r["ori"]), "orientation (grain-ID) field differs"
    assert np.array_equal(g["gid"], r["gid"]), "grain id field differs"
    # envelopes, per cell
    env_cell = np.zeros((len(r["ori"]), 4))
    env_cell[r["grain_cell"]] = r["env"]
    cells = r["grain_cell"]
    assert np.array_equal(g["env"][cells], env_cell[cells]), "FP64 envelopes differ"
    # grain records in grain-id order (grainVector_)
    assert np.array_equal(g["grains"]["env"], r["env"])
    assert np.array_equal(g["grains"]["cell"], r["grain_cell"])
    assert np.array_equal(g["grains"]["orient"], r["grain_orient"])
    # live sites and the erased ones (dT set to 0)
    sc, sd, _ = g["initial_sites"]
    live = dict(zip(g["sites_cell"].tolist(), g["sites_dt"].tolist()))
    nuc = np.full(len(r["ori"]), -1.0)
    nuc[sc] = [live.get(c, 0.0) for c in sc.tolist()]
    assert np.array_equal(nuc, r["nuc_dt"]), "nucleation-site table differs"
    # per-step log
    lg, st = g["log"], r["steps"]
    n = min(len(lg), len(st))
    assert n == r["n_steps"]
    for f in ("time_before", "dt", "n_nucleated", "n_active_before", "n_captured",
              "n_active_after", "n_liquid", "n_wave1", "max_wave"):
        assert np.array_equal(lg[f][:n], st[f][:n].astype(lg[f].dtype)), f
    # live sites at each step's start: all of them first, then each step drops
    # at least its nuclei, and the last step leaves the final live list
    ns, nn = lg["n_sites"][:n].astype(np.int64), lg["n_nucleated"][:n]
    assert ns[0] == len(sc)
    assert (ns[:-1] - ns[1:] >= nn[:-1]).all()
    assert ns[-1] - nn[-1] >= len(g["sites_cell"])
    return r


def test_device_envelope_matches_oracle(amca, oracle):
    """Capture test and child envelope on 400k random inputs, many on or near
    the envelope boundary: identical booleans and identical doubles."""
    rng = np.random.default_rng(20261202)
    n = 400_000
    h = 2e-6
    eul = rng.uniform([0, 0, 0], [2 * np.pi, np.pi, np.pi / 2], size=(512, 3))
    R = oracle.rotations(eul)[rng.integers(0, 512, n)]
    centre = rng.uniform(0, 40 * h, size=(n, 3))
    L = rng.uniform(0.0, 4 * h, size=n)
    L[:1000] = 0.0
    cells = (np.floor(centre / h + rng.integers(-2, 3, size=(n, 3))) + 0.5) * h
    env = np.column_stack([centre, L])
    # a quarter exactly on the boundary: L = ||R^T (x - c)||_1 evaluated in order
    k = n // 4
    d = cells[:k] - centre[:k]
    Rm = R[:k].reshape(-1, 3, 3)
    loc = (d[:, 0:1] * Rm[:, 0, :] + d[:, 1:2] * Rm[:, 1, :]) + d[:, 2:3] * Rm[:, 2, :]
    env[:k, 3] = np.abs(loc).sum(axis=1)
    gi, gc = amca.envelope_selftest(env, R, cells, h, TRUNC)
    oi, oc = oracle.envelope_batch(env, R, cells, h, TRUNC)
    assert np.array_equal(gi, oi)
    assert gi.sum() > n // 4
    assert np.array_equal(gc, oc)


def test_final_grain_files_of_a_deck_run(amca, oracle, tmp_path):
    """A plain deck run from its file and finalised writes finalGrains.vtk next
    to the deck and Results/finalResult.txt (output_data_base_name
    ./Results/Grains): both equal the oracle's writer applied to the oracle's
    own run of the case."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import final_grains as fg
    from amca3d_amd.sim import Simulation
    text, prm = case(**SMALL[1])
    deck = tmp_path / "deck.yaml"
    deck.write_text(text)
    s = Simulation.from_file(str(deck))
    s.initialize()
    s.run(-1)
    s.finalize()
    s.close()
    r = oracle.run_plain(oracle.plain_params(**prm))
    vtk, res = fg.final_grains_text(prm["nx"], prm["ny"], prm["nz"], 0.0, 0.0, 0.0, prm["h"],
                                    r["ori"], None, r["euler"])
    assert (tmp_path / "finalGrains.vtk").read_text() == vtk
    assert (tmp_path / "Results" / "finalResult.txt").read_text() == res


def test_microstructure_file_of_a_plain_deck(amca, oracle, tmp_path):
    """output_microstructure_information on a cellular_automata realm: written
    at the end of the run (CellularAutomataManager.cpp:144-145) with the
    reference's rank suffix; every section equals the oracle's run of the same
    case -- ids, orientation ids, no melted flags, cellNucleationDT_ (-1 off the
    sites, 0 where the nucleation pass dropped a site, the undercooling of the
    sites still listed), Euler angles and grain centres at six digits."""
    import remelt_cases as rc
    from amca3d_amd.sim import Simulation
    text, prm = case(**SMALL[0])
    old = "       options:\n"
    assert old in text
    text = text.replace(old, old + "         - output_microstructure_information:\n"
                                   "             file_name: micro.txt\n", 1)
    deck = tmp_path / "deck.yaml"
    deck.write_text(text)
    s = Simulation.from_file(str(deck))
    s.initialize()
    s.run(-1)
    inf = s.info()
    s.finalize()
    s.close()
    r = oracle.run_plain(oracle.plain_params(**prm))
    m = rc.read_micro_file(tmp_path / "micro0.txt")
    nG = (prm["nx"] + 2) * (prm["ny"] + 2) * (prm["nz"] + 2)
    head = m["header"]
    assert int(head[0]) == nG and int(head[1]) == r["n_orient"] and int(head[2]) == r["n_grains"]
    assert m["n_expected"] == inf["n_surface_sites"] + inf["n_bulk_sites"]
    assert [int(v) for v in head[7:10]] == [prm["nx"] + 2, prm["ny"] + 2, prm["nz"] + 2]
    np.testing.assert_array_equal(m["gid"], r["gid"])
    np.testing.assert_array_equal(m["ori"], r["ori"])
    np.testing.assert_array_equal(m["melted"], np.zeros(nG, np.uint8))
    np.testing.assert_allclose(m["nuc_dt"], r["nuc_dt"], rtol=1e-5, atol=0)
    assert (m["nuc_dt"] == 0).sum() > 0 and (m["nuc_dt"] == -1).sum() > 0
    np.testing.assert_array_equal(m["orient_id"], np.arange(r["n_orient"]))
    np.testing.assert_allclose(m["euler"], r["euler"], rtol=1e-5, atol=1e-9)
    np.testing.assert_array_equal(m["grain_id"], np.arange(r["n_grains"]))
    np.testing.assert_array_equal(m["grain_cell"], r["grain_cell"])
    np.testing.assert_array_equal(m["grain_orient"], r["grain_orient"])
    np.testing.assert_allclose(m["grain_centre"], r["env"][:, :3], rtol=1e-5, atol=1e-12)


def test_vtr_series_of_a_deck_run(amca, oracle, tmp_path):
    """The output block's VTR time series (enabled): a frame after every
    output_frequency-th step and one at the end, each byte-equal to the
    oracle's restatement of VtrManager applied to the oracle's own run
    stopped at that step, and the .pvd listing their times.  The variable
    list holds a name the reference does not know (skipped) and temperature,
    which a realm without a temperature field declares in the .pvtr only."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import vtr_series as vs
    from amca3d_amd.sim import Simulation
    text, prm = case(**SMALL[1])
    n_steps = oracle.run_plain(oracle.plain_params(**prm))["n_steps"]
    freq = n_steps // 4 + 1
    names = ["grain_ID", "no_such_variable", "grain_orientation", "temperature"]
    text = text.replace("      output_frequency: 100\n      output_variables:\n        - grain_ID",
                        f"      output_frequency: {freq}\n      output_variables:\n" +
                        "\n".join(f"        - {v}" for v in names))
    assert f"output_frequency: {freq}" in text
    deck = tmp_path / "deck.yaml"
    deck.write_text(text)
    s = Simulation.from_file(str(deck))
    s.enable_vtr(True)
    s.initialize()
    s.run(-1)
    s.finalize()
    s.close()
    steps = list(range(freq, n_steps + 1, freq)) + [n_steps]  # the last from finalize
    ids = vs.variable_ids(names)
    root = tmp_path / "Results"
    times = []
    for c, st in enumerate(steps):
        r = oracle.run_plain(oracle.plain_params(**dict(prm, max_steps=st)))
        assert r["n_steps"] == st
        times.append(r["time"])
        frame = root / "grainsTimeStates" / f"timeStep_{c:05d}"
        want = vs.vtr_text(prm["nx"], prm["ny"], prm["nz"], 0.0, 0.0, 0.0, prm["h"], ids,
                           r["ori"], r["euler"])
        assert (frame / "Grains00000.vtr").read_text() == want, c
        assert (frame / f"Grains{c:05d}.pvtr").read_text() == \
            vs.pvtr_text(prm["nx"], prm["ny"], prm["nz"], ids, "Grains")
    assert not (root / "grainsTimeStates" / f"timeStep_{len(steps):05d}").exists()
    assert (root / "Grains.pvd").read_text() == vs.pvd_text(times, "Grains")


def test_driver_runs_a_deck(amca, oracle, tmp_path):
    """The amca3d driver (`amca3d -i deck`, the reference's cafe command line)
    as a child process: it runs the deck to the end, writes the VTR frames the
    output block names and finalGrains.vtk, and the files hold the oracle's
    state."""
    import os
    import subprocess
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import final_grains as fg
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "amca3d_amd", "_build", "amca3d")
    text, prm = case(**SMALL[2])
    r = oracle.run_plain(oracle.plain_params(**prm))
    freq = r["n_steps"] // 3 + 1
    text = text.replace("output_frequency: 100", f"output_frequency: {freq}")
    (tmp_path / "deck.yaml").write_text(text)
    p = subprocess.run([exe, "-i", "deck.yaml"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert f"steps {r['n_steps']} " in p.stdout
    frames = sorted((tmp_path / "Results" / "grainsTimeStates").iterdir())
    assert len(frames) == r["n_steps"] // freq + 1
    vtk, res = fg.final_grains_text(prm["nx"], prm["ny"], prm["nz"], 0.0, 0.0, 0.0, prm["h"],
                                    r["ori"], None, r["euler"])
    assert (tmp_path / "finalGrains.vtk").read_text() == vtk
    assert (tmp_path / "Results" / "finalResult.txt").read_text() == res


def test_nucleation_example_parity(amca, oracle):
    """The reference's nucleation example deck (60^3, 270 bulk draws) to full
    solidification."""
    g = run_gpu(amca, decks.nucleation_example_deck())
    prm = dict(nx=60, ny=60, nz=60, h=0.0005, nv=1e7, dTv_mean=3.0, dTv_sd=1.0,
               ns=0.0, dTs_mean=2.0, dTs_sd=0.5, T0=-0.2, Tdot=-20.0, max_dt=1e-3,
               final_time=17.81)
    r = compare(oracle, prm, g)
    assert r["done"] == 1 and g["counters"]["n_liquid"] == 0


SMALL = [
    dict(nx=24, ny=24, nz=24, h=2e-6, surface=(1e9, 1.0, 0.5), faces="z_min",
         bulk=(5e13, 3.0, 1.0), Tdot=-200.0, time_step=1e-4),
    dict(nx=17, ny=23, nz=11, h=2e-6, surface=(2e9, 1.0, 0.5), faces="all",
         bulk=(8e13, 3.0, 1.0), Tdot=-500.0, time_step=1e-4),
    dict(nx=20, ny=9, nz=24, h=5e-6, surface=(0.0, 1.0, 0.5), faces="all",
         bulk=(3e12, 2.0, 1.0), Tdot=-2000.0, time_step=1e-4, T0=-1.0),
]


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 11])
@pytest.mark.parametrize("ci", range(len(SMALL)))
def test_small_grids_all_seeds(amca, oracle, seed, ci):
    text, prm = case(**SMALL[ci], seed=seed)
    g = run_gpu(amca, text)
    compare(oracle, prm, g)


def test_uniform_cooling_96_parity(amca, oracle):
    """The 512^3 uniform-cooling workload's physics (2e-6 m cells, -200 K/s,
    5e13 m^-3 bulk, 1e9 m^-2 on z = 0) on a 96^3 grid, to full solidification."""
    text, prm = case(96, 96, 96, 2e-6, surface=(1e9, 1.0, 0.5), faces="z_min",
                     bulk=(5e13, 3.0, 1.0), Tdot=-200.0, time_step=1e-4, termination_time=1.0)
    g = run_gpu(amca, text)
    r = compare(oracle, prm, g)
    assert r["done"] == 1


def test_ragged_grid_parity(amca, oracle):
    """131 x 97 x 67 cells (ghost-inclusive extents 133, 99, 69: no axis fills
    whole 4-cell bricks), surface sites on all six faces plus the bulk, seed 7,
    to full solidification: bit-exact against the oracle."""
    text, prm = case(131, 97, 67, 2e-6, surface=(1e9, 1.0, 0.5), faces="all",
                     bulk=(5e13, 3.0, 1.0), Tdot=-200.0, time_step=1e-4, termination_time=1.0,
                     seed=7)
    g = run_gpu(amca, text)
    r = compare(oracle, prm, g)
    assert r["done"] == 1


def test_capture_cascade_beyond_wave_one(amca, oracle):
    """144^3 with the 512^3 workload's physics: in step 205 a grain captured
    in that step captures again (wave 2, the path of waves >= 2).  212 steps,
    bit-exact against the oracle, and both step logs record the wave."""
    text, prm = case(144, 144, 144, 2e-6, surface=(1e9, 1.0, 0.5), faces="z_min",
                     bulk=(5e13, 3.0, 1.0), Tdot=-200.0, time_step=1e-4, termination_time=1.0)
    prm["max_steps"] = 212
    g = run_gpu(amca, text, max_steps=212)
    r = compare(oracle, prm, g)
    assert r["n_steps"] == 212
    assert r["steps"]["max_wave"][205] == 2
    assert g["log"]["max_wave"][205] == 2


def test_capture_cascade_to_wave_four(amca, oracle):
    """96^3 with the 512^3 workload's physics and the deck's
    envelope.max_half_diagonal_length at 8 (truncation sqrtf(2) * 8 / 2): larger
    children capture again in the same step, down to wave 4 (two steps reach
    wave 3 or deeper).  The whole run, 433 steps to full solidification, is
    bit-exact against the oracle, and both logs record the waves."""
    text, prm = case(96, 96, 96, 2e-6, surface=(1e9, 1.0, 0.5), faces="z_min",
                     bulk=(5e13, 3.0, 1.0), Tdot=-200.0, time_step=1e-4, termination_time=1.0,
                     max_half_diagonal_length=8.0)
    g = run_gpu(amca, text)
    r = compare(oracle, prm, g)
    assert g["counters"]["done"] == 1
    assert r["steps"]["max_wave"].max() == 4 and (r["steps"]["max_wave"] >= 3).sum() >= 2
    assert np.array_equal(g["log"]["max_wave"], r["steps"]["max_wave"])


def test_given_sites_identity_nucleus(amca, oracle):
    """Low-level ABI with explicit sites and orientation rows; one identity
    nucleus grows to full solidification of a 21^3 box."""
    n = 21
    p = oracle.plain_params(nx=n, ny=n, nz=n, h=1e-6, T0=-5.0, Tdot=-200.0, max_dt=1e-4,
                            final_time=1.0)
    c = oracle.cell_index(p, 11, 11, 11)
    eul = np.zeros((1, 3))
    r = oracle.run_plain(p, given=([c], [0.5], eul))
    d = amca.OwnedDomain(nx=n, ny=n, nz=n, h=1e-6, orient_capacity=1, T0=-5.0, Tdot=-200.0,
                         max_dt=1e-4, final_time=1.0)
    d.set_sites([c], [0.5])
    d.push_orientations(eul, oracle.rotations(eul))
    cnt = d.run(-1)
    out = d.download()
    assert cnt["step"] == r["n_steps"] and cnt["done"] == 1
    assert np.array_equal(out["ori"], r["ori"]) and np.array_equal(out["gid"], r["gid"])
    env_cell = np.zeros((len(r["ori"]), 4))
    env_cell[r["grain_cell"]] = r["env"]
    assert np.array_equal(out["env"][r["grain_cell"]], env_cell[r["grain_cell"]])
    d.close()


def test_rows_must_cover_every_site(amca, oracle):
    """Each live site can take one orientation row, so run, step and run_fixed
    refuse (before launching anything) while the pushed rows do not cover the
    orientations in use plus the live sites; with the missing row pushed the
    run proceeds."""
    from amca3d_amd._lib import AmcaError
    n = 12
    p = oracle.plain_params(nx=n, ny=n, nz=n, h=1e-6, T0=-5.0, Tdot=-200.0, max_dt=1e-4,
                            final_time=1.0)
    cells = [oracle.cell_index(p, 4, 4, 4), oracle.cell_index(p, 8, 8, 8)]
    eul = np.zeros((2, 3))
    d = amca.OwnedDomain(nx=n, ny=n, nz=n, h=1e-6, orient_capacity=2, T0=-5.0, Tdot=-200.0,
                         max_dt=1e-4, final_time=1.0)
    d.set_sites(cells, [0.5, 0.5])
    d.push_orientations(eul[:1], oracle.rotations(eul[:1]))
    for call in (lambda: d.run(-1), d.step, lambda: d.run_fixed(3)):
        with pytest.raises(AmcaError, match="do not cover"):
            call()
    assert d.counters()["step"] == 0
    d.push_orientations(eul[1:], oracle.rotations(eul[1:]))
    cnt = d.run(-1)
    assert cnt["done"] == 1 and cnt["n_orient"] == 2
    d.close()


def test_stepwise_equals_run_and_repeatable(amca, oracle):
    text, prm = case(**SMALL[1], seed=5)
    a = run_gpu(amca, text)
    s = amca.Simulation(text).initialize()
    d = s.domain
    while True:
        c = d.step()
        if c["done"]:
            break
    b = d.download()
    s.close()
    assert np.array_equal(a["ori"], b["ori"]) and np.array_equal(a["gid"], b["gid"])
    assert np.array_equal(a["env"], b["env"])
    a2 = run_gpu(amca, text)
    assert np.array_equal(a["env"], a2["env"]) and np.array_equal(a["gid"], a2["gid"])


@pytest.mark.parametrize("graph_steps", ["0", "3"])
def test_graph_length_changes_nothing(amca, graph_steps, monkeypatch):
    """Whole steps replayed from HIP graphs of 16 steps (the default), of 3, or
    launched one by one (AMCA_DOMAIN_GRAPH_STEPS=0): fields, records and the
    step log are the same bit for bit; run_fixed past termination changes
    nothing."""
    text, prm = case(**SMALL[0])
    a = run_gpu(amca, text)
    monkeypatch.setenv("AMCA_DOMAIN_GRAPH_STEPS", graph_steps)
    b = run_gpu(amca, text)
    for k in ("ori", "gid", "env"):
        assert np.array_equal(a[k], b[k]), k
    for f in a["log"].dtype.names:
        assert np.array_equal(a["log"][f], b["log"][f]), f
    s = amca.Simulation(text).initialize()
    s.domain.run_fixed(a["counters"]["step"] + 37)
    c = s.domain.counters()
    o = s.domain.download()
    s.close()
    assert c["step"] == a["counters"]["step"] and np.array_equal(o["gid"], a["gid"])


def test_graph_length_is_checked(amca, monkeypatch):
    monkeypatch.setenv("AMCA_DOMAIN_GRAPH_STEPS", "sixteen")
    text, _ = case(6, 5, 4, 1e-3)
    with pytest.raises(Exception, match="AMCA_DOMAIN_GRAPH_STEPS"):
        amca.Simulation(text).initialize()


def test_edge_cases(amca, oracle):
    # no sites: every step is empty until the termination time
    text, prm = case(6, 5, 4, 1e-3, termination_time=0.0105, time_step=1e-3)
    g = run_gpu(amca, text)
    r = compare(oracle, prm, g)
    assert g["counters"]["step"] == 11 and g["counters"]["n_liquid"] == 6 * 5 * 4
    # a single cell
    text, prm = case(1, 1, 1, 1e-3, bulk=(1e12, 1.0, 0.1))
    g = run_gpu(amca, text)
    compare(oracle, prm, g)
    assert g["counters"]["n_grains"] == 1 and g["counters"]["n_active"] == 0
    # a line of cells with a saturated site density (every cell a site)
    text, prm = case(40, 1, 1, 1e-4, bulk=(1e13, 3.0, 1.0))
    g = run_gpu(amca, text)
    compare(oracle, prm, g)


def test_full_size_512_invariants(amca):
    """512^3 cells (the benchmark workload) to full solidification, through
    capture cascades of depth 3 and more: every cell solid, ids a permutation
    of 0..N-1 whose records sit at their own cells with their cells'
    orientations, orientations in range and one per nucleus, the step log
    adding up (nuclei + captures = N, active lists chained, liquid count
    falling by exactly the step's new grains)."""
    s = amca.Simulation(decks.uniform_cooling_512_deck()).initialize()
    d = s.domain
    d.enable_log(1 << 12)
    c = s.run(-1)
    n = 512
    N = n ** 3
    assert c["done"] == 1 and c["n_liquid"] == 0 and c["n_grains"] == N
    lg = d.read_log()
    assert len(lg) == c["step"] > 400
    nuc, cap = lg["n_nucleated"].astype(np.int64), lg["n_captured"].astype(np.int64)
    assert nuc.sum() + cap.sum() == N and nuc.sum() == c["n_orient"]
    liq = lg["n_liquid"].astype(np.int64)
    assert (np.diff(np.concatenate([[N], liq])) == -(nuc + cap)).all()
    assert (lg["n_active_before"][1:] == lg["n_active_after"][:-1] + lg["n_nucleated"][1:]).all()
    assert lg["max_wave"].max() >= 3
    out = d.download(env=False)
    ori = out["ori"].reshape(n + 2, n + 2, n + 2)
    gid = out["gid"].reshape(n + 2, n + 2, n + 2)
    inner = (slice(1, -1),) * 3
    assert (ori[inner] >= 0).all() and ori[inner].max() < c["n_orient"]
    g = gid[inner].ravel()
    assert np.bincount(g, minlength=N).max() == 1 and g.min() == 0 and g.max() == N - 1
    gr = d.download_grains()
    assert len(gr["cell"]) == N
    # record g sits at a cell whose grain id is g, with that cell's orientation
    assert np.array_equal(gid.ravel()[gr["cell"]], np.arange(N, dtype=np.int32))
    assert np.array_equal(gr["orient"], ori.ravel()[gr["cell"]])
    assert (gr["env"][:, 3] >= 0).all()
    s.close()
