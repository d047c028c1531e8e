"""Coupled CA-FE input decks (finite_element + cellular_automata_remelting
realms) run through the library's realm layer on the GPU.

The deck names a temperature file in the reference's text frame format; the
realm reads it, draws a random substrate (one grain per interior cell) and the
nucleation sites from its rand() stream, runs the domain to the end and writes
the microstructure information file.  The oracle runs the same case from the
frames parsed here independently, and the two must agree bit for bit; the
written file must hold the device state and, loaded by a second deck, must
give the same run as the oracle started from that state."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
import remelt_cases as rc

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import final_grains as fg  # noqa: E402
import vtr_series as vs  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle")]

UNIT, TUNIT, OFFSET = 1e-3, 1e-3, (0.5, -0.25, 0.0)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def run_deck(deck_path):
    from amca3d_amd.sim import Simulation
    s = Simulation.from_file(deck_path)
    s.initialize()
    d = s.remelt
    d.enable_log(100000)
    c = s.run()
    rc_ = d.counters()
    out = d.download()
    out.update(theta=d.theta(), log=d.read_log(), counters=rc_, run=c, euler=d.euler(),
               sites=d.sites(), info=s.info())
    s.finalize()
    s.close()
    # end-of-run grain files: next to the deck, and in the folder of the
    # realm's output_data_base_name (./Results/Grains)
    folder = os.path.dirname(deck_path)
    with open(os.path.join(folder, "finalGrains.vtk")) as f:
        vtk = f.read()
    with open(os.path.join(folder, "Results", "finalResult.txt")) as f:
        res = f.read()
    out["final_files"] = (vtk, res)
    return out


def check_final_files(dev, case):
    g = case["grid"]
    want = fg.final_grains_text(g["nx"], g["ny"], g["nz"], g["x0"], g["y0"], g["z0"], g["h"],
                                dev["ori"], dev["melted"].astype(np.int32),
                                np.asarray(dev["euler"]).reshape(-1, 3))
    assert dev["final_files"][0] == want[0]
    assert dev["final_files"][1] == want[1]


def compare(orc, dev):
    c = dev["counters"]
    assert c["step"] == orc["n_steps"]
    assert c["n_grains"] == orc["n_grains"]
    assert c["n_orient"] == orc["n_orient"]
    assert bits(np.float64(c["time"])) == bits(np.float64(orc["time"]))
    np.testing.assert_array_equal(dev["ori"], orc["ori"])
    np.testing.assert_array_equal(dev["gid"], orc["gid"])
    np.testing.assert_array_equal(bits(dev["T"]), bits(orc["T"]))
    np.testing.assert_array_equal(dev["melted"], orc["melted"])
    np.testing.assert_array_equal(bits(dev["theta"]), bits(orc["theta"]))
    n = orc["n_grains"]
    np.testing.assert_array_equal(dev["grain_cell"], orc["grain_cell"][:n])
    np.testing.assert_array_equal(dev["grain_orient"], orc["grain_orient"][:n])
    np.testing.assert_array_equal(bits(dev["grain_env"]), bits(orc["grain_env"][:n]))
    np.testing.assert_array_equal(bits(dev["euler"]), bits(orc["euler"][:orc["n_orient"]]))
    lg, so = dev["log"], orc["steps"]
    assert len(lg) == len(so)
    for k in ("n_nucleated", "n_captured", "n_fresh", "n_flagged"):
        np.testing.assert_array_equal(lg[k], so[k], err_msg=k)
    for k in ("time_before", "dt"):
        np.testing.assert_array_equal(bits(lg[k]), bits(so[k]), err_msg=k)


@pytest.fixture(scope="module")
def cafe_run(tmp_path_factory):
    """Deck 1: random substrate and sites; deck 2: loads deck 1's file."""
    tmp = tmp_path_factory.mktemp("cafe")
    nuc = dict(surface=(2e10, 2.0, 0.5), bulk=(2e15, 5.0, 1.0))
    case = rc.track_case(nx=56, ny=20, nz=16, n_frames=16, P=30.0, nucleation=nuc, seed=17)
    rc.write_frames_text(tmp / "frames.txt", case["fe"], unit=UNIT, offset=OFFSET,
                         time_unit=TUNIT)
    (tmp / "deck1.yaml").write_text(rc.cafe_deck(case, mesh="frames.txt", micro_out="grainInfo"))
    dev1 = run_deck(str(tmp / "deck1.yaml"))
    fe = rc.parse_frames_text(tmp / "frames.txt")
    ocase = dict(case, fe=fe, state=None, random_micro=True)
    orc1 = ol.run_remelt(ocase)
    return dict(tmp=tmp, case=case, ocase=ocase, dev1=dev1, orc1=orc1)


def test_deck_with_random_substrate_matches_the_oracle(amca, cafe_run):
    dev, orc = cafe_run["dev1"], cafe_run["orc1"]
    nint = 56 * 20 * 16
    assert dev["info"]["coupled"] == 1 and dev["info"]["n_init_grains"] == nint
    assert orc["steps"]["n_captured"].sum() > 1000
    assert orc["steps"]["n_nucleated"].sum() > 20
    cell, dt = dev["sites"]
    np.testing.assert_array_equal(cell, orc["site_cell"])
    np.testing.assert_array_equal(bits(dt), bits(orc["site_dt"]))
    compare(orc, dev)
    check_final_files(dev, cafe_run["case"])


def test_microstructure_file_holds_the_device_state(amca, cafe_run):
    dev, orc = cafe_run["dev1"], cafe_run["orc1"]
    m = rc.read_micro_file(cafe_run["tmp"] / "grainInfo0.txt")
    n, no = orc["n_grains"], orc["n_orient"]
    np.testing.assert_array_equal(m["gid"], dev["gid"])
    np.testing.assert_array_equal(m["ori"], dev["ori"])
    np.testing.assert_array_equal(m["melted"], dev["melted"])
    sdt = np.full(len(m["gid"]), -1.0)
    sdt[orc["site_cell"]] = orc["site_dt"]
    np.testing.assert_allclose(m["nuc_dt"], sdt, rtol=1e-5, atol=1e-12)
    np.testing.assert_array_equal(m["orient_id"], np.arange(no))
    np.testing.assert_allclose(m["euler"], orc["euler"][:no].reshape(-1, 3), rtol=1e-5, atol=1e-9)
    np.testing.assert_array_equal(m["grain_id"], np.arange(n))
    np.testing.assert_array_equal(m["grain_cell"], orc["grain_cell"][:n])
    np.testing.assert_array_equal(m["grain_orient"], orc["grain_orient"][:n])
    env = orc["grain_env"][:n]
    np.testing.assert_allclose(m["grain_centre"], env[:, :3], rtol=1e-5, atol=1e-12)
    assert m["n_expected"] == len(orc["site_cell"])


@pytest.mark.parametrize("seeds", ["none", "loaded_and_drawn"])
def test_deck_loading_the_written_file(amca, cafe_run, seeds):
    """The written file as the next deck's initial state (default ostream
    precision: the loaded Euler angles and seed undercoolings are those
    six-digit values).  "none": no nucleation; "loaded_and_drawn": the file's
    seeds (load_seeds) with new surface and bulk draws merged into them."""
    tmp, case = cafe_run["tmp"], cafe_run["case"]
    drawn = seeds != "none"
    deck = rc.cafe_deck(case, mesh="frames.txt", micro_load="grainInfo",
                        micro_out=f"grainInfo_{seeds}", seeds=drawn, load_seeds=drawn)
    (tmp / f"deck_{seeds}.yaml").write_text(deck)
    dev = run_deck(str(tmp / f"deck_{seeds}.yaml"))
    m = rc.read_micro_file(tmp / "grainInfo0.txt")
    state = dict(ori=m["ori"], gid=m["gid"], melted=m["melted"], euler=m["euler"],
                 grain_cell=m["grain_cell"], grain_orient=m["grain_orient"])
    ocase = dict(cafe_run["ocase"], random_micro=False, state=state,
                 nucleation=case["nucleation"] if drawn else None,
                 init_site_dt=m["nuc_dt"] if drawn else None)
    orc = ol.run_remelt(ocase)
    assert orc["steps"]["n_captured"].sum() > 0
    compare(orc, dev)
    check_final_files(dev, case)
    out = rc.read_micro_file(tmp / f"grainInfo_{seeds}0.txt")
    if drawn:
        n_loaded = int((m["nuc_dt"] > -1).sum())
        assert len(orc["site_cell"]) > n_loaded
        cell, dt = dev["sites"]
        np.testing.assert_array_equal(cell, orc["site_cell"])
        np.testing.assert_array_equal(bits(dt), bits(orc["site_dt"]))
        assert out["n_expected"] == m["n_expected"] + dev["info"]["n_surface_sites"] + \
            dev["info"]["n_bulk_sites"]
    else:
        assert orc["steps"]["n_nucleated"].sum() == 0
        np.testing.assert_array_equal(out["nuc_dt"], np.full(len(m["gid"]), -1.0))
        assert out["n_expected"] == m["n_expected"]


def test_vtr_series_of_a_coupled_deck(amca, cafe_run):
    """The CA realm's VTR time series (enabled) on deck 1's case: every frame
    (cell temperature truncated to int, ids, Euler angles, melted flags, and
    the debug variable's .pvtr-only declarations) byte-equal to the oracle's
    restatement applied to the oracle's run stopped at that step."""
    from amca3d_amd.sim import Simulation
    tmp, case = cafe_run["tmp"], cafe_run["case"]
    n_steps = cafe_run["orc1"]["n_steps"]
    freq = n_steps // 3 + 1
    names = ["temperature", "grain_ID", "grain_orientation", "melted_flag", "debug"]
    deck = rc.cafe_deck(case, mesh=str(tmp / "frames.txt"))
    old = "      output_frequency: -1\n      output_variables:\n        - grain_orientation"
    assert old in deck
    deck = deck.replace(old, f"      output_frequency: {freq}\n      output_variables:\n" +
                        "\n".join(f"        - {v}" for v in names))
    sub = tmp / "vtr"
    sub.mkdir()
    (sub / "deck.yaml").write_text(deck)
    s = Simulation.from_file(str(sub / "deck.yaml"))
    s.enable_vtr(True)
    s.initialize()
    s.run()
    s.finalize()
    s.close()
    g = case["grid"]
    ids = vs.variable_ids(names)
    steps = list(range(freq, n_steps + 1, freq)) + [n_steps]
    times = []
    for c, st in enumerate(steps):
        orc = ol.run_remelt(dict(cafe_run["ocase"], max_steps=st))
        assert orc["n_steps"] == st
        times.append(orc["time"])
        frame = sub / "Results" / "grainsTimeStates" / f"timeStep_{c:05d}"
        want = vs.vtr_text(g["nx"], g["ny"], g["nz"], g["x0"], g["y0"], g["z0"], g["h"], ids,
                           orc["ori"], orc["euler"][:orc["n_orient"]], T=orc["T"],
                           melted=orc["melted"])
        assert (frame / "Grains00000.vtr").read_text() == want, c
        assert (frame / f"Grains{c:05d}.pvtr").read_text() == \
            vs.pvtr_text(g["nx"], g["ny"], g["nz"], ids, "Grains")
    assert (sub / "Results" / "Grains.pvd").read_text() == vs.pvd_text(times, "Grains")
