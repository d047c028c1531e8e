"""Batches of plain domains on the GPU: every domain of a batch ends bit-equal
to its own single-domain run, whatever the schedule (block split, block order,
steps per graph replay) and after the autotuner has switched schedules
mid-run."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SPECS = [dict(edge=e, seed=s, Tdot=t, density=r) for e, s, t, r in [
    (12, 3, -25.0, 8e15), (40, 0, -1800.0, 1e15), (20, 7, -300.0, 4e15),
    (33, 11, -60.0, 2e15), (16, 5, -1000.0, 6e15), (27, 2, -150.0, 1.5e15)]]


def _single(spec):
    from amca3d_amd import Simulation
    from amca3d_amd.batch import spec_deck
    with Simulation(spec_deck(spec, termination_time=1.0)) as s:
        s.initialize()
        s.domain.enable_log(100000)
        c = s.run()
        d = s.domain
        o = d.download()
        return dict(c=c, ori=o["ori"], gid=o["gid"], env=o["env"], log=d.read_log())


def _batched(specs, schedule=None, autotune=False):
    from amca3d_amd.batch import Batch, simulations
    sims = simulations(specs, termination_time=1.0)
    for s in sims:
        s.domain.enable_log(100000)
    b = Batch([s.domain for s in sims])
    if schedule:
        b.set_schedule(**schedule)
    if autotune:
        b.autotune(trial_graphs=1, rounds=2)
    r = b.run()
    b.close()
    out = []
    for s in sims:
        d = s.domain
        o = d.download()
        out.append(dict(c=d.counters(), ori=o["ori"], gid=o["gid"], env=o["env"],
                        log=d.read_log()))
        s.close()
    return r, out


@pytest.fixture(scope="module")
def singles():
    return [_single(s) for s in SPECS]


def _same(a, b):
    for k in ("step", "n_grains", "n_orient", "n_active", "done"):
        assert a["c"][k] == b["c"][k], k
    assert np.float64(a["c"]["time"]).view(np.uint64) == np.float64(b["c"]["time"]).view(np.uint64)
    np.testing.assert_array_equal(a["ori"], b["ori"])
    np.testing.assert_array_equal(a["gid"], b["gid"])
    np.testing.assert_array_equal(a["env"].view(np.uint64), b["env"].view(np.uint64))
    for f in ("dt", "n_captured", "n_active_after", "n_wave1", "max_wave"):
        np.testing.assert_array_equal(a["log"][f], b["log"][f], err_msg=f)


@pytest.mark.parametrize("schedule", [
    None,
    dict(list_blocks=6, emit_blocks=6, min_blocks=1, interleave=1, steps_per_graph=3),
    dict(list_blocks=4096, emit_blocks=4096, min_blocks=8, interleave=0, steps_per_graph=64),
])
def test_batch_equals_single_runs(amca, singles, schedule):
    r, out = _batched(SPECS, schedule)
    assert r["finished"] == len(SPECS)
    for s, b in zip(singles, out):
        assert s["c"]["done"] == 1
        _same(s, b)


def test_batch_without_graphs(amca, singles, monkeypatch):
    """Replays launched kernel by kernel (AMCA_BATCH_GRAPHS=0) give the same
    runs, maps rebuilt as domains finish included."""
    monkeypatch.setenv("AMCA_BATCH_GRAPHS", "0")
    r, out = _batched(SPECS, dict(list_blocks=64, emit_blocks=64, min_blocks=1, interleave=0,
                                  steps_per_graph=2))
    assert r["finished"] == len(SPECS)
    for s, b in zip(singles, out):
        _same(s, b)


def test_batch_schedule_is_checked(amca):
    from amca3d_amd.batch import Batch, simulations
    sims = simulations(SPECS[:1], termination_time=1.0)
    b = Batch([s.domain for s in sims])
    for bad in (dict(min_blocks=0), dict(min_blocks=4097), dict(list_blocks=0),
                dict(steps_per_graph=0), dict(steps_per_graph=4097)):
        with pytest.raises(Exception, match="bad schedule"):
            b.set_schedule(**bad)
    b.set_schedule(min_blocks=4096, steps_per_graph=2)
    assert b.schedule()["min_blocks"] == 4096
    b.close()
    for s in sims:
        s.close()


def test_batch_graph_setting_is_checked(amca, monkeypatch):
    from amca3d_amd.batch import Batch, simulations
    sims = simulations(SPECS[:1], termination_time=1.0)
    monkeypatch.setenv("AMCA_BATCH_GRAPHS", "2")
    with pytest.raises(Exception, match="AMCA_BATCH_GRAPHS"):
        Batch([s.domain for s in sims])
    for s in sims:
        s.close()


def test_batch_after_autotune(amca, singles):
    r, out = _batched(SPECS, autotune=True)
    assert r["finished"] == len(SPECS)
    for s, b in zip(singles, out):
        _same(s, b)


def test_batch_member_matches_oracle(amca, oracle, singles):
    spec = SPECS[0]
    e = spec["edge"]
    prm = ol.plain_params(nx=e, ny=e, nz=e, h=2e-6, nv=spec["density"], dTv_mean=3.0,
                          dTv_sd=1.0, seed=spec["seed"], Tdot=spec["Tdot"], max_dt=1e-3,
                          final_time=1.0)
    r = ol.run_plain(prm, want_grains=False)
    np.testing.assert_array_equal(singles[0]["ori"], r["ori"])
    assert singles[0]["c"]["step"] == r["n_steps"]


EXTRA = [dict(edge=e, seed=s, Tdot=t, density=r) for e, s, t, r in [
    (18, 13, -40.0, 5e15), (24, 17, -700.0, 3e15), (14, 19, -90.0, 7e15),
    (30, 23, -2500.0, 1e15)]]


@pytest.fixture(scope="module")
def singles10(singles):
    return singles + [_single(s) for s in EXTRA]


def test_batch_rebuilds_maps_as_domains_finish(amca, singles10):
    """Ten domains finishing at different steps, one step per graph replay:
    the block maps are rebuilt over the running domains each time four more
    have finished, and every domain still ends bit-equal to its own run."""
    r, out = _batched(SPECS + EXTRA, dict(list_blocks=64, emit_blocks=64, min_blocks=1,
                                          interleave=1, steps_per_graph=1))
    assert r["finished"] == len(singles10)
    for s, b in zip(singles10, out):
        assert s["c"]["done"] == 1
        _same(s, b)


def test_batch_run_in_chunks(amca, singles10):
    """The same ten domains run one replay per call (maps rebuilt between and
    within calls) end as their own runs do."""
    from amca3d_amd.batch import Batch, simulations
    sims = simulations(SPECS + EXTRA, termination_time=1.0)
    for s in sims:
        s.domain.enable_log(100000)
    b = Batch([s.domain for s in sims])
    b.set_schedule(list_blocks=32, emit_blocks=32, min_blocks=1, interleave=0, steps_per_graph=3)
    calls, fin = 0, 0
    while fin < len(sims):
        r = b.run(max_graphs=1)
        assert r["replays"] == 1
        fin = r["finished"]
        calls += 1
        assert calls < 10000
    b.close()
    for s, single in zip(sims, singles10):
        d = s.domain
        o = d.download()
        _same(single, dict(c=d.counters(), ori=o["ori"], gid=o["gid"], env=o["env"],
                           log=d.read_log()))
        s.close()
