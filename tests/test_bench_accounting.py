"""The measurement's own arithmetic, on the CPU: bench.py's algorithmic-byte
model and whole-step formula on a hand-made step log, its refusal of more
than one GPU, and prof_summary's reading of rocprofv3 CSVs (kernel names,
per-kernel means, the 12-dispatch step window by wave)."""
import csv
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from amca3d_amd import prof_summary as ps  # noqa: E402


def _log():
    # two steps: active grains, wave-1 grains, captures, live sites, liquid cells
    return {"n_active_before": np.array([1000, 2000]), "n_wave1": np.array([100, 300]),
            "n_captured": np.array([150, 400]), "n_sites": np.array([10, 0]),
            "n_liquid": np.array([5000, 4600])}


def test_algorithmic_bytes_unit_figures():
    cells = 64 ** 3
    b = bench.algorithmic_bytes(_log(), cells)
    np.testing.assert_array_equal(b["k_claim"], [48 * 1000 + 40 * 100, 48 * 2000 + 40 * 300])
    np.testing.assert_array_equal(b["k_emit"], [44 * 150, 44 * 400])
    np.testing.assert_array_equal(b["k_nucleate"], [120, 0])
    np.testing.assert_array_equal(b["k_resolve"], [1100 / 8, 2300 / 8])
    np.testing.assert_array_equal(b["k_compact"], [8 * cells / 64 + 4000, 8 * cells / 64 + 8000])
    np.testing.assert_array_equal(b["k_count"], [8 * cells / 64] * 2)
    # grains before the step's captures: cells - (liquid + captures), then wave 1's ids
    np.testing.assert_array_equal(b["k_idscan"], [cells - 5150 + 100, cells - 5000 + 300])
    np.testing.assert_array_equal(b["k_tail"], [120, 120])


def test_step_formula():
    cells = 64 ** 3
    f = bench.step_formula_bytes(_log(), cells)
    np.testing.assert_array_equal(f, [4 * cells + 48 * 1000 + 44 * 150 + 12 * 10,
                                      4 * cells + 48 * 2000 + 44 * 400 + 0])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_short_names():
    assert ps.short_name("amca::k_claim(amca::DevParams, amca::DevBuffers, int)") == "k_claim"
    assert ps.short_name("(anonymous namespace)::k_copy_probe(uint4 const*, uint4*, long long)") \
        == "k_copy_probe"
    assert ps.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_stats_and_counter_means(tmp_path):
    st = tmp_path / "stats.csv"
    _write(st, ["Name", "Calls", "TotalDurationNs", "AverageNs"],
           [["amca::k_claim(amca::DevParams, amca::DevBuffers, int)", 4, 2000000, 500000]])
    s = ps.kernel_stats(str(st))
    assert s == {"k_claim": {"calls": 4, "total_ms": 2.0, "avg_us": 500.0}}
    cc = tmp_path / "cc.csv"
    _write(cc, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           [[1, "amca::k_claim(x)", "WRITE_SIZE", 10.0], [2, "amca::k_claim(x)", "WRITE_SIZE", 30.0],
            [3, "amca::k_emit(x)", "WRITE_SIZE", 7.0], [3, "amca::k_emit(x)", "FETCH_SIZE", 99.0]])
    m = ps.counter_means(str(cc), "WRITE_SIZE")
    assert m == {"k_claim": (20.0, 2), "k_emit": (7.0, 1)}


def test_window_takes_the_last_steps_by_wave():
    rows = [(0, "k_init_grid", 1.0)]
    d = 1
    for step in range(3):
        for q, k in enumerate(ps.STEP_KERNELS):
            rows.append((d, k, 100.0 * step + q))
            d += 1
    w = ps.window_rows(rows, 2)
    assert w["k_claim.w0"] == [103.0, 203.0] and w["k_claim.w1"] == [107.0, 207.0]
    assert w["k_tail"] == [111.0, 211.0]
    with pytest.raises(SystemExit):
        ps.window_rows(rows, 4)  # not enough dispatches
    bad = rows[:-1] + [(d, "k_emit", 0.0)]
    with pytest.raises(SystemExit):
        ps.window_rows(bad, 2)  # a dispatch out of the step's order


def test_batch_workload_specs():
    """The 64-domain batch workload the bench and its GPU test run: 336.8 M cells in cubes
    of 64..256 cells, cooling rates and bulk densities in their stated log-uniform ranges,
    seeds 0..63, the same list on every call."""
    from amca3d_amd.batch import batch_specs, spec_deck
    a, b = batch_specs(), batch_specs()
    assert a == b and len(a) == 64
    assert [s["seed"] for s in a] == list(range(64))
    assert all(64 <= s["edge"] <= 256 for s in a)
    assert all(-2000.0 <= s["Tdot"] <= -20.0 for s in a)
    assert all(1e12 <= s["density"] <= 1e14 for s in a)
    assert sum(s["edge"] ** 3 for s in a) == 336796673
    s5 = a[5]
    deck = spec_deck(s5)
    assert f"random_seed: {s5['seed']}\n" in deck
    assert f"number_cells: [{s5['edge']}, {s5['edge']}, {s5['edge']}]" in deck
    assert f"t_dot: {s5['Tdot']!r}" in deck
