"""Benchmark of the CA step on one MI355X.

Default workload: 512^3 cells of 2e-6 m, uniform cooling -200 K/s from
T0 = -0.2 K, bulk Gaussian nucleation 5e13 m^-3 N(3, 1), surface 1e9 m^-2
N(1, 0.5) on the z = 0 face, RappazGandin kinetics (a1 = -0.544e-4,
a2 = 2.03e-4), FP64 envelopes, maxDt 1e-4 s.  A "step" is one CA step
(nucleation, dt, growth, capture cascade, time advance) over the whole grid.
The timed window is --steps steps centred on the run's largest active front,
after --warmup untimed steps; value = cells x executed steps / wall time of
the window, with the grid resident in HBM.  The full run's figure is reported
beside it.  Other workloads: --workload pbf | batch | surrogate.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "CA cell-updates/sec + % HBM peak; batched+surrogate vs plain GPU at fixed grain-ID error"


def algorithmic_bytes(log, cells):
    """Algorithmic HBM bytes per kernel over a step log, from the unit figures
    of the plain step (48 B per active grain, 44 B per capture, 12 B per live
    site).  k_claim, k_resolve, k_idscan and k_emit run once for wave 0 and
    once for wave 1; the figures are per step, both launches together.

    k_nucleate 12 B per live site (cell index 4 + critical undercooling 8)
    k_count    8 B per 4x4x4 brick (its active word)
    k_compact  8 B per 4x4x4 brick (its active word) and 4 B per active grain
               (its list entry written)
    k_claim    wave 0: 48 B per active grain (centre 24, L 8, grain id 4 and
               orientation 4 read, grown L written 8); wave 1: 40 B per
               wave-1 grain (no growth write)
    k_resolve  1 bit per entry of waves 0 and 1 (its active bit)
    k_idscan   1 B per grain id of the scanned ranges (the capture counts):
               for wave 0 the id range of the grains that claimed a cell, which
               the step log does not hold, so every grain before wave 0's
               children is counted (an upper bound); then wave 0's children
    k_emit     44 B per capture: child centre 24, L 8, grain id 4, orientation
               4, next-wave list entry 4 (captures of waves >= 2, made in
               k_tail, are a handful per step and are counted here too)
    k_tail     120 B per step (the step's log row 56 and counter block 64)
    """
    A = log["n_active_before"].astype(np.float64)
    W1 = log["n_wave1"].astype(np.float64)
    C = log["n_captured"].astype(np.float64)
    S = log["n_sites"].astype(np.float64)
    # grains after nucleation: one per solid cell, N minus the liquid cells
    # before the step's captures
    grains = cells - (log["n_liquid"].astype(np.float64) + C)
    return {
        "k_nucleate": 12.0 * S,
        "k_count": 8.0 * cells / 64.0 * np.ones_like(A),
        "k_compact": 8.0 * cells / 64.0 + 4.0 * A,
        "k_claim": 48.0 * A + 40.0 * W1,
        "k_resolve": (A + W1) / 8.0,
        "k_idscan": grains + W1,
        "k_emit": 44.0 * C,
        "k_tail": 120.0 * np.ones_like(A),
    }


def step_formula_bytes(log, cells):
    """Whole-step algorithmic bytes of the plain step, 4 N + 48 A + 44 C + 12 S
    (N cells, A grains active at the step's start, C captures, S live sites):
    the grain-ID field scan that finds the interface, the active grains'
    records, the captures' records and the sites."""
    return (4.0 * cells + 48.0 * log["n_active_before"].astype(np.float64) +
            44.0 * log["n_captured"].astype(np.float64) +
            12.0 * log["n_sites"].astype(np.float64))


def _oracle_replica(job):
    """One oracle run of the 512^3 workload's physics at edge n, seed `seed`
    (a process of its own: cell-updates, wall seconds, steps)."""
    n, seed = job
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    p = oracle_lib.plain_params(nx=n, ny=n, nz=n, h=2e-6, ns=1e9, dTs_mean=1.0, dTs_sd=0.5,
                                surface_mode=1, nv=5e13, dTv_mean=3.0, dTv_sd=1.0, Tdot=-200.0,
                                max_dt=1e-4, final_time=1.0, seed=seed)
    t0 = time.perf_counter()
    r = oracle_lib.run_plain(p, want_grains=False)
    return n ** 3 * r["n_steps"], time.perf_counter() - t0, r["n_steps"]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n=160, cores=None):
    """The CPU path as a throughput reference: one oracle replica per host core
    (seeds 0..cores-1, the 512^3 workload's physics at n^3, each to full
    solidification, about 15-25 s), run in child processes started before this
    process touches the GPU.  value = all replicas' cell-updates / the wall
    time of the slowest; the one-core figure is replica 0's rate."""
    import multiprocessing as mp
    cores = cores or max(1, min(16, os.cpu_count() or 1))
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        runs = pool.map(_oracle_replica, [(n, seed) for seed in range(cores)])
    wall = time.perf_counter() - t0
    work = sum(r[0] for r in runs)
    return {"value": work / wall, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "one_core_value": runs[0][0] / runs[0][1],
            "sample": f"CPU oracle (sequential restatement, g++ -O2), {cores} replicas in "
                      f"parallel, one per core, seeds 0..{cores - 1}: the same physics at "
                      f"{n}^3 to full solidification ({min(r[2] for r in runs)}-"
                      f"{max(r[2] for r in runs)} steps), {wall:.1f} s wall"}


def copy_probe(mib=2048, reps=20):
    """Measured stream-copy rate of the device (k_copy_probe, read + write)."""
    import ctypes as C
    from amca3d_amd import load
    g, ms = C.c_double(0.0), C.c_double(0.0)
    st = load().amca_copy_probe(mib << 20, reps, C.byref(g), C.byref(ms))
    if st != 0:
        return None
    return {"gbps": g.value, "ms_per_copy": ms.value, "bytes": mib << 20}


def measured_traffic(kernel, summary="pmc_traffic_plain_uniform512.json"):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def remelt_algorithmic_bytes(P, log):
    """Algorithmic bytes per kernel of the coupled step over a step log; X is
    the step's active box in cells (the whole grid without it).

    k_rm_field          24 B per FE node (two frames read, nodal field written;
                        the hot-node test uses the value in registers)
    k_rm_phase          46 B per box cell (element tag 4, parametric coordinates
                        24, T read 8 and written 8, cooling and liquid-neighbour
                        flags 1 + 1) and 8 B per FE node (the nodal field)
    k_rm_liquid         13 B per box cell (T 8, orientation 4, melted flag 1)
    k_rm_nucleate_mint  1 B per box cell (flag) + 20 B per flagged cell (grain
                        id 4, length 8, T 8)
    k_rm_claim          1 B per box cell + 56 B per flagged cell (grain id 4,
                        envelope read 32 and length written 8, T 8, orientation 4)
    k_rm_emit           1 B per box cell + 44 B per capture
    """
    nodes = float(P["nodes"])
    X = log["n_box"].astype(np.float64)
    F = log["n_flagged"].astype(np.float64)
    C = log["n_captured"].astype(np.float64)
    one = np.ones_like(F)
    return {
        "k_rm_field": 24.0 * nodes * one,
        "k_rm_phase": 46.0 * X + 8.0 * nodes,
        "k_rm_liquid": 13.0 * X,
        "k_rm_nucleate_mint": 1.0 * X + 20.0 * F,
        "k_rm_claim": 1.0 * X + 56.0 * F,
        "k_rm_emit": 1.0 * X + 44.0 * C,
    }


def pbf_cpu_baseline(steps=300):
    """The oracle (one core) on a 200 x 80 x 40 sample of the same track physics."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from amca3d_amd import pbf
    w = pbf.track(nx=200, ny=80, nz=40, n_frames=40)
    d = w["desc"]
    case = dict(grid=w["grid"], fe=w["fe"], bin=d["bin"], tol=d["tolerance"],
                max_iter=d["max_iterations"], void_temp=d["void_temperature"],
                physics={k: d[k] for k in ("Tm", "Ts", "Tmin_nuc", "a1", "a2", "a3", "b1", "b2",
                                           "b3")},
                trunc=d["trunc"], tsf=d["time_step_factor"], cutoff_factor=d["cutoff_factor"],
                final_time=d["final_time"], max_steps=steps, initial_melted=0,
                nucleation=dict(surface=w["surface"], bulk=w["bulk"]), seed=w["nuc_seed"],
                state=w["state"])
    t0 = time.perf_counter()
    r = oracle_lib.run_remelt(case)
    wall = time.perf_counter() - t0
    n = 200 * 80 * 40
    return {"value": n * r["n_steps"] / wall, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (sequential restatement, g++ -O2) on the same track physics "
                      f"over 200x80x40 cells, {r['n_steps']} steps, {wall:.1f} s"}


def pbf_main(args):
    """Coupled CA-FE single track: 1000 x 400 x 200 cells of 2.5e-6 m.

    The timed region runs the steps as the library does by default (HIP graphs
    of eight steps, no per-kernel events).  A second domain built from the same
    inputs then runs the same steps with HIP events around every kernel, for
    the per-kernel times and the roofline; its grain-ID field must equal the
    first domain's."""
    from amca3d_amd import pbf
    t_setup = time.perf_counter()
    w = pbf.track(n_frames=args.frames)
    setup_s = time.perf_counter() - t_setup
    g = w["grid"]
    cells = g["nx"] * g["ny"] * g["nz"]

    def run(timing):
        d = pbf.domain(w)
        if args.whole_grid:
            d.set_window(False)
        d.set_graph_steps(args.graph_steps)
        d.enable_log(args.warmup + args.steps + 8)
        if args.warmup > 0:
            d.run(args.warmup)
        if timing:
            d.enable_timing(True)
        c0 = d.counters()
        t0 = time.perf_counter()
        c = d.run(args.steps)
        wall = time.perf_counter() - t0
        log = d.read_log()[int(c0["step"]):int(c["step"])]
        kt = d.kernel_times() if timing else None
        ori = d.download()["ori"]
        d.close()
        return c0, c, wall, log, kt, ori

    c0, c, wall, log, _, ori = run(False)
    executed = int(c["step"] - c0["step"])
    value = cells * executed / wall
    result = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(executed, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Rosenthal frames and Voronoi substrate from the library's host "
                "generators, sites from its rand() stream, seed 0)",
        "config": {"workload": f"coupled CA-FE single track, {g['nx']}x{g['ny']}x{g['nz']} cells "
                               "of 2.5e-6 m, FE nodes every 1e-5 m, Rosenthal 200 W 1 m/s, "
                               "2000 Voronoi grains, bulk 1e13 m^-3 N(5,1)",
                   "cells": cells, "fe_nodes": int(np.prod(w["fe"]["shape"])),
                   "frames": args.frames, "executed_steps": executed,
                   "sim_time_s": c["time"], "captures": int(log["n_captured"].sum()),
                   "nucleated": int(log["n_nucleated"].sum()), "setup_s": setup_s,
                   "value_counts": "every grid cell per step (the step's result equals "
                                   "the whole-grid step's)",
                   "active_box": "off (every step over the whole grid)" if args.whole_grid else
                   {"mean_cells": float(log["n_box"].mean()),
                    "mean_fraction_of_grid": float(log["n_box"].mean()) /
                    ((g["nx"] + 2) * (g["ny"] + 2) * (g["nz"] + 2))}},
    }
    if not args.no_timing:
        _, c2, wall2, log2, kt, ori2 = run(True)
        total = {k: v[0] for k, v in kt.items()}
        dom_k = max(total, key=total.get)
        Pn = dict(nx=g["nx"], ny=g["ny"], nz=g["nz"], nodes=int(np.prod(w["fe"]["shape"])))
        ab = remelt_algorithmic_bytes(Pn, log2)
        roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": dom_k,
                "kernel_ms_total": total[dom_k], "launches": kt[dom_k][1],
                "avg_launch_us": 1e3 * total[dom_k] / max(kt[dom_k][1], 1),
                "events_run_wall_ms": 1e3 * wall2,
                "timed_run_equals_events_run": bool(np.array_equal(ori, ori2)) and
                int(c2["step"]) == int(c["step"])}
        if dom_k in ab and kt[dom_k][1] == len(log2):
            bytes_total = float(ab[dom_k].sum())
            achieved = bytes_total / (total[dom_k] * 1e-3) / 1e9
            roof.update(achieved=achieved, frac=achieved / HBM_PEAK_GBS,
                        bytes_per_launch=bytes_total / max(kt[dom_k][1], 1))
        else:
            roof.update(achieved=None, frac=None)
        roof["traffic"] = measured_traffic(
            dom_k, "pmc_traffic_pbf_track_whole_grid.json" if args.whole_grid
            else "pmc_traffic_pbf_track.json")
        result["roofline"] = roof
        result["kernel_ms"] = {k: round(v, 3) for k, v in total.items()}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = pbf_cpu_baseline()
    return result


def plain_step_bytes(log):
    """Algorithmic bytes of whole plain steps over a step log: 48 B per active
    grain, 40 B per wave-1 grain, 4 B per survivor, 44 B per capture."""
    return float((48.0 * log["n_active_before"] + 40.0 * log["n_wave1"] +
                  4.0 * log["n_active_after"] + 44.0 * log["n_captured"]).sum())


def batch_main(args):
    """64 plain domains of uneven size run as one batch, against the same
    domains run one after another with the single-domain path."""
    from amca3d_amd.batch import Batch, batch_specs, simulations
    specs = batch_specs(args.domains)
    cells = np.array([s["edge"] ** 3 for s in specs], np.float64)
    # sequential single-domain runs (the plain GPU form)
    seq_wall, seq_steps = 0.0, []
    seq_ori = []
    for sim in simulations(specs):
        d = sim.domain
        d.counters()
        t0 = time.perf_counter()
        c = sim.run()
        seq_wall += time.perf_counter() - t0
        seq_steps.append(c["step"])
        seq_ori.append(d.download(env=False)["ori"])
        sim.close()
    seq_value = float((cells * np.array(seq_steps)).sum()) / seq_wall
    # the batch
    sims = simulations(specs)
    for s in sims:
        s.domain.enable_log(1 << 16)
    b = Batch([s.domain for s in sims])
    for s in sims:
        s.domain.counters()
    t0 = time.perf_counter()
    tuned = b.autotune(trial_graphs=1, rounds=2) if not args.no_autotune else b.schedule()
    r = b.run()
    wall = time.perf_counter() - t0
    b.close()
    steps, same, algo = [], True, 0.0
    for i, s in enumerate(sims):
        d = s.domain
        c = d.counters()
        steps.append(c["step"])
        same = same and c["step"] == seq_steps[i] and np.array_equal(d.download(env=False)["ori"],
                                                                      seq_ori[i])
        algo += plain_step_bytes(d.read_log())
        s.close()
    work = float((cells * np.array(steps)).sum())
    value = work / wall
    result = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": int(sum(steps)), "warmup": 0, "ms_per_step": 1e3 * wall / max(max(steps), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (host-drawn sites and orientations per domain, seeds 0..63)",
        "config": {"workload": f"{args.domains} plain domains, cubes of 64..256 cells of 2e-6 m, "
                               "Tdot log-uniform -20..-2000 K/s, bulk log-uniform 1e12..1e14 "
                               "m^-3 N(3,1), master seed 20261202, one hipGraph-replayed batch",
                   "domains": args.domains, "cells": int(cells.sum()),
                   "steps_per_domain_max": int(max(steps)),
                   "schedule": tuned, "replays": r["replays"],
                   "sequential_single_domain_value": seq_value,
                   "batched_over_sequential": value / seq_value,
                   "bit_equal_to_single_runs": bool(same)},
        "roofline": {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": "whole step",
                     "achieved": algo / wall / 1e9, "frac": algo / wall / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
    }
    return result


MLP_FLOPS_PER_ROW = 2 * (5 * 128 + 3 * 128 * 128 + 128)  # 99,840
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_main(args):
    """Batched + surrogate-driven form of the coupled track: the nine scans
    (the track and eight variants) run side by side with temperatures from the
    trained network and the FE fallback, against the same nine scans FE-driven
    (grain-ID mismatch, temperature error of the served cells) and against the
    plain GPU form (the track alone, FE-driven, over the whole grid and with the
    active box)."""
    from amca3d_amd import pbf
    w = pbf.track(n_frames=args.frames)
    g = w["grid"]
    cells = g["nx"] * g["ny"] * g["nz"]
    t0 = time.perf_counter()
    m, desc, losses = pbf.train_surrogate(w)
    train_s = time.perf_counter() - t0
    # surrogate error on the nine evaluation scans, near the melt pool
    rng = np.random.default_rng(11)
    held = []
    for s in pbf.held_out_scans(w):
        n = 50000
        tt = rng.uniform(0, (args.frames - 1) * 1e-5, n)
        pts = np.column_stack([np.clip(s.x_start + s.v * tt + rng.uniform(-1e-3, 1e-4, n), 0,
                                       g["nx"] * g["h"]),
                               np.clip(s.y_c + rng.uniform(-1.5e-4, 1.5e-4, n), 0, g["ny"] * g["h"]),
                               g["nz"] * g["h"] - rng.uniform(0, 1.5e-4, n), tt])
        ts, tf = m.eval(desc, s, pts)
        far = np.abs(tf - 1600.0) >= 50.0
        e = ts - np.minimum(tf, desc.T_clip)
        held.append({"P": s.P, "v": s.v, "rms_K": float(np.sqrt(np.mean(e ** 2))),
                     "rms_outside_band_K": float(np.sqrt(np.mean(e[far] ** 2)))})
    from amca3d_amd import remelt
    scans = pbf.held_out_scans(w)

    def batch(mode):
        """The nine scans run side by side (amca_remelt_run_many)."""
        doms = []
        lat_s = 0.0
        if mode == "surrogate":
            m.enable_timing(True)
        for sc in scans:
            ws = pbf.scan_track(w, sc)
            d = pbf.domain(ws)
            d.set_graph_steps(args.graph_steps)
            d.enable_log(args.steps + 8)
            if mode == "surrogate":
                t1 = time.perf_counter()
                d.use_surrogate(m, desc, sc, band=50.0, resid=2.0)
                lat_s += time.perf_counter() - t1
            d.counters()
            doms.append(d)
        fused = m.fused_times() if mode == "surrogate" else None
        c0 = remelt.run_many(doms, args.warmup) if args.warmup > 0 else None
        t1 = time.perf_counter()
        cs = remelt.run_many(doms, args.steps)
        wall = time.perf_counter() - t1
        steps = [int(c["step"]) - (int(a["step"]) if c0 else 0) for c, a in
                 zip(cs, c0 if c0 else cs)]
        out = dict(value=cells * sum(steps) / wall, wall=wall, steps=steps,
                   ori=[d.download()["ori"] for d in doms], logs=[d.read_log() for d in doms],
                   err=[d.surrogate_error() for d in doms] if mode == "surrogate" else None,
                   lattice_s=lat_s, fused=fused)
        for d in doms:
            d.close()
        return out

    def single(window):
        """The plain GPU form: the config's own scan, one domain, FE-driven."""
        d = pbf.domain(w)
        if not window:
            d.set_window(False)
        d.set_graph_steps(args.graph_steps)
        c0 = d.run(args.warmup) if args.warmup > 0 else d.counters()
        t1 = time.perf_counter()
        c = d.run(args.steps)
        wall = time.perf_counter() - t1
        d.close()
        return cells * (int(c["step"]) - int(c0["step"])) / wall

    plain_whole = single(False)
    plain_box = single(True)
    fe_b = batch("fe")
    sur_b = batch("surrogate")
    mism = [float((a != b).sum()) / a.size for a, b in zip(fe_b["ori"], sur_b["ori"])]
    per_scan = [{"P": sc.P, "v": sc.v, "y_c": sc.y_c, "grain_id_mismatch_fraction": mm,
                 "temperature_rms_served_K": e["rms_served"],
                 "temperature_rms_all_K": e["rms_all"],
                 "served_fraction": e["n_served"] / max(e["n_all"], 1)}
                for sc, mm, e in zip(scans, mism, sur_b["err"])]
    # k_mlp_fused over the lattices, timed by HIP events on the network's stream
    fused_ms, fused_launches, rows = sur_b["fused"]
    lattice = {"rows": rows, "launches": fused_launches, "kernel_ms": fused_ms,
               "avg_launch_us": 1e3 * fused_ms / max(fused_launches, 1),
               "use_surrogate_seconds": sur_b["lattice_s"],
               "tflops": rows * MLP_FLOPS_PER_ROW / max(fused_ms * 1e-3, 1e-12) / 1e12}
    value = sur_b["value"]
    result = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * sur_b["wall"] / max(max(sur_b["steps"]), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Rosenthal FE frames; surrogate trained on 2e7 FE samples of six "
                "held-in scans)",
        "config": {"workload": "batched + surrogate-driven coupled CA-FE: nine scans (the "
                               "1000x400x200 track and eight variants of power, speed and hatch "
                               "offset) side by side, MLP 5-128x4-1 SiLU on BF16 MFMA, FE "
                               "fallback within 50 K of Tm or where node residual > 2 K",
                   "value_counts": "every grid cell of every scan per step",
                   "plain_gpu_whole_grid_value": plain_whole,
                   "plain_gpu_active_box_value": plain_box,
                   "fe_batched_value": fe_b["value"],
                   "over_plain_gpu_whole_grid": value / plain_whole,
                   "over_plain_gpu_active_box": value / plain_box,
                   "surrogate_over_fe_batched": value / fe_b["value"],
                   "grain_id_mismatch_fraction_max": max(mism),
                   "temperature_rms_served_K_max": max(p["temperature_rms_served_K"]
                                                      for p in per_scan),
                   "per_scan": per_scan,
                   "train_seconds": train_s, "train_final_losses": losses,
                   "held_out_scans": held, "lattice_eval": lattice},
        "roofline": {"bound": "mfma", "kernel": "k_mlp_fused (lattice evaluation)",
                     "achieved": lattice.get("tflops"), "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s",
                     "frac": (lattice.get("tflops") or 0.0) / MFMA_BF16_PEAK_TFLOPS,
                     "traffic": None},
    }
    m.close()
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 100 for plain512, 450 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before them (default: 10 for plain512, 20 otherwise)")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sub-results", action="store_true",
                    help="plain512: skip the batched and surrogate sub-results")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP events")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the stream-copy probe (for counter calibration)")
    ap.add_argument("--workload", choices=("plain512", "pbf", "batch", "surrogate"),
                    default="plain512",
                    help="plain512: uniform cooling 512^3 (default); pbf: coupled CA-FE track "
                         "1000x400x200; batch: 64 uneven plain domains in one batch")
    ap.add_argument("--domains", type=int, default=64, help="batch: number of domains")
    ap.add_argument("--no-autotune", action="store_true", help="batch: keep the default schedule")
    ap.add_argument("--frames", type=int, default=100, help="pbf: FE frames (10 us apart)")
    ap.add_argument("--graph-steps", type=int, default=8,
                    help="pbf/surrogate: steps per HIP graph (0: stream launches)")
    ap.add_argument("--whole-grid", action="store_true",
                    help="pbf: every step over the whole grid (no active box)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.workload == "plain512" else 450
    if args.warmup is None:
        args.warmup = 10 if args.workload == "plain512" else 20
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs the CA step on "
              "one GPU only (a single device-resident grid)", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs first, in child processes, before anything here
    # touches the GPU
    cpu = None
    if args.workload == "plain512" and not args.copy_probe and not args.no_cpu_baseline:
        cpu = cpu_baseline()

    if args.copy_probe:
        print(json.dumps({"copy_probe": copy_probe()}))
        return
    if args.workload == "pbf":
        print(json.dumps(pbf_main(args)))
        return
    if args.workload == "batch":
        print(json.dumps(batch_main(args)))
        return
    if args.workload == "surrogate":
        print(json.dumps(surrogate_main(args)))
        return
    # the sub-results run before the headline, so that the headline's window
    # stays the process's last plain steps (the profile summary selects them so)
    sub = None if args.no_sub_results else sub_results(args)
    result = plain_main(args, cpu)
    if sub is not None:
        result["sub_results"] = sub
    print(json.dumps(result))


def sub_results(args):
    """The metric's second half, measured in the same run: the batched +
    surrogate-driven coupled form against the plain GPU form at fixed grain-ID
    error (--workload surrogate, 450 steps after 20), and the batched plain
    form against the same domains one after another (--workload batch)."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.no_cpu_baseline = 450, 20, True
    t0 = time.perf_counter()
    s = surrogate_main(sub)
    c = s["config"]
    sur = {"workload": c["workload"], "value": s["value"], "unit": s["unit"],
           "steps": s["steps"], "warmup": s["warmup"], "ms_per_step": s["ms_per_step"],
           "plain_gpu_whole_grid_value": c["plain_gpu_whole_grid_value"],
           "plain_gpu_active_box_value": c["plain_gpu_active_box_value"],
           "fe_batched_value": c["fe_batched_value"],
           "over_plain_gpu_whole_grid": c["over_plain_gpu_whole_grid"],
           "over_plain_gpu_active_box": c["over_plain_gpu_active_box"],
           "surrogate_over_fe_batched": c["surrogate_over_fe_batched"],
           "grain_id_mismatch_fraction_max": c["grain_id_mismatch_fraction_max"],
           "temperature_rms_served_K_max": c["temperature_rms_served_K_max"],
           "train_seconds": c["train_seconds"], "roofline": s["roofline"],
           "seconds": time.perf_counter() - t0}
    t0 = time.perf_counter()
    b = batch_main(sub)
    c = b["config"]
    bat = {"workload": c["workload"], "value": b["value"], "unit": b["unit"],
           "sequential_single_domain_value": c["sequential_single_domain_value"],
           "batched_over_sequential": c["batched_over_sequential"],
           "bit_equal_to_single_runs": c["bit_equal_to_single_runs"],
           "schedule": c["schedule"], "roofline": b["roofline"],
           "seconds": time.perf_counter() - t0}
    return {"batched_surrogate_pbf": sur, "batched_plain": bat}


def plain_main(args, cpu):
    """The headline: exactly --steps steps of the 512^3 run, timed where the
    active front is largest.

    1. The whole run once, to full solidification: its wall time gives the
       full-run figure and its step log the step with the most active grains.
    2. A second domain from the same deck is fast-forwarded, untimed, so that
       after --warmup more untimed steps the timed window of --steps steps is
       centred on that peak; then the window is timed as the library runs it
       (whole steps replayed from a captured HIP graph).  The window's
       grain-ID field equals the first run's at the same step count by
       construction of the path (the tests check stepwise == run).
    3. Unless --no-timing, a third domain runs the same window with every
       launch bracketed by HIP events on its stream, for the per-kernel
       durations.  Its kernels are the process's last 12 x --steps
       step-kernel dispatches, which is how the profile summary
       (amca3d_amd/prof_summary.py --window-steps) picks them out; with
       --no-timing the last ones are those of step 2's window."""
    import amca3d_amd
    from amca3d_amd import decks
    n = args.size
    cells = n ** 3
    deck = decks.uniform_cooling_512_deck(n=n)

    full = amca3d_amd.Simulation(deck).initialize()
    fd = full.domain
    fd.enable_log(1 << 15)
    fd.counters()
    t0 = time.perf_counter()
    cf = fd.run(-1)
    full_wall = time.perf_counter() - t0
    full_log = fd.read_log()
    full.close()
    total = int(cf["step"])
    peak = int(np.argmax(full_log["n_active_before"]))
    K, W = args.steps, args.warmup
    first = min(max(peak - K // 2, W), max(total - K, W))
    ff = first - W

    def at_window():
        sim = amca3d_amd.Simulation(deck).initialize()
        dom = sim.domain
        dom.enable_log(max(total, first + K) + 8)
        if ff > 0:
            dom.run_fixed(ff)
        if W > 0:
            dom.run_fixed(W)
        return sim, dom

    # the headline: the window as the library runs it (steps replayed from a
    # captured HIP graph, no per-kernel events)
    sim, dom = at_window()
    c0 = dom.counters()  # drains the stream
    t0 = time.perf_counter()
    dom.run_fixed(K)
    wall = time.perf_counter() - t0
    c = dom.counters()
    s0, s1 = int(c0["step"]), int(c["step"])
    executed = s1 - s0
    value = cells * executed / wall
    log = dom.read_log()[s0:s1]
    sim.close()
    step_bytes = float(step_formula_bytes(log, cells).sum())
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / max(executed, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (host-drawn nucleation sites and orientations, seed 0)",
        "config": {"workload": f"plain CA, uniform cooling, {n}^3 cells of 2e-6 m, -200 K/s, "
                               "bulk 5e13 m^-3 N(3,1), surface 1e9 m^-2 N(1,0.5) on z=0",
                   "cells": cells,
                   "window": {"first_step": s0, "last_step": s1 - 1, "executed_steps": executed,
                              "placement": "centred on the step with the most active grains "
                                           f"({peak}) of the {total}-step run",
                              "active_grains_mean": float(log["n_active_before"].mean()),
                              "captures_per_step_mean": float(log["n_captured"].mean()),
                              "liquid_left_at_end": int(c["n_liquid"]),
                              "max_cascade_wave": int(log["max_wave"].max()) if executed else 0},
                   "full_run": {"steps": total, "wall_s": full_wall,
                                "value": cells * total / full_wall,
                                "ms_per_step": 1e3 * full_wall / max(total, 1),
                                "grains": int(cf["n_grains"]),
                                "max_cascade_wave": int(cf["max_wave"])},
                   "step_formula": {"bytes_per_step": step_bytes / max(executed, 1),
                                    "achieved_gbs": step_bytes / wall / 1e9,
                                    "frac": step_bytes / wall / 1e9 / HBM_PEAK_GBS,
                                    "formula": "4 N + 48 A + 44 C + 12 S per step"}},
    }
    if not args.no_timing:
        # per-kernel durations: the same window again on a second domain
        # fast-forwarded to the same step, each launch bracketed by HIP events
        sim, dom = at_window()
        dom.enable_timing(True)
        ct0 = dom.counters()
        dom.run_fixed(K)
        ct1 = dom.counters()
        if (int(ct0["step"]), int(ct1["step"])) != (s0, s1):
            raise SystemExit("the timed replica of the window ran other steps")
        kt = dom.kernel_times()
        sim.close()
        tot = {k: v[0] for k, v in kt.items()}
        dom_k = max(tot, key=tot.get)
        ab = algorithmic_bytes(log, cells)
        per = {}
        for k, (ms, launches) in kt.items():
            if launches == 0:
                continue
            b = float(ab[k].sum()) if k in ab else None
            per[k] = {"ms": round(ms, 4), "launches": launches,
                      "avg_launch_us": round(1e3 * ms / launches, 3),
                      "algorithmic_bytes_per_launch": None if b is None else b / launches,
                      "achieved_gbs": None if b is None else b / (ms * 1e-3) / 1e9}
        b = float(ab[dom_k].sum())
        achieved = b / (tot[dom_k] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom_k),
                "kernel": dom_k, "kernel_ms_total": tot[dom_k], "launches": kt[dom_k][1],
                "avg_launch_us": 1e3 * tot[dom_k] / max(kt[dom_k][1], 1),
                "bytes_per_launch": b / max(kt[dom_k][1], 1),
                "kernel_share_of_wall": tot[dom_k] / (1e3 * wall),
                "kernels_share_of_wall": sum(tot.values()) / (1e3 * wall)}
        cp = copy_probe()
        if cp:
            roof["measured_copy_gbs"] = cp["gbps"]
        result["roofline"] = roof
        result["kernels"] = per
    if cpu is not None:
        result["cpu_baseline"] = cpu
    sim.close()
    return result


if __name__ == "__main__":
    main()
