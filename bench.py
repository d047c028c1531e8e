"""Benchmark of the plain CA step on one MI355X.

Workload: 512^3 cells of 2e-6 m, uniform cooling -200 K/s from T0 = -0.2 K,
bulk Gaussian nucleation 5e13 m^-3 N(3, 1), surface 1e9 m^-2 N(1, 0.5) on the
z = 0 face, RappazGandin kinetics (a1 = -0.544e-4, a2 = 2.03e-4), FP64
envelopes, maxDt 1e-4 s.  A "step" is one CA step (nucleation, dt, growth,
capture cascade, time advance) over the whole grid; value = cells x executed
steps / wall time of the timed region, with the grid resident in HBM.

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


def algorithmic_bytes(log):
    """Algorithmic HBM bytes per kernel over a step log (unit figures of the
    plain step: 48 B per active grain, 44 B per capture, 12 B per live site).

    k_claim    wave 0: 48 B per active grain (centre + L read 32, orientation id
               4, grain id 4, grown L written 8); wave 1: 40 B per wave-1 grain
               (no growth write)
    k_resolve  4 B per survivor-list entry written
    k_emit     44 B per capture: child envelope 32, orientation 4, grain id 4,
               wave-list entry 4 (captures of waves >= 2, made in k_tail, are a
               handful per step and are counted here too)
    k_nucleate 12 B per live site (cell index 4 + critical undercooling 8)
    """
    A = log["n_active_before"].astype(np.float64)
    W1 = log["n_wave1"].astype(np.float64)
    C = log["n_captured"].astype(np.float64)
    Aout = log["n_active_after"].astype(np.float64)
    return {
        "k_claim": 48.0 * A + 40.0 * W1,
        "k_resolve": 4.0 * Aout,
        "k_emit": 44.0 * C,
    }


def cpu_baseline(seconds_target=20.0):
    """The oracle (one core) on the same physics at 160^3, to full solidification
    (about 10 s of CPU work)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    n = 160
    p = oracle_lib.plain_params(nx=n, ny=n, nz=n, h=2e-6, ns=1e9, dTs_mean=1.0, dTs_sd=0.5,
                                surface_mode=1, nv=5e13, dTv_mean=3.0, dTv_sd=1.0, Tdot=-200.0,
                                max_dt=1e-4, final_time=1.0)
    t0 = time.perf_counter()
    r = oracle_lib.run_plain(p, want_grains=False)
    wall = time.perf_counter() - t0
    return {"value": n ** 3 * r["n_steps"] / wall, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"CPU oracle (sequential restatement, g++ -O2) on the same physics at "
                      f"{n}^3, {r['n_steps']} steps to full solidification, {wall:.1f} s"}


def copy_probe(mib=2048, reps=20):
    """Measured stream-copy rate of the device (k_copy_probe, read + write)."""
    import ctypes as C
    from amca3d_amd import load
    g, ms = C.c_double(0.0), C.c_double(0.0)
    st = load().amca_copy_probe(mib << 20, reps, C.byref(g), C.byref(ms))
    if st != 0:
        return None
    return {"gbps": g.value, "ms_per_copy": ms.value, "bytes": mib << 20}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_plain_uniform512.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=450)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP events")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the stream-copy probe (for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs the CA step on "
              "one GPU only (a single device-resident grid)", file=sys.stderr)
        sys.exit(2)

    import amca3d_amd
    from amca3d_amd import decks

    if args.copy_probe:
        print(json.dumps({"copy_probe": copy_probe()}))
        return

    n = args.size
    deck = decks.uniform_cooling_512_deck(n=n)
    if args.warmup > 0:
        with amca3d_amd.Simulation(deck) as w:
            w.initialize()
            w.domain.run_fixed(args.warmup)
    sim = amca3d_amd.Simulation(deck).initialize()
    dom = sim.domain
    dom.enable_log(args.steps + 8)
    if not args.no_timing:
        dom.enable_timing(True)
    dom.counters()  # drains the stream
    t0 = time.perf_counter()
    dom.run_fixed(args.steps)
    wall = time.perf_counter() - t0
    c = dom.counters()
    executed = int(c["step"])
    cells = n ** 3
    value = cells * executed / wall
    log = dom.read_log()
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(executed, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (host-drawn nucleation sites and orientations, seed 0)",
        "config": {"workload": f"plain CA, uniform cooling, {n}^3 cells of 2e-6 m, -200 K/s, "
                               "bulk 5e13 m^-3 N(3,1), surface 1e9 m^-2 N(1,0.5) on z=0",
                   "cells": cells, "executed_steps": executed, "sim_time_s": c["time"],
                   "liquid_left": int(c["n_liquid"]), "grains": int(c["n_grains"]),
                   "max_cascade_wave": int(c["max_wave"])},
    }
    if not args.no_timing:
        kt = dom.kernel_times()
        total = {k: v[0] for k, v in kt.items()}
        dom_k = max(total, key=total.get)
        ab = algorithmic_bytes(log)
        roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": dom_k,
                "kernel_ms_total": total[dom_k], "launches": kt[dom_k][1],
                "avg_launch_us": 1e3 * total[dom_k] / max(kt[dom_k][1], 1),
                "kernel_share_of_wall": total[dom_k] / (1e3 * wall)}
        if dom_k in ab:
            bytes_total = float(ab[dom_k].sum())
            achieved = bytes_total / (total[dom_k] * 1e-3) / 1e9
            roof.update(achieved=achieved, frac=achieved / HBM_PEAK_GBS,
                        bytes_per_launch=bytes_total / max(kt[dom_k][1], 1))
        else:
            roof.update(achieved=None, frac=None)
        roof["traffic"] = measured_traffic(dom_k)
        cp = copy_probe()
        if cp:
            roof["measured_copy_gbs"] = cp["gbps"]
        result["roofline"] = roof
        result["kernel_ms"] = {k: round(v, 3) for k, v in total.items()}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    sim.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
