"""Summarise rocprofv3 output into the short per-kernel tables kept under profiles/.

    python -m amca3d_amd.prof_summary --stats KT/bench_kernel_stats.csv \
        --fetch PMC_F/bench_counter_collection.csv --write PMC_W/bench_counter_collection.csv \
        --copy-fetch CF/copy_counter_collection.csv --copy-write CW/copy_counter_collection.csv \
        --copy-bytes 2147483648 --command "..." --json OUT.json --md OUT.md

FETCH_SIZE / WRITE_SIZE are KiB per dispatch as the counters report them; they
are scaled by the factor that maps the copy probe's counters onto its known
traffic (k_copy_probe reads and writes `--copy-bytes` bytes per dispatch).
"""
import argparse
import csv
import json
import re
from collections import defaultdict


def short_name(name):
    """`amca::k_claim(amca::DevParams, ...)` -> `k_claim`."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0]
    return re.sub(r".*::", "", base)


def kernel_stats(path):
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            out[short_name(r["Name"])] = {"calls": int(r["Calls"]),
                                          "total_ms": float(r["TotalDurationNs"]) * 1e-6,
                                          "avg_us": float(r["AverageNs"]) * 1e-3}
    return out


def counter_means(path, counter):
    acc = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--workload", default="")
    ap.add_argument("--json")
    ap.add_argument("--md")
    a = ap.parse_args()

    st = kernel_stats(a.stats)
    res = {"workload": a.workload, "command": a.command, "kernels": {}}
    fscale = wscale = None
    if a.copy_fetch and a.copy_write and a.copy_bytes > 0:
        cf = counter_means(a.copy_fetch, "FETCH_SIZE")["k_copy_probe"][0]
        cw = counter_means(a.copy_write, "WRITE_SIZE")["k_copy_probe"][0]
        kib = a.copy_bytes / 1024.0
        fscale, wscale = kib / cf, kib / cw
        res["copy_calibration"] = {"bytes": a.copy_bytes, "fetch_kib": cf, "write_kib": cw,
                                   "fetch_scale": fscale, "write_scale": wscale}
    fm = counter_means(a.fetch, "FETCH_SIZE") if a.fetch else {}
    wm = counter_means(a.write, "WRITE_SIZE") if a.write else {}
    for k, s in sorted(st.items(), key=lambda kv: -kv[1]["total_ms"]):
        row = dict(s)
        if k in fm and k in wm and fscale:
            row["fetch_kib"] = fm[k][0]
            row["write_kib"] = wm[k][0]
            row["pmc_dispatches"] = fm[k][1]
            row["bytes_per_launch"] = 1024.0 * (fm[k][0] * fscale + wm[k][0] * wscale)
        res["kernels"][k] = row
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    lines = [f"# {a.workload}", "", "Command:", "", "```", a.command, "```", ""]
    if fscale:
        lines += [f"Copy calibration: k_copy_probe moves {a.copy_bytes:.0f} B each way per "
                  f"dispatch; FETCH_SIZE reads {res['copy_calibration']['fetch_kib']:.1f} KiB, "
                  f"WRITE_SIZE {res['copy_calibration']['write_kib']:.1f} KiB -> scale "
                  f"{fscale:.4f} (fetch), {wscale:.4f} (write).", ""]
    lines += ["| kernel | calls | total ms | avg us | HBM bytes/launch (scaled PMC) |",
              "|---|---|---|---|---|"]
    for k, r in res["kernels"].items():
        b = r.get("bytes_per_launch")
        lines.append(f"| {k} | {r['calls']} | {r['total_ms']:.3f} | {r['avg_us']:.2f} | "
                     f"{'' if b is None else f'{b:.4g}'} |")
    text = "\n".join(lines) + "\n"
    if a.md:
        with open(a.md, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
