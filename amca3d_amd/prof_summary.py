"""Summarise rocprofv3 output into the short per-kernel tables kept under profiles/.

Whole process:

    python -m amca3d_amd.prof_summary --stats KT/bench_kernel_stats.csv \
        --fetch PMC_F/bench_counter_collection.csv --write PMC_W/bench_counter_collection.csv \
        --copy-fetch CF/copy_counter_collection.csv --copy-write CW/copy_counter_collection.csv \
        --copy-bytes 2147483648 --command "..." --json OUT.json --md OUT.md

FETCH_SIZE / WRITE_SIZE are KiB per dispatch as the counters report them; they
are scaled by the factor that maps the copy probe's counters onto its known
traffic (k_copy_probe reads and writes `--copy-bytes` bytes per dispatch).
FETCH_SIZE prices every read request that is not 32 bytes at 64 bytes, while
gfx950 also issues 128-byte reads (streaming reads are all of that size), so
one scale cannot fit kernels whose request sizes differ.  `--rdreq` (a pass
with TCC_EA0_RDREQ_32B_sum, TCC_EA0_RDREQ_64B_sum, TCC_EA0_RDREQ_128B_sum and
TCC_EA0_RDREQ_sum) gives the read bytes from the request counts themselves,
32 a + 64 b + 128 c; when it is given, reads come from it and FETCH_SIZE is
kept as a cross-check.

A window of plain CA steps (bench.py's timed window is the process's last
steps):

    python -m amca3d_amd.prof_summary --trace KT/bench_kernel_trace.csv --window-steps 20 \
        --fetch ... --write ... --l2 PMC_L2/bench_counter_collection.csv ...

keeps the last 12 x window-steps dispatches of the step kernels (k_nucleate;
k_count, k_compact; k_claim, k_resolve, k_idscan, k_emit for wave 0 and again
for wave 1; k_tail: twelve per step, checked), from
the trace and from every counter file, and reports each kernel over the window
and per wave.
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


STEP_KERNELS = ["k_nucleate", "k_count", "k_compact", "k_claim", "k_resolve", "k_idscan",
                "k_emit", "k_claim", "k_resolve", "k_idscan", "k_emit", "k_tail"]
STEP_LABELS = ["k_nucleate", "k_count", "k_compact", "k_claim.w0", "k_resolve.w0", "k_idscan.w0",
               "k_emit.w0", "k_claim.w1", "k_resolve.w1", "k_idscan.w1", "k_emit.w1", "k_tail"]


def window_rows(rows, steps):
    """The last `steps` plain steps among (dispatch id, kernel, value) rows:
    {label: [values]} with labels per position in the step."""
    rows = sorted((r for r in rows if r[1] in STEP_KERNELS), key=lambda r: r[0])
    n = len(STEP_KERNELS) * steps
    if len(rows) < n:
        raise SystemExit(f"only {len(rows)} step-kernel dispatches, need {n}")
    rows = rows[-n:]
    out = defaultdict(list)
    for i, (_, name, v) in enumerate(rows):
        q = i % len(STEP_KERNELS)
        if name != STEP_KERNELS[q]:
            raise SystemExit(f"dispatch {i} of the window is {name}, expected {STEP_KERNELS[q]}")
        out[STEP_LABELS[q]].append(v)
    return out


def trace_rows(path):
    with open(path) as f:
        return [(int(r["Dispatch_Id"]), short_name(r["Kernel_Name"]),
                 (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
                for r in csv.DictReader(f)]


def counter_rows(path, counter):
    with open(path) as f:
        return [(int(r["Dispatch_Id"]), short_name(r["Kernel_Name"]), float(r["Counter_Value"]))
                for r in csv.DictReader(f) if r["Counter_Name"] == counter]


def window_main(a):
    """Per-kernel and per-wave table of the window (see the module docstring)."""
    res = {"workload": a.workload, "command": a.command, "window_steps": a.window_steps,
           "kernels": {}}
    fscale = wscale = None
    if a.copy_fetch and a.copy_write and a.copy_bytes > 0:
        cf = counter_means(a.copy_fetch, "FETCH_SIZE")["k_copy_probe"][0]
        cw = counter_means(a.copy_write, "WRITE_SIZE")["k_copy_probe"][0]
        kib = a.copy_bytes / 1024.0
        fscale, wscale = kib / cf, kib / cw
        res["copy_calibration"] = {"bytes": a.copy_bytes, "fetch_kib": cf, "write_kib": cw,
                                   "fetch_scale": fscale, "write_scale": wscale}
    dur = window_rows(trace_rows(a.trace), a.window_steps)
    fetch = window_rows(counter_rows(a.fetch, "FETCH_SIZE"), a.window_steps) if a.fetch else {}
    write = window_rows(counter_rows(a.write, "WRITE_SIZE"), a.window_steps) if a.write else {}
    rd = {}
    if a.rdreq:
        parts = {sz: window_rows(counter_rows(a.rdreq, f"TCC_EA0_RDREQ_{sz}B_sum"), a.window_steps)
                 for sz in (32, 64, 128)}
        rd = {l: [sum(sz * parts[sz][l][i] for sz in parts) for i in range(len(parts[32][l]))]
              for l in parts[32]}
        rdn = window_rows(counter_rows(a.rdreq, "TCC_EA0_RDREQ_sum"), a.window_steps)
        rd_split = {l: sum(parts[sz][l][i] for sz in parts for i in range(len(parts[32][l])))
                    for l in parts[32]}
        res["rdreq_unsized"] = {l: sum(rdn[l]) - rd_split[l] for l in rdn}
    hit = window_rows(counter_rows(a.l2, "TCC_HIT"), a.window_steps) if a.l2 else {}
    miss = window_rows(counter_rows(a.l2, "TCC_MISS"), a.window_steps) if a.l2 else {}

    def row(labels):
        d = [v for l in labels for v in dur[l]]
        r = {"calls": len(d), "total_ms": sum(d) * 1e-3, "avg_us": sum(d) / len(d)}
        if fscale and all(l in fetch and l in write for l in labels):
            fk = [v for l in labels for v in fetch[l]]
            wk = [v for l in labels for v in write[l]]
            r["fetch_kib"] = sum(fk) / len(fk)
            r["write_kib"] = sum(wk) / len(wk)
            r["bytes_per_launch"] = 1024.0 * (r["fetch_kib"] * fscale + r["write_kib"] * wscale)
        if rd and all(l in rd and l in write for l in labels):
            rb = [v for l in labels for v in rd[l]]
            wk = [v for l in labels for v in write[l]]
            r["read_bytes"] = sum(rb) / len(rb)
            r["write_bytes"] = 1024.0 * sum(wk) / len(wk)
            r["bytes_per_launch"] = r["read_bytes"] + r["write_bytes"]
        if all(l in hit and l in miss for l in labels):
            h = sum(v for l in labels for v in hit[l])
            m = sum(v for l in labels for v in miss[l])
            r["l2_hit_rate"] = h / (h + m) if h + m > 0 else None
        return r

    for k in dict.fromkeys(STEP_KERNELS):
        res["kernels"][k] = row([l for l in STEP_LABELS if l.split(".")[0] == k])
    res["by_wave"] = {l: row([l]) for l in STEP_LABELS}
    step_us = sum(sum(v) for v in dur.values()) / a.window_steps
    res["kernel_us_per_step"] = step_us
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    lines = [f"# {a.workload}", "", "Command:", "", "```", a.command, "```", "",
             f"Window: the last {a.window_steps} steps of the process (bench.py's timed window), "
             f"{len(STEP_KERNELS)} dispatches per step; kernel time {step_us:.1f} us per step.", ""]
    if fscale:
        lines += [f"Copy calibration: k_copy_probe moves {a.copy_bytes:.0f} B each way per "
                  f"dispatch; FETCH_SIZE reads {res['copy_calibration']['fetch_kib']:.1f} KiB, "
                  f"WRITE_SIZE {res['copy_calibration']['write_kib']:.1f} KiB -> scale "
                  f"{fscale:.4f} (fetch), {wscale:.4f} (write).", ""]
    if rd:
        lines += ["Reads from the request counts (32 a + 64 b + 128 c B, TCC_EA0_RDREQ_*B_sum); "
                  "writes from WRITE_SIZE; FETCH_SIZE x scale beside them as a cross-check.", ""]

    def table(title, rows):
        if rd:
            out = [title, "", "| kernel | calls | total ms | avg us | HBM read B/launch "
                              "| HBM write B/launch | FETCH_SIZE x scale | L2 hit rate |",
                   "|---|---|---|---|---|---|---|---|"]
        else:
            out = [title, "", "| kernel | calls | total ms | avg us | HBM bytes/launch (scaled PMC) "
                              "| L2 hit rate |", "|---|---|---|---|---|---|"]
        for k, r in rows.items():
            b, hr = r.get("bytes_per_launch"), r.get("l2_hit_rate")
            hs = '' if hr is None else f'{hr:.3f}'
            if rd:
                fx = r.get("fetch_kib")
                fxs = '' if fx is None or not fscale else f"{1024.0 * fx * fscale:.4g}"
                out.append(f"| {k} | {r['calls']} | {r['total_ms']:.3f} | {r['avg_us']:.2f} | "
                           f"{r.get('read_bytes', 0):.4g} | {r.get('write_bytes', 0):.4g} | "
                           f"{fxs} | {hs} |")
            else:
                out.append(f"| {k} | {r['calls']} | {r['total_ms']:.3f} | {r['avg_us']:.2f} | "
                           f"{'' if b is None else f'{b:.4g}'} | {hs} |")
        return out + [""]
    lines += table("By kernel:", res["kernels"]) + table("By wave:", res["by_wave"])
    text = "\n".join(lines)
    if a.md:
        with open(a.md, "w") as f:
            f.write(text)
    print(text)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats")
    ap.add_argument("--trace")
    ap.add_argument("--window-steps", type=int, default=0)
    ap.add_argument("--l2")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--rdreq")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--workload", default="")
    ap.add_argument("--json")
    ap.add_argument("--md")
    a = ap.parse_args()
    if a.window_steps > 0:
        return window_main(a)

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
    rdm = {}
    if a.rdreq:
        parts = {sz: counter_means(a.rdreq, f"TCC_EA0_RDREQ_{sz}B_sum") for sz in (32, 64, 128)}
        rdm = {k: sum(sz * parts[sz][k][0] for sz in parts if k in parts[sz]) for k in parts[128]}
    for k, s in sorted(st.items(), key=lambda kv: -kv[1]["total_ms"]):
        row = dict(s)
        if k in fm and k in wm and fscale:
            row["fetch_kib"] = fm[k][0]
            row["write_kib"] = wm[k][0]
            row["pmc_dispatches"] = fm[k][1]
            row["bytes_per_launch"] = 1024.0 * (fm[k][0] * fscale + wm[k][0] * wscale)
        if k in rdm and k in wm:
            row["read_bytes"] = rdm[k]
            row["write_bytes"] = 1024.0 * wm[k][0]
            row["bytes_per_launch"] = row["read_bytes"] + row["write_bytes"]
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
    if rdm:
        lines += ["Reads from the request counts (32 a + 64 b + 128 c B, TCC_EA0_RDREQ_*B_sum); "
                  "writes from WRITE_SIZE.", "",
                  "| kernel | calls | total ms | avg us | HBM read B/launch | HBM write B/launch |",
                  "|---|---|---|---|---|---|"]
        for k, r in res["kernels"].items():
            rb, wb = r.get("read_bytes"), r.get("write_bytes")
            lines.append(f"| {k} | {r['calls']} | {r['total_ms']:.3f} | {r['avg_us']:.2f} | "
                         f"{'' if rb is None else f'{rb:.4g}'} | "
                         f"{'' if wb is None else f'{wb:.4g}'} |")
    else:
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
