"""Turn rocprofv3 CSV output into the engine's measurement records.

    python -m gpudbmetalbenchmark_amd.profiling traffic \
        --fetch F_counter_collection.csv --write W_counter_collection.csv \
        --out profiles/traffic.json
    python -m gpudbmetalbenchmark_amd.profiling stats --csv X_kernel_stats.csv

`traffic` reads two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of
`bench.py --probe` (one counter per run: FETCH_SIZE needs three TCC counters
and WRITE_SIZE two, more than one pass can hold together). The probe's plain
copy kernel moves a known number of bytes (`--copy-bytes`, 2 GiB by default,
read once and written once), so the ratio of those bytes to what each counter
reads for the copy is the calibration factor applied to every other kernel.
The result maps each kernel's base name (namespace and template arguments
dropped) to its average HBM bytes per dispatch; bench.py reports it as
`roofline.traffic`.

`stats` prints a `--kernel-trace --stats` summary as a markdown table.

`sq` summarises `rocprofv3 --pmc` passes of SQ counters (one or more
counter_collection CSVs, each with GRBM_GUI_ACTIVE) per kernel: waves per
dispatch, mean resident waves per CU, VALU and LDS busy fractions, the LDS
bank-conflict rate and the fraction of wave time spent waiting. On MI355X
GRBM_GUI_ACTIVE is summed over the 8 XCDs (kernel cycles = value / 8) and
SQ_WAVE_CYCLES counts in units of 4 cycles; a wave64 VALU instruction
occupies its SIMD for 4 cycles, so VALU busy = SQ_INSTS_VALU / (kernel
cycles x CUs).

`roofline` turns `bench.py --pipeline ...` JSON lines into the per-kernel
table of algorithmic bytes, measured HBM traffic and fraction of peak.
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

COPY_KERNEL = "probe_copy"


def base_name(kernel):
    """`void (anonymous namespace)::q1_scan<4>(Q1Cols, ...)` -> `q1_scan`."""
    k = re.sub(r"\(anonymous namespace\)::", "", kernel)
    k = re.sub(r"^void ", "", k)
    out, depth = [], 0
    for ch in k:
        if depth == 0 and ch == "(":
            break
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif depth == 0:
            out.append(ch)
    return "".join(out).split("::")[-1].strip()


def counter_means(path, counter):
    """kernel base name -> (mean counter value per dispatch in KiB, dispatches)."""
    vals = collections.defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            vals[base_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in vals.items()}


def traffic(fetch_csv, write_csv, copy_bytes):
    fetch = counter_means(fetch_csv, "FETCH_SIZE")
    write = counter_means(write_csv, "WRITE_SIZE")
    if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
        raise SystemExit(f"{COPY_KERNEL} missing: profile `bench.py --probe`")
    kib = copy_bytes / 1024.0
    ff = kib / fetch[COPY_KERNEL][0]
    fw = kib / write[COPY_KERNEL][0]
    out = {"_calibration": {"copy_kernel": COPY_KERNEL, "copy_bytes": copy_bytes,
                            "fetch_size_kib": round(fetch[COPY_KERNEL][0], 1),
                            "write_size_kib": round(write[COPY_KERNEL][0], 1),
                            "fetch_factor": round(ff, 4), "write_factor": round(fw, 4)}}
    for k in sorted(set(fetch) | set(write)):
        fb = fetch.get(k, (0.0, 0))[0] * 1024.0 * ff
        wb = write.get(k, (0.0, 0))[0] * 1024.0 * fw
        out[k] = {"fetch_bytes": int(round(fb)), "write_bytes": int(round(wb)),
                  "bytes_per_launch": int(round(fb + wb)),
                  "dispatches": max(fetch.get(k, (0, 0))[1], write.get(k, (0, 0))[1])}
    return out


def stats_table(path, top=40):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((base_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                         float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                         float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    lines = ["| kernel | calls | total us | avg us | min us | max us | % |",
             "|---|---|---|---|---|---|---|"]
    for name, calls, tot, avg, mn, mx, pct in rows[:top]:
        lines.append(f"| `{name}` | {calls} | {tot:.1f} | {avg:.2f} | {mn:.2f} | {mx:.2f} | {pct:.2f} |")
    return "\n".join(lines)


def pmc_means(paths):
    """kernel base name -> {counter: mean value per dispatch}, dispatches."""
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                k = base_name(r["Kernel_Name"])
                vals[k][(os.path.basename(path), r["Counter_Name"])].append(float(r["Counter_Value"]))
    out = {}
    for k, d in vals.items():
        merged = collections.defaultdict(list)
        n = 0
        for (src, c), v in d.items():
            merged[c].append(sum(v) / len(v))
            n = max(n, len(v))
        out[k] = ({c: sum(v) / len(v) for c, v in merged.items()}, n)
    return out


def sq_table(paths, kernels, cus=256, xcds=8):
    m = pmc_means(paths)
    lines = ["| kernel | dispatches | waves / dispatch | kernel cycles | resident waves / CU | "
             "VALU busy | LDS busy | LDS bank-conflict rate | waiting / wave time |",
             "|---|---|---|---|---|---|---|---|---|"]
    for k in kernels:
        if k not in m:
            lines.append(f"| `{k}` | not dispatched | | | | | | | |")
            continue
        c, n = m[k]
        cyc = c.get("GRBM_GUI_ACTIVE", 0.0) / xcds
        def ratio(a, b):
            return f"{a / b:.3f}" if b else "-"
        occ = 4.0 * c.get("SQ_WAVE_CYCLES", 0.0) / (cyc * cus) if cyc else 0.0
        valu = c.get("SQ_INSTS_VALU", c.get("SQ_ACTIVE_INST_VALU", 0.0))
        lines.append(
            f"| `{k}` | {n} | {c.get('SQ_WAVES', 0):.0f} | {cyc:.0f} | {occ:.1f} | "
            f"{ratio(valu, cyc * cus)} | {ratio(c.get('SQ_LDS_IDX_ACTIVE', 0.0), cyc * cus)} | "
            f"{ratio(c.get('SQ_LDS_BANK_CONFLICT', 0.0), c.get('SQ_LDS_IDX_ACTIVE', 0.0))} | "
            f"{ratio(c.get('SQ_WAIT_ANY', 0.0), c.get('SQ_WAVE_CYCLES', 0.0))} |")
    return "\n".join(lines)


def roofline_table(bench_jsonl, traffic_json=None):
    tr = {}
    if traffic_json:
        with open(traffic_json) as f:
            tr = json.load(f)
    lines = ["| pipeline | dominant kernel | algorithmic bytes / launch | HBM traffic / launch | "
             "traffic / algorithmic | kernel avg us | achieved GB/s | frac of 8 TB/s | "
             "pipeline ms / step |",
             "|---|---|---|---|---|---|---|---|---|"]
    with open(bench_jsonl) as f:
        for line in f:
            if not line.startswith("{"):
                continue
            d = json.loads(line)
            r = d.get("roofline") or {}
            k = r.get("kernel")
            if not k:
                continue
            alg = r.get("bytes_per_launch") or 0
            t = (tr.get(k) or {}).get("bytes_per_launch")
            ratio = f"{t / alg:.2f}" if t and alg else "-"
            name = d.get("config", {}).get("pipeline", "?")
            lines.append(f"| {name} | `{k}` | {alg / 1e6:.1f} MB | "
                         f"{(t or 0) / 1e6:.1f} MB | {ratio} | {1000 * r['kernel_ms_avg']:.1f} | "
                         f"{r['achieved']:.0f} | {r['frac']:.3f} | {d.get('ms_per_step', 0):.4f} |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=int, default=1 << 31)
    t.add_argument("--out")
    t.add_argument("--merge", action="store_true",
                   help="update the kernels of an existing --out file, keeping the others")
    s = sub.add_parser("stats")
    s.add_argument("--csv", required=True)
    s.add_argument("--top", type=int, default=40)
    q = sub.add_parser("sq")
    q.add_argument("--pmc", action="append", required=True)
    q.add_argument("--kernels", required=True, help="comma-separated kernel base names")
    q.add_argument("--cus", type=int, default=256)
    q.add_argument("--xcds", type=int, default=8)
    r = sub.add_parser("roofline")
    r.add_argument("--bench", required=True)
    r.add_argument("--traffic")
    a = ap.parse_args(argv)
    if a.cmd == "sq":
        print(sq_table(a.pmc, a.kernels.split(","), a.cus, a.xcds))
        return 0
    if a.cmd == "roofline":
        print(roofline_table(a.bench, a.traffic))
        return 0
    if a.cmd == "traffic":
        res = traffic(a.fetch, a.write, a.copy_bytes)
        if a.merge and a.out and os.path.exists(a.out):
            with open(a.out) as f:
                old = json.load(f)
            res = {**old, **{k: v for k, v in res.items() if not k.startswith("_")}}
        text = json.dumps(res, indent=1, sort_keys=True)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text + "\n")
        else:
            print(text)
    else:
        print(stats_table(a.csv, a.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
