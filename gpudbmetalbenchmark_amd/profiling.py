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
    a = ap.parse_args(argv)
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
