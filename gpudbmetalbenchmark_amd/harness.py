"""Benchmark harness with the reference's CSV contract.

    python -m gpudbmetalbenchmark_amd.harness [sf1|sf10|sf100|all] [--queries q1,q6,...]
                                              [--results DIR] [--data DIR]

The reference's scripts/benchmark_gpu.sh runs its binary per scale factor,
cuts the output into the `Qn | N rows` ... `Total Execution:` blocks and
appends one row per query to results/gpu_results.csv with the header
`timestamp,scale_factor,query,gpu_exec_ms,cpu_post_ms,total_exec_ms,gpu_name,memory_gb`
(scripts/benchmark_gpu.sh:35, :47-68), plus a per-query log. This does the same
over gdb_bench, whose output keeps those lines, and appends two columns the
reference lacks: `gbps` (the pipeline's algorithmic bytes over its GPU time),
`pct_hbm_peak`, and the latency of the timed runs gdb_bench makes per query:
`p50_ms` and `p99_ms` (nearest rank, GPU + CPU post time) and `qps` (queries
per second one at a time).
"""
import argparse
import datetime
import os
import re
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(PKG, "gdb_bench")
HEADER = ("timestamp,scale_factor,query,gpu_exec_ms,cpu_post_ms,total_exec_ms,gpu_name,memory_gb,"
          "gbps,pct_hbm_peak,p50_ms,p99_ms,qps")
HBM_PEAK_GBPS = 8000.0
SF_LABEL = {"sf1": "SF-1", "sf10": "SF-10", "sf100": "SF-100"}


def _metric(block, label):
    m = re.search(re.escape(label) + r":\s+([0-9]+\.[0-9]+)", block)
    return float(m.group(1)) if m else None


def parse_output(text):
    """Per-query records from gdb_bench output: the blocks from `Qn | N rows`
    (or `SF100 Qn | N rows`) to `Total Execution:`, as the reference's
    harness cuts them, plus the device line."""
    dev = re.search(r"^Device: (.*?) \((.*?)\), (\d+) CUs, ([0-9.]+) GB HBM", text, re.M)
    recs = []
    for m in re.finditer(r"^(?:SF100 )?(Q\d+|selection|aggregation|join) \| (\d+) rows\n(.*?Total "
                         r"Execution:[^\n]*)(?:\n\s+Bandwidth:\s+([0-9.]+) GB/s)?"
                         r"(?:\n\s+Latency p50 / p99:\s+([0-9.]+) / ([0-9.]+) ms\s+\((\d+) runs, "
                         r"([0-9.]+) queries/s one at a time\))?", text, re.M | re.S):
        block = m.group(3)
        recs.append(dict(query=m.group(1), rows=int(m.group(2)),
                         gpu_exec_ms=_metric(block, "GPU Execution"),
                         cpu_post_ms=_metric(block, "CPU Post Process") or 0.0,
                         total_exec_ms=_metric(block, "Total Execution") or 0.0,
                         gbps=float(m.group(4)) if m.group(4) else None,
                         p50_ms=float(m.group(5)) if m.group(5) else None,
                         p99_ms=float(m.group(6)) if m.group(6) else None,
                         qps=float(m.group(8)) if m.group(8) else None))
    return dict(gpu_name=dev.group(1) if dev else "unknown",
                memory_gb=float(dev.group(4)) if dev else 0.0, queries=recs)


def csv_rows(parsed, sf_label, timestamp):
    out = []
    for r in parsed["queries"]:
        gbps = r["gbps"] if r["gbps"] is not None else 0.0
        lat = (f"{r['p50_ms']:.3f},{r['p99_ms']:.3f},{r['qps']:.1f}" if r.get("p50_ms") is not None
               else ",,")
        out.append(f"{timestamp},{sf_label},{r['query']},{r['gpu_exec_ms']:.2f},"
                   f"{r['cpu_post_ms']:.2f},{r['total_exec_ms']:.2f},{parsed['gpu_name']},"
                   f"{parsed['memory_gb']:.0f},{gbps:.2f},{100.0 * gbps / HBM_PEAK_GBPS:.2f},{lat}")
    return out


def run(scale, queries, results, data=None, timestamp=None, exe=EXE):
    timestamp = timestamp or datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
    logs = os.path.join(results, "gpu_logs", timestamp)
    os.makedirs(logs, exist_ok=True)
    csv = os.path.join(results, "gpu_results.csv")
    if not os.path.exists(csv):
        with open(csv, "w") as f:
            f.write(HEADER + "\n")
    rows = []
    for q in queries:
        cmd = [exe, scale, q] + (["--data", data] if data else [])
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}):\n{r.stderr}")
        with open(os.path.join(logs, f"{SF_LABEL[scale]}_{q.upper()}.log"), "w") as f:
            f.write(r.stdout)
        rows += csv_rows(parse_output(r.stdout), SF_LABEL[scale], timestamp)
    with open(csv, "a") as f:
        for line in rows:
            f.write(line + "\n")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("scale", nargs="?", default="sf1", choices=["sf1", "sf10", "sf100", "all"])
    ap.add_argument("--queries", default="q1,q3,q6,q9,q13")
    ap.add_argument("--results", default="results")
    ap.add_argument("--data", default=None)
    a = ap.parse_args(argv)
    scales = ["sf1", "sf10", "sf100"] if a.scale == "all" else [a.scale]
    for s in scales:
        for line in run(s, a.queries.split(","), a.results, a.data):
            print(line)
    return 0


if __name__ == "__main__":
    sys.exit(main())
