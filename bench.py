"""Benchmark: one step = one run of a pipeline over tables resident in HBM.

Default workload: Q1 low-cardinality grouped aggregation over a 60,000,000-row
synthetic fact table (7 columns, 26 B/row), the configuration the headline
scan-bandwidth metric is quoted on. `value` is algorithmic bytes per second of
whole-pipeline time (scan + finaliser + 37-value readback), inputs resident.

  python bench.py [--steps K] [--warmup W] [--pipeline q1|q6|q3|q9|q13]

Prints one JSON line. `roofline` covers the dominant kernel, timed with
hipEvents on the engine's stream inside each call; `cpu_baseline` times the
CPU oracle on a bounded sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "scan GB/s (% HBM peak) per pipeline; served QPS at p99, batched vs serial"
HBM_PEAK_GBPS = 8000.0


def load_traffic(name):
    """HBM bytes per launch of `name` measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if a summary exists."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(name, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(n or 1, 16))


def run(args):
    import numpy as np
    from gpudbmetalbenchmark_amd import Engine, Timing, generate

    eng = Engine(0)
    info = eng.device_info()
    pipe = args.pipeline
    tables = {}
    if pipe in ("q1", "q6"):
        n = args.rows or (60_000_000)
        li = generate.lineitem_flat(42, n)
        names = (("l_shipdate", "l_quantity", "l_discount", "l_tax", "l_extendedprice",
                  "l_returnflag", "l_linestatus") if pipe == "q1" else
                 ("l_shipdate", "l_discount", "l_quantity", "l_extendedprice"))
        tables["li"] = eng.table({k: li[k] for k in names})
        host = li
        if pipe == "q1":
            step = lambda tm: eng.q1(tables["li"], 2435, timing=tm)  # noqa: E731
            kernel = "q1_scan"
            workload = f"Q1 grouped aggregation, {n}-row synthetic fact table, 6 groups x 8 aggregates"
        else:
            step = lambda tm: eng.q6(tables["li"], timing=tm)  # noqa: E731
            kernel = "q6_scan"
            workload = f"Q6 filter-scan-sum, {n}-row synthetic fact table"
        rows = n
    else:
        raise SystemExit(f"unknown pipeline {pipe}")

    tm = Timing()
    for _ in range(args.warmup):
        step(tm)
    kernel_ms = []
    gpu_bytes = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(tm)
        kernel_ms.append(tm.main_kernel_ms)
        gpu_bytes = tm.bytes
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    value = gpu_bytes * args.steps / elapsed / 1e9
    kavg = float(np.mean(kernel_ms))
    achieved = tm.main_kernel_bytes / (kavg / 1000.0) / 1e9

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded generator, seed 42)",
        "config": {"workload": workload, "pipeline": pipe, "rows": rows,
                   "bytes_per_step": gpu_bytes, "device": info.name.decode(),
                   "cu_count": info.cu_count, "pct_hbm_peak": round(100 * value / HBM_PEAK_GBPS, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": load_traffic(kernel), "kernel": kernel,
                     "kernel_ms_avg": round(kavg, 5)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pipe, host, args.cpu_seconds)
    for t in tables.values():
        t.free()
    eng.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(pipe, li, seconds):
    """The oracle over a bounded row sample of the same table, all host threads,
    repeated for about `seconds`; reported in the metric's unit (GB/s)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    threads = cpu_threads()
    sample = min(len(li["l_shipdate"]), 20_000_000)
    part = {k: v[:sample] for k, v in li.items()}
    per_row = 26 if pipe == "q1" else 20
    fn = (lambda: O.q1(part, 2435, nthreads=threads)) if pipe == "q1" else \
        (lambda: O.q6(part, nthreads=threads))
    fn()
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": round(sample * per_row * reps / el / 1e9, 3), "unit": "GB/s",
            "cores": threads, "kind": "port",
            "sample": f"first {sample} rows of the same table, {reps} runs in {el:.1f} s "
                      f"(oracle/oracle.c, {threads} threads)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--pipeline", default="q1")
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
