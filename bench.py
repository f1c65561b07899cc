"""Benchmark: one step = one run of a pipeline over one batch of synthetic input.

Default run (no flags): the headline is Q1 low-cardinality grouped aggregation
over a 60,000,000-row synthetic fact table (7 columns, 26 B/row), the
configuration the scan-bandwidth metric is quoted on. `value` is algorithmic
bytes per second of whole-pipeline time (every kernel, the finaliser and the
result readback), with the inputs resident in HBM when the timed region
starts. The same JSON line carries `configs`, one record per benchmark
configuration, each with its own `roofline` and `cpu_baseline`:
  c0_q6_6m        Q6 over 6M rows, a 1 GiB read between steps so the 120 MB of
                  columns come from HBM (the back-to-back, cache-resident rate
                  is reported beside it)
  c2_q3           Q3 over customer 1.5M / orders 15M / fact ~60M
  c3_q9, c3_q13   Q9 (part 2M, partsupp 8M, supplier 100K) and Q13 (orders 15M)
  c3_q9_600m      Q9 over a ~600M-row fact table streamed from pinned host
                  memory (orders 150M, part 20M), with the PCIe rate achieved
  c4_serving      the Poisson serving mix: batched vs serial QPS at p99

  python bench.py [--steps K] [--warmup W] [--quick]
  python bench.py --pipeline q1|q6|q3|q9|q13|select|sum|join|all|serve|tpch|q2,q4,...
                  [--scale sf10|sf1|sf100] [--tune] [--no-cpu] [--probe]

Launch shapes are the engine's defaults, which are the autotuner's measured
best; --tune runs the autotuner (gdb_autotune) on the same tables first and
reports the chosen shape in config.tuning.

--scale sf100 streams the fact table (orders for Q13) from pinned host memory in
chunks (the 600M-row variant); its `value` is then host-to-result throughput.
--pipeline tpch runs the other 17 TPC-H pipelines (Q2, Q4, Q5, Q7, Q8, Q10,
Q11, Q12, Q14-Q22) with the reference SQL's parameters over a seeded
TPC-H-shaped database (SF10: 60M lineitem rows; --scale sf1 for 6M); their
CPU baseline is the numpy oracle (oracle/tpch_oracle.py) on one thread.
`roofline` covers the pipeline's dominant kernel, timed with hipEvents on the
engine stream inside every call; `cpu_baseline` times the CPU oracle
(oracle/oracle.c over every host core) on the same tables or a stated sample.
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
PCIE_PEAK_GBPS = 64.0  # PCIe Gen5 x16, one direction, theoretical

# (customers, orders, parts, suppliers) of the joined tables per scale; the
# fact table has 1..7 rows per order (~4x orders).
JOIN_SCALE = {"sf1": (150_000, 1_500_000, 200_000, 10_000),
              "sf10": (1_500_000, 15_000_000, 2_000_000, 100_000),
              "sf100": (15_000_000, 150_000_000, 20_000_000, 1_000_000)}
FLAT_ROWS = {"sf1": 6_000_000, "sf10": 60_000_000, "sf100": 600_000_000}


def load_traffic(name):
    """HBM bytes per launch of kernel `name` from rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if it was measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(name, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, n or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class Workload:
    """setup() uploads tables; step(timing) runs the pipeline once; oracle()
    runs the CPU restatement once on the same data."""

    def __init__(self, name, kernel, workload, rows, step, oracle, oracle_bytes, tune=None,
                 oracle_desc=None, oracle_threads=None):
        self.name, self.kernel, self.workload, self.rows = name, kernel, workload, rows
        self.step, self.oracle, self.oracle_bytes = step, oracle, oracle_bytes
        self.tune = tune  # runs the autotuner on these tables, keeps the best shape
        self.oracle_desc = oracle_desc      # None: oracle/oracle.c over all host threads
        self.oracle_threads = oracle_threads
        self.oracle_sample = None           # None: the oracle runs the whole workload
        self.eng = None                     # the engine the step runs on (kernel timing switch)


# TPC-H pipelines over the generated TPC-H-shaped database: name -> dominant
# kernel (the one the engine brackets with its kernel events).
TPCH_KERNELS = {"q2": "q2_min", "q4": "q4_late", "q5": "q5_scan", "q7": "q7_scan",
                "q8": "q8_scan", "q10": "q10_probe", "q11": "q11_scan", "q12": "q12_scan",
                "q14": "q14_scan", "q15": "q15_scan", "q16": "q16_scan", "q17": "q17_stats",
                "q18": "q18_owner", "q19": "q19_scan", "q20": "q20_scan", "q21": "q21_owner",
                "q22": "mark_bytes"}
TPCH_SF = {"sf1": 1.0, "sf10": 10.0}


def build_tpch_workloads(eng, names, scale):
    from gpudbmetalbenchmark_amd import tpch
    from gpudbmetalbenchmark_amd.tpchdb import TpchDb
    import tpch_oracle as T
    sf = TPCH_SF[scale]
    tabs = tpch.generate(sf, 11)
    db = TpchDb(eng, tabs)
    nl = len(tabs["lineitem"]["l_orderkey"])
    brass = tpch.codes_like(tpch.P_TYPE, "%BRASS")
    medpol = tpch.codes_like(tpch.P_TYPE, "MEDIUM POLISHED%")
    oracle = {"q2": lambda: T.q2(tabs, 15, brass, 3), "q16": lambda: T.q16(tabs, 45, medpol)}
    out = {}
    for n in names:
        out[n] = Workload(
            n, TPCH_KERNELS[n],
            f"TPC-H {n.upper()} with the reference SQL's parameters over a seeded TPC-H-shaped "
            f"database at SF{sf:g} ({nl} lineitem rows)",
            nl, lambda tm, f=getattr(db, n): f(timing=tm),
            oracle.get(n, lambda f=getattr(T, n): f(tabs)), None,
            oracle_desc="oracle/tpch_oracle.py (numpy, one thread)", oracle_threads=1)
    for w in out.values():
        w.eng = eng
    return out


def build_workloads(eng, names, scale, O):
    from gpudbmetalbenchmark_amd import generate
    stream = scale == "sf100"
    alloc = eng.pinned_empty if stream else None
    th = cpu_threads()
    out = {}
    flat_needed = [n for n in names if n in ("q1", "q6")]
    if flat_needed:
        rows = FLAT_ROWS[scale]
        li = generate.lineitem_flat(42, rows, **({"empty": alloc} if alloc else {}))
        for name in flat_needed:
            if name == "q1":
                t = (eng.host_table if stream else eng.table)(
                    {k: li[k] for k in ("l_shipdate", "l_quantity", "l_discount", "l_tax",
                                        "l_extendedprice", "l_returnflag", "l_linestatus")})
                out[name] = Workload(
                    name, "q1_scan",
                    f"Q1 grouped aggregation, {rows}-row synthetic fact table"
                    f"{' streamed from pinned host memory' if stream else ''}, 6 groups x 8 aggregates",
                    rows, lambda tm, t=t: eng.q1(t, 2435, timing=tm),
                    lambda li=li: O.q1(li, 2435, nthreads=th), 26 * rows,
                    tune=lambda t=t: eng.autotune("q1", [t]))
            else:
                t = (eng.host_table if stream else eng.table)(
                    {k: li[k] for k in ("l_shipdate", "l_discount", "l_quantity",
                                        "l_extendedprice")})
                note = (" (120 MB: resident in the 256 MB Infinity Cache across steps)"
                        if rows == FLAT_ROWS["sf1"] else "")
                out[name] = Workload(
                    name, "q6_scan",
                    f"Q6 filter-scan-sum, {rows}-row synthetic fact table{note}"
                    f"{' streamed from pinned host memory' if stream else ''}",
                    rows, lambda tm, t=t: eng.q6(t, timing=tm),
                    lambda li=li: O.q6(li, nthreads=th), 20 * rows,
                    tune=lambda t=t: eng.autotune("q6", [t]))
    join_needed = [n for n in names if n in ("q3", "q9", "q13", "select", "sum", "join")]
    if join_needed:
        nc, no, npart, ns = JOIN_SCALE[scale]
        seed = 11
        kw = {"empty": alloc} if alloc else {}
        js = dict(customer=generate.customer(seed, nc),
                  orders=generate.orders(seed, no, nc, comments="q13" in names, **kw),
                  lineitem=generate.lineitem_joined(seed, no, npart, ns, **kw),
                  part=generate.part(seed, npart), partsupp=generate.partsupp(seed, npart, ns),
                  supplier=generate.supplier(seed, ns))
        nl = len(js["lineitem"]["l_orderkey"])
        dims = {k: eng.table(js[k]) for k in ("customer", "part", "partsupp", "supplier")}
        orders_res = eng.table({k: v for k, v in js["orders"].items() if k != "o_comment"})
        fact = (eng.host_table if stream else eng.table)(js["lineitem"])
        if "q3" in names:
            out["q3"] = Workload(
                "q3", "q3_probe",
                f"Q3 join+group+top-10: customer {nc}, orders {no}, fact {nl} rows"
                f"{' (fact streamed from pinned host memory)' if stream else ''}",
                nl, lambda tm: eng.q3(dims["customer"], orders_res, fact, 1, 1169, 10, timing=tm),
                lambda: O.q3(js["customer"], js["orders"], js["lineitem"], 1, 1169, 10, nthreads=th),
                5 * nc + 16 * no + 20 * nl,
                tune=lambda: eng.autotune("q3", [dims["customer"], orders_res, fact]))
        if "q9" in names:
            out["q9"] = Workload(
                "q9", "q9_probe",
                f"Q9 multi-way join, part {npart} ('green'), partsupp {4 * npart}, supplier {ns}, "
                f"orders {no}, fact {nl} rows"
                f"{' (fact streamed from pinned host memory)' if stream else ''}",
                nl, lambda tm: eng.q9(dims["part"], dims["supplier"], dims["partsupp"], orders_res,
                                      fact, "green", timing=tm),
                lambda: O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"],
                             js["lineitem"], "green", nthreads=th),
                28 * nl + 59 * npart + 16 * 4 * npart + 8 * ns + 8 * no,
                tune=lambda: eng.autotune("q9", [dims["part"], dims["supplier"], dims["partsupp"],
                                                 orders_res, fact]))
        if "q13" in names:
            otab = (eng.host_table if stream else eng.table)(
                {k: js["orders"][k] for k in ("o_custkey", "o_comment")})
            out["q13"] = Workload(
                "q13", "q13_match",
                f"Q13 left-outer count: orders {no} (80-byte comments), customer {nc}"
                f"{' (orders streamed from pinned host memory)' if stream else ''}",
                no, lambda tm: eng.q13(dims["customer"], otab, timing=tm),
                lambda: O.q13(js["customer"], js["orders"], nthreads=th), 84 * no + 4 * nc,
                tune=lambda: eng.autotune("q13", [dims["customer"], otab]))
        if not stream:
            li = js["lineitem"]
            if "select" in names:
                out["select"] = Workload(
                    "select", "select_lt_bitmap",
                    f"selection l_partkey < {npart // 2} over {nl} rows (device bitmap + count)", nl,
                    lambda tm: eng.select_lt(fact, "l_partkey", npart // 2, bitmap=False,
                                             rowids=False, timing=tm),
                    lambda: O.select_lt(li["l_partkey"], npart // 2), 4 * nl)
            if "sum" in names:
                out["sum"] = Workload(
                    "sum", "sum_scan", f"SUM(l_quantity) over {nl} rows", nl,
                    lambda tm: eng.sum(fact, "l_quantity", timing=tm),
                    lambda: O.sum_col(li["l_quantity"]), 4 * nl)
            if "join" in names:
                out["join"] = Workload(
                    "join", "jd_probe", f"join count orders({no}) x lineitem({nl}) on orderkey",
                    nl, lambda tm: eng.join_count(orders_res, "o_orderkey", fact, "l_orderkey",
                                                  timing=tm),
                    lambda: O.join_count(js["orders"]["o_orderkey"], li["l_orderkey"]),
                    4 * no + 4 * nl,
                    tune=lambda: eng.autotune("join", [orders_res, fact]))
    for w in out.values():
        w.eng = eng
    return out


def measure(w, steps, warmup, tune=False, cpu_seconds=0.0, flush=None):
    """Times `steps` runs of workload `w` after `warmup` untimed ones and
    returns its record: value (algorithmic GB/s of whole-pipeline wall time),
    roofline of the dominant kernel (mean hipEvent time) and, if
    cpu_seconds > 0, the CPU oracle's rate. With `flush` (a callable run
    between steps, outside the timed calls) each step is timed on its own.
    The timed steps run with the engine's kernel timing off: a timing event
    recorded between two kernels leaves the GPU idle for a few microseconds.
    The dominant kernel's time comes from further steps with it on
    (roofline.kernel_steps); a streamed table's per-chunk kernel times come
    from the timed steps themselves."""
    import numpy as np
    from gpudbmetalbenchmark_amd import Timing
    tuning = None
    if tune and w.tune is not None:
        t = w.tune()
        tuning = {k: t[k] for k in ("block", "blocks_per_cu", "unroll", "radix_bits", "candidates")}
    tm = Timing()
    kernel_ms, h2d, elapsed = [], 0, 0.0
    if w.eng is not None:
        w.eng.set_kernel_timing(False)
    try:
        for _ in range(warmup):
            if flush:
                flush()
            w.step(tm)
        t0 = time.perf_counter()
        for _ in range(steps):
            if flush:
                flush()
                t1 = time.perf_counter()
                w.step(tm)
                elapsed += time.perf_counter() - t1
            else:
                w.step(tm)
            if tm.chunks:
                kernel_ms.append(tm.main_kernel_ms)
            h2d = tm.h2d_bytes
        if not flush:
            elapsed = time.perf_counter() - t0
    finally:
        if w.eng is not None:
            w.eng.set_kernel_timing(True)
    kernel_steps = len(kernel_ms)
    if not tm.chunks:
        kernel_steps = min(steps, 10)
        for _ in range(kernel_steps):
            if flush:
                flush()
            w.step(tm)
            kernel_ms.append(tm.main_kernel_ms)
    bytes_step = tm.bytes
    if w.oracle_bytes is None:
        w.oracle_bytes = bytes_step
    value = bytes_step * steps / elapsed / 1e9
    kavg = float(np.mean(kernel_ms))
    # Streamed tables: the kernel time and bytes are summed over the step's chunks.
    achieved = tm.main_kernel_bytes / (kavg / 1000.0) / 1e9 if kavg > 0 else 0.0
    rec = {
        "workload": w.workload, "pipeline": w.name, "rows": w.rows, "value": round(value, 2),
        "unit": "GB/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(elapsed * 1000.0 / steps, 4), "bytes_per_step": bytes_step,
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBPS, 2), "tuning": tuning,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": load_traffic(w.kernel), "kernel": w.kernel,
                     "kernel_ms_avg": round(kavg, 5), "kernel_steps": kernel_steps,
                     "bytes_per_launch": tm.main_kernel_bytes},
    }
    if tm.chunks:
        rate = h2d * steps / elapsed / 1e9
        rec["roofline"]["per"] = f"step ({tm.chunks} chunk launches)"
        rec["streaming"] = {"chunks": tm.chunks, "h2d_bytes": h2d, "h2d_gbps": round(rate, 2),
                            "pcie_peak_gbps": PCIE_PEAK_GBPS,
                            "pcie_frac": round(rate / PCIE_PEAK_GBPS, 4)}
    if cpu_seconds > 0:
        rec["cpu_baseline"] = cpu_baseline(w, cpu_seconds)
    return rec


def run_one(w, args, info):
    """One JSON line for one pipeline (the --pipeline modes)."""
    rec = measure(w, args.steps, args.warmup, args.tune,
                  0.0 if args.no_cpu else args.cpu_seconds)
    out = {"metric": METRIC, "value": rec["value"], "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
           "data": "synthetic (seeded generator)",
           "config": {"workload": w.workload, "pipeline": w.name, "scale": args.scale,
                      "rows": w.rows, "bytes_per_step": rec["bytes_per_step"],
                      "device": (info.name or info.arch).decode(errors="replace"),
                      "cu_count": info.cu_count, "pct_hbm_peak": rec["pct_hbm_peak"],
                      "tuning": rec["tuning"]},
           "roofline": rec["roofline"]}
    if "streaming" in rec:
        out["streaming"] = rec["streaming"]
    if args.hbm_probe:
        out["hbm_probe"] = args.hbm_probe
    if "cpu_baseline" in rec:
        out["cpu_baseline"] = rec["cpu_baseline"]
    return out


def cpu_baseline(w, seconds):
    """The oracle over the same tables (or the workload's stated sample), on
    every host core, repeated for about `seconds` after one untimed run;
    reported in the metric's unit (GB/s)."""
    threads = w.oracle_threads or cpu_threads()
    desc = w.oracle_desc or f"oracle/oracle.c, {threads} threads"
    sample = w.oracle_sample or f"the full {w.rows}-row workload"
    w.oracle()
    reps, t0 = 0, time.perf_counter()
    while True:
        w.oracle()
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": round(w.oracle_bytes * reps / el / 1e9, 3), "unit": "GB/s",
            "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{sample}, {reps} runs in {el:.1f} s ({desc})"}


def serve_sweep(eng, flat, t, seconds, serial_rates, mults, max_batch=64, slots=2):
    """Config 4: Poisson arrivals drawn 50/30/20 from the Q6 / Q1 / Q3
    templates. Serial plain execution is swept first; its highest sustained
    rate sets the p99 bound, then the batched loop is swept and its highest
    sustained rate within that p99 is the result."""
    from gpudbmetalbenchmark_amd import Server
    index = {}

    def sweep(mode, rates, p99_bound=None):
        best, points = None, []
        with Server(eng, flat, t["customer"], t["orders"], t["lineitem"], mode=mode,
                    max_batch=max_batch, slots=slots) as srv:
            st0 = srv.stats()
            if mode == "batched":
                index.update(bytes=st0["index_bytes"], build_ms=round(st0["index_build_ms"], 1))
            srv.loadgen(rate=min(rates), seconds=0.3, seed=999)  # warm-up
            for i, r in enumerate(rates):
                st = srv.loadgen(rate=r, seconds=seconds, seed=100 + i)
                ok = st["qps"] >= 0.9 * r and st["completed"] == st["submitted"]
                ok = ok and (p99_bound is None or st["p99_ms"] <= p99_bound)
                points.append(dict(rate=r, qps=round(st["qps"], 1), p50_ms=round(st["p50_ms"], 3),
                                   p99_ms=round(st["p99_ms"], 3),
                                   mean_batch=round(st["mean_batch"], 2), sustained=ok))
                print(json.dumps({"serve_point": mode, **points[-1]}), file=sys.stderr, flush=True)
                if ok and (best is None or st["qps"] > best["qps"]):
                    best = points[-1]
                elif not ok and st["qps"] < 0.8 * r:
                    break
        return best, points

    s_best, s_pts = sweep("serial", serial_rates)
    bound = s_best["p99_ms"] if s_best else None
    base = s_best["qps"] if s_best else 500.0
    b_best, b_pts = sweep("batched", [int(base * m) for m in mults], bound)
    # Strict comparison: serial's best sustained point without queueing (p50
    # within 3x its p50 at the lowest rate), and the batched points whose p99
    # is no higher.
    strict = None
    p50_idle = s_pts[0]["p50_ms"] if s_pts else 0.0
    calm = [x for x in s_pts if x["sustained"] and x["p50_ms"] <= 3 * p50_idle]
    if calm:
        s_calm = max(calm, key=lambda x: x["qps"])
        ok = [x for x in b_pts if x["sustained"] and x["p99_ms"] <= s_calm["p99_ms"]]
        b_calm = max(ok, key=lambda x: x["qps"]) if ok else None
        strict = {"serial": s_calm, "batched": b_calm,
                  "speedup_qps": round(b_calm["qps"] / s_calm["qps"], 2) if b_calm else None}
    return {"serial_best": s_best, "batched_best": b_best, "p99_bound_ms": bound,
            "speedup_qps": round(b_best["qps"] / s_best["qps"], 2) if (b_best and s_best) else None,
            "strict_p99": strict, "q3_index": index,
            "serial_points": s_pts, "batched_points": b_pts}


SERVE_WORKLOAD = ("Poisson serving mix 50/30/20 Q6/Q1/Q3 with seeded parameters; 60M-row flat "
                  "fact table and customer 1.5M / orders 15M / ~60M-row joined fact resident "
                  "in HBM; serial = each request alone through gdb_q6/q1/q3, batched = fused "
                  "flat passes and Q3 batches over join indexes built at server start")


def serve_tables(eng):
    from gpudbmetalbenchmark_amd import generate
    li = generate.lineitem_flat(42, FLAT_ROWS["sf10"])
    flat = eng.table({k: li[k] for k in FLAT_COLS})
    del li
    nc, no, npart, ns = JOIN_SCALE["sf10"]
    js = dict(customer=generate.customer(11, nc), orders=generate.orders(11, no, nc),
              lineitem=generate.lineitem_joined(11, no, npart, ns))
    t = {k: eng.table(v) for k, v in js.items()}
    return flat, t


def run_serve(args, info):
    from gpudbmetalbenchmark_amd import Engine
    eng = Engine(0)
    flat, t = serve_tables(eng)
    r = serve_sweep(eng, flat, t, args.serve_seconds,
                    [500, 1000, 1500, 2000, 3000, 4000, 5000, 6000],
                    (1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 18, 20, 24, 30))
    b_best = r["batched_best"]
    out = {
        "metric": "served QPS at p99 (batched vs serial), config-4 mix", "value":
        b_best["qps"] if b_best else 0.0, "unit": "req/s", "n_gpus": 1,
        "steps": len(r["batched_points"]), "warmup": 1, "ms_per_step": args.serve_seconds * 1000.0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": SERVE_WORKLOAD,
                   "device": (info.name or info.arch).decode(errors="replace"), **r},
    }
    for x in list(t.values()) + [flat]:
        x.free()
    eng.close()
    print(json.dumps(out), flush=True)


FLAT_COLS = ("l_shipdate", "l_quantity", "l_discount", "l_tax", "l_extendedprice",
             "l_returnflag", "l_linestatus")


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_default(args, info, O):
    """The headline (Q1, 60M rows) plus one record per benchmark configuration."""
    from gpudbmetalbenchmark_amd import Engine, generate
    th = cpu_threads()
    quick = args.quick
    side = dict(steps=max(3, args.steps // (4 if quick else 1)), warmup=min(args.warmup, 3))
    cpu_side = 0.0 if args.no_cpu else (1.0 if quick else 3.0)
    configs = {}
    eng = Engine(0)
    # ---- config 1 (headline) and config 0 -------------------------------------
    wl = build_workloads(eng, ["q1"], "sf10", O)
    say("c1 q1 60M")
    # CPU baselines run after the GPU timings of their phase: an all-core oracle
    # run just before a timed GPU step slowed that step's host side.
    head = measure(wl["q1"], args.steps, args.warmup, args.tune)
    say("c0 q6 6M")
    li6 = generate.lineitem_flat(42, FLAT_ROWS["sf1"])
    q6t = eng.table({k: li6[k] for k in ("l_shipdate", "l_discount", "l_quantity",
                                          "l_extendedprice")})
    flusher = eng.table({"x": __import__("numpy").zeros(1 << 28, "int32")})  # 1 GiB
    w6 = Workload("q6", "q6_scan",
                  "Q6 filter-scan-sum, 6000000-row synthetic fact table (120 MB of columns); a "
                  "1 GiB column read between steps evicts them from the 256 MB Infinity Cache",
                  FLAT_ROWS["sf1"], lambda tm: eng.q6(q6t, timing=tm),
                  lambda: O.q6(li6, nthreads=th), 20 * FLAT_ROWS["sf1"])
    w6.eng = eng
    c0 = measure(w6, side["steps"] * 2, side["warmup"], flush=lambda: eng.sum(flusher, "x"))
    res = measure(w6, side["steps"] * 2, side["warmup"])
    c0["cache_resident"] = {"value": res["value"], "ms_per_step": res["ms_per_step"],
                            "kernel_ms_avg": res["roofline"]["kernel_ms_avg"],
                            "note": "back-to-back steps, columns left in the Infinity Cache"}
    if not args.no_cpu:
        say("c1/c0 cpu baselines")
        head["cpu_baseline"] = cpu_baseline(wl["q1"], args.cpu_seconds)
        c0["cpu_baseline"] = cpu_baseline(w6, cpu_side)
    del wl
    configs["c0_q6_6m"] = c0
    flusher.free()
    q6t.free()
    del li6
    # ---- configs 2 and 3 at the 60M-row scale --------------------------------
    say("c2/c3 tables")
    wl = build_workloads(eng, ["q3", "q9", "q13"], "sf10", O)
    c23 = (("c2_q3", "q3"), ("c3_q9", "q9"), ("c3_q13", "q13"))
    for key, name in c23:
        say(f"{key}")
        configs[key] = measure(wl[name], side["steps"], side["warmup"], args.tune)
    if cpu_side > 0:
        say("c2/c3 cpu baselines")
        for key, name in c23:
            configs[key]["cpu_baseline"] = cpu_baseline(wl[name], cpu_side)
    del wl
    eng.close()
    # ---- config 3, the 600M-row streamed variant -----------------------------
    if not args.no_stream:
        say("c3 q9 600M streamed")
        eng = Engine(0)
        configs["c3_q9_600m"] = run_streamed_q9(eng, O, quick, cpu_side)
        eng.close()
    # ---- config 4, serving ------------------------------------------------------
    if not args.no_serve:
        say("c4 serving")
        eng = Engine(0)
        flat, t = serve_tables(eng)
        r = serve_sweep(eng, flat, t, 0.6 if quick else 1.0,
                        [1000, 2000, 3000, 4000, 5000],
                        (2, 4, 8, 12, 14, 16, 18, 20, 22, 24))
        b, s = r["batched_best"], r["serial_best"]
        configs["c4_serving"] = {
            "workload": SERVE_WORKLOAD, "value": b["qps"] if b else 0.0, "unit": "req/s",
            "serial_qps": s["qps"] if s else None, "speedup_qps": r["speedup_qps"],
            "batched_p50_ms": b["p50_ms"] if b else None, "batched_p99_ms": b["p99_ms"] if b else None,
            "serial_p50_ms": s["p50_ms"] if s else None, "serial_p99_ms": s["p99_ms"] if s else None,
            "p99_bound_ms": r["p99_bound_ms"], "strict_p99": r["strict_p99"],
            "q3_index": r["q3_index"], "batched_points": r["batched_points"],
            "serial_points": r["serial_points"]}
        eng.close()
    out = {
        "metric": METRIC, "value": head["value"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": head["workload"], "pipeline": "q1", "scale": "sf10",
                   "rows": head["rows"], "bytes_per_step": head["bytes_per_step"],
                   "device": (info.name or info.arch).decode(errors="replace"),
                   "cu_count": info.cu_count, "pct_hbm_peak": head["pct_hbm_peak"],
                   "tuning": head["tuning"],
                   "host": {"cpu_model": cpu_model(), "cpu_threads": th}},
        "roofline": head["roofline"],
    }
    if "cpu_baseline" in head:
        out["cpu_baseline"] = head["cpu_baseline"]
    out["configs"] = configs
    return out


def run_streamed_q9(eng, O, quick, cpu_side):
    """Q9 over the ~600M-row fact table of 150M orders, streamed from pinned
    host memory; dimensions (orders 150M, part 20M, partsupp 80M, supplier 1M)
    resident. The CPU baseline runs the oracle on the first tenth of the fact
    rows (a bounded sample of the same workload)."""
    from gpudbmetalbenchmark_amd import generate
    nc, no, npart, ns = JOIN_SCALE["sf100"]
    seed = 11
    fact = generate.lineitem_joined(seed, no, npart, ns, empty=eng.pinned_empty)
    fact = {k: v for k, v in fact.items() if k != "l_shipdate"}
    n = len(fact["l_orderkey"])
    js = dict(orders=generate.orders(seed, no, nc), part=generate.part(seed, npart),
              partsupp=generate.partsupp(seed, npart, ns), supplier=generate.supplier(seed, ns))
    dims = {k: eng.table(v) for k, v in js.items()}
    host = eng.host_table(fact)
    h2d_ceiling = eng.h2d_probe(fact["l_extendedprice"][: 1 << 27], reps=3)  # 1 GiB
    m = n // 10
    sample = {k: v[:m] for k, v in fact.items()}
    w = Workload("q9", "q9_probe",
                 f"Q9 multi-way join, fact {n} rows (28 B/row) streamed from pinned host memory "
                 f"in chunks; part {npart} ('green'), partsupp {4 * npart}, supplier {ns}, "
                 f"orders {no} resident",
                 n, lambda tm: eng.q9(dims["part"], dims["supplier"], dims["partsupp"],
                                      dims["orders"], host, "green", timing=tm),
                 lambda: O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"], sample,
                              "green", nthreads=cpu_threads()),
                 28 * m + 59 * npart + 16 * 4 * npart + 8 * ns + 8 * no)
    w.oracle_sample = f"the first {m} fact rows (a tenth) against the full dimensions"
    w.eng = eng
    rec = measure(w, 2 if quick else 4, 1, cpu_seconds=cpu_side)
    rec["streaming"]["h2d_probe_gbps"] = round(h2d_ceiling, 2)
    rec["streaming"]["frac_of_probe"] = round(rec["streaming"]["h2d_gbps"] / h2d_ceiling, 4)
    host.free()
    for t in dims.values():
        t.free()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--pipeline", default="default")
    ap.add_argument("--scale", default="sf10", choices=["sf1", "sf10", "sf100"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on the tables first instead of the default shapes")
    ap.add_argument("--no-tune", dest="tune", action="store_false")
    ap.add_argument("--quick", action="store_true", help="default run with fewer side steps")
    ap.add_argument("--no-stream", action="store_true", help="default run without c3_q9_600m")
    ap.add_argument("--no-serve", action="store_true", help="default run without c4_serving")
    ap.add_argument("--probe", action="store_true", help="also measure plain read/copy GB/s")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.pipeline == "serve":
        from gpudbmetalbenchmark_amd import Engine
        e = Engine(0)
        info = e.device_info()
        e.close()
        run_serve(args, info)
        return
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from gpudbmetalbenchmark_amd import Engine
    if args.pipeline == "default":
        e = Engine(0)
        info = e.device_info()
        e.close()
        print(json.dumps(run_default(args, info, O)), flush=True)
        return
    names = (["q1", "q6", "q3", "q9", "q13", "select", "sum", "join"]
             if args.pipeline == "all" else
             list(TPCH_KERNELS) if args.pipeline == "tpch" else args.pipeline.split(","))
    tpch_names = [n for n in names if n in TPCH_KERNELS]
    names = [n for n in names if n not in TPCH_KERNELS]
    if args.scale == "sf100":
        names = [n for n in names if n in ("q1", "q6", "q3", "q9", "q13")]
        tpch_names = []
        args.tune = False
    eng = Engine(0)
    info = eng.device_info()
    args.hbm_probe = None
    if args.probe:
        r, c = eng.bandwidth_probe(1 << 31, 10)
        args.hbm_probe = {"read_gbps": round(r, 1), "copy_gbps": round(c, 1)}
    if names:
        wl = build_workloads(eng, names, args.scale, O)
        for n in names:
            print(json.dumps(run_one(wl[n], args, info)), flush=True)
        del wl
    if tpch_names:
        wl = build_tpch_workloads(eng, tpch_names, args.scale)
        for n in tpch_names:
            print(json.dumps(run_one(wl[n], args, info)), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
