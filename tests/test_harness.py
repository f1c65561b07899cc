"""CPU: the harness cuts gdb_bench output the way the reference's
scripts/benchmark_gpu.sh does and writes its CSV schema."""
from gpudbmetalbenchmark_amd import harness

SAMPLE = """Device: AMD Instinct MI355X (gfx950:sramecc+:xnack-), 256 CUs, 309 GB HBM
Data: seeded synthetic tables
--- Running TPC-H Query 6 Benchmark ---
TPC-H Query 6 Result:
Total Revenue: $1234.5678

Q6 | 6000000 rows
  CPU Parsing (.tbl):      15.20 ms
  GPU Execution:            0.05 ms
  CPU Post Process:         0.00 ms
  Total Execution:          0.05 ms  (GPU + CPU post)
  Bandwidth:             2400.00 GB/s
  Latency p50 / p99:       0.050 / 0.061 ms  (10 runs, 19230.8 queries/s one at a time)

--- Running TPC-H Query 9 Benchmark ---

SF100 Q9 | 599988546 rows
  CPU Parsing (.tbl):     100.00 ms
  GPU Execution:          298.29 ms
  CPU Post Process:         0.01 ms
  Total Execution:        298.30 ms  (GPU + CPU post)
  Bandwidth:               68.62 GB/s
"""


def test_parse_blocks():
    p = harness.parse_output(SAMPLE)
    assert p["gpu_name"] == "AMD Instinct MI355X" and p["memory_gb"] == 309
    q = {r["query"]: r for r in p["queries"]}
    assert q["Q6"]["gpu_exec_ms"] == 0.05 and q["Q6"]["rows"] == 6000000 and q["Q6"]["gbps"] == 2400.0
    assert q["Q9"]["total_exec_ms"] == 298.30 and q["Q9"]["cpu_post_ms"] == 0.01
    assert (q["Q6"]["p50_ms"], q["Q6"]["p99_ms"], q["Q6"]["qps"]) == (0.05, 0.061, 19230.8)
    assert q["Q9"]["p50_ms"] is None


def test_csv_schema():
    rows = harness.csv_rows(harness.parse_output(SAMPLE), "SF-1", "20260101_000000")
    assert harness.HEADER.startswith(
        "timestamp,scale_factor,query,gpu_exec_ms,cpu_post_ms,total_exec_ms,gpu_name,memory_gb")
    assert rows[0] == ("20260101_000000,SF-1,Q6,0.05,0.00,0.05,AMD Instinct MI355X,309,2400.00,30.00,"
                       "0.050,0.061,19230.8")
    assert rows[1].endswith(",,,")
    assert len(rows) == 2 and all(len(r.split(",")) == len(harness.HEADER.split(",")) for r in rows)
