"""GPU: tables parsed from .tbl files run through the pipelines and give the
oracle's results on the arrays they were written from."""
import numpy as np
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import generate, tbl

pytestmark = pytest.mark.gpu


def test_tbl_files_through_the_pipelines(engine, tmp_path):
    seed = 13
    flat = generate.lineitem_flat(seed, 40_000, receipt=True)
    js = generate.join_set(seed, 3_000, 20_000, 4_000, 200)
    parsed = {}
    for name, cols in [("flat", flat)] + [(k, v) for k, v in js.items()]:
        table = "lineitem" if name == "flat" else name
        p = str(tmp_path / f"{name}.tbl")
        tbl.dump(p, table, cols)
        parsed[name] = tbl.load(p, table, columns=list(cols))
    t = engine.table({k: parsed["flat"][k] for k in ("l_shipdate", "l_quantity", "l_discount", "l_tax",
                                                     "l_extendedprice", "l_returnflag",
                                                     "l_linestatus")})
    assert engine.q6(t) == O.q6(flat)
    got1, exp1 = engine.q1(t), O.q1(flat)
    strip = lambda rows: [{k: r[k] for k in r if not k.startswith("avg")} for r in rows]
    assert strip(got1) == strip(exp1)
    d = {k: engine.table(parsed[k]) for k in ("customer", "orders", "lineitem", "part", "partsupp",
                                              "supplier")}
    assert engine.q3(d["customer"], d["orders"], d["lineitem"]) == \
        O.q3(js["customer"], js["orders"], js["lineitem"], 1, 1169, 10)
    assert engine.q9(d["part"], d["supplier"], d["partsupp"], d["orders"], d["lineitem"]) == \
        O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"], js["lineitem"], "green")
    assert engine.q13(d["customer"], d["orders"]) == O.q13(js["customer"], js["orders"])
    for x in list(d.values()) + [t]:
        x.free()


def test_cli_reads_tbl_directory(tmp_path):
    """gdb_bench --data DIR: the reference's `<scale> <query>` CLI over dbgen-format files."""
    import os
    import re
    import subprocess
    seed = 17
    js = generate.join_set(seed, 2_000, 10_000, 3_000, 100)
    li = dict(js["lineitem"])
    extra = generate.lineitem_flat(seed, len(li["l_orderkey"]), receipt=True)
    for k in ("l_tax", "l_returnflag", "l_linestatus", "l_receiptdate"):
        li[k] = extra[k]
    js["lineitem"] = li
    for name, cols in js.items():
        tbl.dump(str(tmp_path / f"{name}.tbl"), name, cols)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "gpudbmetalbenchmark_amd", "gdb_bench")
    r = subprocess.run([exe, "sf1", "q6", "--data", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Total Revenue: \$(-?\d+)\.(\d{4})", r.stdout)
    rev = O.q6(li)["revenue"]
    assert m and int(m.group(1)) * 10000 + int(m.group(2)) == rev, r.stdout
    assert "Q6 |" in r.stdout and "GPU Execution:" in r.stdout
    r = subprocess.run([exe, "sf1", "q3", "--data", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    top = O.q3(js["customer"], js["orders"], li, 1, 1169, 10)
    keys = [int(x) for x in re.findall(r"^\|\s+(\d+) \|", r.stdout, re.M)]
    assert keys == [row["orderkey"] for row in top["rows"]], r.stdout


def test_harness_writes_the_reference_csv(tmp_path):
    from gpudbmetalbenchmark_amd import harness
    rows = harness.run("sf1", ["q6"], str(tmp_path), timestamp="t0")
    assert len(rows) == 1 and rows[0].startswith("t0,SF-1,Q6,")
    csv = (tmp_path / "gpu_results.csv").read_text().splitlines()
    assert csv[0] == harness.HEADER and csv[1] == rows[0]
    assert (tmp_path / "gpu_logs" / "t0" / "SF-1_Q6.log").exists()


def test_cli_runs_the_tpch_pipelines_over_a_tbl_database(tmp_path):
    """gdb_bench --data DIR all over a whole TPC-H database in dbgen format:
    every one of the 22 registry names runs, each prints the reference's
    `Qn | N rows` header, and printed results equal the oracle's."""
    import os
    import re
    import subprocess
    import tpch_oracle as T
    from gpudbmetalbenchmark_amd import tbl, tpch
    db = tpch.generate(0.02, 4)
    tbl.dump_database(str(tmp_path), db)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "gpudbmetalbenchmark_amd", "gdb_bench")
    print("case: gdb_bench sf1 all --data <tpch sf0.02>", flush=True)
    r = subprocess.run([exe, "sf1", "all", "--data", str(tmp_path)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout
    for q in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22):
        assert re.search(rf"^Q{q} \| \d+ rows$", out, re.M), f"Q{q} header missing"
    promo, total = T.q14(db)
    m = re.search(r"Promo Revenue = ([0-9.]+)%", out)
    assert m and abs(float(m.group(1)) - round(100.0 * promo / total, 2)) < 0.006
    m = re.search(r"avg_yearly = ([0-9.]+)", out)
    assert m and abs(float(m.group(1)) - T.q17(db) / 100.0 / 7.0) < 0.006
    for code, cnt, bal in T.q22(db):
        assert f"  {code} numcust {cnt} totacctbal {bal / 100.0:.2f}" in out
    for prio, cnt in T.q4(db):
        assert re.search(rf"^  {re.escape(tpch.O_ORDERPRIORITY[prio])}\s+{cnt}$", out, re.M)
    rows = T.q21(db)
    assert f"TPC-H Q21: {min(len(rows), 100)} of" in out
    for sk, n in rows[:5]:
        assert f"  Supplier#{sk:09d} numwait {n}" in out
