"""GPU: the serving mix at its full size, checked against the oracle.

Tables as the serving benchmark holds them resident in HBM: a 60M-row flat fact
table for the Q6 / Q1 templates, and customer 1.5M / orders 15M / a ~60M-row
joined fact table for the Q3 template. Requests are drawn 50/30/20 from the
three templates with seeded parameters and arrive in waves while earlier
batches are still on the device, so batches form and overlap as they do under
load. Every response must equal the oracle run on the same rows (and the same
request run alone through the plain pipeline), in batched mode and in serial
mode.
"""
import os
import time

import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import Server, generate, loadgen_request, request

pytestmark = pytest.mark.gpu

FLAT_COLS = ("l_shipdate", "l_quantity", "l_discount", "l_tax", "l_extendedprice",
             "l_returnflag", "l_linestatus")
THREADS = max(1, min(len(os.sched_getaffinity(0)), 16))


def strip(rows):
    return [{k: r[k] for k in r if not k.startswith("avg")} for r in rows]


@pytest.fixture(scope="module")
def full(engine):
    li = generate.lineitem_flat(42, 60_000_000)
    js = dict(customer=generate.customer(11, 1_500_000),
              orders=generate.orders(11, 15_000_000, 1_500_000),
              lineitem=generate.lineitem_joined(11, 15_000_000, 2_000_000, 100_000))
    flat = engine.table({k: li[k] for k in FLAT_COLS})
    t = {k: engine.table(v) for k, v in js.items()}
    print(f"[serve-full] tables resident: flat {len(li['l_shipdate'])}, "
          f"fact {len(js['lineitem']['l_orderkey'])}", flush=True)
    yield li, js, flat, t
    flat.free()
    for x in t.values():
        x.free()


_ORACLE = {}


def oracle(li, js, kind, p):
    key = (kind, tuple(sorted(p.items())))
    if key not in _ORACLE:
        _ORACLE[key] = _oracle(li, js, kind, p)
    return _ORACLE[key]


def _oracle(li, js, kind, p):
    if kind == "q6":
        return O.q6(li, nthreads=THREADS, **p)
    if kind == "q1":
        return strip(O.q1(li, p["cutoff"], nthreads=THREADS))
    return O.q3(js["customer"], js["orders"], js["lineitem"], p["segment"], p["date"],
                p["top_k"], nthreads=THREADS)


@pytest.mark.parametrize("mode", ["batched", "serial"])
def test_mix_60m_equals_oracle(engine, full, mode):
    li, js, flat, t = full
    reqs = [loadgen_request(2024, i) for i in range(96)]
    reqs += [("q3", dict(segment=0, date=2400, top_k=100)), ("q1", dict(cutoff=2525)),
             ("q6", dict(date_lo=0, date_hi=2526, disc_lo=0, disc_hi=10, qty_lt=51))]
    got = {}
    with Server(engine, flat, t["customer"], t["orders"], t["lineitem"], mode=mode,
                max_batch=64, slots=2) as srv:
        ids = {}
        for w in range(0, len(reqs), 11):  # waves arrive while earlier batches run
            for i in range(w, min(w + 11, len(reqs))):
                k, p = reqs[i]
                ids[srv.submit(request(k, **p))] = i
            for r in srv.poll(256, 0):
                got[ids[r["id"]]] = r
            time.sleep(0.0005)
        while len(got) < len(reqs):
            for r in srv.poll(256, 1000):
                got[ids[r["id"]]] = r
        st = srv.stats()
    print(f"[serve-full] {mode}: {st['batches']} batches, mean {st['mean_batch']:.1f}", flush=True)
    assert st["completed"] == len(reqs)
    if mode == "batched":
        assert st["batches"] < len(reqs)
    for i, (k, p) in enumerate(reqs):
        assert got[i]["status"] == 0, (i, k, p)
        res = strip(got[i]["result"]) if k == "q1" else got[i]["result"]
        assert res == oracle(li, js, k, p), (mode, i, k, p)
    # The same requests alone through the plain pipelines (a sample of each kind).
    seen = set()
    for i, (k, p) in enumerate(reqs):
        if k in seen:
            continue
        seen.add(k)
        if k == "q6":
            alone = engine.q6(flat, **p)
        elif k == "q1":
            alone = strip(engine.q1(flat, **p))
        else:
            alone = engine.q3(t["customer"], t["orders"], t["lineitem"], p["segment"], p["date"],
                              p["top_k"])
        res = strip(got[i]["result"]) if k == "q1" else got[i]["result"]
        assert res == alone, (i, k, p)
