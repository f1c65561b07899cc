"""GPU: the autotuner picks a valid launch shape and results do not depend on
the shape (every candidate is bit-exact against the oracle)."""
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import GdbError, generate
from gpudbmetalbenchmark_amd.abi import GDB_E_INVALID

pytestmark = pytest.mark.gpu


def test_autotune_q6_q1_keeps_results(engine):
    li = generate.lineitem_flat(17, 2_000_003)
    t = engine.table(li)
    exp6, exp1 = O.q6(li, nthreads=8), O.q1(li, nthreads=8)
    saved = engine.get_tuning("q6"), engine.get_tuning("q1")
    try:
        best = engine.autotune("q6", [t], reps=2)
        assert best["candidates"] > 1 and best["ms"] > 0
        assert engine.get_tuning("q6")["blocks_per_cu"] == best["blocks_per_cu"]
        assert engine.q6(t) == exp6
        best = engine.autotune("q1", [t], reps=2)
        assert best["block"] in (128, 256)
        got = engine.q1(t)
        assert [{k: r[k] for k in r if not k.startswith("avg")} for r in got] == \
            [{k: r[k] for k in r if not k.startswith("avg")} for r in exp1]
        for shape in ((64, 1, 1), (128, 3, 2), (256, 0, 4)):
            engine.set_tuning("q1", *shape)
            assert engine.q1(t) == got, shape
    finally:
        for s in saved:
            engine.set_tuning(s["pipeline"], s["block"], s["blocks_per_cu"], s["unroll"])


def test_set_tuning_rejects_bad_shapes(engine):
    with pytest.raises(GdbError) as e:
        engine.set_tuning("q1", 96, 1, 1)
    assert e.value.code == GDB_E_INVALID
