"""GPU: gdb_q6_noalloc, the allocation-free variant of gdb_q6, called through
the C ABI. Its results equal gdb_q6's bit for bit on two calls in a row (the
second on a smaller table, so the shared scratch holds the first call's
partials when the second starts), with every device allocation of the context
made to fail meanwhile (gdb_debug_fail_allocs): the call allocates nothing."""
import pytest

from gpudbmetalbenchmark_amd import Engine, GdbError, Timing, generate, lib
from gpudbmetalbenchmark_amd.abi import GDB_E_UNSUPPORTED

pytestmark = pytest.mark.gpu

Q6_COLS = ("l_shipdate", "l_discount", "l_quantity", "l_extendedprice")
PARAMS = [dict(), dict(date_lo=0, date_hi=2526, disc_lo=0, disc_hi=10, qty_lt=51),
          dict(date_lo=400, date_hi=401, disc_lo=3, disc_hi=3, qty_lt=10)]


def test_q6_noalloc_equals_q6_twice_second_smaller():
    with Engine(0) as e:
        big = generate.lineitem_flat(77, 6_000_003)
        small = generate.lineitem_flat(78, 100_001)
        tb = e.table({k: big[k] for k in Q6_COLS})
        ts = e.table({k: small[k] for k in Q6_COLS})
        exp = [[e.q6(t, **p) for p in PARAMS] for t in (tb, ts)]
        used = e.hbm_used()
        lib().gdb_debug_fail_allocs(e._ctx, 1 << 20)  # any allocation now fails
        try:
            tm = Timing()
            for i, t in enumerate((tb, ts)):  # big, then smaller
                for j, p in enumerate(PARAMS):
                    assert e.q6_noalloc(t, timing=tm, **p) == exp[i][j], (i, j)
                    assert tm.main_kernel_ms >= 0
            # the smaller table again after the big one, and the reverse
            assert e.q6_noalloc(ts) == exp[1][0]
            assert e.q6_noalloc(tb) == exp[0][0]
        finally:
            lib().gdb_debug_fail_allocs(e._ctx, 0)
        assert e.hbm_used() == used
        empty = e.table({k: big[k][:0] for k in Q6_COLS})
        assert e.q6_noalloc(empty) == dict(revenue=0, rows=0)
        host = e.host_table({k: small[k] for k in Q6_COLS})
        with pytest.raises(GdbError) as err:
            e.q6_noalloc(host)
        assert err.value.code == GDB_E_UNSUPPORTED
