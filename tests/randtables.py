"""Random tables for the property tests: numpy builders driven by a seed and a
few shape choices (which the tests draw with hypothesis), covering what the
seeded generator never produces: empty and one-row tables, sparse, negative,
permuted and extreme keys, keys spread over the whole int32 range, orphan foreign keys, constant columns (every row
tied), group codes outside the TPC-H sets, and text fields with bytes left
after their terminating NUL."""
import numpy as np

COLOURS = [b"green", b"blue", b"red", b"forest", b"lime", b"olive", b"azure", b"greenish",
           b"gr", b"een", b"ivory"]
WORDS = [b"special", b"requests", b"spec", b"ial", b"re", b"quests", b"pending", b"s", b"sp",
         b"deposits", b"furiously"]


def _text(rng, n, width, words, garbage):
    out = np.zeros((n, width), np.uint8)
    for i in range(n):
        s = b" ".join(words[j] for j in rng.integers(0, len(words), rng.integers(0, 9)))[:width]
        out[i, :len(s)] = np.frombuffer(s, np.uint8)
        if garbage and len(s) + 1 < width and rng.random() < 0.5:
            tail = words[rng.integers(0, len(words))][:width - len(s) - 1]
            out[i, len(s) + 1:len(s) + 1 + len(tail)] = np.frombuffer(tail, np.uint8)
    return out


def _keys(rng, n, mode):
    """n unique int32 keys laid out as `mode`."""
    base = np.arange(1, n + 1, dtype=np.int64)
    if mode == "dense":
        k = base
    elif mode == "from0":
        k = base - 1
    elif mode == "sparse":
        k = base * 37 + rng.integers(0, 37)
    elif mode == "negative":
        k = base - (1 << 30)
    elif mode == "extreme":
        k = np.int64(2**31 - 1) - (n - base) * 1000
    elif mode == "permuted":
        k = rng.permutation(base)
    elif mode == "spread":  # across the whole int32 range, rows in random order
        k = rng.permutation(np.unique(rng.integers(-2**31, 2**31, 2 * n + 16)))[:n]
        assert len(k) == n
    else:
        raise ValueError(mode)
    return k.astype(np.int32)


KEY_MODES = ("dense", "from0", "sparse", "negative", "extreme", "permuted", "spread")


def flat(seed, n, constant=False, wide=False):
    """A flat fact table for Q6 / Q1. wide: quantities up to 127, discounts
    and taxes up to 100 and prices up to 2^32 - 1 (the Q1 exact domain)."""
    rng = np.random.default_rng(seed)
    if constant:
        n1 = max(n, 0)
        t = dict(l_shipdate=np.full(n1, 1000, np.int32), l_quantity=np.full(n1, 20, np.int32),
                 l_discount=np.full(n1, 6, np.int32), l_tax=np.full(n1, 3, np.int32),
                 l_extendedprice=np.full(n1, 123456, np.int64),
                 l_returnflag=np.full(n1, ord("A"), np.uint8),
                 l_linestatus=np.full(n1, ord("F"), np.uint8))
        return t
    qmax, dmax, pmax = (127, 100, 2**32) if wide else (50, 10, 10_500_000)
    return dict(
        l_shipdate=rng.integers(-60, 2700, n).astype(np.int32),
        l_quantity=rng.integers(0, qmax + 1, n).astype(np.int32),
        l_discount=rng.integers(0, dmax + 1, n).astype(np.int32),
        l_tax=rng.integers(0, min(dmax, 8) + 1, n).astype(np.int32),
        l_extendedprice=rng.integers(0, pmax, n).astype(np.int64),
        l_returnflag=rng.choice(np.frombuffer(b"ANRX", np.uint8), n),
        l_linestatus=rng.choice(np.frombuffer(b"FOZ", np.uint8), n))


def joined(seed, nc, no, nparts, nsupp, cmode="dense", omode="dense", orphans=True,
           garbage=True, text=True):
    """customer / orders / lineitem / part / partsupp / supplier with unique
    keys laid out as cmode (customers) and omode (orders). Orders may name
    customers that do not exist and fact rows orders that do not exist
    (orphans); every fact row's (part, supplier) is a partsupp row, as TPC-H
    requires. text=False leaves the comments and names empty (faster)."""
    rng = np.random.default_rng(seed)
    ck = _keys(rng, nc, cmode)
    cust = dict(c_custkey=ck, c_mktsegment=rng.integers(0, 5, nc).astype(np.uint8))
    ok = _keys(rng, no, omode)
    oc = rng.choice(ck, no) if nc else np.zeros(no, np.int32)
    if orphans and no:
        m = rng.random(no) < 0.05
        oc = np.where(m, np.int32(2**31 - 5), oc)
    odate = rng.integers(0, 2406, no).astype(np.int32)
    orders = dict(o_orderkey=ok, o_custkey=oc.astype(np.int32), o_orderdate=odate,
                  o_shippriority=rng.integers(0, 3, no).astype(np.int32),
                  o_comment=_text(rng, no, 80, WORDS, garbage) if text else
                  np.zeros((no, 80), np.uint8))
    pk = rng.permutation(np.arange(1, nparts + 1)).astype(np.int32)
    part = dict(p_partkey=pk, p_name=_text(rng, nparts, 55, COLOURS, garbage))
    supp = dict(s_suppkey=np.arange(1, nsupp + 1, dtype=np.int32),
                s_nationkey=rng.integers(0, 25, nsupp).astype(np.int32))
    per = min(4, nsupp)
    ps_pk = np.repeat(np.arange(1, nparts + 1), per).astype(np.int32)
    ps_sk = np.concatenate([rng.choice(np.arange(1, nsupp + 1), per, replace=False)
                            for _ in range(nparts)]).astype(np.int32) if nparts and per else \
        np.zeros(0, np.int32)
    partsupp = dict(ps_partkey=ps_pk, ps_suppkey=ps_sk,
                    ps_supplycost=rng.integers(100, 100_000, len(ps_pk)).astype(np.int64))
    lines = rng.integers(0, 8, no)
    m = int(lines.sum())
    owner = np.repeat(np.arange(no), lines)
    lok = ok[owner] if no else np.zeros(0, np.int32)
    lship = odate[owner] + rng.integers(1, 122, m) if no else np.zeros(0, np.int32)
    if orphans and m:
        o = rng.random(m) < 0.03
        lok = np.where(o, np.int32(-2**31), lok)
    if nparts and per:
        row = rng.integers(0, len(ps_pk), m)
        lpk, lsk = ps_pk[row], ps_sk[row]
    else:
        lpk = lsk = np.zeros(m, np.int32)
    li = dict(l_orderkey=lok.astype(np.int32), l_partkey=lpk.astype(np.int32),
              l_suppkey=lsk.astype(np.int32), l_quantity=rng.integers(1, 51, m).astype(np.int32),
              l_discount=rng.integers(0, 11, m).astype(np.int32),
              l_extendedprice=rng.integers(90_000, 10_500_000, m).astype(np.int64),
              l_shipdate=lship.astype(np.int32))
    return dict(customer=cust, orders=orders, lineitem=li, part=part, partsupp=partsupp,
                supplier=supp)
