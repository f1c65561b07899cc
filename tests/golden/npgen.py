"""Independent numpy restatement of the engine's seeded table generator.

The golden fixtures are built from these arrays (not from libgdb.so), so a
fixture pins two things: the generator in libgdb.so must reproduce the column
digests recorded here, and the oracle must reproduce the SQL results sqlite3
computes on the same columns.

Distributions are those of the benchmark configurations (flat fact table,
customer, orders, joined fact table, part, partsupp, supplier); day 0 is
1992-01-01. The counter-based draws are SplitMix64 outputs keyed by
(seed, stream, key); a uniform integer in [lo, hi] takes the top 32 bits
through Lemire's multiply-shift.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GAMMA = 0x9E3779B97F4A7C15

STREAMS = dict(LI_SHIP=1, LI_QTY=2, LI_DISC=3, LI_RETAIL=4, LI_TAX=5, LI_RECEIPT=6,
               LI_RFLAG=7, C_SEG=10, O_CUST=20, O_DATE=21, O_NLINES=22, O_CLEN=23,
               O_CWORD=24, J_SHIP=30, J_QTY=31, J_DISC=32, J_RETAIL=33, J_PART=34,
               J_SUPPIDX=35, P_NAME=40, PS_COST=50, S_NATION=60)

PNAME_WORDS = (
    "almond antique aquamarine azure beige bisque black blanched blue blush brown "
    "burlywood burnished chartreuse chiffon chocolate coral cornflower cornsilk cream "
    "cyan dark deep dim dodger drab firebrick floral forest frosted gainsboro ghost "
    "goldenrod green grey honeydew hot indian ivory khaki lace lavender lawn lemon light "
    "lime linen magenta maroon medium metallic midnight mint misty moccasin navajo navy "
    "olive orange orchid pale papaya peach peru pink plum powder puff purple red rose "
    "rosy royal saddle salmon sandy seashell sienna sky slate smoke snow spring steel tan "
    "thistle tomato turquoise violet wheat white yellow").split()
assert len(PNAME_WORDS) == 92

COMMENT_WORDS = (
    "furiously carefully quickly slyly blithely regular final ironic pending express "
    "bold even unusual silent daring fluffy deposits accounts packages requests special "
    "theodolites instructions foxes pinto beans asymptotes dependencies excuses "
    "platelets ideas courts dolphins sleep wake haggle nag use boost affix detect "
    "integrate cajole among across above after according").split()
assert len(COMMENT_WORDS) == 48


def _mix64(z):
    with np.errstate(over="ignore"):
        z = (z + np.uint64(GAMMA)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def stream_base(seed, stream):
    return _mix64(np.uint64(seed) ^ (np.uint64(STREAMS[stream]) << np.uint64(32)))


def draw(seed, stream, keys):
    keys = np.asarray(keys, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix64(stream_base(seed, stream) + keys * np.uint64(GAMMA))


def uni(r, lo, hi):
    span = np.uint64(hi - lo + 1)
    return (np.int64(lo) + ((r >> np.uint64(32)) * span >> np.uint64(32)).astype(np.int64))


def lineitem_flat(seed, n, row0=0):
    r = np.arange(row0, row0 + n, dtype=np.uint64)
    ship = uni(draw(seed, "LI_SHIP", r), 0, 2525).astype(np.int32)
    qty = uni(draw(seed, "LI_QTY", r), 1, 50).astype(np.int32)
    disc = uni(draw(seed, "LI_DISC", r), 0, 10).astype(np.int32)
    tax = uni(draw(seed, "LI_TAX", r), 0, 8).astype(np.int32)
    price = qty.astype(np.int64) * uni(draw(seed, "LI_RETAIL", r), 90000, 209900)
    receipt = (ship + uni(draw(seed, "LI_RECEIPT", r), 1, 30)).astype(np.int32)
    ls = np.where(ship < 1263, ord("F"), ord("O")).astype(np.uint8)
    coin = uni(draw(seed, "LI_RFLAG", r), 0, 1)
    rf = np.where(receipt < 1263, np.where(coin == 1, ord("R"), ord("A")), ord("N")).astype(np.uint8)
    return dict(l_shipdate=ship, l_quantity=qty, l_discount=disc, l_tax=tax,
                l_extendedprice=price, l_returnflag=rf, l_linestatus=ls, l_receiptdate=receipt)


def customer(seed, n):
    i = np.arange(n, dtype=np.uint64)
    return dict(c_custkey=(np.arange(n) + 1).astype(np.int32),
                c_mktsegment=uni(draw(seed, "C_SEG", i), 0, 4).astype(np.uint8))


def orderkey_of(i):
    i = np.asarray(i, dtype=np.int64)
    return (((i >> 3) << 5) + (i & 7) + 1).astype(np.int32)


def order_comments(seed, idx):
    out = np.zeros((len(idx), 80), dtype=np.uint8)
    words = uni(draw(seed, "O_CWORD", np.repeat(np.asarray(idx, np.uint64) * np.uint64(16), 16)
                     + np.tile(np.arange(16, dtype=np.uint64), len(idx))), 0, 47).reshape(-1, 16)
    cut = uni(draw(seed, "O_CLEN", idx), 20, 80)
    for j in range(len(idx)):
        s = " ".join(COMMENT_WORDS[w] for w in words[j]).encode()
        s = s[:min(int(cut[j]), len(s))]
        out[j, :len(s)] = np.frombuffer(s, dtype=np.uint8)
    return out


def orders(seed, n, ncust, comments=False):
    i = np.arange(n, dtype=np.uint64)
    t = dict(o_orderkey=orderkey_of(np.arange(n)),
             o_custkey=uni(draw(seed, "O_CUST", i), 1, ncust).astype(np.int32),
             o_orderdate=uni(draw(seed, "O_DATE", i), 0, 2405).astype(np.int32),
             o_shippriority=np.zeros(n, dtype=np.int32))
    if comments:
        t["o_comment"] = order_comments(seed, i)
    return t


def supplier_of(p, j, S):
    return (((p - 1) + j * (S // 4)) % S + 1).astype(np.int32)


def lineitem_joined(seed, norders, nparts, nsupp):
    oi = np.arange(norders, dtype=np.uint64)
    nl = uni(draw(seed, "O_NLINES", oi), 1, 7)
    order_of_row = np.repeat(np.arange(norders, dtype=np.int64), nl)
    starts = np.cumsum(nl) - nl
    line = np.arange(len(order_of_row)) - np.repeat(starts, nl)
    key = (order_of_row * 8 + line).astype(np.uint64)
    odate = uni(draw(seed, "O_DATE", oi), 0, 2405)
    qty = uni(draw(seed, "J_QTY", key), 1, 50).astype(np.int32)
    part = uni(draw(seed, "J_PART", key), 1, nparts)
    sidx = uni(draw(seed, "J_SUPPIDX", key), 0, 3)
    return dict(
        l_orderkey=orderkey_of(order_of_row),
        l_partkey=part.astype(np.int32),
        l_suppkey=supplier_of(part, sidx, nsupp),
        l_quantity=qty,
        l_discount=uni(draw(seed, "J_DISC", key), 0, 10).astype(np.int32),
        l_extendedprice=qty.astype(np.int64) * uni(draw(seed, "J_RETAIL", key), 90000, 209900),
        l_shipdate=(odate[order_of_row] + uni(draw(seed, "J_SHIP", key), 1, 121)).astype(np.int32))


def part(seed, n):
    names = np.zeros((n, 55), dtype=np.uint8)
    base = stream_base(seed, "P_NAME")
    for i in range(n):
        picked, k = [], 0
        while len(picked) < 5:
            with np.errstate(over="ignore"):
                r = _mix64(base + np.uint64(i * 64 + k) * np.uint64(GAMMA))
            c = int(uni(np.array([r], dtype=np.uint64), 0, 91)[0])
            k += 1
            if c not in picked:
                picked.append(c)
        s = " ".join(PNAME_WORDS[c] for c in picked).encode()
        names[i, :len(s)] = np.frombuffer(s, dtype=np.uint8)
    return dict(p_partkey=(np.arange(n) + 1).astype(np.int32), p_name=names)


def partsupp(seed, nparts, nsupp):
    r = np.arange(nparts * 4, dtype=np.int64)
    p = r // 4 + 1
    return dict(ps_partkey=p.astype(np.int32), ps_suppkey=supplier_of(p, r % 4, nsupp),
                ps_supplycost=uni(draw(seed, "PS_COST", r.astype(np.uint64)), 100, 100000))


def supplier(seed, n):
    i = np.arange(n, dtype=np.uint64)
    return dict(s_suppkey=(np.arange(n) + 1).astype(np.int32),
                s_nationkey=uni(draw(seed, "S_NATION", i), 0, 24).astype(np.int32))
