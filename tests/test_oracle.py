"""CPU tests: pin the oracle (against the reference's own compiled MurmurHash2_64.cu, hand
vectors, closed forms and brute force) before it is trusted as the checker."""
import datetime
import os
import struct

import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- MurmurHash64A (MurmurHash2_64.cu:63-104) --------------------------------------------
def _ref_murmur():
    import ctypes as C

    if not os.path.exists(O.REF_MURMUR):
        pytest.skip("oracle/_ref not built (reference absent)")
    lib = C.CDLL(O.REF_MURMUR)
    f = getattr(lib, "_Z13MurmurHash64APKvij")
    f.restype = C.c_uint64
    f.argtypes = [C.c_char_p, C.c_int, C.c_uint32]
    return f


def test_murmur_matches_reference_build():
    ref = _ref_murmur()
    r = np.random.default_rng(0)
    for n in range(0, 65):
        for _ in range(8):
            b = r.integers(0, 256, n, dtype=np.uint8).tobytes()
            seed = int(r.integers(0, 1 << 32))
            assert O.murmur64a(b, seed) == ref(b, n, seed)
            assert O.murmur64a(b, 100) == ref(b, n, 100)


def test_murmur_golden_vectors():
    """Vectors written by tests/golden/make_murmur_vectors.py with the reference's own
    compiled MurmurHash64A; they pin the oracle where oracle/_ref is absent."""
    path = os.path.join(GOLDEN, "murmur64a.txt")
    n = 0
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            hexkey, seed, want = line.split()
            key = bytes.fromhex(hexkey) if hexkey != "-" else b""
            assert O.murmur64a(key, int(seed)) == int(want, 16)
            assert L.murmur64a(key, int(seed)) == int(want, 16)
            n += 1
    assert n >= 50


def test_murmur_closed_form_single_block():
    """len=8, seed=0, key=0: h = 8m, then the empty-key finalisation (hand evaluation)."""
    m = 0xC6A4A7935BD1E995
    M = (1 << 64) - 1
    h = (8 * m) & M
    k = 0
    k = (k * m) & M
    k ^= k >> 47
    k = (k * m) & M
    h ^= k
    h = (h * m) & M
    h ^= h >> 47
    h = (h * m) & M
    h ^= h >> 47
    assert O.murmur64a(b"\0" * 8, 0) == h


# ---- generators ---------------------------------------------------------------------------
def test_day_to_ymd_matches_datetime():
    d0 = datetime.date(1992, 1, 1)
    for d in range(0, 2557 + 130):
        x = d0 + datetime.timedelta(days=d)
        assert O.day_to_ymd(d) == x.year * 10000 + x.month * 100 + x.day
    assert O.day_to_ymd(1169) == 19950315  # 1169/2557 of the days precede 1995-03-15


def test_generator_ranges():
    key, date, price, disc = O.gen_flat(1, 100000)
    assert key.min() >= 0 and key.max() < 250000
    assert date.min() >= 19920101 and date.max() <= 19981231
    assert price.min() >= 90000 and price.max() <= 10500000
    assert set(np.unique(disc)) == set(range(11))
    sel = (date < 19950315).mean()
    assert abs(sel - 1169 / 2557) < 0.01
    ok, oc, od, op = O.gen_orders(1, 40000, 1500)
    assert np.all(np.diff(ok) > 0)                       # unique, ascending
    blocks = (ok - 1) // 32
    assert np.all(np.bincount(blocks) == 4)              # 4 used keys per 32-key block
    assert oc.min() >= 1 and oc.max() <= 1500 and np.all(op == 0)
    lk, lp, ld, ls = O.gen_fact(1, 40000)
    assert np.all(np.diff(lk) >= 0)
    per = np.unique(lk, return_counts=True)[1]
    assert per.min() >= 1 and per.max() <= 7 and len(per) == 40000
    # shipdate - orderdate in [1, 121] days
    d0 = datetime.date(1992, 1, 1)
    ymd2day = {O.day_to_ymd(d): d for d in range(0, 2700)}
    odate = dict(zip(ok.tolist(), od.tolist()))
    gap = np.array([ymd2day[s] - ymd2day[odate[k]] for k, s in zip(lk[:5000], ls[:5000])])
    assert gap.min() >= 1 and gap.max() <= 121


def test_dict_first_seen():
    codes_of, codes = O.dict_first_seen(np.array([3, 3, 1, 4, 1, 0]), 5)
    assert list(codes_of) == [3, 1, -1, 0, 2]
    assert list(codes) == [0, 0, 1, 2, 1, 3]


# ---- LIKE (filter.cu:67-129) ------------------------------------------------------------------
def _like_regex(pattern):
    import re

    out = []
    for ch in pattern:
        out.append(".*" if ch == "%" else "[^.]" if ch == "?" else re.escape(ch))
    return re.compile("".join(out), re.S)


def test_like_matches_wildcard_regex():
    """With a width far above the matcher's work (at most (#'%' + 1) * (len + 1)^2 steps),
    the position budget never binds and LIKE is the wildcard language: '%' = any run,
    '?' = one character other than '.'."""
    rng = np.random.default_rng(7)
    for _ in range(4000):
        v = "".join(rng.choice(list("ab."), size=rng.integers(0, 8)))
        p = "".join(rng.choice(list("ab.?%"), size=rng.integers(0, 6)))
        want = _like_regex(p).fullmatch(v) is not None
        assert O.like(v, 4096, p) == want, (v, p)


@pytest.mark.parametrize("value,width,pattern,want", [
    # traced by hand: the position counter j (filter.cu:83) is shared by every restart after
    # a '%' and bounded by the column width, so late restarts run out of budget
    ("BUILDING", 8, "%ING", False),   # j reaches 8 on the 'N' of the last attempt
    ("BUILDING", 10, "%ING", True),   # a wider column leaves room for the 'G'
    ("aaab", 4, "%ab", False),        # third attempt starts with j == 4
    ("aaab", 5, "%ab", False),        # third attempt matches 'a' with j == 4, then j == 5
    ("aaab", 6, "%ab", True),
    ("ab", 2, "%b", True),            # second attempt matches 'b' with j == 1
    ("BUILDING", 8, "BUILD%", True),  # a trailing '%' run matches at once
    ("BUILDING", 8, "BUILDING%%", True),  # the value ends first; trailing '%' are skipped
    ("a.c", 3, "a?c", False),         # '?' never matches '.'
    ("abc", 3, "a?c", True),
    ("", 4, "%", True),
    ("", 4, "", True),
    ("abc", 3, "", False),            # pattern exhausted before the value: mismatch
    ("abcdef", 3, "abc", True),       # only the first width bytes are the value
])
def test_like_position_budget(value, width, pattern, want):
    assert O.like(value, width, pattern) == want


# ---- codecs -------------------------------------------------------------------------------
def test_pack_hand_example():
    # codes 5, 2, 7 at 3 bits, LSB first: 101 | 010 | 111 -> 0b111_010_101
    w = O.pack(np.array([105, 102, 107]), 100, 3)
    assert int(w[0]) == 0b111010101
    assert [O.unpack(w, i, 100, 3) for i in range(3)] == [105, 102, 107]


@pytest.mark.parametrize("bits", [1, 3, 17, 31, 32, 33, 47, 63, 64])
def test_pack_round_trip(bits):
    r = np.random.default_rng(bits)
    base = -12345
    hi = (1 << bits) - 1 if bits < 64 else (1 << 63) - 1
    codes = r.integers(0, hi, 999, dtype=np.uint64, endpoint=True)
    v = (codes.astype(np.uint64) + np.uint64(base & ((1 << 64) - 1))).view(np.int64)
    w = O.pack(v, base, bits)
    got = np.array([O.unpack(w, i, base, bits) for i in range(len(v))], np.int64)
    assert np.array_equal(got, v)


def test_pfor_file_layout():
    """Header fields at the offsets compress.cu:511-522 writes and :180-190 reads."""
    v = np.array([1000, 1003, 1001, 1255, 1000], np.int64)
    buf = O.pfor_encode(v)
    cnt, lo, hi, comp = struct.unpack_from("<IqqI", buf, 0)
    assert (cnt, lo, hi, comp) == (5, 1000, 1255, 0)       # 8-bit codes: 255 range -> 8 bits
    assert list(buf[24:29]) == [0, 3, 1, 255, 0]
    cnt2, rec, bits, olo, fit, start = struct.unpack_from("<IIIqIq", buf, 24 + cnt)
    assert (cnt2, rec, bits, olo, fit, start) == (5, 5, 8, 1000, 8, 0)
    assert np.array_equal(O.pfor_decode(buf, 5), v)
    w = np.array([0, 1 << 20], np.int64)                   # 21 bits -> rounded up to 32
    b2 = O.pfor_encode(w)
    assert struct.unpack_from("<I", b2, 24 + 8 + 8)[0] == 32


def test_binary_table_layouts():
    """data.dictionary (save_col_data, cm.cu:4057-4080) and .header (writeHeader,
    cm.cu:1251-1263), byte by byte."""
    dd = O.data_dictionary_bytes({"t": {"b": (2, 10), "a": (0, 2)}})
    assert dd == (struct.pack("<Q", 1) + struct.pack("<Q", 1) + b"t" + struct.pack("<Q", 2)
                  + struct.pack("<Q", 1) + b"a" + struct.pack("<II", 0, 2)
                  + struct.pack("<Q", 1) + b"b" + struct.pack("<II", 2, 10))
    assert O.parse_data_dictionary(dd) == {"t": {"a": (0, 2), "b": (2, 10)}}
    h = O.header_bytes(3000, 3, 1300, 2600)
    assert len(h) == 20 and struct.unpack("<QIII", h) == (3000, 3, 1300, 2600)


def test_pfor_delta_round_trip():
    r = np.random.default_rng(2)
    v = np.cumsum(r.integers(0, 50, 1000)).astype(np.int64) + 7
    buf = O.pfor_encode(v, delta=True)
    cnt, lo, hi, comp = struct.unpack_from("<IqqI", buf, 0)
    assert comp == 1 and lo == v[0] and hi == v[-1]
    assert np.array_equal(O.pfor_decode(buf, len(v)), v)


# ---- predicates and expressions -----------------------------------------------------------
NAME, NUM, STR, CMP, AND, OR = L.RPN_NAME, L.RPN_NUMBER, L.RPN_STRING, L.RPN_CMP, L.RPN_AND, L.RPN_OR
ADD, MINUS, MUL, DIV = L.RPN_ADD, L.RPN_MINUS, L.RPN_MUL, L.RPN_DIV


def test_filter_plain_comparisons_brute_force():
    r = np.random.default_rng(1)
    a, b = r.integers(-50, 50, 500), r.integers(-50, 50, 500)
    ops = {1: np.less, 2: np.greater, 3: np.not_equal, 4: np.equal, 5: np.less_equal, 6: np.greater_equal}
    for code, f in ops.items():
        m = O.filter_mask([(NAME, 0, 0), (NUM, 0, 7), (CMP, code, 0)], [a, b], [0, 0])
        assert np.array_equal(m, f(a, 7))
        # NUMBER CMP NAME reverses the operator (reverse_op, cm.cu:3166)
        m = O.filter_mask([(NUM, 0, 7), (NAME, 0, 0), (CMP, code, 0)], [a, b], [0, 0])
        assert np.array_equal(m, f(7, a))
        m = O.filter_mask([(NAME, 0, 0), (NAME, 1, 0), (CMP, code, 0)], [a, b], [0, 0])
        assert np.array_equal(m, f(a, b))
    m = O.filter_mask([(NAME, 0, 0), (NUM, 0, 0), (CMP, 1, 0), (NAME, 1, 0), (NUM, 0, 10), (CMP, 2, 0),
                       (OR, 0, 0), (NAME, 0, 0), (NUM, 0, -40), (CMP, 6, 0), (AND, 0, 0)], [a, b], [0, 0])
    assert np.array_equal(m, ((a < 0) | (b > 10)) & (a >= -40))


def test_filter_decimal_scaling():
    price = np.array([49999, 50000, 50001, 100, 999999])  # decimal(.,2): 499.99 .. 9999.99
    # price > 500.00 : column scale 2, literal precision 2
    m = O.filter_mask([(NAME, 0, 0), (NUM, 2, 50000), (CMP, 2, 0)], [price], [2])
    assert list(m) == [False, False, True, False, True]
    # price > 500.5 : literal precision 1 is scaled to 2 (filter.cu:538-548)
    m = O.filter_mask([(NAME, 0, 0), (NUM, 1, 5005), (CMP, 2, 0)], [price], [2])
    assert list(m) == list(price > 50050)
    # price >= 5 : integer literal scaled by 10^2
    m = O.filter_mask([(NAME, 0, 0), (NUM, 0, 5), (CMP, 6, 0)], [price], [2])
    assert list(m) == list(price >= 500)
    # 5 <= price, reversed operand order: the reference scales by the raw precisions
    # (compare(t, n1, op, p1, p2), filter.cu:598) so the column is multiplied by 10^2
    m = O.filter_mask([(NUM, 0, 5), (NAME, 0, 0), (CMP, 5, 0)], [price], [2])
    assert list(m) == list(price * 100 >= 5)


def test_expression_revenue_trace():
    """SUM(price*(1-discount)) with decimal(.,2) operands: scale 4, exact (select.cu:747-803)."""
    price = np.array([90000, 10500000, 123456])
    disc = np.array([0, 10, 5])
    v, p = O.eval_expr([(NAME, 0, 0), (NUM, 0, 1), (NAME, 1, 0), (MINUS, 0, 0), (MUL, 0, 0)],
                       [price, disc], [2, 2])
    assert p == 4
    assert list(v) == list(price * (100 - disc))


def test_expression_scaling_quirks():
    """Hand-derived from cm.cu:2480-2546: NAME op NUMBER with a scaled column scales the
    constant by 10^p twice for ADD/MINUS; MUL uses the raw constant."""
    price = np.array([100, 250])  # 1.00, 2.50
    v, p = O.eval_expr([(NAME, 0, 0), (NUM, 0, 1), (ADD, 0, 0)], [price], [2])
    assert p == 2 and list(v) == [100 + 10000, 250 + 10000]
    v, p = O.eval_expr([(NUM, 0, 1), (NAME, 0, 0), (ADD, 0, 0)], [price], [2])
    assert p == 2 and list(v) == [100 + 100, 250 + 100]
    v, p = O.eval_expr([(NAME, 0, 0), (NUM, 0, 3), (MUL, 0, 0)], [price], [2])
    assert p == 2 and list(v) == [300, 750]
    v, p = O.eval_expr([(NAME, 0, 0), (NAME, 1, 0), (MINUS, 0, 0)], [price, np.array([1, 2])], [2, 0])
    # precision_func(0 (top), 2): res = 0, top operand scaled by 10^2 (zone_map.cu:668-690)
    assert p == 0 and list(v) == [100 - 100, 250 - 200]
    v, p = O.eval_expr([(NUM, 0, 10), (NUM, 0, 3), (MINUS, 0, 0)], [price], [0])
    assert list(v) == [3 - 10, 3 - 10]  # constant folding uses the top of stack on the left


def test_zone_map_soundness():
    """For quirk-free predicates 'A' means every value in [min,max] passes, 'N' none."""
    r = np.random.default_rng(3)
    for _ in range(300):
        lo = int(r.integers(-100, 100))
        hi = lo + int(r.integers(0, 60))
        c = int(r.integers(-120, 160))
        code = int(r.integers(1, 7))
        ops = [(NAME, 0, 0), (NUM, 0, c), (CMP, code, 0)]
        z = O.zone_map_check(ops, [lo], [hi], [0])
        vals = np.arange(lo, hi + 1)
        m = O.filter_mask(ops, [vals], [0])
        if z == "A":
            assert m.all()
        elif z == "N":
            assert not m.any()
        assert L.zone_map_check(ops, [lo], [hi], [0]) == z


def test_zone_map_quirks():
    ops_eq = [(NAME, 0, 0), (NUM, 0, 500), (CMP, 4, 0)]
    assert O.zone_map_check(ops_eq, [0], [10], [0]) == "R"     # '==' is never 'N' (zone_map.cu:169)
    assert O.zone_map_check(ops_eq, [500], [500], [0]) == "A"
    assert O.zone_map_check([(NAME, 0, 0), (NUM, 0, 5), (CMP, 3, 0)], [0], [3], [0]) == "R"
    # constant comparisons: host_compare(int,int) yields 'A' whenever the constants differ
    assert O.zone_map_check([(NUM, 0, 4), (NUM, 0, 3), (CMP, 1, 0)], [0], [0], [0]) == "A"
    assert O.zone_map_check([(NUM, 0, 3), (NUM, 0, 3), (CMP, 1, 0)], [0], [0], [0]) == "N"
    # arithmetic makes the segment 'R' here (the reference's interval arithmetic is unsound)
    assert O.zone_map_check([(NAME, 0, 0), (NUM, 0, 1), (ADD, 0, 0), (NUM, 0, 0), (CMP, 2, 0)],
                            [5], [9], [0]) == "R"
    assert O.zone_map_check(ops_eq, [0], [10], [0], not_compressed=True) == "R"


# ---- group-by / join / order -------------------------------------------------------------
def test_groupby_brute_force():
    r = np.random.default_rng(7)
    k1, k2 = r.integers(0, 20, 3000), r.integers(-3, 3, 3000)
    v = r.integers(-1000, 1000, 3000)
    aggs = [L.AGG_SUM, L.AGG_COUNT, L.AGG_MIN, L.AGG_MAX]
    gk, gv = O.groupby([k1, k2], [v, v, v, v], aggs)
    d = {}
    for a, b, x in zip(k1, k2, v):
        d.setdefault((a, b), []).append(x)
    keys = sorted(d)
    assert list(zip(gk[0], gk[1])) == keys
    assert list(gv[0]) == [sum(d[k]) for k in keys]
    assert list(gv[1]) == [len(d[k]) for k in keys]
    assert list(gv[2]) == [min(d[k]) for k in keys]
    assert list(gv[3]) == [max(d[k]) for k in keys]
    # several segments: same groups, ascending MurmurHash64A(keys, 16, 100) order
    sk, sv = O.groupby([k1, k2], [v, v, v, v], aggs, seg_bounds=[0, 1000, 1700, 3000])
    h = [O.murmur64a(struct.pack("<qq", a, b), 100) for a, b in zip(sk[0], sk[1])]
    assert h == sorted(h)
    got = {(a, b): (s, c, mn, mx) for a, b, s, c, mn, mx in zip(sk[0], sk[1], *sv)}
    assert got == {k: (sum(d[k]), len(d[k]), min(d[k]), max(d[k])) for k in keys}


def test_join_nested_loop():
    r = np.random.default_rng(8)
    lk, rk = r.integers(0, 40, 700), r.integers(0, 40, 300)
    bounds = [0, 250, 700]
    ol, orr = O.join_sortmerge(lk, rk, bounds)
    want = sorted((i, j) for i in range(700) for j in range(300) if lk[i] == rk[j])
    assert sorted(zip(ol.tolist(), orr.tolist())) == want
    # order: per left segment, by left key, then left row, then right row
    for s in range(2):
        sel = (ol >= bounds[s]) & (ol < bounds[s + 1])
        seq = list(zip(lk[ol[sel]], ol[sel], orr[sel]))
        assert seq == sorted(seq)


@pytest.mark.parametrize("nkeys", [1, 2])
def test_semi_anti_brute_force(nkeys):
    """Semi = the rows of a side with some equal-keyed row on the other side, anti = the
    rest; left rows in row order, right rows in the stable order of the right's first key."""
    r = np.random.default_rng(21 + nkeys)
    for trial in range(30):
        nl, nr = int(r.integers(0, 60)), int(r.integers(0, 60))
        lk = [r.integers(0, 12, nl) for _ in range(nkeys)]
        rk = [r.integers(0, 12, nr) for _ in range(nkeys)]
        ltup = [tuple(int(k[i]) for k in lk) for i in range(nl)]
        rtup = [tuple(int(k[i]) for k in rk) for i in range(nr)]
        rset, lset = set(rtup), set(ltup)
        bounds = [0, nl // 3, nl] if trial % 2 else None
        want_l = [i for i in range(nl) if ltup[i] in rset]
        want_r_rows = sorted(range(nr), key=lambda i: (int(rk[0][i]), i))
        assert list(O.semi_anti(lk, rk, "1", bounds)) == want_l
        assert list(O.semi_anti(lk, rk, "3", bounds)) == [i for i in range(nl) if ltup[i] not in rset]
        assert list(O.semi_anti(lk, rk, "2", bounds)) == [i for i in want_r_rows if rtup[i] in lset]
        assert list(O.semi_anti(lk, rk, "4", bounds)) == [i for i in want_r_rows if rtup[i] not in lset]


@pytest.mark.parametrize("width", [1, 3, 8, 10, 101, 255])
def test_str_sort_python_sorted(width):
    """char ORDER BY key: Python's stable sorted() on the NUL-padded bytes (unsigned), from a
    shuffled starting order; descending keeps ties in their order too."""
    r = np.random.default_rng(width)
    alpha = ["", "A", "AB", "ABC", "B", "a", "\x7f", "\u00e9", "Z Z", "AB C"]
    vals = [alpha[i] for i in r.integers(0, len(alpha), 300)]
    start = r.permutation(300).astype(np.uint32)
    padded = [v.encode()[:width].ljust(width, b"\0") for v in vals]
    asc = sorted(start.tolist(), key=lambda i: padded[i])
    assert list(O.str_sort(vals, width, start)) == asc
    neg = sorted(start.tolist(), key=lambda i: bytes(255 - b for b in padded[i]))
    assert list(O.str_sort(vals, width, start, desc=True)) == neg


def test_str_sort_unsupported_width_keeps_order():
    perm = np.array([3, 1, 2, 0], np.uint32)
    assert list(O.str_sort(["d", "c", "b", "a"], 120, perm)) == [3, 1, 2, 0]


def test_order_python_sorted():
    r = np.random.default_rng(9)
    a, b = r.integers(0, 10, 2000), r.integers(0, 5, 2000)
    perm = O.order([a, b], [1, 0])
    want = sorted(range(2000), key=lambda i: (-a[i], b[i]))  # Python's sort is stable
    assert list(perm) == want


def test_parallel_cpu_q3_matches_oracle():
    """The Thrust/OpenMP q3 run used as the benchmark's CPU baseline (and as the checker of
    the full-size GPU answer) returns the oracle's rows and pass counts."""
    import q3_reference as Q3

    t = Q3.tables(31, 6000, 60000)
    for seg, date in [("BUILDING", 19950315), ("FURNITURE", 19950310), ("AUTOMOBILE", 19930101)]:
        want, _, stats = Q3.run(t, segment=seg, date=date)
        got, used, st = Q3.run_cpu_parallel(t, 4, segment=seg, date=date, with_stats=True)
        assert got == want and used == 4 and st == stats


def test_parallel_cpu_flat_matches_oracle():
    """The flat filter + GROUP BY CPU baseline returns the oracle's groups in key order."""
    import q3_reference as Q3

    cols = O.gen_flat(5, 200_000)
    sel = cols[1] < Q3.DATE
    rev = cols[2][sel] * (100 - cols[3][sel])
    wk, wv = O.groupby([cols[0][sel]], [rev, rev], [L.AGG_SUM, L.AGG_COUNT])
    k, s, c = Q3.run_flat_cpu(cols, threads=4)
    assert np.array_equal(k, wk[0]) and np.array_equal(s, wv[0]) and np.array_equal(c, wv[1])
    assert Q3.run_flat_cpu(cols, want=False) == len(wk[0])


# ---- compaction (count + copy_if, cm.cu:3648-3650) ----------------------------------------
@pytest.mark.parametrize("density", [0.0, 0.003, 0.5, 0.97, 1.0])
def test_compact_matches_flatnonzero(density):
    r = np.random.default_rng(int(density * 1000) + 5)
    for n in (1, 63, 64, 65, 4095, 4096, 100_003):
        mask = r.random(n) < density
        for row0 in (0, 1, 1_000_000, (1 << 32) - n - 1):
            got = O.compact(mask, row0)
            want = (np.flatnonzero(mask) + row0).astype(np.uint32)
            assert got.dtype == np.uint32 and np.array_equal(got, want), (n, row0)


def test_compact_edge_masks():
    assert len(O.compact(np.zeros(0, np.uint8), 7)) == 0
    assert len(O.compact(np.zeros(1000, np.uint8), 7)) == 0
    assert np.array_equal(O.compact(np.ones(1000, np.uint8), 7), np.arange(7, 1007, dtype=np.uint32))
    last = np.zeros(1000, np.uint8)
    last[-1] = 1
    assert list(O.compact(last, 0)) == [999]  # the reference's star_join drops this row
    first = np.zeros(1000, np.uint8)
    first[0] = 1
    assert list(O.compact(first, 3)) == [3]


# ---- FP64 Kahan sum: the checker of FP64 aggregates ----------------------------------------
def test_kahan_sum_within_error_bound_of_fsum():
    import math

    u = 2.0 ** -53
    r = np.random.default_rng(17)
    cases = [
        r.uniform(-1, 1, 10_000) * 10.0 ** r.integers(-8, 9, 10_000),  # mixed magnitudes
        np.concatenate([[1e16], r.uniform(0, 1, 5000), [-1e16]]),  # cancellation
        r.uniform(900, 105000, 100_000).round(2) * (1 - r.integers(0, 11, 100_000) / 100),
    ]
    for v in cases:
        got, want = O.kahan_sum(v), math.fsum(v)
        bound = (2 * u + len(v) * u * u) * float(np.abs(v).sum())
        assert abs(got - want) <= bound, (got, want, bound)


def test_kahan_sum_beats_naive_where_naive_fails():
    import math

    # 1 + 10^6 * 1e-16: a running double sum never leaves 1.0, the compensated one does
    v = np.concatenate([[1.0], np.full(1_000_000, 1e-16)])
    naive = 0.0
    for x in v[:2000]:
        naive += x
    assert naive == 1.0  # the naive sum stalls
    got, want = O.kahan_sum(v), math.fsum(v)
    assert abs(want - (1.0 + 1e-10)) < 1e-20
    assert abs(got - want) <= 4 * 2.0 ** -53 * want
    assert abs(got - 1.0) > 9e-11  # compensation kept the small terms


def test_kahan_sum_exact_cases():
    assert O.kahan_sum(np.zeros(0)) == 0.0
    assert O.kahan_sum(np.arange(1, 100_001, dtype=np.float64)) == 100_000 * 100_001 / 2
    r = np.random.default_rng(3)
    x = r.normal(size=5000) * 1e8
    v = np.concatenate([x, -x, [0.125]])
    r.shuffle(v)
    assert abs(O.kahan_sum(v) - 0.125) <= 2 * 2.0 ** -53 * float(np.abs(v).sum())


# ---- f64 (type 1) operands: CudaSet's float overloads (cm.cu:2178-2437, 2440-2773) ---------
def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


FCODES = {1: lambda x, y: (y - x) > 1e-8, 2: lambda x, y: (x - y) > 1e-8,
          3: lambda x, y: ((x - y) > 1e-8) | ((x - y) < -1e-8),
          4: lambda x, y: ((x - y) < 1e-8) & ((x - y) > -1e-8),
          5: lambda x, y: ((y - x) > 1e-8) | (((x - y) < 1e-8) & ((x - y) > -1e-8)),
          6: lambda x, y: ((x - y) > 1e-8) | (((x - y) < 1e-8) & ((x - y) > -1e-8))}


def test_f64_compare_closed_form():
    """EPSILON comparisons (f_less .. f_not_equal_to, cm.cu:81-134) of an f64 column with a
    FLOAT, with another f64 column, with an int column (plain cast, long_to_float_type) and
    with a decimal NUMBER, written out with numpy; values sit inside, on and outside the
    1e-8 band around the literal."""
    r = np.random.default_rng(31)
    lit = 2.5
    f = lit + r.choice([0.0, 5e-9, -5e-9, 1e-8, -1e-8, 2e-8, -2e-8, 1.0, -1.0], 500)
    g = f + r.choice([0.0, 3e-9, -3e-9, 3e-8], 500)
    k = r.integers(-5, 6, 500)  # int column (decimal scale 2: 2.5 is stored as 250)
    cols = [_bits(f), _bits(g), k]
    dec = [-1, -1, 2]
    FL = (12, 0, int(_bits([lit])[0]))
    for code, fn in FCODES.items():
        # column OP FLOAT, FLOAT OP column (the left operand is the deeper one)
        assert np.array_equal(O.filter_mask([(1, 0, 0), FL, (4, code, 0)], cols, dec), fn(f, lit))
        assert np.array_equal(O.filter_mask([FL, (1, 0, 0), (4, code, 0)], cols, dec), fn(lit, f))
        assert np.array_equal(O.filter_mask([(1, 0, 0), (1, 1, 0), (4, code, 0)], cols, dec), fn(f, g))
        # an int column converts by a plain cast of the stored integer
        assert np.array_equal(O.filter_mask([(1, 2, 0), (1, 0, 0), (4, code, 0)], cols, dec),
                              fn(k.astype(np.float64), f))
        # a NUMBER by its decimal value: 25 at precision 1 is 2.5
        assert np.array_equal(O.filter_mask([(1, 0, 0), (2, 1, 25), (4, code, 0)], cols, dec), fn(f, 2.5))


def test_f64_arithmetic_closed_form():
    """Double arithmetic in the stack's order (left = deeper operand), bit-equal to numpy's
    IEEE operations, with FLOAT FLOAT folded."""
    r = np.random.default_rng(32)
    f = r.normal(0, 100, 300)
    k = r.integers(-1000, 1000, 300)
    cols = [_bits(f), k]
    dec = [-1, 2]
    FL = lambda x: (12, 0, int(_bits([x])[0]))  # noqa: E731
    cases = [
        ([(1, 0, 0), FL(1.5), (7, 0, 0)], f + 1.5),
        ([FL(1.5), (1, 0, 0), (8, 0, 0)], 1.5 - f),
        ([(1, 0, 0), (1, 1, 0), (9, 0, 0)], f * k.astype(np.float64)),
        ([(1, 1, 0), (1, 0, 0), (10, 0, 0)], k.astype(np.float64) / f),
        ([(1, 0, 0), (2, 2, 125), (8, 0, 0)], f - 1.25),
        ([(1, 0, 0), FL(2.0), (9, 0, 0), (1, 1, 0), (7, 0, 0)], f * 2.0 + k.astype(np.float64)),
        ([FL(1.5), FL(2.25), (9, 0, 0)], np.full(300, 1.5 * 2.25)),
    ]
    for ops, want in cases:
        got, prec = O.eval_expr(ops, cols, dec)
        assert prec == -1
        assert np.array_equal(got, _bits(want)), ops
    # an int-only expression keeps the int path and its scale
    got, prec = O.eval_expr([(1, 1, 0), (2, 0, 3), (9, 0, 0)], cols, dec)
    assert prec == 2 and np.array_equal(got, k * 3)
