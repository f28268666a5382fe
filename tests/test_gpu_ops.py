"""GPU parity tests: every operator through the C ABI against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
import q3_reference as Q3
from alenka_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.int64)).to(ctx.device)


def host(t):
    return t.cpu().numpy()


# ---- generators --------------------------------------------------------------------------
def test_generators_match_oracle(ctx):
    from alenka_amd import device as D

    for got, want in zip(D.gen_flat(ctx, 7, 5000), O.gen_flat(7, 5000)):
        assert np.array_equal(host(got), want)
    for got, want in zip(D.gen_dim(ctx, 7, 3000), O.gen_dim(7, 3000)):
        assert np.array_equal(host(got), want)
    for got, want in zip(D.gen_orders(ctx, 7, 4001, 3000), O.gen_orders(7, 4001, 3000)):
        assert np.array_equal(host(got), want)
    for got, want in zip(D.gen_fact(ctx, 7, 4001), O.gen_fact(7, 4001)):
        assert np.array_equal(host(got), want)
    code_of, codes = D.dict_first_seen(ctx, D.gen_dim(ctx, 7, 3000)[1], 5)
    w_code_of, w_codes = O.dict_first_seen(O.gen_dim(7, 3000)[1], 5)
    assert code_of == list(w_code_of)
    assert np.array_equal(host(codes), w_codes)


# ---- codec -------------------------------------------------------------------------------
CODEC_CASES = [
    ("small", lambda r, n: r.integers(0, 11, n)),
    ("dates", lambda r, n: r.integers(19920101, 19981231, n)),
    ("negative", lambda r, n: r.integers(-(1 << 40), 1 << 20, n)),
    ("constant", lambda r, n: np.full(n, -5)),
    ("full64", lambda r, n: r.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)),
    ("bits33", lambda r, n: r.integers(0, 1 << 33, n)),
]


@pytest.mark.parametrize("name,gen", CODEC_CASES)
@pytest.mark.parametrize("n,seg_rows", [(1, 1000), (10007, 1000), (70000, 65536)])
def test_encode_decode(ctx, name, gen, n, seg_rows):
    from alenka_amd import device as D

    r = np.random.default_rng(n + seg_rows)
    v = np.asarray(gen(r, n), np.int64)
    col = D.Column.encode(ctx, dev(ctx, v), seg_rows)
    assert np.array_equal(host(col.decode(ctx)), v)
    words = host(col.words).view(np.uint32)
    for s in range(col.nsegs):
        sg = col.hsegs[s]
        part = v[s * seg_rows: s * seg_rows + sg.nrows]
        assert sg.vmin == part.min() and sg.vmax == part.max()
        assert sg.bits == O.load().orc_bits_for_range(int(part.min()), int(part.max()))
        if sg.bits:
            nw = (sg.nrows * sg.bits + 31) // 32
            want = O.pack(part, sg.base, sg.bits)[:nw]
            assert np.array_equal(words[sg.word_off: sg.word_off + nw], want)
    rows = np.sort(r.choice(n, size=min(n, 333), replace=False)).astype(np.int32)
    got = D.gather(ctx, col, torch.as_tensor(rows).to(ctx.device))
    assert np.array_equal(host(got), v[rows])


# ---- scan / filter -----------------------------------------------------------------------
NAME, NUM, CMP, AND, OR = L.RPN_NAME, L.RPN_NUMBER, L.RPN_CMP, L.RPN_AND, L.RPN_OR
PREDICATES = [
    ("date_lt", [(NAME, 1, 0), (NUM, 0, 19950315), (CMP, 1, 0)]),
    ("range_and", [(NAME, 1, 0), (NUM, 0, 19940101), (CMP, 6, 0), (NAME, 1, 0), (NUM, 0, 19950101),
                   (CMP, 1, 0), (AND, 0, 0), (NAME, 3, 0), (NUM, 0, 5), (CMP, 5, 0), (AND, 0, 0)]),
    ("or", [(NAME, 0, 0), (NUM, 0, 1000), (CMP, 1, 0), (NAME, 3, 0), (NUM, 0, 10), (CMP, 4, 0), (OR, 0, 0)]),
    ("neq", [(NAME, 3, 0), (NUM, 0, 0), (CMP, 3, 0)]),
    ("decimal_const", [(NAME, 2, 0), (L.RPN_NUMBER, 2, 50000), (CMP, 2, 0)]),  # price > 500.00
    ("reverse", [(NUM, 0, 19950315), (NAME, 1, 0), (CMP, 2, 0)]),  # 19950315 > date
    ("arith", [(NAME, 2, 0), (NUM, 0, 1), (NAME, 3, 0), (L.RPN_MINUS, 0, 0), (L.RPN_MUL, 0, 0),
               (NUM, 0, 500000000), (CMP, 2, 0)]),
    ("col_col", [(NAME, 3, 0), (NAME, 0, 0), (CMP, 1, 0)]),  # disc*100 < key (scales as cm.cu:2299)
    ("const_const", [(NUM, 0, 3), (NUM, 0, 4), (CMP, 1, 0)]),
]


@pytest.mark.parametrize("name,ops", PREDICATES)
@pytest.mark.parametrize("zone_maps", [False, True])
def test_scan_filter(ctx, name, ops, zone_maps):
    from alenka_amd import device as D

    n, seg_rows = 50000, 8192
    cols = O.gen_flat(11, n)
    # sort a copy of the date column by segment so zone maps classify segments A / N / R
    cols[1] = np.sort(cols[1])
    dec = [0, 0, 2, 2]
    dcols = [D.Column.encode(ctx, dev(ctx, c), seg_rows, decimals=d) for c, d in zip(cols, dec)]
    prog, _ = L.compile_rpn(ops, dec)
    modes = None
    if zone_maps:
        modes = []
        for s in range(dcols[0].nsegs):
            mins = [c.hsegs[s].vmin for c in dcols]
            maxs = [c.hsegs[s].vmax for c in dcols]
            m = L.zone_map_check(ops, mins, maxs, dec)
            assert m == O.zone_map_check(ops, mins, maxs, dec)
            modes.append(m)
    got = host(D.scan_filter(ctx, dcols, prog, modes)).astype(np.uint32)
    if modes is None:
        want = O.compact(O.filter_mask(ops, cols, dec))
    else:
        # the reference's filter_op: 'A' keeps the whole segment, 'N' none, 'R' evaluates
        parts = []
        for s, m in enumerate(modes):
            lo, hi = s * seg_rows, min(n, (s + 1) * seg_rows)
            if m == "A":
                parts.append(np.arange(lo, hi, dtype=np.uint32))
            elif m == "R":
                seg = [c[lo:hi] for c in cols]
                parts.append(O.compact(O.filter_mask(ops, seg, dec), lo))
        want = np.concatenate(parts) if parts else np.zeros(0, np.uint32)
    assert np.array_equal(got, want)


def test_eval_revenue(ctx):
    from alenka_amd import device as D

    n = 40000
    cols = O.gen_flat(3, n)
    dec = [0, 0, 2, 2]
    dcols = [D.Column.encode(ctx, dev(ctx, c), 4096, decimals=d) for c, d in zip(cols, dec)]
    ops = [(NAME, 2, 0), (NUM, 0, 1), (NAME, 3, 0), (L.RPN_MINUS, 0, 0), (L.RPN_MUL, 0, 0)]
    prog, prec = L.compile_rpn(ops, dec, mode=1)
    want, wprec = O.eval_expr(ops, cols, dec)
    assert prec == wprec == 4
    rows = np.arange(0, n, 3, dtype=np.int32)
    got = host(D.eval_rows(ctx, dcols, prog, torch.as_tensor(rows).to(ctx.device)))
    assert np.array_equal(got, want[rows])
    assert np.array_equal(want, cols[2] * (100 - cols[3]))  # closed form of the trace


# ---- group-by ----------------------------------------------------------------------------
def test_groupby_flat_query(ctx):
    """Flat 1M-row filter + GROUP BY key: SUM(price*(1-discount)), COUNT(key), one segment."""
    from alenka_amd import device as D

    n = 200000
    key, date, price, disc = O.gen_flat(42, n)
    sel = date < 19950315
    rev = price[sel] * (100 - disc[sel])
    k = key[sel]
    ok, ov = D.groupby(ctx, [dev(ctx, k)], [dev(ctx, rev), dev(ctx, rev)], [L.AGG_SUM, L.AGG_COUNT])
    wk, wv = O.groupby([k], [rev, rev], [L.AGG_SUM, L.AGG_COUNT])
    assert np.array_equal(host(ok[0]), wk[0])
    assert np.array_equal(host(ov[0]), wv[0])
    assert np.array_equal(host(ov[1]), wv[1])
    # FP64 check of the exact decimal sums: within 1e-12 of a Kahan-summed float oracle
    g = {int(a): i for i, a in enumerate(wk[0][:50])}
    for kk, i in g.items():
        m = k == kk
        f = O.kahan_sum((price[sel][m] / 100.0) * (1.0 - disc[sel][m] / 100.0))
        assert abs(wv[0][i] / 1e4 - f) <= 1e-12 * abs(f)


@pytest.mark.parametrize("nkeys", [1, 2, 3])
def test_groupby_hash_order(ctx, nkeys):
    """Multi-segment output order: ascending MurmurHash64A of the key tuple (merge.cu)."""
    from alenka_amd import device as D

    r = np.random.default_rng(nkeys)
    n = 30000
    keys = [r.integers(0, 300, n), r.integers(-50, 50, n), r.integers(19920101, 19920131, n)][:nkeys]
    v = r.integers(-1000, 1000, n)
    bounds = [0, 10000, 20000, n]
    aggs = [L.AGG_SUM, L.AGG_COUNT, L.AGG_MIN, L.AGG_MAX]
    ok, ov = D.groupby(ctx, [dev(ctx, x) for x in keys], [dev(ctx, v)] * 4, aggs, order=L.ORDER_HASH,
                       seed=100)
    wk, wv = O.groupby(keys, [v] * 4, aggs, seg_bounds=bounds, seed=100)
    for a, b in zip(ok, wk):
        assert np.array_equal(host(a), b)
    for a, b in zip(ov, wv):
        assert np.array_equal(host(a), b)


def test_groupby_wide_keys(ctx):
    """Key tuples wider than 63 bits take the sort-based path."""
    from alenka_amd import device as D

    r = np.random.default_rng(5)
    n = 20000
    k1 = r.integers(-(1 << 62), 1 << 62, 200)[r.integers(0, 200, n)]
    k2 = r.integers(0, 1 << 20, 50)[r.integers(0, 50, n)]
    v = r.integers(0, 1 << 30, n)
    for order in (L.ORDER_KEYS, L.ORDER_HASH):
        ok, ov = D.groupby(ctx, [dev(ctx, k1), dev(ctx, k2)], [dev(ctx, v), dev(ctx, v)],
                           [L.AGG_SUM, L.AGG_COUNT], order=order)
        wk, wv = O.groupby([k1, k2], [v, v], [L.AGG_SUM, L.AGG_COUNT],
                           seg_bounds=[0, 1, n] if order == L.ORDER_HASH else None)
        for a, b in zip(ok + ov, wk + wv):
            assert np.array_equal(host(a), b)


def _gb_check(ctx, keys, vals, aggs, order=None, bounds=None):
    from alenka_amd import device as D

    order = L.ORDER_KEYS if order is None else order
    ok, ov = D.groupby(ctx, [dev(ctx, x) for x in keys], [dev(ctx, v) for v in vals], aggs, order=order, seed=100)
    wk, wv = O.groupby(keys, vals, aggs, seg_bounds=bounds, seed=100)
    for a, b in zip(ok + ov, wk + wv):
        assert np.array_equal(host(a), b)
    return len(wk[0])


AGGS4 = [L.AGG_SUM, L.AGG_COUNT, L.AGG_MIN, L.AGG_MAX]


@pytest.mark.parametrize("n", [1, 1000, 3_000_007])
def test_groupby_no_keys(ctx, n):
    """No GROUP BY: one group, a reduction over all rows (reduce, select.cu:549), against
    numpy's sum / min / max / len; f64 SUM within 1e-12 of math.fsum, f64 MIN / MAX exact."""
    import math

    from alenka_amd import device as D

    r = np.random.default_rng(n)
    v = r.integers(-(1 << 40), 1 << 40, n)
    ok, ov = D.groupby(ctx, [], [dev(ctx, v)] * 4, AGGS4)
    assert ok == [] and all(x.numel() == 1 for x in ov)
    assert [int(host(x)[0]) for x in ov] == [int(v.sum()), n, int(v.min()), int(v.max())]
    f = r.standard_normal(n) * 1e6
    fb = torch.as_tensor(f.view(np.int64)).to(ctx.device)
    _, fv = D.groupby(ctx, [], [fb] * 3, [L.AGG_FSUM, L.AGG_FMIN, L.AGG_FMAX])
    got = [host(x).view(np.float64)[0] for x in fv]
    want = math.fsum(f)
    assert abs(got[0] - want) <= 1e-12 * max(1.0, np.abs(f).sum())
    assert got[1] == f.min() and got[2] == f.max()


def test_groupby_constant_key(ctx):
    """Keys that never vary: one group with that key."""
    n = 500_001
    r = np.random.default_rng(7)
    assert _gb_check(ctx, [np.full(n, -42, np.int64)], [r.integers(-100, 100, n)] * 4, AGGS4) == 1


@pytest.mark.parametrize("case", ["g10", "g1000_2keys", "dense_2e20", "zipf_sparse", "retry_3m", "clustered"])
def test_groupby_paths(ctx, case):
    """The group-by paths against the oracle: direct-indexed LDS slots (10 groups), hashed
    LDS slots over direct-indexed globals (1000 groups of 2 keys), direct-indexed globals
    (2^20 codes), the hash table with hot keys (Zipf over 40-bit keys), more than 2^20
    groups (the first table overflows; the rows are aggregated again in a 2n table), and
    sorted input (runs of equal keys in adjacent lanes); key order and hash order."""
    r = np.random.default_rng(len(case))
    if case == "g10":
        n = 2_000_003
        keys = [r.integers(0, 10, n) * 7 + 19920101]
    elif case == "g1000_2keys":
        n = 1_000_000
        keys = [r.integers(0, 40, n), r.integers(-12, 13, n)]
    elif case == "dense_2e20":
        n = 1 << 20
        keys = [r.integers(0, 1 << 20, n)]
    elif case == "zipf_sparse":
        n = 2_000_000
        z = np.minimum(r.zipf(1.2, n), 1 << 20)
        keys = [z * 1_000_003 + 5]
    elif case == "retry_3m":
        n = 3_000_000
        keys = [r.integers(0, 1 << 40, n)]
    else:
        n = 1_500_000
        keys = [np.sort(r.integers(0, 100_000, n))]
    v = r.integers(-(1 << 30), 1 << 30, n)
    g = _gb_check(ctx, keys, [v] * 4, AGGS4)
    assert g > 1
    bounds = [0, n // 3, 2 * n // 3, n]
    _gb_check(ctx, keys, [v, v], [L.AGG_SUM, L.AGG_COUNT], order=L.ORDER_HASH, bounds=bounds)


# ---- join --------------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["dense", "dups", "sparse", "wide", "dense_big"])
def test_hash_join(ctx, case):
    """dense: unique keys over a range of <= 64 per build row (bitmap + rank index); dups:
    repeated keys (sorted runs in the hash table); sparse / wide: unique keys over a range
    too wide for the bitmap (hash table), the second beyond 2^32; dense_big: 2M probe rows."""
    from alenka_amd import device as D

    r = np.random.default_rng(9 + len(case))
    nb, npb = 20000, 50000
    dups = case == "dups"
    if dups:
        build = r.integers(0, 5000, nb)
    elif case == "sparse":
        build = r.permutation(1 << 24)[:nb].astype(np.int64) * 7 - 300000
    elif case == "wide":
        build = r.integers(-(1 << 40), 1 << 40, nb)
        build = np.unique(build)[:nb]
        nb = build.size
    else:
        if case == "dense_big":
            nb, npb = 500000, 2000000
        build = r.permutation(1 << 20 if nb < 100000 else 1 << 22)[:nb].astype(np.int64) - 300000
    lo, hi = int(build.min()) - 1000, int(build.max()) + 1000
    probe = np.concatenate([build[r.integers(0, nb, npb // 2)], r.integers(lo, hi, npb - npb // 2)])
    probe = r.permutation(probe)
    pi, bi = D.hash_join(ctx, dev(ctx, build), dev(ctx, probe))
    pi, bi = host(pi).astype(np.uint32), host(bi).astype(np.uint32)
    # expected: probe-row order, then ascending build row
    if not dups:  # unique build keys: one pair per probe row whose key is a build key
        pos = {int(b): j for j, b in enumerate(build)}
        want_p = np.array([i for i, p in enumerate(probe) if int(p) in pos], np.uint32)
        want_b = np.array([pos[int(probe[i])] for i in want_p], np.uint32)
        assert np.array_equal(pi, want_p) and np.array_equal(bi, want_b)
    else:
        exp = []
        where = {}
        for j, b in enumerate(build):
            where.setdefault(int(b), []).append(j)
        for i, p in enumerate(probe):
            for j in where.get(int(p), []):
                exp.append((i, j))
        exp = np.array(exp, np.uint32).reshape(-1, 2)
        assert np.array_equal(np.stack([pi, bi], 1), exp)
    if case != "dense_big":  # same multiset as the reference's sort-merge join
        ol, orr = O.join_sortmerge(probe, build)
        assert sorted(zip(ol.tolist(), orr.tolist())) == sorted(zip(pi.tolist(), bi.tolist()))


@pytest.mark.parametrize("width", [3, 8, 25])
def test_like_dict_matches_oracle(ctx, width):
    """LIKE over a dictionary on the device against the oracle's restatement of gpu_regex
    (filter.cu:67-129), with entries that fill the width (where the position budget binds)."""
    from alenka_amd import device as D

    r = np.random.default_rng(width)
    alpha = list("ab.")
    entries = ["".join(r.choice(alpha, size=int(r.integers(0, width + 1)))) for _ in range(777)]
    entries += ["BUILDING"[:width], "aaab"[:width], ""]
    pats = ["%", "", "a%", "%b", "%a%b%", "a?b", "?%", "%.%", "%ab", "a.b%", "%%a", "b?%?a"]
    for pat in pats:
        bits = D.like_dict(ctx, entries, width, pat)
        got = [(int(bits[i >> 5]) >> (i & 31)) & 1 == 1 for i in range(len(entries))]
        want = [O.like(e, width, pat) for e in entries]
        assert got == want, pat


@pytest.mark.parametrize("nrows,nidx", [(1, 0), (31, 5), (32, 64), (100003, 250000), (5000, 0)])
def test_join_rows(ctx, nrows, nidx):
    """Semi/anti reduction: rows that occur among a join side's pair indices (with
    duplicates), or do not, in ascending order - the oracle's sort + unique and set
    difference (operators.cu:1357-1399)."""
    from alenka_amd import device as D

    r = np.random.default_rng(nrows + nidx)
    idx = r.integers(0, nrows, nidx).astype(np.int32)
    t = torch.as_tensor(idx if nidx else np.zeros(1, np.int32)).to(ctx.device)[:nidx]
    semi = np.unique(idx)
    anti = np.setdiff1d(np.arange(nrows), semi)
    assert np.array_equal(host(D.join_rows(ctx, t, nrows, False)), semi.astype(np.int32))
    assert np.array_equal(host(D.join_rows(ctx, t, nrows, True)), anti.astype(np.int32))
    if len(anti) > 1:
        with pytest.raises(L.AlnError) as e:
            D.join_rows(ctx, t, nrows, True, cap=len(anti) - 1)
        assert e.value.code == L.ALN_E_CAPACITY


# ---- sort / top-k ------------------------------------------------------------------------
@pytest.mark.parametrize("wide", [False, True])
def test_sort_and_topk(ctx, wide):
    from alenka_amd import device as D

    r = np.random.default_rng(21 + wide)
    n = 60000
    if wide:
        keys = [r.integers(-(1 << 62), 1 << 62, n), r.integers(0, 1 << 40, n)]
    else:
        keys = [r.integers(0, 1000, n), r.integers(19920101, 19981231, n)]
    desc = [1, 0]
    perm = host(D.sort_keys(ctx, [dev(ctx, k) for k in keys], desc)).astype(np.uint32)
    want = O.order(keys, desc)
    assert np.array_equal(perm, want)
    for k in (1, 10, 100):
        top = host(D.topk(ctx, [dev(ctx, x) for x in keys], desc, k)).astype(np.uint32)
        assert np.array_equal(top, want[:k])


def test_radix_sort_pairs(ctx):
    from alenka_amd import device as D

    r = np.random.default_rng(3)
    n = 100003
    k = r.integers(0, 1 << 63, n, dtype=np.int64).astype(np.uint64) | np.uint64(1 << 63)
    kt = torch.as_tensor(k.view(np.int64)).to(ctx.device)
    vt = torch.arange(n, dtype=torch.int32, device=ctx.device)
    D.radix_sort_pairs(ctx, kt, vt, 64)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(host(kt).view(np.uint64), k[order])
    assert np.array_equal(host(vt), order.astype(np.int32))


RADIX_CASES = [(1, 64, "u"), (2, 64, "u"), (4095, 64, "u"), (4096, 17, "u"), (4097, 8, "u"), (100003, 5, "u"),
               ((1 << 20) + 3, 64, "u"), ((1 << 20) + 3, 40, "few"), (1 << 20, 64, "const"), ((1 << 20) + 5, 64, "hi"),
               (300001, 64, "hi")]


@pytest.mark.parametrize("n,bits,kind", RADIX_CASES)
def test_radix_sort_pairs_shapes(ctx, n, bits, kind):
    """Stable sort by the low `bits` bits: tile-boundary sizes, one-pass sorts (sorted through
    a temporary), digits that every item shares (skipped passes at 2^20 items and up), all
    keys equal, and bits above `bits` that must not take part."""
    from alenka_amd import device as D

    r = np.random.default_rng(n + bits)
    if kind == "u":
        k = r.integers(0, 1 << 63, n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + r.integers(0, 2, n).astype(np.uint64)
    elif kind == "few":  # 16 values, digits 0-1 and 3-4 shared by every key
        k = r.integers(0, 16, n).astype(np.uint64) << np.uint64(20)
    elif kind == "const":
        k = np.full(n, 0x0123456789ABCDEF, np.uint64)
    else:  # only the top digit varies
        k = r.integers(0, 256, n).astype(np.uint64) << np.uint64(56)
    kt = torch.as_tensor(k.view(np.int64)).to(ctx.device)
    vt = torch.arange(n, dtype=torch.int32, device=ctx.device)
    D.radix_sort_pairs(ctx, kt, vt, bits)
    low = k & np.uint64((1 << bits) - 1 if bits < 64 else 0xFFFFFFFFFFFFFFFF)
    order = np.argsort(low, kind="stable")
    assert np.array_equal(host(vt), order.astype(np.int32))
    assert np.array_equal(host(kt).view(np.uint64), k[order])


@pytest.mark.parametrize("n,shape", [(60000, "wide_small_last"), (60000, "small_first_wide"),
                                     ((1 << 20) + 7, "wide_small_last"), ((1 << 20) + 7, "three")])
def test_sort_keys_lsd_shapes(ctx, n, shape):
    """ORDER BY keys whose codes exceed 64 bits: one sort per key, last key first, each one
    reading its key through the permutation so far; a key of at most 8 bits is a one-pass
    sort whose permutation is both read and written."""
    from alenka_amd import device as D

    r = np.random.default_rng(n + len(shape))
    full = r.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)
    small = r.integers(-100, 100, n)
    if shape == "wide_small_last":
        keys, desc = [full, small], [0, 1]
    elif shape == "small_first_wide":
        keys, desc = [small, full], [1, 0]
    else:
        keys, desc = [small, full, r.integers(0, 1 << 40, n)], [0, 1, 1]
    perm = host(D.sort_keys(ctx, [dev(ctx, k) for k in keys], desc)).astype(np.uint32)
    assert np.array_equal(perm, O.order(keys, desc))


@pytest.mark.parametrize("case", ["uniform_desc", "u32_keys", "u28_keys", "tiny_buckets", "overflow", "dups", "two_keys"])
def test_sort_keys_bucket_path(ctx, case):
    """ORDER BY of 2^20 rows and more with at least four varying digits: two passes on the
    top two digits, then each bucket sorted in LDS; buckets too large for a block fall back
    to the digit passes.  Cases: uniform 64-bit keys DESC; 32- and 28-bit keys (many small
    buckets per block); >256 buckets in some blocks (two ordinal passes); buckets of 4096
    rows (fallback); 300 distinct values (fallback); two keys packed into one code."""
    from alenka_amd import device as D

    r = np.random.default_rng(len(case) * 7)
    desc = [0]
    if case == "uniform_desc":
        n = 3_000_001
        keys, desc = [r.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)], [1]
    elif case == "u32_keys":
        n = 2_000_003
        keys = [r.integers(0, 1 << 32, n)]
    elif case == "u28_keys":
        n = 1_500_000
        keys = [r.integers(0, 1 << 28, n) - (1 << 27)]
    elif case == "tiny_buckets":
        n = 1_100_000
        top = np.concatenate([r.integers(0, 60000, 100_000), 60000 + r.integers(0, 5536, n - 100_000)])
        keys = [(top << 24) | r.integers(0, 1 << 24, n)]
    elif case == "overflow":
        n = 4_000_000
        keys = [(r.integers(0, 4, n) << 40) | r.integers(0, 1 << 32, n)]
    elif case == "dups":
        n = 2_000_000
        keys = [r.integers(-(1 << 62), 1 << 62, 300)[r.integers(0, 300, n)]]
    else:
        n = 1_200_000
        keys, desc = [r.integers(0, 1 << 20, n), r.integers(-(1 << 30), 1 << 30, n)], [1, 0]
    perm = host(D.sort_keys(ctx, [dev(ctx, k) for k in keys], desc)).astype(np.uint32)
    assert np.array_equal(perm, O.order(keys, desc))


def test_hash_rows(ctx):
    from alenka_amd import device as D

    r = np.random.default_rng(4)
    a, b = r.integers(-(1 << 62), 1 << 62, 1000), r.integers(0, 100, 1000)
    h = host(D.hash_rows(ctx, [dev(ctx, a), dev(ctx, b)], 100)).view(np.uint64)
    for i in range(0, 1000, 37):
        tup = np.array([a[i], b[i]], np.int64).tobytes()
        assert int(h[i]) == O.murmur64a(tup, 100)


# ---- fused q3 pipeline -------------------------------------------------------------------
def encode_q3(ctx, t, seg_rows):
    from alenka_amd import device as D

    enc = lambda a, d=0: D.Column.encode(ctx, dev(ctx, a), seg_rows, decimals=d)
    dim = [enc(t["c_key"]), enc(t["c_seg"])]
    mid = [enc(t["o_key"]), enc(t["o_cust"]), enc(t["o_date"]), enc(t["o_prio"])]
    fact = [enc(t["l_okey"]), enc(t["l_price"], 2), enc(t["l_disc"], 2), enc(t["l_ship"])]
    return dim, mid, fact


def q3_args(t, dim, mid, fact):
    building = int(t["code_of"][Q3.SEGMENTS.index("BUILDING")])
    D_ = dict(cols=dim, ranges=[(1, building, building)])
    M_ = dict(cols=mid, ranges=[(2, -(1 << 63), Q3.DATE - 1)])
    F_ = dict(cols=fact, ranges=[(3, Q3.DATE + 1, (1 << 63) - 1)])
    spec = dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1)
    return D_, M_, F_, spec


@pytest.mark.parametrize("join", ["bitmap", "hash"])
@pytest.mark.parametrize("scale,seg_rows", [(0.02, 4096), (0.1, 65536)])
def test_star_q3(ctx, scale, seg_rows, join):
    from alenka_amd import device as D

    ndim, norders = int(150000 * scale), int(1500000 * scale)
    t = Q3.tables(1234, ndim, norders)
    top, groups, stats = Q3.run(t)
    dim, mid, fact = encode_q3(ctx, t, seg_rows)
    D_, M_, F_, spec = q3_args(t, dim, mid, fact)
    rows, g, st = D.star_agg(ctx, D_, M_, F_, spec, limit=10, want_groups=True, force_hash=join == "hash")
    assert rows == top
    if join == "bitmap":  # a prepared query gives the same answer on every run
        q = D.StarQuery(ctx, D_, M_, F_, spec, limit=10)
        for _ in range(2):
            r2, _, st2 = q.run()
            assert r2 == top and st2.fact_pass == stats["fact_pass"]
    assert st.dim_pass == stats["dim_pass"] and st.mid_pass == stats["mid_pass"]
    assert st.fact_pass == stats["fact_pass"] and st.groups == stats["groups"]
    assert np.array_equal(host(g), groups)


def star_expected(t, dr, mr, fr, limit=10, coef=(0, 1, 100, -1)):
    """Oracle answer of a star query with range lists (col, lo, hi) on the q3 tables: the same
    sort-merge joins and sort-based group-by as the q3 reference run, SUM((a0 + a1 price) *
    (b0 + b1 disc)) with coef = (a0, a1, b0, b1) (q3: price * (100 - disc))."""
    cols = dict(d=[t["c_key"], t["c_seg"]], m=[t["o_key"], t["o_cust"], t["o_date"], t["o_prio"]],
                f=[t["l_okey"], t["l_price"], t["l_disc"], t["l_ship"]])

    def ids(k, ranges):
        m = np.ones(len(cols[k][0]), bool)
        for c, lo, hi in ranges:
            m &= (cols[k][c] >= lo) & (cols[k][c] <= hi)
        return O.compact(m)

    c_ids, o_ids, l_ids = ids("d", dr), ids("m", mr), ids("f", fr)
    pl, _ = O.join_sortmerge(t["o_cust"][o_ids], t["c_key"][c_ids])
    olc1 = o_ids[pl]
    ql, qr = O.join_sortmerge(t["l_okey"][l_ids], t["o_key"][olc1])
    lrow, orow = l_ids[ql], olc1[qr]
    a0, a1, b0, b1 = coef
    rev = (a0 + a1 * t["l_price"][lrow]) * (b0 + b1 * t["l_disc"][lrow])
    gk, gv = O.groupby([t["o_key"][orow], t["o_date"][orow], t["o_prio"][orow]], [rev, rev],
                       [L.AGG_SUM, L.AGG_COUNT])
    perm = O.order([gv[0], gk[1]], [1, 0])
    top = [(int(gk[0][i]), int(gk[1][i]), int(gk[2][i]), int(gv[0][i]), int(gv[1][i])) for i in perm[:limit]]
    groups = np.stack([gk[0], gk[1], gk[2], gv[0], gv[1]], axis=1) if len(gk[0]) else np.zeros((0, 5), np.int64)
    return top, groups, len(c_ids), len(olc1), len(lrow)


@pytest.mark.parametrize("join", ["bitmap", "hash"])
@pytest.mark.parametrize("case", ["wide", "multi", "zone", "dense"])
def test_star_scan_shapes(ctx, case, join):
    """Segments the scans' fast decoder cannot take (codes wider than 31 bits, several ranges
    on a table) run the generic staged decoder, zone maps skip ('N') or pass ('A') whole
    segments, and with no filter every fact row joins (more joined rows per wave than its
    deferred slots hold, so the scan adds the rest itself); every shape gives the oracle's
    answer."""
    from alenka_amd import device as D

    t = Q3.tables(77, 3000, 30000)
    seg_rows = 4096
    enc = lambda a, d=0, fb=0: D.Column.encode(ctx, dev(ctx, a), seg_rows, decimals=d, force_bits=fb)
    wide = case == "wide"
    dim = [enc(t["c_key"], fb=36 if wide else 0), enc(t["c_seg"])]
    mid = [enc(t["o_key"], fb=40 if wide else 0), enc(t["o_cust"], fb=36 if wide else 0), enc(t["o_date"]),
           enc(t["o_prio"])]
    fact = [enc(t["l_okey"], fb=40 if wide else 0), enc(t["l_price"], 2), enc(t["l_disc"], 2), enc(t["l_ship"])]
    building = int(t["code_of"][Q3.SEGMENTS.index("BUILDING")])
    lo, hi = -(1 << 63), (1 << 63) - 1
    if case == "multi":
        dr = [(1, building, building), (0, lo, 2400)]
        mr = [(2, lo, Q3.DATE - 1), (2, 19930101, hi)]
        fr = [(3, Q3.DATE + 1, hi), (2, 2, 8)]
    elif case == "dense":  # every row of every table passes
        dr, mr, fr = [(1, lo, hi)], [(2, lo, hi)], [(3, lo, hi)]
    elif case == "zone":  # fact and mid ranges on the (clustered) order key itself
        kmid = int(np.median(t["o_key"]))
        dr = [(1, building, building)]
        mr = [(0, lo, kmid)]
        fr = [(0, int(t["l_okey"][len(t["l_okey"]) // 5]), kmid)]
    else:
        dr, mr, fr = [(1, building, building)], [(2, lo, Q3.DATE - 1)], [(3, Q3.DATE + 1, hi)]
    top, groups, dpass, mpass, fpass = star_expected(t, dr, mr, fr)
    spec = dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1)
    rows, g, st = D.star_agg(ctx, dict(cols=dim, ranges=dr), dict(cols=mid, ranges=mr), dict(cols=fact, ranges=fr),
                             spec, limit=10, want_groups=True, force_hash=join == "hash")
    assert rows == top
    assert (st.dim_pass, st.mid_pass, st.fact_pass) == (dpass, mpass, fpass)
    assert np.array_equal(host(g), groups)


@pytest.mark.parametrize("join", ["bitmap", "hash"])
def test_star_q3_orders_shuffled(ctx, join):
    """The orders table's rows in random order (the answer does not depend on it): every
    mid tile's passing keys then span most of the key range, so the top-k reads the rows of
    its candidate groups from every tile instead of the few a key-ordered table leaves."""
    from alenka_amd import device as D

    t = Q3.tables(4321, 3000, 30000)
    perm = np.random.default_rng(5).permutation(len(t["o_key"]))
    t2 = dict(t)
    for k in ("o_key", "o_cust", "o_date", "o_prio"):
        t2[k] = t[k][perm]
    top, groups, stats = Q3.run(t)
    assert Q3.run(t2)[0] == top
    dim, mid, fact = encode_q3(ctx, t2, 4096)
    D_, M_, F_, spec = q3_args(t2, dim, mid, fact)
    rows, g, st = D.star_agg(ctx, D_, M_, F_, spec, limit=10, want_groups=True, force_hash=join == "hash")
    assert rows == top
    assert (st.dim_pass, st.mid_pass, st.fact_pass) == (stats["dim_pass"], stats["mid_pass"], stats["fact_pass"])
    assert np.array_equal(host(g), groups)


@pytest.mark.parametrize("join", ["bitmap", "hash"])
@pytest.mark.parametrize("case", ["dense", "filtered"])
def test_star_topk_ties(ctx, case, join):
    """SUM(1) per joined row: a group's sum is its row count (1 to 7), so many groups tie at
    the top-k threshold and the order falls to (o_date, o_key).  Every row passing gives
    thousands of candidate groups at the threshold (the top-k then looks up every group's mid
    row); the q3 filters give fewer (each candidate's row found by its code)."""
    from alenka_amd import device as D

    t = Q3.tables(91, 3000, 30000)
    seg_rows = 4096
    dim, mid, fact = encode_q3(ctx, t, seg_rows)
    lo, hi = -(1 << 63), (1 << 63) - 1
    building = int(t["code_of"][Q3.SEGMENTS.index("BUILDING")])
    if case == "dense":
        dr, mr, fr = [(1, lo, hi)], [(2, lo, hi)], [(3, lo, hi)]
    else:
        dr, mr, fr = [(1, building, building)], [(2, lo, Q3.DATE - 1)], [(3, Q3.DATE + 1, hi)]
    top, groups, dpass, mpass, fpass = star_expected(t, dr, mr, fr, coef=(1, 0, 1, 0))
    kth = np.sort(groups[:, 3])[::-1][9]  # the 10th largest sum: groups at or above it are candidates
    assert int((groups[:, 3] >= kth).sum()) > (2048 if case == "dense" else 10)
    spec = dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=1, a1=0, b0=1, b1=0)
    for want in (False, True):  # the top-k alone, and with the group list
        rows, g, st = D.star_agg(ctx, dict(cols=dim, ranges=dr), dict(cols=mid, ranges=mr),
                                 dict(cols=fact, ranges=fr), spec, limit=10, want_groups=want,
                                 force_hash=join == "hash")
        assert rows == top
        assert (st.dim_pass, st.mid_pass, st.fact_pass) == (dpass, mpass, fpass)
        if want:
            assert np.array_equal(host(g), groups)


@pytest.mark.parametrize("join", ["bitmap", "hash"])
def test_star_ranges_reused_then_changed(ctx, join):
    """A call on the same key columns as the last one reuses their key ranges and checks them
    on the device.  Here the dimension key and the orders' foreign key are rewritten in place
    (same buffers and segment count, keys moved up by 100000, new zone maps): the reused
    ranges are stale (with them no key would fall in the dimension bitmap), so the query must
    run again with fresh ranges and give the oracle's answer for the new contents."""
    import ctypes

    import torch
    from alenka_amd import device as D

    t = Q3.tables(1234, 3000, 30000)
    seg_rows = 4096
    t2 = dict(t)
    t2["c_key"] = t["c_key"] + 100000
    t2["o_cust"] = t["o_cust"] + 100000
    enc = lambda a, d=0, fb=0: D.Column.encode(ctx, dev(ctx, a), seg_rows, decimals=d, force_bits=fb)
    dim = [enc(t["c_key"], fb=20), enc(t["c_seg"])]
    mid = [enc(t["o_key"]), enc(t["o_cust"], fb=20), enc(t["o_date"]), enc(t["o_prio"])]
    fact = [enc(t["l_okey"]), enc(t["l_price"], 2), enc(t["l_disc"], 2), enc(t["l_ship"])]
    D_, M_, F_, spec = q3_args(t, dim, mid, fact)
    force = join == "hash"
    for tab in (t, t2):
        if tab is t2:  # the new contents into the same buffers
            for old, new in ((dim[0], enc(t2["c_key"], fb=20)), (mid[1], enc(t2["o_cust"], fb=20))):
                assert old.words.numel() == new.words.numel() and old.nsegs == new.nsegs
                old.words.copy_(new.words)
                old.dsegs.copy_(new.dsegs)
                ctypes.memmove(old.hsegs, new.hsegs, ctypes.sizeof(old.hsegs))
            torch.cuda.synchronize()
        top, groups, stats = Q3.run(tab)
        for _ in range(2):  # the first call after the change, and one reusing its ranges
            rows, g, st = D.star_agg(ctx, D_, M_, F_, spec, limit=10, want_groups=True, force_hash=force)
            assert rows == top
            assert (st.dim_pass, st.mid_pass, st.fact_pass) == (stats["dim_pass"], stats["mid_pass"],
                                                                 stats["fact_pass"])
            assert np.array_equal(host(g), groups)


@pytest.mark.parametrize("scale,seg_rows", [(0.02, 4096), (0.1, 65536)])
def test_star_q3_batch(ctx, scale, seg_rows):
    """q3 with different substitution parameters (segment, date) answered by one pass of each
    table: every query equals its own oracle run and its own single-query star_agg."""
    from alenka_amd import device as D

    ndim, norders = int(150000 * scale), int(1500000 * scale)
    t = Q3.tables(4321, ndim, norders)
    dim, mid, fact = encode_q3(ctx, t, seg_rows)
    params = [("BUILDING", 19950315), ("MACHINERY", 19950301), ("AUTOMOBILE", 19950331),
              ("BUILDING", 19920201), ("HOUSEHOLD", 19980801)]
    qs = []
    for seg_name, date in params:
        code = int(t["code_of"][Q3.SEGMENTS.index(seg_name)])
        qs.append((dict(cols=dim, ranges=[(1, code, code)]), dict(cols=mid, ranges=[(2, -(1 << 63), date - 1)]),
                   dict(cols=fact, ranges=[(3, date + 1, (1 << 63) - 1)]),
                   dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1)))
    got = D.star_agg_batch(ctx, qs, limit=10)
    for (seg_name, date), q, (rows, st) in zip(params, qs, got):
        top, groups, stats = Q3.run(t, segment=seg_name, date=date)
        assert rows == top, (seg_name, date)
        assert st.dim_pass == stats["dim_pass"] and st.mid_pass == stats["mid_pass"]
        assert st.fact_pass == stats["fact_pass"]
        alone, _, st1 = D.star_agg(ctx, *q, limit=10)
        assert alone == rows
        assert (st1.mid_probe, st1.fact_probe) == (st.mid_probe, st.fact_probe)


@pytest.mark.parametrize("force_hash", [False, True])
def test_star_duplicate_keys_rejected(ctx, force_hash):
    """A build side with a repeated key is not a star join: the fused pipeline refuses it
    (ALN_E_UNSUPPORTED) so the caller falls back to the unfused operators."""
    from alenka_amd import device as D

    t = Q3.tables(55, 3000, 30000)
    o_key = t["o_key"].copy()
    passing = np.nonzero(t["o_date"] < Q3.DATE)[0]
    o_key[passing[1]] = o_key[passing[0]]  # two passing orders share a key
    t2 = dict(t, o_key=o_key)
    dim, mid, fact = encode_q3(ctx, t2, 4096)
    D_, M_, F_, spec = q3_args(t2, dim, mid, fact)
    # every dimension row passes, so both orders reach the mid bitmap
    D_ = dict(cols=dim, ranges=[])
    with pytest.raises(L.AlnError) as e:
        D.star_agg(ctx, D_, M_, F_, spec, limit=10, force_hash=force_hash)
    assert e.value.code == L.ALN_E_UNSUPPORTED


# ---- the benchmark's segment layout: 1,000,000-row segments ------------------------------
@pytest.mark.parametrize("join", ["bitmap", "hash"])
def test_star_q3_million_row_segments(ctx, join):
    """q3 at the 6M-row base size (dimension 150k, orders 1.5M, fact ~6M rows) in the
    benchmark's 1,000,000-row segments: six fact segments, each 244 full 4096-row tiles and
    a 576-row tail tile, and the resolve grid built from that shape; groups, pass counts and
    top-10 against the oracle, then three runs of a prepared query as the benchmark loop
    makes them."""
    from alenka_amd import device as D

    t = Q3.tables(2024, 150000, 1500000)
    top, groups, stats = Q3.run(t)
    assert len(t["l_okey"]) > 5_000_000 and stats["fact_pass"] > 10_000
    dim, mid, fact = encode_q3(ctx, t, 1_000_000)
    assert fact[0].nsegs == 6 and all(fact[0].hsegs[s].nrows % 4096 == 576 for s in range(5))
    D_, M_, F_, spec = q3_args(t, dim, mid, fact)
    rows, g, st = D.star_agg(ctx, D_, M_, F_, spec, limit=10, want_groups=True, force_hash=join == "hash")
    assert rows == top
    assert (st.dim_pass, st.mid_pass, st.fact_pass, st.groups) == (
        stats["dim_pass"], stats["mid_pass"], stats["fact_pass"], stats["groups"])
    assert np.array_equal(host(g), groups)
    if join == "bitmap":
        q = D.StarQuery(ctx, D_, M_, F_, spec, limit=10)
        for _ in range(3):
            r2, _, st2 = q.run()
            assert r2 == top and (st2.mid_pass, st2.fact_pass) == (stats["mid_pass"], stats["fact_pass"])


@pytest.mark.parametrize("name,ops", [PREDICATES[0], PREDICATES[1], PREDICATES[2], PREDICATES[6]])
@pytest.mark.parametrize("zone_maps", [False, True])
def test_scan_filter_million_row_segments(ctx, name, ops, zone_maps):
    """aln_scan_filter over 3.5 segments of 1,000,000 rows (the benchmark's segment size):
    exact row ids against the oracle's compaction, with and without zone-map skipping."""
    from alenka_amd import device as D

    n, seg_rows = 3_500_000, 1_000_000
    cols = O.gen_flat(12, n)
    cols[1] = np.sort(cols[1])  # date sorted, so zone maps give A / N / R segments
    dec = [0, 0, 2, 2]
    dcols = [D.Column.encode(ctx, dev(ctx, c), seg_rows, decimals=d) for c, d in zip(cols, dec)]
    assert dcols[0].nsegs == 4
    prog, _ = L.compile_rpn(ops, dec)
    modes = None
    if zone_maps:
        modes = [L.zone_map_check(ops, [c.hsegs[s].vmin for c in dcols], [c.hsegs[s].vmax for c in dcols], dec)
                 for s in range(4)]
    got = host(D.scan_filter(ctx, dcols, prog, modes)).astype(np.uint32)
    parts = []
    for s in range(4):
        lo, hi = s * seg_rows, min(n, (s + 1) * seg_rows)
        m = "R" if modes is None else modes[s]
        if m == "A":
            parts.append(np.arange(lo, hi, dtype=np.uint32))
        elif m == "R":
            parts.append(O.compact(O.filter_mask(ops, [c[lo:hi] for c in cols], dec), lo))
    want = np.concatenate(parts)
    assert np.array_equal(got, want)
    if modes is None:  # the whole-table mask gives the same ids
        assert np.array_equal(got, O.compact(O.filter_mask(ops, cols, dec)))
