"""f64 (type 1, d_columns_float) columns on the GPU: double-precision predicates and
expressions (CudaSet's float overloads, cm.cu:2178-2437, 2440-2773: EPSILON comparisons,
plain casts of int columns) through aln_scan_filter / aln_eval against the oracle bit for bit;
SUM / MIN / MAX of doubles against math.fsum and numpy; ORDER BY and GROUP BY on f64 keys
(update_permutation on d_columns_float, sorts.cu:17-86); and the operator layer: FLOAT
literals in scripts, f64 columns through FILTER, SELECT, GROUP BY, ORDER BY, joins (payload
gathers), STORE's "%.2f" text (cm.cu:1569) and binary segment files (comp_type 3)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NAME, NUM, CMP, AND, OR, FLT = L.RPN_NAME, L.RPN_NUMBER, L.RPN_CMP, L.RPN_AND, L.RPN_OR, L.RPN_FLOAT
ADD, MINUS, MUL, DIV = L.RPN_ADD, L.RPN_MINUS, L.RPN_MUL, L.RPN_DIV


def fb(x):
    return int(np.array([x], np.float64).view(np.int64)[0])


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def dev(ctx, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.int64)).to(ctx.device)


def table(ctx, n, seg_rows, seed):
    from alenka_amd import device as D

    r = np.random.default_rng(seed)
    f = np.round(r.normal(10, 30, n), 3) + r.choice([0, 5e-9, -5e-9, 2e-8], n)
    g = r.normal(0, 1e6, n)
    k = r.integers(-500, 500, n)
    d = r.integers(0, 10000, n)  # decimal(2)
    host = [bits(f), bits(g), k, d]
    dec = [L.DEC_F64, L.DEC_F64, 0, 2]
    cols = [D.Column.encode(ctx, dev(ctx, c), seg_rows, decimals=max(x, 0)) for c, x in zip(host, dec)]
    return host, dec, cols, (f, g, k, d)


PREDICATES = [
    [(NAME, 0, 0), (FLT, 0, fb(10.0)), (CMP, 2, 0)],
    [(FLT, 0, fb(10.0)), (NAME, 0, 0), (CMP, 6, 0)],
    [(NAME, 0, 0), (NAME, 2, 0), (CMP, 5, 0)],
    [(NAME, 0, 0), (NUM, 2, 1050), (CMP, 4, 0), (NAME, 1, 0), (FLT, 0, fb(0.0)), (CMP, 1, 0), (OR, 0, 0)],
    [(NAME, 0, 0), (FLT, 0, fb(2.0)), (MUL, 0, 0), (NAME, 3, 0), (CMP, 3, 0),
     (NAME, 2, 0), (NUM, 0, 100), (CMP, 1, 0), (AND, 0, 0)],
    [(NAME, 1, 0), (NAME, 0, 0), (DIV, 0, 0), (FLT, 0, fb(-1e3)), (CMP, 2, 0)],
]


@pytest.mark.parametrize("pi", range(len(PREDICATES)))
@pytest.mark.parametrize("n,seg_rows", [(70001, 16384), (2_100_000, 1_000_000)])
def test_f64_scan_matches_oracle(ctx, pi, n, seg_rows):
    from alenka_amd import device as D

    host, dec, cols, _ = table(ctx, n, seg_rows, 40 + pi)
    ops = PREDICATES[pi]
    prog, _ = L.compile_rpn(ops, dec)
    got = D.scan_filter(ctx, cols, prog).cpu().numpy().astype(np.uint32)
    want = O.compact(O.filter_mask(ops, host, dec))
    assert got.size > 0 and np.array_equal(got, want)


EXPRS = [
    [(NAME, 0, 0), (FLT, 0, fb(1.5)), (ADD, 0, 0)],
    [(NAME, 0, 0), (NAME, 2, 0), (MUL, 0, 0), (NAME, 1, 0), (MINUS, 0, 0)],
    [(NAME, 1, 0), (NAME, 0, 0), (DIV, 0, 0)],
    [(NUM, 1, 25), (NAME, 0, 0), (MINUS, 0, 0), (NAME, 3, 0), (DIV, 0, 0)],
]


@pytest.mark.parametrize("ei", range(len(EXPRS)))
def test_f64_eval_bit_exact(ctx, ei):
    """IEEE double +, -, *, / on the device equal the CPU's bit for bit (division included)."""
    from alenka_amd import device as D

    host, dec, cols, _ = table(ctx, 50000, 8192, 60 + ei)
    prog, prec = L.compile_rpn(EXPRS[ei], dec, mode=1)
    assert prec == L.DEC_F64
    got = D.eval_rows(ctx, cols, prog).cpu().numpy()
    want, wp = O.eval_expr(EXPRS[ei], host, dec)
    assert wp == L.DEC_F64
    assert np.array_equal(got.view(np.float64), want.view(np.float64), equal_nan=True)


def test_f64_group_aggregates(ctx):
    """SUM of doubles within 1e-12 relative of math.fsum (the order of additions is not
    fixed), MIN / MAX and COUNT exact, per int key and over f64 keys in value order."""
    from alenka_amd import device as D

    r = np.random.default_rng(7)
    n = 400_000
    k = r.integers(0, 1000, n)
    v = r.uniform(0.5, 1e4, n)
    fk = r.choice([-3.5, -0.25, 0.0, 1.0, 2.5, 1e9], n)
    for keys, kv in ((dev(ctx, k), k), (None, fk)):
        if keys is None:  # f64 keys: sort by value (order-preserving keys, mapped back)
            raw = dev(ctx, bits(fk))
            keys = torch.empty_like(raw)
            L.check(ctx.lib.aln_f64_order_keys(ctx.h, D._p(raw), n, D._p(keys)), "order keys")
        vals = dev(ctx, bits(v))
        (ok,), (s, mn, mx, c) = D.groupby(ctx, [keys], [vals, vals, vals, vals],
                                          [L.AGG_FSUM, L.AGG_FMIN, L.AGG_FMAX, L.AGG_COUNT])
        ok = ok.cpu().numpy()
        if kv is fk:
            back = torch.empty(ok.size, dtype=torch.int64, device=ctx.device)
            L.check(ctx.lib.aln_f64_unorder_keys(ctx.h, D._p(dev(ctx, ok)), ok.size, D._p(back)), "unorder")
            ok = back.cpu().numpy().view(np.float64)
        want_keys = np.unique(kv)
        assert np.array_equal(ok, want_keys)
        s, mn, mx, c = (x.cpu().numpy() for x in (s, mn, mx, c))
        for i, key in enumerate(want_keys):
            sel = v[kv == key]
            exact = math.fsum(sel)
            assert abs(s.view(np.float64)[i] - exact) <= 1e-12 * abs(exact)
            assert mn.view(np.float64)[i] == sel.min() and mx.view(np.float64)[i] == sel.max()
            assert c[i] == sel.size


def test_f64_order_keys_sort_as_doubles(ctx):
    from alenka_amd import device as D

    r = np.random.default_rng(8)
    x = np.concatenate([r.normal(0, 1e3, 5000), [0.0, -0.0, np.inf, -np.inf, 1e-300, -1e-300, 5e-324]])
    r.shuffle(x)
    raw = dev(ctx, bits(x))
    keys = torch.empty_like(raw)
    L.check(ctx.lib.aln_f64_order_keys(ctx.h, D._p(raw), x.size, D._p(keys)), "order keys")
    for desc in (0, 1):
        perm = D.sort_keys(ctx, [keys], [desc]).cpu().numpy()
        want = np.argsort(-x if desc else x, kind="stable")
        # -0.0 and +0.0 are equal doubles: the sort is stable between them
        assert np.array_equal(x[perm], x[want])
        assert np.array_equal(perm, want)


@pytest.fixture(scope="module")
def ops(ctx):
    from alenka_amd.ops import Ops

    o = Ops(0)
    yield o
    o.close()


def test_f64_operator_statements(ops, tmp_path):
    r = np.random.default_rng(9)
    n = 60000
    k = r.integers(0, 300, n)
    f = np.round(r.uniform(-50, 50, n), 2)
    price = r.integers(100, 100000, n)
    # one segment: GROUP BY returns groups in key order (several would give the hash order)
    ops.table("T", n, 65536, [("k", k, 0), ("f", f, "f64"), ("price", price, 2)])
    dk = np.arange(0, 300, 3)
    w = np.round(r.uniform(0, 9, dk.size), 3)
    ops.table("D", dk.size, 16384, [("dk", dk, 0), ("w", w, "f64")])
    out = str(tmp_path / "o.txt")
    script = (
        "F1 := FILTER T BY f > 1.5E1;\n"
        "S := SELECT k AS k, SUM(f) AS s, MIN(f) AS mn, MAX(f) AS mx, AVG(f) AS av, COUNT(k) AS c "
        "FROM F1 GROUP BY k;\n"
        "O := ORDER S BY s DESC;\n"
        f"STORE O INTO '{out}' USING ('|') LIMIT 5;\n"
        "G := SELECT f AS f, SUM(price) AS sp FROM T GROUP BY f;\n"
        "P := SELECT f * 2.5E0 AS g, f - price AS h FROM T;\n"
        "J := SELECT k AS k, w AS w FROM T JOIN D ON k = dk;\n"
        "Q := FILTER T BY f <= -2.5E1 AND price > 500.00;\n")
    assert ops.alenkaExecute(script) == 0, ops.lib.alenka_last_error()
    # FILTER with an APPROXNUM (15.0, EPSILON comparison) then GROUP BY k
    m = (f - 15.0) > 1e-8
    keys = np.unique(k[m])
    assert np.array_equal(ops.fetch("S", "k"), keys)
    s = ops.fetch_f64("S", "s")
    for i, key in enumerate(keys):
        sel = f[m & (k == key)]
        assert abs(s[i] - math.fsum(sel)) <= 1e-12 * max(1.0, abs(math.fsum(sel)))
        assert ops.fetch_f64("S", "mn")[i] == sel.min() and ops.fetch_f64("S", "mx")[i] == sel.max()
        assert abs(ops.fetch_f64("S", "av")[i] - math.fsum(sel) / sel.size) <= 1e-12 * abs(math.fsum(sel) / sel.size)
    assert np.array_equal(ops.fetch("S", "c"), np.array([np.sum(m & (k == x)) for x in keys]))
    assert [c[1] for c in ops.columns("S")] == [0, 1, 1, 1, 1, 0]
    # ORDER BY a double, DESC, and its "%.2f" text
    o_s = ops.fetch_f64("O", "s")
    assert np.all(o_s[:-1] >= o_s[1:])
    lines = open(out).read().split("\n")
    assert len(lines) == 5
    first = lines[0].split("|")
    assert first[1] == "%.2f" % o_s[0] and first[0] == str(ops.fetch("O", "k")[0])
    # GROUP BY an f64 key: groups in ascending double order, exact int sums
    fk = np.unique(f)
    assert np.array_equal(ops.fetch_f64("G", "f"), fk)
    assert np.array_equal(ops.fetch("G", "sp"), np.array([price[f == x].sum() for x in fk]))
    # projections: f * 2.5 (double), f - price (int column cast as stored: cents)
    assert np.array_equal(ops.fetch_f64("P", "g"), f * 2.5)
    assert np.array_equal(ops.fetch_f64("P", "h"), f - price.astype(np.float64))
    # join: the f64 payload gathered through the pairs, which come out in ascending left key
    # (then left row) order within the left's one segment (operators.cu:1184-1240)
    jk = k[np.isin(k, dk)]
    jk = jk[np.argsort(jk, kind="stable")]
    assert np.array_equal(ops.fetch("J", "k"), jk)
    assert np.array_equal(ops.fetch_f64("J", "w"), w[jk // 3])
    # f64 and decimal predicates together
    q = (((-25.0 - f) > 1e-8) | (np.abs(f + 25.0) < 1e-8)) & (price > 50000)
    assert ops.rows("Q") == int(q.sum())
    # float join keys are refused, as the reference refuses non-decimal float keys
    assert ops.alenkaExecute("X := SELECT k AS k FROM T JOIN D ON f = w;") == L.ALN_E_UNSUPPORTED


def test_f64_binary_round_trip(ops, tmp_path):
    r = np.random.default_rng(10)
    n = 5000
    f = r.normal(0, 1e5, n)
    k = r.integers(0, 100, n)
    ops.table("B", n, 4096, [("k", k, 0), ("f", f, "f64")])
    prefix = str(tmp_path / "btab")
    ops.lib.alenka_set_process_count(2000)
    ops.emit_store_binary("B", prefix, 0)
    ops.check(None, "store binary")
    seg0 = open(f"{prefix}.f.0", "rb").read()
    assert len(seg0) == 8 + 8 * 2000
    assert int.from_bytes(seg0[:4], "little") == 2000 and int.from_bytes(seg0[-4:], "little") == 3
    assert np.array_equal(np.frombuffer(seg0[4:-4], np.float64), f[:2000])
    ops.emit_load_binary("B2", prefix, 0)
    ops.check(None, "load binary")
    assert np.array_equal(ops.fetch_f64("B2", "f"), f)
    assert np.array_equal(ops.fetch("B2", "k"), k)
    ops.lib.alenka_set_process_count(1000000000)
    for p in os.listdir(tmp_path):
        os.remove(tmp_path / p)
