"""The reference's scripts replayed through the operator layer (emit_* over the C ABI),
issuing the calls in the order the reference's parser (bison.y:127-287) issues them, and
checked against the CPU oracle."""
import os

import numpy as np
import pytest

import oracle_lib as O
import q3_reference as Q3
from alenka_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops(ctx):
    from alenka_amd.ops import Ops

    o = Ops(0)
    yield o
    o.close()


def flat_query(ops, seg_rows, n=300000, seed=42):
    key, date, price, disc = O.gen_flat(seed, n)
    ops.table("T", n, seg_rows, [("key", key, 0), ("date", date, 0), ("price", price, 2),
                                 ("discount", disc, 2)])
    # F := FILTER T BY date < 19950315;
    ops.emit_name("date"); ops.emit_number(19950315); ops.emit_cmp(1); ops.emit_filter("F", "T")
    ops.check(None, "filter")
    # R := SELECT key AS k, SUM(price*(1-discount)) AS rev, COUNT(key) AS cnt FROM F GROUP BY key;
    ops.emit_name("key"); ops.emit_sel_name("k")
    ops.emit_name("price"); ops.emit_number(1); ops.emit_name("discount"); ops.emit_minus()
    ops.emit_mul(); ops.emit_sum(); ops.emit_sel_name("rev")
    ops.emit_name("key"); ops.emit_count(); ops.emit_sel_name("cnt")
    ops.emit_name("key")
    ops.emit_select("R", "F", 1)
    ops.check(None, "select")
    return key, date, price, disc


@pytest.mark.parametrize("seg_rows", [1000000, 65536])
def test_flat_filter_groupby(ops, seg_rows):
    n = 300000
    key, date, price, disc = flat_query(ops, seg_rows, n)
    sel = date < 19950315
    assert ops.rows("F") == int(sel.sum())
    rev = price[sel] * (100 - disc[sel])
    nsegs = (n + seg_rows - 1) // seg_rows
    bounds = [min(n, s * seg_rows) for s in range(nsegs + 1)]
    # group-by input segments are the source table's (emit_select loops over them)
    seg_of = np.searchsorted(bounds, np.nonzero(sel)[0], side="right") - 1
    fb = [int(np.searchsorted(seg_of, s)) for s in range(nsegs)] + [len(seg_of)]
    wk, wv = O.groupby([key[sel]], [rev, rev], [L.AGG_SUM, L.AGG_COUNT],
                       seg_bounds=fb if nsegs > 1 else None, seed=100)
    assert np.array_equal(ops.fetch("R", "k"), wk[0])
    assert np.array_equal(ops.fetch("R", "rev"), wv[0])
    assert np.array_equal(ops.fetch("R", "cnt"), wv[1])
    cols = ops.columns("R")
    assert [c[0] for c in cols] == ["k", "rev", "cnt"]
    assert [c[2] for c in cols] == [0, 4, 0]  # decimal scale of SUM(price*(1-discount))


def test_filter_chain_and_zone_maps(ops):
    n = 100000
    key, date, price, disc = O.gen_flat(5, n)
    order = np.argsort(date, kind="stable")  # clustered dates: zone maps skip segments
    key, date, price, disc = key[order], date[order], price[order], disc[order]
    ops.table("Z", n, 8192, [("key", key, 0), ("date", date, 0), ("price", price, 2),
                             ("discount", disc, 2)])
    ops.emit_name("date"); ops.emit_number(19940101); ops.emit_cmp(6); ops.emit_filter("Z1", "Z")
    ops.emit_name("discount"); ops.emit_decimal("0.05"); ops.emit_cmp(5); ops.emit_filter("Z2", "Z1")
    ops.check()
    want = (date >= 19940101) & (disc <= 5)
    assert ops.rows("Z2") == int(want.sum())
    assert np.array_equal(ops.fetch("Z2", "key"), key[want])


def q3_script(ops, t, seg_rows, out_path):
    nd, no, nl = len(t["c_key"]), len(t["o_key"]), len(t["l_okey"])
    ops.table("customer", nd, seg_rows, [("c_custkey", t["c_key"], 0), ("c_mktsegment", t["c_seg_str"], 10)])
    ops.table("orders", no, seg_rows, [("o_orderkey", t["o_key"], 0), ("o_custkey", t["o_cust"], 0),
                                       ("o_orderdate", t["o_date"], 0), ("o_shippriority", t["o_prio"], 0)])
    ops.table("lineitem", nl, seg_rows, [("orderkey", t["l_okey"], 0), ("price", t["l_price"], 2),
                                         ("discount", t["l_disc"], 2), ("shipdate", t["l_ship"], 0)])
    # OFI := FILTER orders BY o_orderdate < 19950315;
    ops.emit_name("o_orderdate"); ops.emit_number(19950315); ops.emit_cmp(1); ops.emit_filter("OFI", "orders")
    # CF := FILTER customer BY c_mktsegment == "BUILDING";
    ops.emit_name("c_mktsegment"); ops.emit_string('"BUILDING"'); ops.emit_cmp(4); ops.emit_filter("CF", "customer")
    # LF := FILTER lineitem BY shipdate > 19950315;
    ops.emit_name("shipdate"); ops.emit_number(19950315); ops.emit_cmp(2); ops.emit_filter("LF", "lineitem")
    ops.check(None, "filters")
    # OLC := SELECT ... FROM LF JOIN OFI ON orderkey = o_orderkey JOIN CF ON o_custkey = c_custkey;
    for c in ("o_orderkey", "o_orderdate", "o_shippriority", "price", "discount"):
        ops.emit_name(c); ops.emit_sel_name(c)
    ops.emit_name("orderkey"); ops.emit_name("o_orderkey"); ops.emit_eq()
    ops.emit_name("o_custkey"); ops.emit_name("c_custkey"); ops.emit_eq()
    ops.emit_join_tab("CF", b"I"); ops.emit_join_tab("OFI", b"I")
    ops.emit_join("OLC", "LF", 1, 0, -1)
    ops.check(None, "join")
    # F := SELECT o_orderkey AS o_orderkey1, o_orderdate AS orderdate1, o_shippriority AS
    #      shippriority1, SUM(price*(1-discount)) AS sum_revenue, COUNT(o_orderkey) AS cnt
    #      FROM OLC GROUP BY o_orderkey, o_orderdate, o_shippriority;
    ops.emit_name("o_orderkey"); ops.emit_sel_name("o_orderkey1")
    ops.emit_name("o_orderdate"); ops.emit_sel_name("orderdate1")
    ops.emit_name("o_shippriority"); ops.emit_sel_name("shippriority1")
    ops.emit_name("price"); ops.emit_number(1); ops.emit_name("discount"); ops.emit_minus(); ops.emit_mul()
    ops.emit_sum(); ops.emit_sel_name("sum_revenue")
    ops.emit_name("o_orderkey"); ops.emit_count(); ops.emit_sel_name("cnt")
    ops.emit_name("o_orderkey"); ops.emit_name("o_orderdate"); ops.emit_name("o_shippriority")
    ops.emit_select("F", "OLC", 3)
    ops.check(None, "group by")
    # RES := ORDER F BY sum_revenue DESC, orderdate1 ASC;
    ops.emit_var_desc("sum_revenue"); ops.emit_var_asc("orderdate1"); ops.emit_order("RES", "F", 2, 0)
    ops.check(None, "order")
    # STORE RES INTO 'mytest.txt' USING ('|') LIMIT 10;
    ops.emit_limit(10); ops.emit_store("RES", out_path, "|")
    ops.check(None, "store")


def test_q3_script(ops, tmp_path):
    t = Q3.tables(77, 6000, 60000)
    code_to_str = {int(c): Q3.SEGMENTS[v] for v, c in enumerate(t["code_of"]) if c >= 0}
    t["c_seg_str"] = [code_to_str[int(c)] for c in t["c_seg"]]
    top, groups, stats = Q3.run(t)
    out = str(tmp_path / "q3.txt")
    q3_script(ops, t, 16384, out)
    assert ops.rows("OLC") == stats["fact_pass"]
    got = np.stack([ops.fetch("F", c) for c in ("o_orderkey1", "orderdate1", "shippriority1",
                                                 "sum_revenue", "cnt")], axis=1)
    assert np.array_equal(got, groups)
    res = [tuple(int(ops.fetch("RES", c)[i]) for c in ("o_orderkey1", "orderdate1", "shippriority1",
                                                       "sum_revenue", "cnt")) for i in range(10)]
    assert res == top
    text = open(out).read().split("\n")
    assert len(text) == 10
    k, d, p, s, c = top[0]
    frac = str(s)
    while len(frac) <= 4:
        frac = "0" + frac
    assert text[0] == f"{k}|{d}|{p}|{frac[:-4]}.{frac[-4:]}|{c}|"


def test_errors_do_not_exit(ops):
    ops.emit_name("nothing"); ops.emit_number(1); ops.emit_cmp(1); ops.emit_filter("X", "no_such_table")
    assert ops.lib.alenka_status() == L.ALN_E_ARG
    assert b"couldn't find" in ops.lib.alenka_last_error()


SCRIPT_Q3 = """
-- TPC-H q3 written in the reference's statement language
ord_f := FILTER orders BY o_orderdate < {date};
cust_f := FILTER customer BY c_mktsegment == "{segment}";
li_f := FILTER lineitem BY shipdate > {date};
joined := SELECT o_orderkey AS o_orderkey, o_orderdate AS o_orderdate,
                 o_shippriority AS o_shippriority, price AS price, discount AS discount
          FROM li_f JOIN ord_f ON orderkey = o_orderkey
                    JOIN cust_f ON o_custkey = c_custkey;
grouped := SELECT o_orderkey AS k, o_orderdate AS d, o_shippriority AS p,
                  SUM(price * (1 - discount)) AS rev, COUNT(o_orderkey) AS n
           FROM joined GROUP BY o_orderkey, o_orderdate, o_shippriority;
ranked := ORDER grouped BY rev DESC, d ASC;
STORE ranked INTO '{out}' USING ('|') LIMIT 10;
"""


@pytest.mark.parametrize("segment,date", [("BUILDING", 19950315), ("MACHINERY", 19950303)])
def test_script_text_q3(ops, tmp_path, segment, date):
    """The same q3 as a script (alenkaExecute): parsed into the parser's emit_* calls."""
    t = Q3.tables(91, 6000, 60000)
    code_to_str = {int(c): Q3.SEGMENTS[v] for v, c in enumerate(t["code_of"]) if c >= 0}
    t["c_seg_str"] = [code_to_str[int(c)] for c in t["c_seg"]]
    top, groups, stats = Q3.run(t, segment=segment, date=date)
    seg_rows = 16384
    ops.table("customer", len(t["c_key"]), seg_rows, [("c_custkey", t["c_key"], 0), ("c_mktsegment", t["c_seg_str"], 10)])
    ops.table("orders", len(t["o_key"]), seg_rows, [("o_orderkey", t["o_key"], 0), ("o_custkey", t["o_cust"], 0),
                                                    ("o_orderdate", t["o_date"], 0), ("o_shippriority", t["o_prio"], 0)])
    ops.table("lineitem", len(t["l_okey"]), seg_rows, [("orderkey", t["l_okey"], 0), ("price", t["l_price"], 2),
                                                       ("discount", t["l_disc"], 2), ("shipdate", t["l_ship"], 0)])
    out = str(tmp_path / "q3.txt")
    rc = ops.alenkaExecute(SCRIPT_Q3.format(date=date, segment=segment, out=out))
    assert rc == 0, ops.lib.alenka_last_error()
    assert ops.rows("joined") == stats["fact_pass"]
    got = np.stack([ops.fetch("grouped", c) for c in ("k", "d", "p", "rev", "n")], axis=1)
    assert np.array_equal(got, groups)
    res = [tuple(int(ops.fetch("ranked", c)[i]) for c in ("k", "d", "p", "rev", "n")) for i in range(10)]
    assert res == top
    assert len(open(out).read().split("\n")) == 10


def test_script_errors(ops):
    assert ops.alenkaExecute("x := LOAD 'f.tbl' USING ('|') AS (a{1}:int);") == L.ALN_E_UNSUPPORTED
    assert ops.alenkaExecute("x := FILTER ;") == L.ALN_E_ARG
    assert b"statement 1" in ops.lib.alenka_last_error()
    assert ops.alenkaExecute("y := FILTER no_such_table BY a < 1; z := FILTER y BY b > 2;") != 0
    assert b"couldn't find" in ops.lib.alenka_last_error()


# ---- LIKE over char columns (filter.cu:67-129 matched per dictionary entry, 452-499) ----------
P_TYPES = [f"{a} {b} {c}" for a in ("STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO")
           for b in ("ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED")
           for c in ("TIN", "NICKEL", "BRASS", "STEEL", "COPPER")]


@pytest.fixture(scope="module")
def part_table(ops):
    rng = np.random.default_rng(11)
    n = 60000
    key = np.arange(n, dtype=np.int64) * 3 + 1
    ty = [P_TYPES[i] for i in rng.integers(0, len(P_TYPES), n)]
    size = rng.integers(1, 51, n).astype(np.int64)
    return key, ty, size


@pytest.mark.parametrize("pattern", ["%BRASS", "PROMO%", "%POLISHED%", "?MALL%", "%", "LARGE BRUSHED STEEL",
                                     "%Z%", "ST%ED%TIN", "%O?%L"])
def test_script_filter_like(ops, part_table, pattern):
    """FILTER ... BY p_type LIKE pattern AND p_size == 15 (the shape of the reference's q2)."""
    key, ty, size = part_table
    ops.table("part", len(key), 8192, [("p_partkey", key, 0), ("p_type", ty, 25), ("p_size", size, 0)])
    rc = ops.alenkaExecute(f'PF := FILTER part BY p_type LIKE "{pattern}" AND p_size == 15;')
    assert rc == 0, ops.lib.alenka_last_error()
    match = {t: O.like(t, 25, pattern) for t in set(ty)}
    want = np.array([match[t] for t in ty]) & (size == 15)
    assert ops.rows("PF") == int(want.sum())
    assert np.array_equal(ops.fetch("PF", "p_partkey"), key[want])


def test_emit_like_or_and_position_budget(ops):
    """emit_* calls for `v LIKE "%ING" OR v LIKE "a?c"` on a char[8] column whose values fill
    the width: the matcher's position budget (the column width) rejects BUILDING."""
    vals = ["BUILDING", "aaab", "a.c", "abc", "RING", "xING", "", "ING.ING."]
    rng = np.random.default_rng(3)
    pick = rng.integers(0, len(vals), 20000)
    col = [vals[i] for i in pick]
    ids = np.arange(len(col), dtype=np.int64)
    ops.table("V", len(col), 4096, [("id", ids, 0), ("v", col, 8)])
    ops.emit_name("v"); ops.emit_string(b"%ING"); ops.emit_cmp(7)
    ops.emit_name("v"); ops.emit_string(b"a?c"); ops.emit_cmp(7)
    ops.emit_or()
    ops.emit_filter("V1", "V")
    ops.check(None, "filter")
    match = [O.like(v, 8, "%ING") or O.like(v, 8, "a?c") for v in vals]
    assert match == [False, False, False, True, True, True, False, False]
    want = np.array([match[i] for i in pick])
    assert ops.rows("V1") == int(want.sum())
    assert np.array_equal(ops.fetch("V1", "id"), ids[want])


# ---- semi / anti joins (operators.cu:1357-1399, 1473-1530) -------------------------------------
def semi_anti_tables(ops, seed=17):
    r = np.random.default_rng(seed)
    nc, no = 3000, 40000
    c_key = r.permutation(np.arange(1, nc + 1)).astype(np.int64)
    c_nat = r.integers(0, 25, nc).astype(np.int64)
    c_prio = r.integers(0, 5, nc).astype(np.int64)
    o_key = np.arange(no, dtype=np.int64) * 7 + 3
    o_cust = r.integers(1, 4500, no).astype(np.int64)  # a third match no customer
    o_prio = r.integers(0, 5, no).astype(np.int64)
    ops.table("customer", nc, 1024, [("c_custkey", c_key, 0), ("c_nationkey", c_nat, 0), ("c_prio", c_prio, 0)])
    ops.table("orders", no, 16384, [("o_orderkey", o_key, 0), ("o_custkey", o_cust, 0), ("o_prio", o_prio, 0)])
    return dict(c_key=c_key, c_nat=c_nat, c_prio=c_prio, o_key=o_key, o_cust=o_cust, o_prio=o_prio)


@pytest.mark.parametrize("words,kind", [("LEFT SEMI", "1"), ("LEFT ANTI", "3"), ("RIGHT SEMI", "2"),
                                        ("RIGHT ANTI", "4")])
@pytest.mark.parametrize("two_keys", [False, True])
def test_script_semi_anti_join(ops, words, kind, two_keys):
    t = semi_anti_tables(ops)
    cond = "o_custkey = c_custkey" + (" AND o_prio = c_prio" if two_keys else "")
    left = kind in "13"
    sel = "o_orderkey AS k, o_custkey AS c" if left else "c_custkey AS k, c_nationkey AS c"
    script = (f"C := FILTER customer BY c_nationkey < 20;\n"
              f"S := SELECT {sel} FROM orders {words} JOIN C ON {cond};")
    assert ops.alenkaExecute(script) == 0, ops.lib.alenka_last_error()
    keep = t["c_nat"] < 20
    ck, cp, cn = t["c_key"][keep], t["c_prio"][keep], t["c_nat"][keep]
    lk = [t["o_cust"]] + ([t["o_prio"]] if two_keys else [])
    rk = [ck] + ([cp] if two_keys else [])
    rows = O.semi_anti(lk, rk, kind)
    assert ops.rows("S") == len(rows)
    if left:
        assert np.array_equal(ops.fetch("S", "k"), t["o_key"][rows])
        assert np.array_equal(ops.fetch("S", "c"), t["o_cust"][rows])
    else:
        assert np.array_equal(ops.fetch("S", "k"), ck[rows])
        assert np.array_equal(ops.fetch("S", "c"), cn[rows])


def test_script_inner_join_two_keys(ops):
    """`ON a = b AND c = d` (parsed as a = ((b AND c) = d), bison.y:42-49): pairs of the first
    key's join where the second keys are equal too, in the pair order (operators.cu:1288-1348)."""
    t = semi_anti_tables(ops, seed=29)
    rc = ops.alenkaExecute("S := SELECT o_orderkey AS k, c_nationkey AS n FROM orders "
                           "JOIN customer ON o_custkey = c_custkey AND o_prio = c_prio;")
    assert rc == 0, ops.lib.alenka_last_error()
    no = len(t["o_key"])
    bounds = list(range(0, no, 16384)) + [no]
    ol, orr = O.join_sortmerge(t["o_cust"], t["c_key"], bounds)
    keep = t["o_prio"][ol] == t["c_prio"][orr]
    assert ops.rows("S") == int(keep.sum())
    assert np.array_equal(ops.fetch("S", "k"), t["o_key"][ol[keep]])
    assert np.array_equal(ops.fetch("S", "n"), t["c_nat"][orr[keep]])
    rc = ops.alenkaExecute("S := SELECT o_orderkey AS k FROM orders JOIN customer "
                           "ON o_custkey = c_custkey AND o_prio = c_prio AND o_orderkey = c_nationkey;")
    assert rc == L.ALN_E_UNSUPPORTED


@pytest.mark.parametrize("script", [False, True])
def test_anti_join_in_a_chain(ops, script):
    """`nation JOIN customer ON ... LEFT ANTI JOIN orders ON ...`: the parser hands over the
    innermost join first (bison.y:260-275), so customer LEFT ANTI JOIN orders runs first and
    nation is joined with the customers that have no order."""
    t = semi_anti_tables(ops, seed=5)
    nat = np.arange(0, 25, dtype=np.int64)
    ops.table("nation", 25, 64, [("n_nationkey", nat, 0), ("n_regionkey", nat % 5, 0)])
    if script:
        rc = ops.alenkaExecute("S := SELECT c_custkey AS k, n_regionkey AS rg FROM nation "
                               "JOIN customer ON n_nationkey = c_nationkey "
                               "LEFT ANTI JOIN orders ON c_custkey = o_custkey;")
        assert rc == 0, ops.lib.alenka_last_error()
    else:
        ops.emit_name("c_custkey"); ops.emit_sel_name("k")
        ops.emit_name("n_regionkey"); ops.emit_sel_name("rg")
        ops.emit_name("n_nationkey"); ops.emit_name("c_nationkey"); ops.emit_eq()
        ops.emit_name("c_custkey"); ops.emit_name("o_custkey"); ops.emit_eq()
        ops.emit_join_tab(b"orders", b"3")
        ops.emit_join_tab(b"customer", b"I")
        ops.emit_join("S", "nation", 1, 0, -1)
        ops.check(None, "join")
    anti = O.semi_anti([t["c_key"]], [t["o_cust"]], "3")
    ln, rc_ = O.join_sortmerge(nat, t["c_nat"][anti])
    assert ops.rows("S") == len(ln)
    assert np.array_equal(ops.fetch("S", "k"), t["c_key"][anti][rc_])
    assert np.array_equal(ops.fetch("S", "rg"), (nat % 5)[ln])


def test_semi_anti_with_an_empty_side(ops):
    """No pairs: semi joins are empty, anti joins keep every row of their side."""
    t = semi_anti_tables(ops, seed=3)
    script = ("E := FILTER customer BY c_nationkey > 100;\n"
              "S1 := SELECT o_orderkey AS k FROM orders LEFT SEMI JOIN E ON o_custkey = c_custkey;\n"
              "S3 := SELECT o_orderkey AS k FROM orders LEFT ANTI JOIN E ON o_custkey = c_custkey;\n"
              "S4 := SELECT c_custkey AS k FROM customer RIGHT ANTI JOIN E ON c_custkey = c_custkey;")
    assert ops.alenkaExecute(script) == 0, ops.lib.alenka_last_error()
    assert ops.rows("S1") == 0
    assert np.array_equal(ops.fetch("S3", "k"), t["o_key"])
    assert ops.rows("S4") == 0


def test_unsupported_join_kinds(ops):
    semi_anti_tables(ops)
    for words in ("LEFT", "RIGHT", "OUTER"):
        rc = ops.alenkaExecute(f"S := SELECT o_orderkey AS k FROM orders {words} JOIN customer ON o_custkey = c_custkey;")
        assert rc == L.ALN_E_UNSUPPORTED, words


# ---- ORDER BY on char columns (update_char_permutation cm.cu:4218-4260, str_sort
# strings_sort_device.cu:125-137) ---------------------------------------------------------------
def reference_order(keys, n):
    """The reference's ORDER BY: one stable sort per key, the last key first
    (operators.cu:1754-1762); keys are (values, width or None, desc)."""
    perm = np.arange(n, dtype=np.uint32)
    for vals, width, desc in reversed(keys):
        if width is None:
            perm = perm[O.order([np.asarray(vals)[perm]], [int(desc)])]
        else:
            perm = O.str_sort(vals, width, perm, desc)
    return perm


@pytest.mark.parametrize("width", [10, 4, 120])
def test_script_order_by_char(ops, width):
    r = np.random.default_rng(width)
    words = ["PROMO", "AUTOMOBILE", "BUILDING", "", "MACHINERY", "A", "Z", "HOUSEHOLD", "FURNITURE", "AB"]
    n = 5000
    name = [words[i][:width] for i in r.integers(0, len(words), n)]
    grp = r.integers(0, 7, n).astype(np.int64)
    ids = np.arange(n, dtype=np.int64)
    ops.table("T", n, 1024, [("id", ids, 0), ("name", name, width), ("grp", grp, 0)])
    assert ops.alenkaExecute("O1 := ORDER T BY name ASC, grp DESC;\nO2 := ORDER T BY grp ASC, name DESC;") == 0, \
        ops.lib.alenka_last_error()
    want1 = reference_order([(name, width, False), (grp, None, True)], n)
    want2 = reference_order([(grp, None, False), (name, width, True)], n)
    assert np.array_equal(ops.fetch("O1", "id"), ids[want1])
    assert np.array_equal(ops.fetch("O2", "id"), ids[want2])
    if width == 120:  # no string sort at this width: the int key alone orders the rows
        assert np.array_equal(want2, np.argsort(grp, kind="stable"))
