"""The reference's operator API (alenkaExecute -> emit_filter / emit_join / emit_select /
emit_order / emit_store) reaching the fused star pipeline: STORE ... LIMIT 10 of the q3
statements runs one aln_star_agg top-k, a read of the GROUP BY runs its group output, and
every other shape or a refusal by the fused kernels runs the statements one by one.  Each
path is checked against the CPU oracle or against the statement-by-statement run."""
import numpy as np
import pytest

import q3_reference as Q3

pytestmark = pytest.mark.gpu

FILTERS = """
ord_f := FILTER orders BY o_orderdate < {date};
cust_f := FILTER customer BY c_mktsegment == "{segment}";
li_f := FILTER lineitem BY shipdate > {date};
joined := SELECT o_orderkey AS o_orderkey, o_orderdate AS o_orderdate,
                 o_shippriority AS o_shippriority, price AS price, discount AS discount
          FROM li_f JOIN ord_f ON orderkey = o_orderkey
                    JOIN cust_f ON o_custkey = c_custkey;
"""
GROUP = """
grouped := SELECT o_orderkey AS k, o_orderdate AS d, o_shippriority AS p,
                  SUM(price * (1 - discount)) AS rev, COUNT(o_orderkey) AS n
           FROM joined GROUP BY o_orderkey, o_orderdate, o_shippriority;
ranked := ORDER grouped BY {order};
STORE ranked INTO '{out}' USING ('|') LIMIT {limit};
"""
Q3_ORDER = "rev DESC, d ASC"


def path(ops):
    return ops.lib.alenka_last_path().decode()


def load(ops, t, seg_rows, dup_customer=False):
    code_to_str = {int(c): Q3.SEGMENTS[v] for v, c in enumerate(t["code_of"]) if c >= 0}
    c_key, c_seg = t["c_key"], [code_to_str[int(c)] for c in t["c_seg"]]
    if dup_customer:  # a repeated customer row: the dimension's join key is not unique
        i = next(j for j, s in enumerate(c_seg) if s == "BUILDING")
        c_key = np.append(c_key, c_key[i])
        c_seg = c_seg + [c_seg[i]]
    ops.table("customer", len(c_key), seg_rows, [("c_custkey", c_key, 0), ("c_mktsegment", c_seg, 10)])
    ops.table("orders", len(t["o_key"]), seg_rows, [("o_orderkey", t["o_key"], 0), ("o_custkey", t["o_cust"], 0),
                                                    ("o_orderdate", t["o_date"], 0), ("o_shippriority", t["o_prio"], 0)])
    ops.table("lineitem", len(t["l_okey"]), seg_rows, [("orderkey", t["l_okey"], 0), ("price", t["l_price"], 2),
                                                       ("discount", t["l_disc"], 2), ("shipdate", t["l_ship"], 0)])


def run(ops, tmp_path, name, date=Q3.DATE, segment="BUILDING", order=Q3_ORDER, limit=10, unfused=False):
    """The q3 statements; unfused: the join is read between its statement and the GROUP BY,
    so every statement runs on its own (the reference's execution)."""
    out = str(tmp_path / name)
    rc = ops.alenkaExecute(FILTERS.format(date=date, segment=segment))
    assert rc == 0, ops.lib.alenka_last_error()
    if unfused:
        ops.rows("joined")
        assert path(ops) == "unfused"
    rc = ops.alenkaExecute(GROUP.format(order=order, out=out, limit=limit))
    assert rc == 0, ops.lib.alenka_last_error()
    return out, open(out).read()


def text_rows(rows):
    lines = []
    for k, d, p, s, c in rows:
        frac = str(s).rjust(5, "0")
        lines.append(f"{k}|{d}|{p}|{frac[:-4]}.{frac[-4:]}|{c}|")
    return "\n".join(lines)


@pytest.fixture(scope="module")
def ops(ctx):
    from alenka_amd.ops import Ops

    o = Ops(0)
    yield o
    o.close()


@pytest.mark.parametrize("seed,nd,no,seg_rows", [(91, 6000, 60000, 16384), (93, 150000, 1500000, 1000000)])
def test_q3_statements_run_fused(ops, tmp_path, seed, nd, no, seg_rows):
    """q3 at the benchmark's segment size (6M fact rows, 1M-row segments: six fact segments
    ending in a partial tile) and a small one: STORE LIMIT 10 is one aln_star_agg top-k, then
    the GROUP BY read in full is its group output, both equal to the oracle."""
    t = Q3.tables(seed, nd, no)
    top, groups, stats = Q3.run(t)
    load(ops, t, seg_rows)
    _, text = run(ops, tmp_path, "q3.txt")
    assert path(ops) == "aln_star_agg top-k"
    assert text == text_rows(top)
    got = np.stack([ops.fetch("grouped", c) for c in ("k", "d", "p", "rev", "n")], axis=1)
    assert path(ops) == "aln_star_agg groups"
    assert np.array_equal(got, groups)
    res = [tuple(int(ops.fetch("ranked", c)[i]) for c in ("k", "d", "p", "rev", "n")) for i in range(10)]
    assert res == top
    assert ops.rows("joined") == stats["fact_pass"]


@pytest.mark.parametrize("order,limit", [(Q3_ORDER, 10), (Q3_ORDER + ", k ASC", 7), ("rev DESC, d ASC, n DESC", 10),
                                         ("rev DESC", 10), ("n DESC, rev DESC", 10), ("d ASC, rev DESC", 25),
                                         (Q3_ORDER, 600)])
def test_fused_and_unfused_agree(ops, tmp_path, order, limit):
    """The same statements run fused (where the ORDER BY's keys give the fused top-k's order
    and LIMIT fits) and one by one give the same text."""
    t = Q3.tables(95, 6000, 60000)
    load(ops, t, 16384)
    _, fused = run(ops, tmp_path, "a.txt", order=order, limit=limit)
    fused_path = path(ops)
    _, plain = run(ops, tmp_path, "b.txt", order=order, limit=limit, unfused=True)
    assert fused == plain
    takes_topk = order in (Q3_ORDER, Q3_ORDER + ", k ASC") and limit <= 512
    assert fused_path == ("aln_star_agg top-k" if takes_topk else "aln_star_agg groups")


def test_duplicate_dimension_keys_run_unfused(ops, tmp_path):
    """A customer key that occurs twice: the fused kernels refuse it (each such order joins
    twice), and STORE runs the statements one by one."""
    t = Q3.tables(97, 6000, 60000)
    load(ops, t, 16384, dup_customer=True)
    _, fused = run(ops, tmp_path, "a.txt")
    assert path(ops) == "unfused"
    _, plain = run(ops, tmp_path, "b.txt", unfused=True)
    assert fused == plain


def test_deferred_join_reads_its_inputs_as_issued(ops, tmp_path):
    """A variable the deferred join reads is assigned again before the join runs: the join
    still reads the filter it was issued with."""
    t = Q3.tables(99, 6000, 60000)
    top, _, _ = Q3.run(t)
    load(ops, t, 16384)
    out = str(tmp_path / "q3.txt")
    assert ops.alenkaExecute(FILTERS.format(date=Q3.DATE, segment="BUILDING")) == 0
    assert ops.alenkaExecute("li_f := FILTER lineitem BY shipdate < 0;\nord_f := FILTER orders BY o_orderdate < 0;") == 0
    assert ops.alenkaExecute(GROUP.format(order=Q3_ORDER, out=out, limit=10)) == 0, ops.lib.alenka_last_error()
    assert open(out).read() == text_rows(top)
    assert ops.rows("li_f") == 0


def test_other_group_by_shapes_run_unfused(ops, tmp_path):
    """GROUP BY o_orderdate over the join (not the orders key): statement by statement,
    equal to a numpy group-by of the oracle's join rows."""
    t = Q3.tables(101, 6000, 60000)
    _, groups, _ = Q3.run(t)
    load(ops, t, 16384)
    assert ops.alenkaExecute(FILTERS.format(date=Q3.DATE, segment="BUILDING")) == 0
    assert ops.alenkaExecute("g2 := SELECT o_orderdate AS d, SUM(price * (1 - discount)) AS rev, "
                             "COUNT(o_orderkey) AS n FROM joined GROUP BY o_orderdate;") == 0
    d, rev, n = (ops.fetch("g2", c) for c in ("d", "rev", "n"))
    assert path(ops) == "unfused"
    want_d = np.unique(groups[:, 1])
    want_rev = np.array([groups[groups[:, 1] == x, 3].sum() for x in want_d])
    want_n = np.array([groups[groups[:, 1] == x, 4].sum() for x in want_d])
    order = np.argsort(d, kind="stable")
    assert np.array_equal(d[order], want_d)
    assert np.array_equal(rev[order], want_rev)
    assert np.array_equal(n[order], want_n)
