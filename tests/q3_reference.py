"""The TPC-H q3 pipeline of the reference's readme (readme.md:48-80) evaluated with the
CPU oracle's operators, step by step as the reference script runs them:

  OFI := FILTER orders BY o_orderdate < 19950315;
  CF  := FILTER customer BY c_mktsegment == "BUILDING";
  LF  := FILTER lineitem BY shipdate > 19950315;
  OLC := SELECT ... FROM LF JOIN OFI ON orderkey = o_orderkey JOIN CF ON o_custkey = c_custkey;
  F   := SELECT o_orderkey, o_orderdate, o_shippriority, SUM(price*(1-discount)), COUNT(o_orderkey)
         FROM OLC GROUP BY o_orderkey, o_orderdate, o_shippriority;
  RES := ORDER F BY sum_revenue DESC, orderdate1 ASC;  ... LIMIT 10
"""
import numpy as np

import oracle_lib as O
from alenka_amd import _lib as L

SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"]
DATE = 19950315


def tables(seed, ndim, norders):
    c_key, c_seg_idx = O.gen_dim(seed, ndim)
    code_of, c_seg = O.dict_first_seen(c_seg_idx, len(SEGMENTS))
    o_key, o_cust, o_date, o_prio = O.gen_orders(seed, norders, ndim)
    l_okey, l_price, l_disc, l_ship = O.gen_fact(seed, norders)
    return dict(c_key=c_key, c_seg=c_seg, code_of=code_of, o_key=o_key, o_cust=o_cust, o_date=o_date,
                o_prio=o_prio, l_okey=l_okey, l_price=l_price, l_disc=l_disc, l_ship=l_ship)


def run(t, limit=10, segment="BUILDING", date=DATE):
    """The q3 script on the oracle; segment and date are q3's substitution parameters."""
    building = int(t["code_of"][SEGMENTS.index(segment)])
    cf = O.filter_mask([(L.RPN_NAME, 0, 0), (L.RPN_STRING, 0, building), (L.RPN_CMP, 4, 0)],
                       [t["c_seg"]], [0])
    ofi = O.filter_mask([(L.RPN_NAME, 0, 0), (L.RPN_NUMBER, 0, date), (L.RPN_CMP, 1, 0)],
                        [t["o_date"]], [0])
    lf = O.filter_mask([(L.RPN_NAME, 0, 0), (L.RPN_NUMBER, 0, date), (L.RPN_CMP, 2, 0)],
                       [t["l_ship"]], [0])
    c_ids = O.compact(cf)
    o_ids = O.compact(ofi)
    l_ids = O.compact(lf)
    # OLC1 := OFI JOIN CF ON o_custkey = c_custkey (left = orders, right = customers)
    pl, pr = O.join_sortmerge(t["o_cust"][o_ids], t["c_key"][c_ids])
    olc1 = o_ids[pl]
    # OLC := LF JOIN OLC1 ON orderkey = o_orderkey
    ql, qr = O.join_sortmerge(t["l_okey"][l_ids], t["o_key"][olc1])
    lrow = l_ids[ql]
    orow = olc1[qr]
    rev, prec = O.eval_expr([(L.RPN_NAME, 0, 0), (L.RPN_NUMBER, 0, 1), (L.RPN_NAME, 1, 0),
                             (L.RPN_MINUS, 0, 0), (L.RPN_MUL, 0, 0)],
                            [t["l_price"][lrow], t["l_disc"][lrow]], [2, 2])
    assert prec == 4
    gk, gv = O.groupby([t["o_key"][orow], t["o_date"][orow], t["o_prio"][orow]],
                       [rev, rev], [L.AGG_SUM, L.AGG_COUNT])
    perm = O.order([gv[0], gk[1]], [1, 0])
    top = [(int(gk[0][i]), int(gk[1][i]), int(gk[2][i]), int(gv[0][i]), int(gv[1][i]))
           for i in perm[:limit]]
    groups = np.stack([gk[0], gk[1], gk[2], gv[0], gv[1]], axis=1) if len(gk[0]) else np.zeros((0, 5), np.int64)
    stats = dict(dim_pass=int(cf.sum()), mid_pass=int(len(olc1)), fact_pass=int(len(lrow)),
                 groups=int(len(gk[0])))
    return top, groups, stats


def _cpu_lib():
    import ctypes as C
    import os

    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "libq3cpu.so"))
    P = C.POINTER(C.c_int64)
    lib.q3cpu_run_stats.restype = C.c_int
    lib.q3cpu_run_stats.argtypes = [P, P, C.c_uint64, P, P, P, P, C.c_uint64, P, P, P, P, C.c_uint64,
                                    C.c_int64, C.c_int64, C.c_int, C.c_int, P, P]
    lib.flatcpu_run.restype = C.c_int64
    lib.flatcpu_run.argtypes = [P, P, P, P, C.c_uint64, C.c_int64, C.c_int, P, P, P, C.c_uint64]
    lib.q3cpu_threads.restype = C.c_int
    return lib, P


def run_cpu_parallel(t, threads=0, limit=10, segment="BUILDING", date=DATE, with_stats=False):
    """The same q3 through oracle/libq3cpu.so (the reference's Thrust algorithms on the OpenMP
    backend); returns (top rows, threads used) and, with_stats, the pass counts as run()
    names them."""
    lib, P = _cpu_lib()
    arr = {k: np.ascontiguousarray(t[k], np.int64) for k in ("c_key", "c_seg", "o_key", "o_cust", "o_date",
                                                             "o_prio", "l_okey", "l_price", "l_disc", "l_ship")}
    p = lambda k: arr[k].ctypes.data_as(P)
    top = np.zeros((max(1, limit), 5), np.int64)
    st = np.zeros(4, np.int64)
    code = int(t["code_of"][SEGMENTS.index(segment)])
    n = lib.q3cpu_run_stats(p("c_key"), p("c_seg"), len(arr["c_key"]), p("o_key"), p("o_cust"), p("o_date"),
                            p("o_prio"), len(arr["o_key"]), p("l_okey"), p("l_price"), p("l_disc"), p("l_ship"),
                            len(arr["l_okey"]), code, date, threads, limit, top.ctypes.data_as(P),
                            st.ctypes.data_as(P))
    used = threads if threads > 0 else lib.q3cpu_threads()
    rows = [tuple(int(x) for x in top[i]) for i in range(n)]
    if with_stats:
        return rows, used, dict(dim_pass=int(st[0]), mid_pass=int(st[1]), fact_pass=int(st[2]), groups=int(st[3]))
    return rows, used


def run_flat_cpu(cols, date_lt=DATE, threads=0, want=True):
    """The flat query (WHERE date < date_lt GROUP BY key, SUM(price*(1-discount)), COUNT(*))
    through oracle/libq3cpu.so; cols = [key, date, price, disc].  Returns (keys, sums,
    counts) in key order, or only the group count when want is False."""
    lib, P = _cpu_lib()
    a = [np.ascontiguousarray(c, np.int64) for c in cols]
    n = len(a[0])
    if not want:
        return lib.flatcpu_run(*[x.ctypes.data_as(P) for x in a], n, date_lt, threads, None, None, None, 0)
    out = [np.zeros(max(1, n), np.int64) for _ in range(3)]
    g = lib.flatcpu_run(*[x.ctypes.data_as(P) for x in a], n, date_lt, threads,
                        *[o.ctypes.data_as(P) for o in out], n)
    return tuple(o[:g] for o in out)
