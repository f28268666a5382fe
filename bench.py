#!/usr/bin/env python3
"""Benchmark of the relational hot path on one MI355X.

Workload (the 600M-row compressed star pipeline): dimension 15M rows (dense int32 key,
char[10] segment dictionary-encoded), orders 150M rows (sparse unique key, 4 of every 32;
FK into the dimension; orderdate; priority 0), fact ~600M rows (1-7 per order, in
order-key order; price, discount decimal(.,2); shipdate = orderdate + 1..121 days).  Every
column is stored frame-of-reference + bit-packed in 1,000,000-row segments carrying
min/max zone maps, resident in HBM.  One step is the TPC-H q3 pipeline:
  filter dimension (segment == "BUILDING"), orders (orderdate < 19950315), fact
  (shipdate > 19950315); join fact -> orders -> dimension; group by (orderkey, orderdate,
  priority) SUM(price*(1-discount)), COUNT; ORDER BY revenue DESC, orderdate ASC LIMIT 10.

Prints one JSON line: fact rows/s over the whole step (inputs already in HBM), the
fraction of HBM peak the dominant kernel (the fact phase) reaches on the bytes it must
move, the same q3 on the host cores over the same tables (rocThrust's OpenMP backend),
whose answer checks the GPU's, and the 1M-row flat and 6M-row star configurations timed
on both.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fact rows/s + % HBM peak, 600M-row join+group-by; batched vs plain q/s at p99"
SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"]
DATE = 19950315


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_tables(ctx, D, seed, scale, seg_rows, host=None):
    """The three q3 tables, encoded in HBM.  With a dict `host`, the raw int64 columns are
    also copied to host memory under the oracle's names (the CPU check's input)."""
    import torch

    ndim, norders = int(150000 * scale), int(1500000 * scale)
    t0 = time.time()
    keep = (lambda name, col: host.__setitem__(name, col.cpu().numpy())) if host is not None else (lambda *a: None)
    c_key, c_seg_idx = D.gen_dim(ctx, seed, ndim)
    code_of, c_seg = D.dict_first_seen(ctx, c_seg_idx, len(SEGMENTS))
    del c_seg_idx
    keep("c_key", c_key)
    keep("c_seg", c_seg)
    dim = [D.Column.encode(ctx, c_key, seg_rows), D.Column.encode(ctx, c_seg, seg_rows)]
    del c_key, c_seg
    mid = []
    for name, col in zip(("o_key", "o_cust", "o_date", "o_prio"), D.gen_orders(ctx, seed, norders, ndim)):
        keep(name, col)
        mid.append(D.Column.encode(ctx, col, seg_rows))
        del col
    fact = []
    okey, price, disc, ship = D.gen_fact(ctx, seed, norders)
    nfact = okey.numel()
    for name, col, dec in (("l_okey", okey, 0), ("l_price", price, 2), ("l_disc", disc, 2), ("l_ship", ship, 0)):
        keep(name, col)
        fact.append(D.Column.encode(ctx, col, seg_rows, decimals=dec))
    del okey, price, disc, ship, col
    if host is not None:
        host["code_of"] = list(code_of)
    torch.cuda.empty_cache()
    log(f"tables: dim {ndim}, orders {norders}, fact {nfact} rows; built in {time.time() - t0:.1f}s")
    return dim, mid, fact, code_of, nfact


def q3_spec(code_of, dim, mid, fact):
    b = int(code_of[SEGMENTS.index("BUILDING")])
    D_ = dict(cols=dim, ranges=[(1, b, b)])
    M_ = dict(cols=mid, ranges=[(2, -(1 << 63), DATE - 1)])
    F_ = dict(cols=fact, ranges=[(3, DATE + 1, (1 << 63) - 1)])
    spec = dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1)
    return D_, M_, F_, spec


def scanned_segments(t):
    """Segments the zone maps do not skip (the kernels' own rule: any range disjoint from
    the segment's [min, max] skips it)."""
    cols, segs = t["cols"], []
    for s in range(cols[0].nsegs):
        skip = False
        for c, lo, hi in t["ranges"]:
            g = cols[c].hsegs[s]
            if g.vmax < lo or g.vmin > hi:
                skip = True
        if not skip:
            segs.append(s)
    return segs


def scan_bytes(t, used_cols):
    segs = scanned_segments(t)
    return sum(t["cols"][c].packed_bytes(segs) for c in used_cols), len(segs)


def star_sources_hash():
    """sha256 over the sources the star kernels are compiled from."""
    import hashlib

    h = hashlib.sha256()
    for f in ("star.hip", "stage.h", "aln_internal.h"):
        with open(os.path.join(ROOT, "alenka_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def cpu_info():
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    return model, threads


def median_time(fn, runs, budget_s=None):
    """1 warm-up call (its result is returned), then `runs` timed calls (fewer if budget_s
    seconds are used up); returns (median seconds, runs timed, warm-up result)."""
    first = fn()
    ts = []
    t_end = None if budget_s is None else time.time() + budget_s
    while len(ts) < runs and (t_end is None or time.time() < t_end or not ts):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], len(ts), first


def cpu_q3(t, threads, runs, budget_s=None):
    """q3 on the host cores through oracle/libq3cpu.so (the reference's Thrust algorithms on
    rocThrust's OpenMP backend): (median s, runs, (top rows, pass counts))."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import q3_reference as Q3

    def go():
        rows, _, st = Q3.run_cpu_parallel(t, threads, with_stats=True)
        return rows, st

    return median_time(go, runs, budget_s)


def flat_gpu_query(ctx, D, L, cols, date_lt=DATE):
    """The flat query on the GPU through the per-statement operators: scan + compaction of
    date < date_lt, key and SUM(price*(1-discount)) of the passing rows, GROUP BY key in key
    order (one segment)."""
    prog = [(L.I_COL, 1, 1), (L.I_CONST, 0, date_lt), (L.I_CMP, 1, 0)]
    rows = D.scan_filter(ctx, cols, prog)
    key = D.eval_rows(ctx, cols, [(L.I_COL, 0, 1)], rows)
    rev = D.eval_rows(ctx, cols, [(L.I_COL, 2, 1), (L.I_CONST, 0, 100), (L.I_COL, 3, 1), (L.I_SUB, 0, 0),
                                  (L.I_MUL, 0, 0)], rows)
    return D.groupby(ctx, [key], [rev, rev], [L.AGG_SUM, L.AGG_COUNT], order=L.ORDER_KEYS)


def check_star(ctx, D, D_, M_, F_, spec, rows, cpu_rows, cpu_stats):
    """The GPU's top-10 (from the timed runs) and one more run's pass and group counts
    against the CPU port's on the same tables."""
    _, g, st = D.star_agg(ctx, D_, M_, F_, spec, limit=10, want_groups=True,
                          groups_cap=max(16, 4 * cpu_stats["groups"]))
    del g
    return bool(rows == cpu_rows and (st.dim_pass, st.mid_pass, st.fact_pass, st.groups) == (
        cpu_stats["dim_pass"], cpu_stats["mid_pass"], cpu_stats["fact_pass"], cpu_stats["groups"]))


def config0_leg(ctx, D, seed, reps, cpu_threads, cpu_runs):
    """Flat filter + GROUP BY over a 1M-row table in one segment: GPU (operators, inputs
    in HBM) and CPU (Thrust/OpenMP port), results checked against each other."""
    import numpy as np
    import torch

    from alenka_amd import _lib as L

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import q3_reference as Q3

    n = 1_000_000
    raw = D.gen_flat(ctx, seed, n)
    cols = [D.Column.encode(ctx, c, n, decimals=d) for c, d in zip(raw, (0, 0, 2, 2))]
    host = [c.cpu().numpy() for c in raw]
    del raw
    t_gpu, (gk, gv) = timed(lambda: flat_gpu_query(ctx, D, L, cols), reps)
    out = {"rows": n, "query": "WHERE date < 19950315 GROUP BY key: SUM(price*(1-discount)), COUNT(*)",
           "groups": int(gk[0].numel()),
           "gpu": {"ms": t_gpu * 1e3, "rows_per_s": n / t_gpu,
                   "path": "aln_scan_filter -> aln_eval x2 -> aln_groupby (key order)"}}
    if cpu_threads:
        t_cpu, nr, (ck, cs, cc) = median_time(lambda: Q3.run_flat_cpu(host, threads=cpu_threads), cpu_runs)
        out["cpu"] = {"ms": t_cpu * 1e3, "rows_per_s": n / t_cpu, "threads": cpu_threads, "runs": nr}
        out["checked"] = bool(np.array_equal(gk[0].cpu().numpy(), ck) and np.array_equal(gv[0].cpu().numpy(), cs)
                              and np.array_equal(gv[1].cpu().numpy(), cc))
    del cols, gk, gv
    torch.cuda.empty_cache()
    return out


def config1_leg(ctx, D, seed, seg_rows, steps, cpu_threads, cpu_runs):
    """The q3 star pipeline at the 6M-row base size (the same generator, scale 1): GPU
    (aln_star_agg, compressed columns in HBM) and CPU (Thrust/OpenMP port), checked."""
    import torch

    host = {}
    dim, mid, fact, code_of, nfact = build_tables(ctx, D, seed, 1.0, seg_rows, host=host)
    D_, M_, F_, spec = q3_spec(code_of, dim, mid, fact)
    q = D.StarQuery(ctx, D_, M_, F_, spec, limit=10)
    t_gpu, (rows, _, _) = timed(q.run, steps)
    out = {"fact_rows": nfact, "orders_rows": mid[0].nrows, "dim_rows": dim[0].nrows, "seg_rows": seg_rows,
           "gpu": {"ms": t_gpu * 1e3, "fact_rows_per_s": nfact / t_gpu, "path": "aln_star_agg"}}
    if cpu_threads:
        t_cpu, nr, (crows, cst) = cpu_q3(host, cpu_threads, cpu_runs)
        out["cpu"] = {"ms": t_cpu * 1e3, "fact_rows_per_s": nfact / t_cpu, "threads": cpu_threads, "runs": nr}
        out["checked"] = check_star(ctx, D, D_, M_, F_, spec, rows, crows, cst)
    del dim, mid, fact, q
    torch.cuda.empty_cache()
    return out


API_Q3 = """ord_f := FILTER orders BY o_orderdate < {date};
cust_f := FILTER customer BY c_mktsegment == "BUILDING";
li_f := FILTER lineitem BY shipdate > {date};
joined := SELECT o_orderkey AS o_orderkey, o_orderdate AS o_orderdate, o_shippriority AS o_shippriority,
                 price AS price, discount AS discount
          FROM li_f JOIN ord_f ON orderkey = o_orderkey JOIN cust_f ON o_custkey = c_custkey;
grouped := SELECT o_orderkey AS k, o_orderdate AS d, o_shippriority AS p,
                  SUM(price * (1 - discount)) AS rev, COUNT(o_orderkey) AS n
           FROM joined GROUP BY o_orderkey, o_orderdate, o_shippriority;
ranked := ORDER grouped BY rev DESC, d ASC;
STORE ranked INTO '{out}' USING ('|') LIMIT 10;
"""


def api_leg(ctx, D, seed, scale, seg_rows, steps, warmup, want_rows, star_ms):
    """The same q3 through the reference's operator API: the three tables registered with
    alenka_table_add_*, and each step one alenkaExecute of the q3 script text (parse, three
    FILTERs, the JOIN, GROUP BY, ORDER BY, STORE ... LIMIT 10 into a file).  The deferred
    statements run as one aln_star_agg top-k at the STORE; the file must hold the fused
    step's rows."""
    import ctypes as C
    import tempfile

    import numpy as np
    import torch

    from alenka_amd.ops import Ops

    ops = Ops(0)
    t0 = time.time()
    ndim, norders = int(150000 * scale), int(1500000 * scale)

    def add(table, name, col, dec=0):
        ops.check(ops.lib.alenka_table_add_int_device(table.encode(), name.encode(), C.c_void_p(col.data_ptr()),
                                                      dec, 1), name)

    key, seg_idx = D.gen_dim(ctx, seed, ndim)
    ops.check(ops.lib.alenka_table_create(b"customer", ndim, seg_rows), "customer")
    add("customer", "c_custkey", key)
    names = np.array([s.encode().ljust(10, b"\0") for s in SEGMENTS], dtype="S10")
    chars = names[seg_idx.cpu().numpy()].tobytes()
    ops.check(ops.lib.alenka_table_add_char(b"customer", b"c_mktsegment", chars, 10), "c_mktsegment")
    del key, seg_idx, chars
    ocols = D.gen_orders(ctx, seed, norders, ndim)
    ops.check(ops.lib.alenka_table_create(b"orders", norders, seg_rows), "orders")
    for name, col in zip(("o_orderkey", "o_custkey", "o_orderdate", "o_shippriority"), ocols):
        add("orders", name, col)
    del ocols, col
    nfact = D.gen_fact_rows(ctx, seed, norders)
    ops.check(ops.lib.alenka_table_create(b"lineitem", nfact, seg_rows), "lineitem")
    for i, (name, dec) in enumerate((("orderkey", 0), ("price", 2), ("discount", 2), ("shipdate", 0))):
        col = D.gen_fact(ctx, seed, norders, cols=tuple(j == i for j in range(4)))[i]
        add("lineitem", name, col, dec)
        del col
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"api: tables registered through alenka_table_add_* in {time.time() - t0:.1f}s")
    out = os.path.join(tempfile.gettempdir(), f"alenka_q3_{os.getpid()}.txt")
    script = API_Q3.format(date=DATE, out=out)

    def step():
        rc = ops.alenkaExecute(script)
        if rc != 0:
            raise RuntimeError(f"alenkaExecute: {ops.lib.alenka_last_error().decode()}")

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t1) * 1000.0 / steps
    path = ops.lib.alenka_last_path().decode()
    with open(out) as f:
        text = f.read()
    os.remove(out)
    want = []
    for k, d, p, rev, n in want_rows:
        frac = str(rev).rjust(5, "0")
        want.append(f"{k}|{d}|{p}|{frac[:-4]}.{frac[-4:]}|{n}|")
    ops.close()
    torch.cuda.empty_cache()
    return {"ms": ms, "fact_rows_per_s": nfact / (ms / 1000.0), "vs_fused_step": ms / star_ms, "path": path,
            "checked": text == "\n".join(want), "steps": steps,
            "route": "alenkaExecute(q3 script) -> emit_filter x3, emit_join, emit_select, emit_order, emit_store"}


def batch_queries(seed, nq, nkeys, price_lo, price_span, ndays):
    """nq filter + GROUP BY queries over the fact table [okey, price, disc, ship, zipf]:
    selectivity log-uniform over [1%, 90%] (a price range, or every 4th query a ship-date
    range), group cardinality log-uniform over [10, nkeys] (key = zipf // div), each
    computing SUM(price * (1 - discount)) and COUNT(*)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    qs = []
    for q in range(nq):
        sel = float(np.exp(rng.uniform(np.log(0.01), np.log(0.9))))
        card = float(np.exp(rng.uniform(np.log(10), np.log(nkeys))))
        div = max(1, int(round(nkeys / card)))
        if q % 4 == 3:
            w = max(1, int(sel * ndays))
            d0 = int(rng.integers(0, ndays - w + 1))
            lo, hi, pred = ymd(d0), ymd(d0 + w - 1), 3
        else:
            w = max(1, int(sel * price_span))
            lo = price_lo + int(rng.integers(0, price_span - w + 1))
            hi, pred = lo + w - 1, 1
        qs.append(dict(pred_col=pred, key_col=4, x_col=1, y_col=2, lo=lo, hi=hi, key_div=div,
                       a0=0, a1=1, b0=100, b1=-1, sel=sel, card=card))
    return qs


def ymd(day):
    """yyyymmdd of `day` days after 1992-01-01."""
    import datetime

    d = datetime.date(1992, 1, 1) + datetime.timedelta(days=int(day))
    return d.year * 10000 + d.month * 100 + d.day


def plain_query(ctx, D, L, cols, q):
    """One query through the per-statement operators: scan_filter -> eval -> group-by."""
    p = q["pred_col"]
    prog = [(L.I_COL, p, 1), (L.I_CONST, 0, q["lo"]), (L.I_CMP, 6, 0),
            (L.I_COL, p, 1), (L.I_CONST, 0, q["hi"]), (L.I_CMP, 5, 0), (L.I_AND, 0, 0)]
    rows = D.scan_filter(ctx, cols, prog)
    key = D.eval_rows(ctx, cols, [(L.I_COL, q["key_col"], 1), (L.I_CONST, 0, q["key_div"]), (L.I_DIV, 0, 0)], rows)
    rev = D.eval_rows(ctx, cols, [(L.I_COL, q["x_col"], 1), (L.I_CONST, 0, 100), (L.I_COL, q["y_col"], 1),
                                  (L.I_SUB, 0, 0), (L.I_MUL, 0, 0)], rows)
    return D.groupby(ctx, [key], [rev, rev], [L.AGG_SUM, L.AGG_COUNT])


def batch_leg(ctx, D, fact, nfact, seed, nq, steps, sort_rows):
    """The batched form: nq queries in one persistent launch (captured in a HIP graph, shape
    chosen by the autotuner) against the same queries one at a time through the operators,
    plus the ORDER BY of a sort_rows-row (int64 key, int32 payload) result."""
    import numpy as np
    import torch

    from alenka_amd import _lib as L

    nkeys = 10_000_000
    t0 = time.time()
    z = D.gen_zipf(ctx, seed, nfact, nkeys, 1.1)
    zc = D.Column.encode(ctx, z, fact[0].seg_rows)
    del z
    torch.cuda.empty_cache()
    cols = list(fact) + [zc]
    qs = batch_queries(seed + 3, nq, nkeys, 90000, 10410001, 2557 + 121)
    log(f"batch: zipf column built in {time.time() - t0:.1f}s; {nq} queries")
    b = D.Batch(ctx, cols, qs, use_graph=True)
    shape, tune_ms = b.autotune(reps=2, use_graph=True)
    log(f"batch: autotuned hot_groups={shape[0]} blocks_per_cu={shape[1]} ({tune_ms:.2f} ms)")
    b.run()
    torch.cuda.synchronize()
    lat = []
    for _ in range(steps):
        t1 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    # every query of the batch against its plain run: same groups, sums and counts
    plain_lat, mismatches = [], 0
    for i, q in enumerate(qs):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pk, pv = plain_query(ctx, D, L, cols, q)
        torch.cuda.synchronize()
        plain_lat.append(time.perf_counter() - t1)
        bk, bs, bc = b.results(i, max(1, pk[0].numel()))
        same = (bk.numel() == pk[0].numel() and torch.equal(bk, pk[0]) and torch.equal(bs, pv[0])
                and torch.equal(bc, pv[1]))
        mismatches += 0 if same else 1
        del pk, pv, bk, bs, bc
    b.close()
    del cols
    torch.cuda.empty_cache()
    lat_ms = np.array(lat) * 1e3
    plain_ms = np.array(plain_lat) * 1e3
    batched_qps = nq / (lat_ms.mean() / 1e3)
    plain_qps = nq / (plain_ms.sum() / 1e3)
    # ORDER BY of a sort_rows-row result: uniform int64 keys, then signed 32-bit values
    order_by = {}
    for kind in ("int64", "int32"):
        keys = sort_input(ctx, D, kind, seed, sort_rows)
        st, ok = time_sort(ctx, D, keys, 1, 3)
        del keys
        torch.cuda.empty_cache()
        sort_s = float(np.median(st)) / 1e3
        order_by[kind] = {"rows": sort_rows, "ms": sort_s * 1e3, "rows_per_s": sort_rows / sort_s,
                          "algorithmic_gbs": 24 * sort_rows / sort_s / 1e9, "sorted_stable": ok}
    return zc, {
        "queries": nq,
        "selectivity_range": [min(q["sel"] for q in qs), max(q["sel"] for q in qs)],
        "group_cardinality_range": [min(q["card"] for q in qs), max(q["card"] for q in qs)],
        "zipf_s": 1.1, "zipf_keys": nkeys,
        "tuning": {"hot_groups": shape[0], "blocks_per_cu": shape[1], "graph": True},
        "batched_qps": batched_qps,
        "batched_ms_per_batch": {"mean": float(lat_ms.mean()), "p50": float(np.percentile(lat_ms, 50)),
                                 "p99": float(np.percentile(lat_ms, 99))},
        "plain_qps": plain_qps,
        "plain_ms_per_query": {"mean": float(plain_ms.mean()), "p50": float(np.percentile(plain_ms, 50)),
                               "p99": float(np.percentile(plain_ms, 99))},
        "speedup_qps": batched_qps / plain_qps,
        "results_match_plain": mismatches == 0,
        "order_by": order_by,
    }


def serve_tables(ctx, D, dim, mid, fact, zc, seed, scales):
    """Device-resident tables of the served mix at each scale: flat tables of 1M x scale
    rows, and prefixes (whole segments) of the 600M-row star tables: orders 1.5M x scale,
    fact 6M x scale (with its Zipf column); the dimension table is used whole."""
    import torch

    seg = fact[0].seg_rows
    flat, facts, stars = {}, {}, {}
    for s in scales:
        n = 1_000_000 * s
        cols = []
        for c, dec in zip(D.gen_flat(ctx, seed + s, n), (0, 0, 2, 2)):
            cols.append(D.Column.encode(ctx, c, seg, decimals=dec))
            del c
        flat[s] = cols
        nf = -(-6_000_000 * s // seg)
        nm = -(-1_500_000 * s // seg)
        facts[s] = [c.prefix(nf) for c in list(fact) + [zc]]
        stars[s] = (dim, [c.prefix(nm) for c in mid], [c.prefix(nf) for c in fact])
    torch.cuda.empty_cache()
    return flat, facts, stars


def serve_trace(seed, rate, duration, scales, seg_codes, nkeys):
    """Poisson arrivals of requests drawn uniformly from the four templates (flat filter +
    GROUP BY; the star pipeline at base and at compressed large scale, with TPC-H q3's
    substitution parameters drawn per request: a market segment of the five and a date in
    1995-03-01..31; filter + GROUP BY on the fact table with Zipf keys), each at a random
    scale."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        tmpl = int(rng.integers(0, 4))
        s = int(scales[rng.integers(0, len(scales))])
        if tmpl == 0:
            day = int(rng.integers(365, 2557 - 365))
            q = dict(pred_col=1, key_col=0, x_col=2, y_col=3, lo=0, hi=ymd(day) - 1, key_div=1,
                     a0=0, a1=1, b0=100, b1=-1)
        elif tmpl == 3:
            q = batch_queries(int(rng.integers(1 << 30)), 1, nkeys, 90000, 10410001, 2557 + 121)[0]
        else:
            q = dict(seg=int(seg_codes[rng.integers(0, len(seg_codes))]),
                     date=19950300 + int(rng.integers(1, 32)))
        out.append((t, tmpl, s, q))


def serve_run(D, srv, tids, stars, trace):
    """Submits the trace at its arrival times from one client thread while a second
    thread waits the requests in order; returns per-request latencies (ms) and the span."""
    import threading

    import numpy as np

    spec = dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1)
    ids = [None] * len(trace)
    ready = [threading.Event() for _ in trace]
    lat = np.zeros(len(trace))
    status = np.zeros(len(trace), np.int64)
    t_last = [0.0]

    def waiter():
        for i in range(len(trace)):
            ready[i].wait()
            r = srv.wait(ids[i], check=False)
            lat[i], status[i] = r.latency_ms, r.status
            srv.release(ids[i])
            t_last[0] = time.perf_counter()

    w = threading.Thread(target=waiter)
    w.start()
    t0 = time.perf_counter()
    for i, (t, tmpl, s, q) in enumerate(trace):
        d = t0 + t - time.perf_counter()
        if d > 0:
            time.sleep(d)
        if tmpl in (1, 2):
            dm, md, fc = stars[s]
            ids[i] = srv.submit_star(dict(cols=dm, ranges=[(1, q["seg"], q["seg"])]),
                                     dict(cols=md, ranges=[(2, -(1 << 63), q["date"] - 1)]),
                                     dict(cols=fc, ranges=[(3, q["date"] + 1, (1 << 63) - 1)]), spec, limit=10)
        else:
            ids[i] = srv.submit(tids[(tmpl, s)], q)
        ready[i].set()
    w.join()
    return lat, status, t_last[0] - t0


def serve_point(D, kw, budget_gb, flat, facts, stars, scales, seg_codes, seed, rate, duration):
    """One run of a form over a Poisson trace: achieved q/s and latency percentiles."""
    import numpy as np
    import torch

    trace = serve_trace(seed, rate, duration, scales, seg_codes, 10_000_000)
    srv = D.Server(budget_bytes=int(budget_gb) << 30, **kw)
    try:
        tids = {}
        for s in scales:
            tids[(0, s)] = srv.add_table(flat[s])
            tids[(3, s)] = srv.add_table(facts[s])
        torch.cuda.synchronize()
        lat, status, span = serve_run(D, srv, tids, stars, trace)
        st = srv.stats()
    finally:
        srv.close()
    return {"offered_qps": rate, "requests": len(trace), "qps": len(trace) / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "failed": int((status != 0).sum()), "batches": st["batches"],
            "batched_queries": st["batched_queries"], "max_concurrent": st["max_concurrent"]}


def serve_leg(ctx, D, dim, mid, fact, zc, code_of, seed, rate, duration, budget_gb):
    """The served form against serial plain execution (one query at a time through the
    operators) on Poisson traces of the four templates.  Protocol: each form's saturation
    throughput from a burst; the latency target is the plain form's p99 at 80 % of its
    saturation; each form's max sustained q/s is the highest offered rate (a sweep around
    its saturation, then two bisection steps between the last rate within the target and
    the first beyond it) whose p99 stays within that target.  Also both forms at `rate`."""
    import torch

    scales = [1, 10, 100]
    flat, facts, stars = serve_tables(ctx, D, dim, mid, fact, zc, seed, scales)
    codes = [int(c) for c in code_of if c >= 0]
    forms = {"served": dict(streams=3, max_batch=64, plain=False),
             "plain": dict(streams=1, max_batch=1, plain=True)}
    pt = lambda f, r, d, sd: serve_point(D, forms[f], budget_gb, flat, facts, stars, scales, codes, sd, r, d)
    out = {"templates": ["flat filter+group-by", "star q3", "star q3 (large)", "fact filter+group-by (Zipf)"],
           "scales": scales, "budget_gb": budget_gb}
    # every point holds at least `nreq` requests so that its p99 is not one request's latency
    nreq = 300
    span = lambda r: min(max(duration, nreq / r), 15.0)
    sat = {}
    for f in forms:
        burst = pt(f, 1e6, nreq / 1e6, seed + 21)  # nreq requests arriving at once
        sat[f] = burst["qps"]
        out[f + "_saturation"] = burst
        log(f"serve {f} saturation: {burst['qps']:.1f} q/s")
    slo_point = pt("plain", 0.8 * sat["plain"], span(0.8 * sat["plain"]), seed + 22)
    slo = slo_point["p99_ms"]
    out["slo"] = {"p99_ms": slo, "from": "plain form at 80% of its saturation", "point": slo_point}
    log(f"serve SLO p99 {slo:.1f} ms")
    for f in forms:
        best = None
        ok_lo, bad_hi = 0.0, None  # offered fractions of saturation within / beyond the target
        for k, frac in enumerate((0.5, 0.7, 0.85, 1.0, 1.2)):
            r = pt(f, frac * sat[f], span(frac * sat[f]), seed + 30 + k)
            log(f"serve {f} at {r['offered_qps']:.1f} q/s: {r['qps']:.1f} q/s, p99 {r['p99_ms']:.1f} ms")
            ok = r["p99_ms"] <= slo and r["failed"] == 0
            if ok and (best is None or r["qps"] > best["qps"]):
                best = r
            if ok and bad_hi is None:
                ok_lo = frac
            elif not ok and bad_hi is None:
                bad_hi = frac
        # two bisection steps between the last rate within the target and the first beyond it
        for k in range(2 if bad_hi is not None else 0):
            frac = (ok_lo + bad_hi) / 2
            r = pt(f, frac * sat[f], span(frac * sat[f]), seed + 40 + k)
            log(f"serve {f} at {r['offered_qps']:.1f} q/s: {r['qps']:.1f} q/s, p99 {r['p99_ms']:.1f} ms")
            if r["p99_ms"] <= slo and r["failed"] == 0:
                ok_lo = frac
                if best is None or r["qps"] > best["qps"]:
                    best = r
            else:
                bad_hi = frac
        out[f + "_at_slo"] = best
    if out["served_at_slo"] and out["plain_at_slo"]:
        out["speedup_qps_at_equal_p99"] = out["served_at_slo"]["qps"] / out["plain_at_slo"]["qps"]
    for f in forms:
        out[f + "_at_rate"] = pt(f, rate, span(rate), seed + 11)
        log(f"serve {f} at {rate} q/s: {out[f + '_at_rate']}")
    del flat, facts, stars
    torch.cuda.empty_cache()
    return out


P_TYPES = [f"{a} {b} {c}" for a in ("STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO")
           for b in ("ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED")
           for c in ("TIN", "NICKEL", "BRASS", "STEEL", "COPPER")]


def timed(fn, reps):
    """Median wall time (s) of fn() over reps runs, synchronised on both sides."""
    import torch

    out, ts = None, []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], out


def widened_rows(ctx, D, seed, n=100_000_000, reps=5):
    """The operators around the hot path on 100M-row synthetic inputs, each checked and timed
    (inputs resident in HBM; medians of `reps` runs, one synchronising C call each):
    LIKE over a char(25) column (dictionary match + set-membership scan), a semi join's row
    reduction, the char ORDER BY key map, and a segment file read into HBM."""
    import numpy as np
    import torch

    from alenka_amd import _lib as L

    out = {}
    g = torch.Generator(device="cuda").manual_seed(seed)
    # LIKE '%BRASS' over dictionary codes of 150 part types (the shape of q2's filter)
    codes = torch.randint(0, len(P_TYPES), (n,), generator=g, device="cuda", dtype=torch.int64)
    col = D.Column.encode(ctx, codes, 1_000_000)
    words = D.like_dict(ctx, P_TYPES, 25, "%BRASS")
    dset = torch.as_tensor(words.view(np.int32)).to("cuda")
    prog, _ = L.compile_rpn([(L.RPN_NAME, 0, 0), (L.RPN_CODESET, len(P_TYPES), dset.data_ptr()), (L.RPN_CMP, 7, 0)], [0])
    t, rows = timed(lambda: D.scan_filter(ctx, [col], prog), reps)
    per_code = torch.bincount(codes, minlength=len(P_TYPES)).cpu().numpy()
    want = int(sum(int(per_code[i]) for i, p in enumerate(P_TYPES) if p.endswith("BRASS")))
    assert rows.numel() == want, "LIKE scan disagrees with the expected count"
    packed = int(col.words.numel()) * 4
    out["like_filter"] = {"rows": n, "passing": want, "ms": t * 1e3, "rows_per_s": n / t,
                          "algorithmic_gbs": (packed + 4 * want) / t / 1e9,
                          "bytes": "packed codes + 4 B per passing row id"}
    del col, rows
    # semi join reduction: 100M probe rows, 10M unique build keys, pairs -> distinct probe rows
    probe = torch.randint(0, 20_000_000, (n,), generator=g, device="cuda", dtype=torch.int64)
    build = torch.randperm(20_000_000, generator=g, device="cuda")[:10_000_000].to(torch.int64)
    tj, (pi, bi) = timed(lambda: D.hash_join(ctx, build, probe), 3)
    tr, semi = timed(lambda: D.join_rows(ctx, pi, n, False), reps)
    assert semi.numel() == pi.numel(), "unique build keys: every matched probe row once"
    out["semi_join"] = {"probe_rows": n, "build_rows": 10_000_000, "pairs": int(pi.numel()),
                        "hash_join_ms": tj * 1e3, "row_reduction_ms": tr * 1e3,
                        "rows_per_s": n / (tj + tr),
                        "reduction_gbs": (4 * pi.numel() + n / 8 + 4 * semi.numel()) / tr / 1e9}
    del probe, build, pi, bi, semi
    # char ORDER BY key map: dictionary rank per row
    ranks = torch.as_tensor(np.argsort(np.argsort(np.array(P_TYPES, dtype=object), kind="stable"),
                                       kind="stable").astype(np.int64)).to("cuda")
    rk = ctx.empty(n)

    def lookup():
        L.check(ctx.lib.aln_dict_lookup(ctx.h, codes.data_ptr(), n, ranks.data_ptr(), ranks.numel(), rk.data_ptr()),
                "aln_dict_lookup")
        return rk

    tl, r = timed(lookup, reps)
    assert torch.equal(r[:n], ranks[codes]), "rank map disagrees"
    out["char_order_key"] = {"rows": n, "ms": tl * 1e3, "rows_per_s": n / tl, "algorithmic_gbs": 16 * n / tl / 1e9}
    del rk, r
    # a segment file (the reference's FOR layout, 32-bit codes) read into HBM
    vals = torch.randint(0, 1 << 30, (n // 4,), generator=g, device="cuda", dtype=torch.int64)
    blob = D.segfile_encode(ctx, vals)
    ts, (dec, lo, hi) = timed(lambda: D.segfile_decode(ctx, blob), 3)
    assert torch.equal(dec, vals), "segment file round trip"
    out["segment_file_read"] = {"rows": n // 4, "file_bytes": len(blob), "ms": ts * 1e3, "rows_per_s": (n // 4) / ts,
                                "gbs_pcie_inclusive": len(blob) / ts / 1e9,
                                "note": "host file bytes copied over PCIe, then decoded on the device"}
    return out


def sort_input(ctx, D, kind, seed, rows):
    """ORDER BY keys: "int64" uniform over the whole range; "int32" uniform signed 32-bit
    values (in the int64 column type), which cross zero."""
    import torch

    if kind == "int64":
        return D.gen_sort_input(ctx, seed, rows, payload=False)[0]
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randint(-(1 << 31), 1 << 31, (rows,), generator=g, device="cuda", dtype=torch.int64)


def time_sort(ctx, D, keys, warmup, steps):
    """Per-call ms of sort_keys over `keys`, and whether the last permutation is a stable
    sort: keys in order, equal keys in row order, every row once (compared, not subtracted:
    int64 differences can wrap)."""
    import torch

    for _ in range(warmup):
        D.sort_keys(ctx, [keys])
    ts = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        perm = D.sort_keys(ctx, [keys])
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    p = perm.long()
    sk = keys[p]
    ok = bool((sk[1:] >= sk[:-1]).all().item())
    ok = ok and bool(((sk[1:] != sk[:-1]) | (p[1:] > p[:-1])).all().item())
    ok = ok and bool((torch.bincount(p, minlength=keys.numel()) == 1).all().item())
    return ts, ok


def op_leg(ctx, D, op, seed, steps, warmup, sort_rows, sort_kinds=("int64", "int32")):
    """One operator alone, for profiling it: `sort` = ORDER BY of sort_rows rows (int64
    keys, then signed 32-bit keys; the permutation is the result); `join` = 100M probe rows
    x 10M unique build keys over a 20M range (the semi join's inner join).  Prints one JSON
    line of per-call times and returns."""
    import numpy as np
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    if op == "sort":
        rec = {"op": "sort", "rows": sort_rows, "steps": steps, "warmup": warmup}
        for kind in sort_kinds:
            keys = sort_input(ctx, D, kind, seed, sort_rows)
            ts, ok = time_sort(ctx, D, keys, warmup, steps)
            rec[kind] = {"ms": ts, "ms_median": float(np.median(ts)), "checked_sorted_stable": ok}
            del keys
        print(json.dumps(rec), flush=True)
        return
    probe = torch.randint(0, 20_000_000, (100_000_000,), generator=g, device="cuda", dtype=torch.int64)
    build = torch.randperm(20_000_000, generator=g, device="cuda")[:10_000_000].to(torch.int64)
    fn = lambda: D.hash_join(ctx, build, probe)  # noqa: E731
    for _ in range(warmup):
        out = fn()
    ts = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    pi, bi = out
    ok = bool((build[bi.long()] == probe[pi.long()]).all().item())
    rec = {"op": "join", "probe_rows": 100_000_000, "build_rows": 10_000_000, "pairs": int(pi.numel()),
           "checked_keys_equal": ok}
    rec.update({"steps": steps, "warmup": warmup, "ms": ts, "ms_median": float(np.median(ts))})
    print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=float, default=100.0, help="x the 6M-row base (100 = 600M fact rows)")
    ap.add_argument("--seg-rows", type=int, default=1000000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs-full", type=int, default=3,
                    help="timed CPU runs of the full-size q3 (after the run that checks the GPU's answer)")
    ap.add_argument("--cpu-budget-s", type=float, default=60.0, help="time cap on those runs (at least one)")
    ap.add_argument("--no-configs", action="store_true", help="skip the 1M flat and 6M star configurations")
    ap.add_argument("--no-api", action="store_true", help="skip q3 through the operator API (alenkaExecute)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-queries leg")
    ap.add_argument("--batch-queries", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--sort-rows", type=int, default=100_000_000)
    ap.add_argument("--sort-kinds", default="int64,int32", help="key kinds of --op sort: int64, int32")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-mix leg")
    ap.add_argument("--no-widened", action="store_true", help="skip the operators around the hot path")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain device copy of 1 GiB (the known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--op", choices=("sort", "join"), default=None,
                    help="time one operator alone (the ORDER BY or the 100M x 10M join) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (collectives and "
              "sharding are out of scope); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from alenka_amd import device as D

    ctx = D.Context(0)
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")  # 1 GiB
        src.fill_(1)
        for _ in range(3):
            dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
        log("copy calibration: 3 x clone of 1 GiB")
    if args.op:
        op_leg(ctx, D, args.op, args.seed, args.steps, args.warmup, args.sort_rows, args.sort_kinds.split(","))
        return
    model, threads = cpu_info()
    host = None if args.no_cpu_baseline else {}
    dim, mid, fact, code_of, nfact = build_tables(ctx, D, args.seed, args.scale, args.seg_rows, host=host)
    D_, M_, F_, spec = q3_spec(code_of, dim, mid, fact)
    run = D.StarQuery(ctx, D_, M_, F_, spec, limit=10).run

    for i in range(args.warmup):
        rows, _, st = run()
    log(f"warmup done ({args.warmup} steps); top row {rows[0] if rows else None}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        rows, _, st = run()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step")

    # dominant kernel: the fact phase, timed with HIP events on the context stream
    ctx.profile(True)
    ctx.profile_reset()
    nprof = max(3, min(10, args.steps))
    for i in range(nprof):
        rows, _, st = run()
    kernels = {}
    for k in ("star_dim", "star_mid", "star_build", "star_fact", "star_topk"):
        ms, n = ctx.profile_get(k)
        kernels[k] = dict(total_ms=ms, launches=n, avg_ms=(ms / n if n else 0.0))
    ctx.profile(False)
    t_fact = kernels["star_fact"]["avg_ms"] / 1000.0

    # the same q3 on the host cores over the same tables: its answer checks the GPU's, its
    # time is the full-size CPU baseline
    full = None
    groups = None
    if host is not None:
        log(f"cpu q3 over the full tables ({threads} threads) ...")
        t_cpu, nr, (crows, cst) = cpu_q3(host, threads, args.cpu_runs_full, budget_s=args.cpu_budget_s)
        checked = check_star(ctx, D, D_, M_, F_, spec, rows, crows, cst)
        groups = cst["groups"]
        full = {"ms": t_cpu * 1e3, "fact_rows_per_s": nfact / t_cpu, "threads": threads, "runs": nr,
                "checked": checked, "cpu_top1": list(crows[0]) if crows else None, "cpu_counts": cst}
        log(f"cpu q3: {t_cpu * 1e3:.0f} ms, checked {checked}")
        del host
        assert checked, "the GPU's q3 answer differs from the CPU port's on the same tables"

    # Algorithmic bytes of the fact phase: what it must move from HBM whatever the kernel -
    # the packed words of the columns read for every row of the scanned segments (join key
    # and predicate columns), one 4-byte word per payload column per joined row, the join
    # index once (mid bitmap over the key range + its rank directory, 16 B per 512 bits) and
    # each touched group's SUM (8 B) and COUNT (4 B) read and written.
    fact_cols = sorted({spec["f_fk"]} | {c for c, _, _ in F_["ranges"]})
    fact_scan, fact_segs = scan_bytes(F_, fact_cols)
    kmin = min(mid[0].hsegs[i].vmin for i in range(mid[0].nsegs))
    kmax = max(mid[0].hsegs[i].vmax for i in range(mid[0].nsegs))
    bitmap_bytes = (kmax - kmin + 1 + 31) // 32 * 4
    join_index = bitmap_bytes + (kmax - kmin + 1 + 511) // 512 * 16
    touched = groups if groups is not None else st.fact_pass
    b_fact = fact_scan + 4 * 2 * st.fact_pass + join_index + 2 * 12 * touched
    achieved = b_fact / t_fact / 1e9 if t_fact > 0 else 0.0
    # the survey's per-unit counts for the whole step (8 B per probe row, whatever serves it)
    all_fact_scan, _ = scan_bytes(F_, [spec["f_fk"], spec["f_x"], spec["f_y"], 3])
    mid_scan, mid_segs = scan_bytes(M_, [0, 1, 2, 3])
    dim_scan, dim_segs = scan_bytes(D_, [0, 1])
    b_pipeline = (dim_scan + mid_scan + all_fact_scan + 8 * st.dim_pass + 8 * st.mid_pass
                  + 8 * st.mid_probe + 8 * st.fact_probe + 16 * st.fact_pass + 10 * 40)
    pipeline_gbs = b_pipeline / (ms_per_step / 1000.0) / 1e9

    api = None
    if not args.no_api:
        api = api_leg(ctx, D, args.seed, args.scale, args.seg_rows, args.steps, args.warmup, rows, ms_per_step)
        log(f"api: q3 through alenkaExecute {api['ms']:.3f} ms/step ({api['vs_fused_step']:.2f}x the fused "
            f"step), path {api['path']}, checked {api['checked']}")

    configs = {"2": {"fact_rows": nfact, "gpu": {"ms": ms_per_step, "fact_rows_per_s": nfact / (ms_per_step / 1e3),
                                                  "path": "aln_star_agg"}}}
    if full is not None:
        configs["2"]["cpu"] = {k: full[k] for k in ("ms", "fact_rows_per_s", "threads", "runs")}
        configs["2"]["checked"] = full["checked"]
    if not args.no_configs:
        cthreads = 0 if args.no_cpu_baseline else threads
        configs["0"] = config0_leg(ctx, D, args.seed, 20, cthreads, 5)
        log(f"config 0: {configs['0']}")
        configs["1"] = config1_leg(ctx, D, args.seed, args.seg_rows, 20, cthreads, 5)
        log(f"config 1: {configs['1']}")

    batched = served = None
    if not args.no_batch:
        zc, batched = batch_leg(ctx, D, fact, nfact, args.seed, args.batch_queries, args.batch_steps,
                                args.sort_rows)
        log(f"batch: {batched['batched_qps']:.1f} q/s batched vs {batched['plain_qps']:.1f} q/s plain")
        if not args.no_serve:
            served = serve_leg(ctx, D, dim, mid, fact, zc, code_of, args.seed, args.serve_rate,
                               args.serve_seconds, args.serve_budget_gb)
        del zc
    widened = None
    if not args.no_widened:
        try:
            widened = widened_rows(ctx, D, args.seed)
            log("widened: " + ", ".join(f"{k} {v.get('ms', v.get('row_reduction_ms', 0)):.2f} ms"
                                        for k, v in widened.items()))
        except Exception as e:  # a failed check is reported in the line, not lost with it
            widened = {"error": f"{type(e).__name__}: {e}"}
            log(f"widened: {widened['error']}")

    # HBM bytes per fact phase, from the two counter passes (FETCH_SIZE, WRITE_SIZE) of this
    # workload, scaled by a plain copy's counts.  The counters cannot be read from inside a
    # run, so the figure comes from the committed profile, and only while the kernels' sources
    # hash to what that profile was measured on (otherwise it is stale and left out).
    traffic = traffic_src = None
    tf = os.path.join(ROOT, "profiles", "star_fact_phase_600M_hbm_traffic.json")
    if (args.scale, args.seg_rows, args.seed) == (100.0, 1000000, 2024) and os.path.exists(tf):
        with open(tf) as f:
            prof = json.load(f)
        if prof.get("sources_sha256") == star_sources_hash():
            traffic = prof["traffic_bytes"]
            traffic_src = ("profiles/star_fact_phase_600M_hbm_traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE of "
                           "this workload, kernel sources sha256 " + prof["sources_sha256"][:12] + ")")
        else:
            traffic_src = "stale: the committed counter profile was taken on other kernel sources"

    value = nfact / (ms_per_step / 1000.0)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "fact rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded SplitMix64 generator, resident in HBM)",
        "config": {
            "workload": ("q3 star pipeline: filter x3 -> join fact->orders->dimension -> group-by "
                         "SUM/COUNT -> top-10, compressed FOR+bit-packed columns, 1M-row segments"),
            "fact_rows": nfact, "orders_rows": mid[0].nrows, "dim_rows": dim[0].nrows,
            "seg_rows": args.seg_rows, "seed": args.seed,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "star_fact phase: k_star_fact_bm + k_fact_resolve",
            "kernel_avg_ms": kernels["star_fact"]["avg_ms"],
            "algorithmic_bytes": b_fact,
            "bytes_definition": ("packed words of the fact columns read for every row of the scanned segments "
                                 "(join key, shipdate) + 4 B per payload column per joined row + the join index "
                                 "once (mid bitmap + rank directory) + 12 B SUM/COUNT read and written per group"),
            "fact_segments_scanned": fact_segs,
        },
        "pipeline": {
            "algorithmic_bytes": b_pipeline,
            "bytes_definition": ("per-unit counts for the whole step: packed bytes of every referenced column + 8 B "
                                 "per build row + 8 B per probe row (whatever serves it) + 16 B per aggregated row"),
            "gbs": pipeline_gbs,
            "frac_of_peak": pipeline_gbs / HBM_PEAK_GBS,
            # per step: every launch under the phase's name (build and top-k are several)
            "kernels_ms": {k: round(v["total_ms"] / nprof, 4) for k, v in kernels.items()},
            "counts": dict(dim_pass=st.dim_pass, mid_probe=st.mid_probe, mid_pass=st.mid_pass,
                           fact_probe=st.fact_probe, fact_pass=st.fact_pass),
            "segments_scanned": dict(dim=dim_segs, mid=mid_segs, fact=fact_segs),
            "top1": list(rows[0]) if rows else None,
        },
        "checked": None if full is None else full["checked"],
        "cpu_baseline": None if full is None else {
            "value": full["fact_rows_per_s"], "unit": "fact rows/s", "cores": full["threads"], "kind": "port",
            "sample": (f"the whole workload (fact {nfact} rows): q3 through oracle/libq3cpu.so, the "
                       f"reference's Thrust algorithms on rocThrust's OpenMP backend, {full['threads']} threads on "
                       f"{model}; median of {full['runs']} runs after one that checked the GPU's top-10 and "
                       f"pass counts")},
        "api": api,
        "configs": configs,
        "batched": batched,
        "served": served,
        "widened": widened,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
