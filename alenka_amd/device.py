"""Python handles over the C ABI: a context bound to torch's current HIP stream, the
segmented bit-packed column layout, and one function per operator entry point.

torch provides device memory and the stream; every computation runs in
libalenka_hip.so.  Shapes are checked here before any launch.
"""
import ctypes as C

import numpy as np

import torch

from . import _lib as L


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Context:
    """An aln_ctx on `device`, enqueueing on torch's current stream of that device."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("alenka_amd needs a HIP device; none is visible")
        self.lib = L.load()
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        self.stream = torch.cuda.current_stream(self.device)
        h = C.c_void_p()
        L.check(self.lib.aln_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(h)),
                "aln_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.aln_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        L.check(self.lib.aln_sync(self.h), "aln_sync")

    # ---- profiling -------------------------------------------------------------------
    def profile(self, on=True):
        L.check(self.lib.aln_profile_enable(self.h, int(on)))

    def profile_reset(self):
        L.check(self.lib.aln_profile_reset(self.h))

    def profile_get(self, kernel):
        ms = C.c_double()
        n = C.c_uint64()
        L.check(self.lib.aln_profile_get(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def empty(self, n, dtype=torch.int64):
        return torch.empty(max(1, int(n)), dtype=dtype, device=self.device)


class Column:
    """One column in the segmented FOR bit-packed layout, resident in HBM."""

    def __init__(self, words, dsegs, hsegs, nrows, seg_rows, decimals, keep=None):
        self.words = words        # torch int32 tensor (the 32-bit words)
        self.dsegs = dsegs        # torch uint8 tensor holding nsegs aln_seg records
        self.hsegs = hsegs        # ctypes array of Seg (host copy of the zone maps)
        self.nrows = int(nrows)
        self.seg_rows = int(seg_rows)
        self.decimals = int(decimals)
        self.keep = keep          # tensors a plain column aliases

    @property
    def nsegs(self):
        return len(self.hsegs)

    @classmethod
    def encode(cls, ctx, vals, seg_rows, decimals=0, force_bits=0):
        """FOR + dense bit-packing of an int64 device tensor, seg_rows rows per segment."""
        assert vals.dtype == torch.int64 and vals.is_cuda and vals.is_contiguous()
        n = vals.numel()
        nsegs = (n + seg_rows - 1) // seg_rows
        hsegs = (L.Seg * max(1, nsegs))()
        words = C.c_uint64()
        L.check(ctx.lib.aln_encode_plan(ctx.h, _p(vals), n, seg_rows, force_bits, hsegs,
                                        C.byref(words)), "aln_encode_plan")
        dwords = ctx.empty(words.value, torch.int32)
        dsegs = ctx.empty(max(1, nsegs) * C.sizeof(L.Seg), torch.uint8)
        L.check(ctx.lib.aln_encode_pack(ctx.h, _p(vals), n, seg_rows, hsegs, _p(dwords), _p(dsegs)),
                "aln_encode_pack")
        hs = (L.Seg * nsegs)()
        C.memmove(hs, hsegs, C.sizeof(L.Seg) * nsegs)
        return cls(dwords, dsegs, hs, n, seg_rows, decimals)

    @classmethod
    def plain(cls, ctx, vals, seg_rows, decimals=0):
        """An int64 device tensor used in place as a 64-bit column (no copy)."""
        assert vals.dtype == torch.int64 and vals.is_cuda and vals.is_contiguous()
        n = vals.numel()
        nsegs = (n + seg_rows - 1) // seg_rows
        dsegs = ctx.empty(max(1, nsegs) * C.sizeof(L.Seg), torch.uint8)
        col = L.Col()
        L.check(ctx.lib.aln_plain_col(ctx.h, _p(vals), n, seg_rows, decimals, _p(dsegs),
                                      C.byref(col)), "aln_plain_col")
        hs = (L.Seg * nsegs)()
        for s in range(nsegs):
            r0 = s * seg_rows
            hs[s].word_off, hs[s].base, hs[s].bits = 2 * r0, 0, 64
            hs[s].nrows = min(seg_rows, n - r0)
            hs[s].vmin, hs[s].vmax = -(1 << 63), (1 << 63) - 1
        return cls(vals.view(torch.int32), dsegs, hs, n, seg_rows, decimals, keep=vals)

    def prefix(self, nsegs):
        """The first nsegs segments as a column of their own (shares the device data)."""
        nsegs = max(1, min(int(nsegs), self.nsegs))
        hs = (L.Seg * nsegs)()
        C.memmove(hs, self.hsegs, C.sizeof(L.Seg) * nsegs)
        rows = sum(hs[s].nrows for s in range(nsegs))
        return Column(self.words, self.dsegs, hs, rows, self.seg_rows, self.decimals, keep=self)

    def desc(self):
        c = L.Col()
        c.words = self.words.data_ptr()
        c.segs = self.dsegs.data_ptr()
        c.nsegs = self.nsegs
        c.decimals = self.decimals
        return c

    def packed_bytes(self, segs=None):
        """Stored bytes of the given segments (all by default): ceil(nrows*bits/8) each."""
        idx = range(self.nsegs) if segs is None else segs
        return sum((self.hsegs[s].nrows * self.hsegs[s].bits + 7) // 8 for s in idx)

    def decode(self, ctx, begin=0, end=None):
        end = self.nrows if end is None else end
        out = ctx.empty(end - begin)
        d = self.desc()
        L.check(ctx.lib.aln_decode(ctx.h, C.byref(d), self.seg_rows, begin, end, _p(out)),
                "aln_decode")
        return out[: end - begin]


def _cols_array(cols):
    arr = (L.Col * len(cols))()
    for i, c in enumerate(cols):
        arr[i] = c.desc()
    return arr


def segfile_decode(ctx, data: bytes):
    """One segment file in the reference's format (FOR / FOR-DELTA) -> (int64 device tensor,
    header lower, header upper)."""
    rows, lo, hi = C.c_uint64(), C.c_int64(), C.c_int64()
    rc = ctx.lib.aln_segfile_decode(ctx.h, data, len(data), None, 0, C.byref(rows), C.byref(lo), C.byref(hi))
    if rc not in (L.ALN_OK, L.ALN_E_CAPACITY):
        L.check(rc, "aln_segfile_decode")
    out = ctx.empty(rows.value)
    L.check(ctx.lib.aln_segfile_decode(ctx.h, data, len(data), _p(out), rows.value, C.byref(rows), C.byref(lo),
                                       C.byref(hi)), "aln_segfile_decode")
    return out[: rows.value], lo.value, hi.value


def segfile_encode(ctx, vals):
    """The FOR segment file the reference's pfor_compress writes for these values (bytes)."""
    n = C.c_uint64()
    L.check(ctx.lib.aln_segfile_encode(ctx.h, _p(vals), vals.numel(), None, 0, C.byref(n)), "aln_segfile_encode")
    buf = C.create_string_buffer(n.value)
    L.check(ctx.lib.aln_segfile_encode(ctx.h, _p(vals), vals.numel(), buf, n.value, C.byref(n)),
            "aln_segfile_encode")
    return buf.raw[: n.value]


def scan_filter(ctx, cols, prog, seg_modes=None, cap=None):
    """Ascending global row ids of the rows that pass `prog` (compiled predicate)."""
    seg_rows = cols[0].seg_rows
    arr = _cols_array(cols)
    ins = L.insn_array(prog)
    n = cols[0].nrows
    cap = n if cap is None else cap
    out = ctx.empty(cap, torch.int32)
    cnt = C.c_uint64()
    modes = None if seg_modes is None else "".join(seg_modes).encode()
    L.check(ctx.lib.aln_scan_filter(ctx.h, arr, len(cols), seg_rows, ins, len(prog), modes,
                                    _p(out), cap, C.byref(cnt)), "aln_scan_filter")
    return out[: cnt.value]


def gather(ctx, col, rows=None, n=None):
    n = rows.numel() if rows is not None else (col.nrows if n is None else n)
    out = ctx.empty(n)
    d = col.desc()
    L.check(ctx.lib.aln_gather(ctx.h, C.byref(d), col.seg_rows, _p(rows), n, _p(out)), "aln_gather")
    return out[:n]


def eval_rows(ctx, cols, prog, rows=None, n=None):
    n = rows.numel() if rows is not None else (cols[0].nrows if n is None else n)
    out = ctx.empty(n)
    arr = _cols_array(cols)
    ins = L.insn_array(prog)
    L.check(ctx.lib.aln_eval(ctx.h, arr, len(cols), cols[0].seg_rows, ins, len(prog), _p(rows), n,
                             _p(out)), "aln_eval")
    return out[:n]


def apply_perm(ctx, vals, perm):
    out = ctx.empty(vals.numel())
    L.check(ctx.lib.aln_apply_perm(ctx.h, _p(vals), _p(perm), vals.numel(), _p(out)), "aln_apply_perm")
    return out[: vals.numel()]


def hash_join(ctx, build_keys, probe_keys, cap=None):
    """(probe_idx, build_idx) int32 tensors in probe-row order."""
    nb, npb = build_keys.numel(), probe_keys.numel()
    cap = max(npb, 1) if cap is None else cap
    while True:
        pi = ctx.empty(cap, torch.int32)
        bi = ctx.empty(cap, torch.int32)
        cnt = C.c_uint64()
        rc = ctx.lib.aln_hash_join(ctx.h, _p(build_keys), nb, _p(probe_keys), npb, _p(pi), _p(bi), cap,
                                   C.byref(cnt))
        if rc == L.ALN_E_CAPACITY and cnt.value > cap:
            cap = cnt.value
            continue
        L.check(rc, "aln_hash_join")
        return pi[: cnt.value], bi[: cnt.value]


def like_dict(ctx, entries, width, pattern):
    """Bit set (numpy uint32 words) of the entries (str list, NUL-padded to width bytes)
    that match the LIKE pattern."""
    import numpy as np

    buf = b"".join(e.encode()[:width].ljust(width, b"\0") for e in entries) or b"\0"
    d = torch.frombuffer(bytearray(buf), dtype=torch.uint8).to(ctx.device)
    n = len(entries)
    words = ctx.empty((n + 63) // 64 * 2 + 1, torch.int32)
    L.check(ctx.lib.aln_like_dict(ctx.h, _p(d), n, width, pattern.encode(), _p(words)), "aln_like_dict")
    return words.cpu().numpy().view(np.uint32)


def join_rows(ctx, idx, nrows, anti, cap=None):
    """Rows 0..nrows-1 that occur in idx (anti=False) or do not (anti=True), ascending
    (int32 tensor)."""
    cap = max(nrows, 1) if cap is None else cap
    out = ctx.empty(max(cap, 1), torch.int32)
    cnt = C.c_uint64()
    L.check(ctx.lib.aln_join_rows(ctx.h, _p(idx), idx.numel(), nrows, int(anti), _p(out), cap, C.byref(cnt)),
            "aln_join_rows")
    return out[: cnt.value]


def groupby(ctx, keys, vals, aggs, order=L.ORDER_KEYS, seed=100, cap=None, n=None):
    """Returns (out_keys list, out_vals list) trimmed to the group count.  No keys: one
    group over n rows (n defaults to the first value column's length)."""
    if n is None:
        n = keys[0].numel() if keys else next(v.numel() for v in vals if v is not None)
    cap = max(n, 1) if cap is None else cap
    ok = [ctx.empty(cap) for _ in keys]
    ov = [ctx.empty(cap) for _ in vals]
    kp = (C.c_void_p * max(1, len(keys)))(*[k.data_ptr() for k in keys])
    vp = (C.c_void_p * max(1, len(vals)))(*[(v.data_ptr() if v is not None else 0) for v in vals])
    okp = (C.c_void_p * max(1, len(ok)))(*[k.data_ptr() for k in ok])
    ovp = (C.c_void_p * max(1, len(ov)))(*[v.data_ptr() for v in ov])
    ag = (C.c_int32 * max(1, len(aggs)))(*aggs)
    g = C.c_uint64()
    L.check(ctx.lib.aln_groupby(ctx.h, kp, len(keys), vp, ag, len(vals), n, order, seed, okp, ovp, cap,
                                C.byref(g)), "aln_groupby")
    return [k[: g.value] for k in ok], [v[: g.value] for v in ov]


def sort_keys(ctx, keys, desc=None):
    n = keys[0].numel()
    desc = desc or [0] * len(keys)
    perm = ctx.empty(n, torch.int32)
    kp = (C.c_void_p * len(keys))(*[k.data_ptr() for k in keys])
    dp = (C.c_int32 * len(keys))(*desc)
    L.check(ctx.lib.aln_sort_keys(ctx.h, kp, dp, len(keys), n, _p(perm)), "aln_sort_keys")
    return perm[:n]


def topk(ctx, keys, desc, k):
    n = keys[0].numel()
    out = ctx.empty(max(k, 1), torch.int32)
    kp = (C.c_void_p * len(keys))(*[x.data_ptr() for x in keys])
    dp = (C.c_int32 * len(keys))(*desc)
    nout = C.c_uint32()
    L.check(ctx.lib.aln_topk(ctx.h, kp, dp, len(keys), n, k, _p(out), C.byref(nout)), "aln_topk")
    return out[: nout.value]


def radix_sort_pairs(ctx, keys_u64, vals_u32, bits):
    L.check(ctx.lib.aln_radix_sort_pairs(ctx.h, _p(keys_u64), _p(vals_u32), keys_u64.numel(), bits),
            "aln_radix_sort_pairs")


def hash_rows(ctx, keys, seed):
    n = keys[0].numel()
    out = ctx.empty(n)
    kp = (C.c_void_p * len(keys))(*[k.data_ptr() for k in keys])
    L.check(ctx.lib.aln_hash_rows(ctx.h, kp, len(keys), n, seed, _p(out)), "aln_hash_rows")
    ctx.sync()
    return out[:n]


# ---- synthetic workloads -----------------------------------------------------------------
def gen_flat(ctx, seed, n):
    t = [ctx.empty(n) for _ in range(4)]
    L.check(ctx.lib.aln_gen_flat(ctx.h, seed, n, *[_p(x) for x in t]), "aln_gen_flat")
    return [x[:n] for x in t]


def gen_dim(ctx, seed, n):
    t = [ctx.empty(n) for _ in range(2)]
    L.check(ctx.lib.aln_gen_dim(ctx.h, seed, n, *[_p(x) for x in t]), "aln_gen_dim")
    return [x[:n] for x in t]


def gen_orders(ctx, seed, n, ndim):
    t = [ctx.empty(n) for _ in range(4)]
    L.check(ctx.lib.aln_gen_orders(ctx.h, seed, n, ndim, *[_p(x) for x in t]), "aln_gen_orders")
    return [x[:n] for x in t]


def gen_fact_rows(ctx, seed, norders):
    n = C.c_uint64()
    L.check(ctx.lib.aln_gen_fact_rows(ctx.h, seed, norders, C.byref(n)), "aln_gen_fact_rows")
    return n.value


def gen_fact(ctx, seed, norders, cols=(True, True, True, True)):
    n = gen_fact_rows(ctx, seed, norders)
    t = [ctx.empty(n) if want else None for want in cols]
    L.check(ctx.lib.aln_gen_fact(ctx.h, seed, norders, *[_p(x) for x in t]), "aln_gen_fact")
    return [x[:n] if x is not None else None for x in t]


def gen_zipf(ctx, seed, n, nkeys, s):
    out = ctx.empty(n)
    L.check(ctx.lib.aln_gen_zipf(ctx.h, seed, n, nkeys, float(s), _p(out)), "aln_gen_zipf")
    return out[:n]


def gen_sort_input(ctx, seed, n, payload=True):
    keys = ctx.empty(n)
    pay = ctx.empty(n, torch.int32) if payload else None
    L.check(ctx.lib.aln_gen_sort_input(ctx.h, seed, n, _p(keys), _p(pay)), "aln_gen_sort_input")
    return keys[:n], (pay[:n] if payload else None)


def dict_first_seen(ctx, vals, nvals):
    code_of = (C.c_int64 * nvals)()
    out = ctx.empty(vals.numel())
    L.check(ctx.lib.aln_dict_first_seen(ctx.h, _p(vals), vals.numel(), nvals, code_of, _p(out)),
            "aln_dict_first_seen")
    return list(code_of), out[: vals.numel()]


def star_agg(ctx, dim, mid, fact, spec, limit=10, want_groups=False, groups_cap=None,
             force_hash=False):
    """Fused filter-join-join-aggregate-top-k.  dim/mid/fact: dicts with 'cols' (list of
    Column), 'ranges' (list of (col, lo, hi)) and optional 'modes'.  spec: dict of the
    column slots d_key, m_key, m_fk, m_g1, m_g2, f_fk, f_x, f_y and a0, a1, b0, b1."""
    a, keep = star_args(dim, mid, fact, spec, limit, (1 if want_groups else 0) | (2 if force_hash else 0))
    top = (L.StarRow * max(1, limit))()
    ntop = C.c_uint32()
    st = L.StarStats()
    groups = None
    cap = 0
    if want_groups:
        cap = groups_cap or mid["cols"][0].nrows
        groups = ctx.empty(max(1, cap) * C.sizeof(L.StarRow), torch.uint8)
    L.check(ctx.lib.aln_star_agg(ctx.h, C.byref(a), top, C.byref(ntop), _p(groups), cap, C.byref(st)),
            "aln_star_agg")
    rows = [(top[i].key, top[i].g1, top[i].g2, top[i].sum, top[i].count) for i in range(ntop.value)]
    g = None
    if want_groups:
        g = groups[: st.groups * C.sizeof(L.StarRow)].view(torch.int64).view(-1, 5)
    return rows, g, st


class StarQuery:
    """A prepared star_agg call: the argument block is built once and each run() is one
    C call (for repeated execution of the same query, e.g. a benchmark loop)."""

    def __init__(self, ctx, dim, mid, fact, spec, limit=10):
        self.ctx = ctx
        self.args, self._keep = star_args(dim, mid, fact, spec, limit, 0)
        self.limit = limit
        self.top = (L.StarRow * max(1, limit))()
        self.ntop = C.c_uint32()
        self.st = L.StarStats()
        self._fn = ctx.lib.aln_star_agg
        self._a = C.byref(self.args)
        self._nt = C.byref(self.ntop)
        self._s = C.byref(self.st)

    def run(self):
        """Returns (rows, None, stats) as star_agg does (no group output)."""
        L.check(self._fn(self.ctx.h, self._a, self.top, self._nt, None, 0, self._s), "aln_star_agg")
        rows = [(self.top[i].key, self.top[i].g1, self.top[i].g2, self.top[i].sum, self.top[i].count)
                for i in range(self.ntop.value)]
        return rows, None, self.st


def star_agg_batch(ctx, queries, limit=10):
    """Several star pipelines over the same tables in one pass of each table.  queries:
    list of (dim, mid, fact, spec) as for star_agg.  Returns a list of (rows, stats)."""
    n = len(queries)
    args = (L.StarArgs * n)()
    keep = []
    for i, (dm, md, fc, spec) in enumerate(queries):
        a, k = star_args(dm, md, fc, spec, limit, 0)
        args[i] = a
        keep.append(k)
    tops = [(L.StarRow * max(1, limit))() for _ in range(n)]
    tp = (C.POINTER(L.StarRow) * n)(*[C.cast(t, C.POINTER(L.StarRow)) for t in tops])
    ntop = (C.c_uint32 * n)()
    st = (L.StarStats * n)()
    L.check(ctx.lib.aln_star_agg_batch(ctx.h, args, n, tp, ntop, st), "aln_star_agg_batch")
    out = []
    for i in range(n):
        rows = [(tops[i][r].key, tops[i][r].g1, tops[i][r].g2, tops[i][r].sum, tops[i][r].count)
                for r in range(ntop[i])]
        out.append((rows, st[i]))
    return out


def star_args(dim, mid, fact, spec, limit, flags):
    """aln_star_args for three table scans; returns (args, host arrays to keep alive)."""
    keep = []

    def scan(t):
        cols = _cols_array(t["cols"])
        rg = (L.Range * max(1, len(t["ranges"])))()
        for i, (c, lo, hi) in enumerate(t["ranges"]):
            rg[i].col, rg[i].lo, rg[i].hi = c, lo, hi
        ts = L.TableScan()
        ts.cols = cols
        ts.ncols = len(t["cols"])
        ts.seg_rows = t["cols"][0].seg_rows
        modes = t.get("modes")
        ts.seg_modes = None if modes is None else "".join(modes).encode()
        ts.ranges = rg
        ts.nranges = len(t["ranges"])
        keep.extend([cols, rg, ts.seg_modes])
        return ts

    a = L.StarArgs()
    a.dim, a.mid, a.fact = scan(dim), scan(mid), scan(fact)
    for k in ("d_key", "m_key", "m_fk", "m_g1", "m_g2", "f_fk", "f_x", "f_y", "a0", "a1", "b0", "b1"):
        setattr(a, k, spec[k])
    a.limit = limit
    a.flags = flags
    return a, keep


class Batch:
    """A prepared batch of filter + GROUP BY queries over one table (aln_batch_*).
    queries: list of dicts with pred_col, key_col, x_col, y_col, lo, hi, key_div, a0, a1,
    b0, b1 (column slots index `cols`)."""

    def __init__(self, ctx, cols, queries, hot_groups=-1, blocks_per_cu=0, use_graph=False):
        self.ctx = ctx
        self.cols = list(cols)
        self._arr = _cols_array(self.cols)
        qs = (L.BatchQuery * len(queries))()
        for i, q in enumerate(queries):
            for k in ("pred_col", "key_col", "x_col", "y_col", "lo", "hi", "key_div", "a0", "a1", "b0", "b1"):
                setattr(qs[i], k, int(q[k]))
        self.nq = len(queries)
        self.tuning = L.BatchTuning(hot_groups, blocks_per_cu, 1 if use_graph else 0)
        h = C.c_void_p()
        L.check(ctx.lib.aln_batch_create(ctx.h, self._arr, len(self.cols), self.cols[0].seg_rows, qs,
                                         self.nq, C.byref(self.tuning), C.byref(h)), "aln_batch_create")
        self.h = h

    def set_tuning(self, hot_groups=-1, blocks_per_cu=0, use_graph=False):
        self.tuning = L.BatchTuning(hot_groups, blocks_per_cu, 1 if use_graph else 0)
        L.check(self.ctx.lib.aln_batch_set_tuning(self.h, C.byref(self.tuning)), "aln_batch_set_tuning")

    def autotune(self, reps=3, use_graph=False):
        best = L.BatchTuning(0, 0, 1 if use_graph else 0)
        ms = C.c_double()
        L.check(self.ctx.lib.aln_batch_autotune(self.h, reps, C.byref(best), C.byref(ms)), "aln_batch_autotune")
        self.tuning = best
        return (best.hot_groups, best.blocks_per_cu), ms.value

    def run(self):
        L.check(self.ctx.lib.aln_batch_run(self.h), "aln_batch_run")

    def results(self, q, cap):
        keys, sums, cnts = self.ctx.empty(cap), self.ctx.empty(cap), self.ctx.empty(cap)
        n = C.c_uint64()
        L.check(self.ctx.lib.aln_batch_results(self.h, q, _p(keys), _p(sums), _p(cnts), cap, C.byref(n)),
                "aln_batch_results")
        k = n.value
        return keys[:k], sums[:k], cnts[:k]

    def close(self):
        if self.h:
            self.ctx.lib.aln_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The served form (aln_server_*): a request loop over device-resident tables with
    continuous batching.  Columns passed to add_table must stay alive while it runs."""

    def __init__(self, device=0, streams=2, max_batch=64, plain=False, budget_bytes=200 << 30,
                 hot_groups=-1, blocks_per_cu=0, use_graph=False):
        self.lib = L.load()
        cfg = L.ServerConfig(streams, max_batch, 1 if plain else 0, 0, budget_bytes,
                             L.BatchTuning(hot_groups, blocks_per_cu, 1 if use_graph else 0))
        h = C.c_void_p()
        L.check(self.lib.aln_server_create(device, C.byref(cfg), C.byref(h)), "aln_server_create")
        self.h = h
        self.tables = []

    def add_table(self, cols):
        arr = _cols_array(cols)
        tid = C.c_int32()
        L.check(self.lib.aln_server_add_table(self.h, arr, len(cols), cols[0].seg_rows, C.byref(tid)),
                "aln_server_add_table")
        self.tables.append(list(cols))
        return tid.value

    def submit(self, table, q):
        bq = L.BatchQuery()
        for k in ("pred_col", "key_col", "x_col", "y_col", "lo", "hi", "key_div", "a0", "a1", "b0", "b1"):
            setattr(bq, k, int(q[k]))
        rid = C.c_uint64()
        L.check(self.lib.aln_server_submit(self.h, table, C.byref(bq), C.byref(rid)), "aln_server_submit")
        return rid.value

    def submit_star(self, dim, mid, fact, spec, limit=10):
        a, keep = star_args(dim, mid, fact, spec, limit, 0)
        rid = C.c_uint64()
        L.check(self.lib.aln_server_submit_star(self.h, C.byref(a), C.byref(rid)), "aln_server_submit_star")
        return rid.value

    def wait(self, rid, check=True):
        r = L.ServerResult()
        rc = self.lib.aln_server_wait(self.h, rid, C.byref(r))
        if check:
            L.check(rc, "aln_server_wait")
        return r

    def fetch(self, rid):
        r = self.wait(rid)
        n = max(1, r.ngroups)
        k, s, c = (np.empty(n, np.int64) for _ in range(3))
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        L.check(self.lib.aln_server_fetch(self.h, rid, ptr(k), ptr(s), ptr(c), n), "aln_server_fetch")
        g = r.ngroups
        return k[:g], s[:g], c[:g]

    def fetch_star(self, rid):
        r = self.wait(rid)
        cap = max(1, r.ngroups)
        top = (L.StarRow * cap)()
        n = C.c_uint32()
        L.check(self.lib.aln_server_fetch_star(self.h, rid, top, cap, C.byref(n)), "aln_server_fetch_star")
        return [(top[i].key, top[i].g1, top[i].g2, top[i].sum, top[i].count) for i in range(n.value)]

    def release(self, rid):
        L.check(self.lib.aln_server_release(self.h, rid), "aln_server_release")

    def stats(self):
        st = L.ServerStats()
        L.check(self.lib.aln_server_stats(self.h, C.byref(st)), "aln_server_stats")
        return {k: getattr(st, k) for k, _ in L.ServerStats._fields_}

    def close(self):
        if self.h:
            self.lib.aln_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
