"""numpy wrappers of the CPU oracle (oracle/liboracle.so): the checker of every parity test."""
import ctypes as C
import os
import subprocess

import numpy as np

from alenka_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF_MURMUR = os.path.join(ROOT, "oracle", "_ref", "libmurmur_ref.so")

_o = None


def load():
    global _o
    if _o is None:
        src = os.path.join(ROOT, "oracle", "oracle.c")
        if not os.path.exists(ORACLE) or os.path.getmtime(ORACLE) < os.path.getmtime(src):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        _o = C.CDLL(ORACLE)
        _o.orc_murmur64a.restype = C.c_uint64
        _o.orc_murmur64a.argtypes = [C.c_char_p, C.c_int, C.c_uint32]
        _o.orc_day_to_ymd.restype = C.c_int32
        _o.orc_gen_fact_rows.restype = C.c_uint64
        _o.orc_bits_for_range.restype = C.c_uint32
        _o.orc_bits_for_range.argtypes = [C.c_int64, C.c_int64]
        _o.orc_unpack.restype = C.c_int64
        _o.orc_pfor_encode.restype = C.c_uint64
        _o.orc_pfor_delta_encode.restype = C.c_uint64
        _o.orc_pfor_decode.restype = C.c_int64
        _o.orc_compact.restype = C.c_uint64
        _o.orc_groupby_sorted.restype = C.c_uint64
        _o.orc_groupby_segmented.restype = C.c_uint64
        _o.orc_join_sortmerge.restype = C.c_uint64
        _o.orc_like.restype = C.c_int
        _o.orc_like.argtypes = [C.c_char_p, C.c_uint32, C.c_char_p]
    return _o


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _ptrs(arrs):
    return (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])


def murmur64a(b: bytes, seed: int) -> int:
    return load().orc_murmur64a(b, len(b), seed)


def day_to_ymd(d):
    return load().orc_day_to_ymd(C.c_int32(d))


def gen_flat(seed, n):
    o = load()
    cols = [np.empty(n, np.int64) for _ in range(4)]
    o.orc_gen_flat(C.c_uint64(seed), C.c_uint64(n), *[_ptr(c) for c in cols])
    return cols


def gen_dim(seed, n):
    cols = [np.empty(n, np.int64) for _ in range(2)]
    load().orc_gen_dim(C.c_uint64(seed), C.c_uint64(n), *[_ptr(c) for c in cols])
    return cols


def gen_orders(seed, n, ndim):
    cols = [np.empty(n, np.int64) for _ in range(4)]
    load().orc_gen_orders(C.c_uint64(seed), C.c_uint64(n), C.c_uint64(ndim), *[_ptr(c) for c in cols])
    return cols


def gen_fact(seed, norders):
    o = load()
    n = o.orc_gen_fact_rows(C.c_uint64(seed), C.c_uint64(norders))
    cols = [np.empty(n, np.int64) for _ in range(4)]
    o.orc_gen_fact(C.c_uint64(seed), C.c_uint64(norders), *[_ptr(c) for c in cols])
    return cols


def like(value, width, pattern):
    """LIKE on one char[width] value (NUL-padded to width) - filter.cu:67-129."""
    v = value.encode() if isinstance(value, str) else bytes(value)
    buf = v[:width].ljust(width, b"\0") + b"\0"
    return bool(load().orc_like(buf, width, pattern.encode()))


def dict_first_seen(vals, nvals):
    code_of = np.empty(nvals, np.int64)
    codes = np.empty(len(vals), np.int64)
    v = np.ascontiguousarray(vals, np.int64)
    load().orc_dict_first_seen(_ptr(v), C.c_uint64(len(v)), C.c_int64(nvals), _ptr(code_of), _ptr(codes))
    return code_of, codes


def pack(vals, base, bits):
    v = np.ascontiguousarray(vals, np.int64)
    words = np.zeros(max(1, (len(v) * bits + 31) // 32), np.uint32)
    load().orc_pack(_ptr(v), C.c_uint64(len(v)), C.c_int64(base), C.c_uint32(bits), _ptr(words))
    return words


def unpack(words, i, base, bits):
    return load().orc_unpack(_ptr(words), C.c_uint64(i), C.c_int64(base), C.c_uint32(bits))


def pfor_encode(vals, delta=False):
    o = load()
    v = np.ascontiguousarray(vals, np.int64)
    fn = o.orc_pfor_delta_encode if delta else o.orc_pfor_encode
    need = fn(_ptr(v), C.c_uint32(len(v)), None, C.c_uint64(0))
    buf = np.zeros(need, np.uint8)
    fn(_ptr(v), C.c_uint32(len(v)), _ptr(buf), C.c_uint64(need))
    return buf.tobytes()


def pfor_decode(buf: bytes, cap):
    out = np.empty(max(1, cap), np.int64)
    b = np.frombuffer(buf, np.uint8).copy()
    n = load().orc_pfor_decode(_ptr(b), C.c_uint64(len(b)), _ptr(out), C.c_uint64(cap))
    if n < 0:
        raise ValueError("bad segment buffer")
    return out[:n]


# ---- the reference's binary table files (written here as the reference writes them) ----------
def data_dictionary_bytes(dd):
    """save_col_data (cm.cu:4057-4080): u64 table count; per table (sorted) u64 name length,
    name, u64 column count; per column (sorted) u64 name length, name, u32 type, u32 length."""
    import struct

    out = [struct.pack("<Q", len(dd))]
    for t in sorted(dd):
        out += [struct.pack("<Q", len(t)), t.encode(), struct.pack("<Q", len(dd[t]))]
        for c in sorted(dd[t]):
            typ, length = dd[t][c]
            out += [struct.pack("<Q", len(c)), c.encode(), struct.pack("<II", typ, length)]
    return b"".join(out)


def parse_data_dictionary(b):
    """load_col_data (cm.cu:4082-4114): the inverse of data_dictionary_bytes."""
    import struct

    p, out = 8, {}
    for _ in range(struct.unpack_from("<Q", b)[0]):
        ln = struct.unpack_from("<Q", b, p)[0]
        t = b[p + 8:p + 8 + ln].decode()
        p += 8 + ln
        nc = struct.unpack_from("<Q", b, p)[0]
        p += 8
        out[t] = {}
        for _ in range(nc):
            ln = struct.unpack_from("<Q", b, p)[0]
            c = b[p + 8:p + 8 + ln].decode()
            p += 8 + ln
            out[t][c] = struct.unpack_from("<II", b, p)
            p += 8
    assert p == len(b)
    return out


def header_bytes(total, nseg, max_recs, max_payload):
    """writeHeader (cm.cu:1251-1263): u64 rows, u32 segments, u32 most rows, u32 largest payload."""
    import struct

    return struct.pack("<QIII", total, nseg, max_recs, max_payload)


def write_binary_table(prefix, columns, seg_sizes, hash_seed=100):
    """Files of one table as the reference's text loader leaves them (CudaSet::compress,
    cm.cu:1034-1240, compress_char 1765-1844): columns are (name, kind, values, arg) with kind
    "int" (arg: True for FOR-DELTA segments), "decimal" (arg: digits) or "char" (values: str
    list, arg: width).  Returns the data dictionary entry {column: (type, length)}."""
    import struct

    bounds = np.concatenate([[0], np.cumsum(seg_sizes)]).astype(np.int64)
    entry = {}
    for name, kind, vals, arg in columns:
        base = f"{prefix}.{name}"
        largest = 0
        if kind == "char":
            width = arg
            code_of, dict_bytes, codes = {}, [], []
            for v in vals:
                b = v.encode()[:width].ljust(width, b"\0")
                if b not in code_of:
                    code_of[b] = len(code_of)
                    dict_bytes.append(b)
                codes.append(code_of[b])
            open(base, "wb").write(b"".join(dict_bytes))
            codes = np.array(codes, np.int64)
            for g in range(len(seg_sizes)):
                seg = codes[bounds[g]:bounds[g + 1]]
                f = pfor_encode(seg)
                largest = max(largest, struct.unpack_from("<I", f)[0])
                open(f"{base}.{g}.idx", "wb").write(f)
                hs = [murmur64a(dict_bytes[c], hash_seed) // 2 for c in seg]
                open(f"{base}.{g}.hash", "wb").write(struct.pack("<I", len(seg)) + struct.pack(f"<{len(hs)}Q", *hs))
            entry[name] = (2, width)
        else:
            v = np.ascontiguousarray(vals, np.int64)
            for g in range(len(seg_sizes)):
                f = pfor_encode(v[bounds[g]:bounds[g + 1]], delta=(kind == "int" and bool(arg)))
                largest = max(largest, struct.unpack_from("<I", f)[0])
                open(f"{base}.{g}", "wb").write(f)
            entry[name] = (0, arg if kind == "decimal" else 0)
        open(f"{base}.header", "wb").write(header_bytes(int(bounds[-1]), len(seg_sizes), int(max(seg_sizes)), largest))
    return entry


def filter_mask(ops, cols, col_dec):
    cols = [np.ascontiguousarray(c, np.int64) for c in cols]
    n = len(cols[0]) if cols else 0
    mask = np.zeros(max(1, n), np.uint8)
    dec = (C.c_int32 * max(1, len(col_dec)))(*col_dec)
    rc = load().orc_filter(L.rpn_array(ops), len(ops), _ptrs(cols), dec, C.c_uint64(n), _ptr(mask))
    if rc != 0:
        raise L.AlnError(rc, "orc_filter")
    return mask[:n].astype(bool)


def eval_expr(ops, cols, col_dec):
    cols = [np.ascontiguousarray(c, np.int64) for c in cols]
    n = len(cols[0])
    out = np.empty(max(1, n), np.int64)
    prec = C.c_int32()
    dec = (C.c_int32 * max(1, len(col_dec)))(*col_dec)
    rc = load().orc_eval(L.rpn_array(ops), len(ops), _ptrs(cols), dec, C.c_uint64(n), _ptr(out),
                         C.byref(prec))
    if rc != 0:
        raise L.AlnError(rc, "orc_eval")
    return out[:n], prec.value


def zone_map_check(ops, mins, maxs, col_dec, not_compressed=False):
    n = max(1, len(col_dec))
    mn = (C.c_int64 * n)(*mins)
    mx = (C.c_int64 * n)(*maxs)
    dec = (C.c_int32 * n)(*col_dec)
    out = C.create_string_buffer(1)
    rc = load().orc_zone_map_check(L.rpn_array(ops), len(ops), mn, mx, dec, int(not_compressed), out)
    if rc != 0:
        raise L.AlnError(rc, "orc_zone_map_check")
    return out.raw.decode()


def compact(mask, row0=0):
    m = np.ascontiguousarray(mask, np.uint8)
    ids = np.empty(max(1, len(m)), np.uint32)
    k = load().orc_compact(_ptr(m), C.c_uint64(len(m)), C.c_uint32(row0), _ptr(ids))
    return ids[:k]


def groupby(keys, vals, aggs, seg_bounds=None, seed=100):
    keys = [np.ascontiguousarray(k, np.int64) for k in keys]
    vals = [np.ascontiguousarray(v, np.int64) for v in vals]
    n = len(keys[0])
    ok = [np.empty(max(1, n), np.int64) for _ in keys]
    ov = [np.empty(max(1, n), np.int64) for _ in vals]
    ag = (C.c_int32 * max(1, len(aggs)))(*aggs)
    o = load()
    if seg_bounds is None or len(seg_bounds) <= 2:
        g = o.orc_groupby_sorted(_ptrs(keys), len(keys), _ptrs(vals), ag, len(vals), C.c_uint64(n),
                                 _ptrs(ok), _ptrs(ov))
    else:
        sb = np.ascontiguousarray(seg_bounds, np.uint64)
        g = o.orc_groupby_segmented(_ptrs(keys), len(keys), _ptrs(vals), ag, len(vals), C.c_uint64(n),
                                    _ptr(sb), len(sb) - 1, C.c_uint32(seed), _ptrs(ok), _ptrs(ov))
    return [k[:g] for k in ok], [v[:g] for v in ov]


def join_sortmerge(lkeys, rkeys, lseg_bounds=None):
    lk = np.ascontiguousarray(lkeys, np.int64)
    rk = np.ascontiguousarray(rkeys, np.int64)
    o = load()
    if lseg_bounds is None:
        sb, ns = np.array([0, len(lk)], np.uint64), 1
    else:
        sb, ns = np.ascontiguousarray(lseg_bounds, np.uint64), len(lseg_bounds) - 1
    need = o.orc_join_sortmerge(_ptr(lk), C.c_uint64(len(lk)), _ptr(rk), C.c_uint64(len(rk)), _ptr(sb),
                                ns, None, None, C.c_uint64(0))
    ol = np.empty(max(1, need), np.uint32)
    orr = np.empty(max(1, need), np.uint32)
    o.orc_join_sortmerge(_ptr(lk), C.c_uint64(len(lk)), _ptr(rk), C.c_uint64(len(rk)), _ptr(sb), ns,
                         _ptr(ol), _ptr(orr), C.c_uint64(need))
    return ol[:need], orr[:need]


def str_sort(values, width, perm, desc=False):
    """Stable sort of the row ids `perm` by char[width] values (str list), unsigned byte order."""
    buf = b"".join(v.encode()[:width].ljust(width, b"\0") for v in values)
    p = np.ascontiguousarray(perm, np.uint32).copy()
    load().orc_str_sort(buf, C.c_uint32(width), C.c_uint64(len(p)), int(desc), _ptr(p))
    return p


def semi_anti(lkeys, rkeys, kind, lseg_bounds=None):
    """kind '1' LEFT SEMI, '2' RIGHT SEMI, '3' LEFT ANTI, '4' RIGHT ANTI; lkeys/rkeys: lists
    of equally long key columns (first = the sort-merge key).  Returns the kept rows."""
    lk = [np.ascontiguousarray(k, np.int64) for k in lkeys]
    rk = [np.ascontiguousarray(k, np.int64) for k in rkeys]
    nl, nr = len(lk[0]), len(rk[0])
    if lseg_bounds is None:
        sb, ns = np.array([0, nl], np.uint64), 1
    else:
        sb, ns = np.ascontiguousarray(lseg_bounds, np.uint64), len(lseg_bounds) - 1
    out = np.empty(max(1, nl, nr), np.uint32)
    o = load()
    o.orc_semi_anti.restype = C.c_uint64
    n = o.orc_semi_anti(_ptrs(lk), _ptrs(rk), len(lk), C.c_uint64(nl), C.c_uint64(nr), _ptr(sb), ns,
                        C.c_char(kind.encode()), _ptr(out))
    return out[:n]


def order(keys, desc):
    keys = [np.ascontiguousarray(k, np.int64) for k in keys]
    n = len(keys[0])
    perm = np.empty(max(1, n), np.uint32)
    d = (C.c_int32 * len(keys))(*desc)
    load().orc_order(_ptrs(keys), d, len(keys), C.c_uint64(n), _ptr(perm))
    return perm[:n]


def kahan_sum(v):
    a = np.ascontiguousarray(v, np.float64)
    out = C.c_double()
    load().orc_kahan_sum_f64(_ptr(a), C.c_uint64(len(a)), C.byref(out))
    return out.value
