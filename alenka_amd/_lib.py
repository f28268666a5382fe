"""ctypes binding of libalenka_hip.so (the C ABI declared in include/alenka_hip.h).

The library is loaded from the package directory only; a missing or unloadable library
raises immediately (there is no fallback implementation).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalenka_hip.so")

ALN_OK = 0
ALN_E_ARG = -1
ALN_E_HIP = -2
ALN_E_NOMEM = -3
ALN_E_UNSUPPORTED = -4
ALN_E_CAPACITY = -5
SCAN_NOALLOC_MAX_TILES = 262144  # ALN_SCAN_NOALLOC_MAX_TILES: tiles of 4096 rows

RPN_NAME, RPN_NUMBER, RPN_STRING, RPN_CMP, RPN_AND, RPN_OR = 1, 2, 3, 4, 5, 6
RPN_ADD, RPN_MINUS, RPN_MUL, RPN_DIV, RPN_CODESET, RPN_FLOAT = 7, 8, 9, 10, 11, 12
I_COL, I_CONST, I_ADD, I_SUB, I_MUL, I_DIV, I_CMP, I_AND, I_OR, I_INSET = range(1, 11)
I_I2F, I_FADD, I_FSUB, I_FMUL, I_FDIV, I_FCMP = range(11, 17)
DEC_F64 = -1  # col_dec of an f64 column slot
AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX, AGG_FSUM, AGG_FMIN, AGG_FMAX = 0, 2, 4, 5, 6, 7, 8
ORDER_KEYS, ORDER_HASH, ORDER_NONE = 0, 1, 2
MAX_INSN, MAX_STACK, MAX_COLS, MAX_KEYS, MAX_AGGS, MAX_RANGES = 64, 16, 16, 4, 8, 8


class AlnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"alenka status {code}: {msg}")
        self.code = code


class Rpn(C.Structure):
    _fields_ = [("type", C.c_int32), ("arg", C.c_int32), ("val", C.c_int64)]


class Insn(C.Structure):
    _fields_ = [("op", C.c_int32), ("a", C.c_int32), ("k", C.c_int64)]


class Seg(C.Structure):
    _fields_ = [("word_off", C.c_uint64), ("base", C.c_int64), ("vmin", C.c_int64),
                ("vmax", C.c_int64), ("bits", C.c_uint32), ("nrows", C.c_uint32)]


class Col(C.Structure):
    _fields_ = [("words", C.c_void_p), ("segs", C.c_void_p), ("nsegs", C.c_uint32),
                ("decimals", C.c_int32)]


class Range(C.Structure):
    _fields_ = [("col", C.c_int32), ("pad", C.c_int32), ("lo", C.c_int64), ("hi", C.c_int64)]


class TableScan(C.Structure):
    _fields_ = [("cols", C.POINTER(Col)), ("ncols", C.c_int32), ("seg_rows", C.c_uint32),
                ("seg_modes", C.c_char_p), ("ranges", C.POINTER(Range)), ("nranges", C.c_int32)]


class StarArgs(C.Structure):
    _fields_ = [("dim", TableScan), ("mid", TableScan), ("fact", TableScan),
                ("d_key", C.c_int32),
                ("m_key", C.c_int32), ("m_fk", C.c_int32), ("m_g1", C.c_int32), ("m_g2", C.c_int32),
                ("f_fk", C.c_int32), ("f_x", C.c_int32), ("f_y", C.c_int32),
                ("a0", C.c_int64), ("a1", C.c_int64), ("b0", C.c_int64), ("b1", C.c_int64),
                ("limit", C.c_uint32), ("flags", C.c_uint32)]


class StarRow(C.Structure):
    _fields_ = [("key", C.c_int64), ("g1", C.c_int64), ("g2", C.c_int64), ("sum", C.c_int64),
                ("count", C.c_int64)]


class StarStats(C.Structure):
    _fields_ = [("dim_pass", C.c_uint64), ("mid_probe", C.c_uint64), ("mid_pass", C.c_uint64),
                ("fact_probe", C.c_uint64),
                ("fact_pass", C.c_uint64), ("groups", C.c_uint64), ("ms_dim", C.c_double), ("ms_mid", C.c_double),
                ("ms_fact", C.c_double), ("ms_topk", C.c_double)]


class BatchQuery(C.Structure):
    _fields_ = [("pred_col", C.c_int32), ("key_col", C.c_int32), ("x_col", C.c_int32), ("y_col", C.c_int32),
                ("lo", C.c_int64), ("hi", C.c_int64), ("key_div", C.c_int64),
                ("a0", C.c_int64), ("a1", C.c_int64), ("b0", C.c_int64), ("b1", C.c_int64)]


class BatchTuning(C.Structure):
    _fields_ = [("hot_groups", C.c_int32), ("blocks_per_cu", C.c_int32), ("use_graph", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("streams", C.c_int32), ("max_batch", C.c_int32), ("plain", C.c_int32), ("reserved", C.c_int32),
                ("budget_bytes", C.c_uint64), ("tune", BatchTuning)]


class ServerResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("batch_size", C.c_int32), ("reserved", C.c_int32),
                ("batch", C.c_uint64), ("ngroups", C.c_uint64), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "batches", "batched_queries",
                                          "max_concurrent", "pending", "mem_in_use")]


P = C.c_void_p
U64, I64, U32, I32 = C.c_uint64, C.c_int64, C.c_uint32, C.c_int32
PU64, PI64, PU32, PI32 = C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "aln_last_error": (C.c_char_p, []),
    "aln_version": (C.c_char_p, []),
    "aln_create": (C.c_int, [C.c_int, P, C.POINTER(P)]),
    "aln_destroy": (None, [P]),
    "aln_set_stream": (C.c_int, [P, P]),
    "aln_sync": (C.c_int, [P]),
    "aln_profile_enable": (C.c_int, [P, C.c_int]),
    "aln_profile_reset": (C.c_int, [P]),
    "aln_profile_get": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double), PU64]),
    "aln_compile_rpn": (C.c_int, [C.POINTER(Rpn), C.c_int, PI32, C.c_int, C.c_int,
                                  C.POINTER(Insn), C.POINTER(C.c_int), PI32]),
    "aln_zone_map_check": (C.c_int, [C.POINTER(Rpn), C.c_int, PI64, PI64, PI32, C.c_int,
                                     C.c_int, C.c_char_p]),
    "aln_encode_plan": (C.c_int, [P, P, U64, U32, U32, C.POINTER(Seg), PU64]),
    "aln_encode_pack": (C.c_int, [P, P, U64, U32, C.POINTER(Seg), P, P]),
    "aln_plain_col": (C.c_int, [P, P, U64, U32, I32, P, C.POINTER(Col)]),
    "aln_decode": (C.c_int, [P, C.POINTER(Col), U32, U64, U64, P]),
    "aln_segfile_decode": (C.c_int, [P, C.c_char_p, U64, P, U64, PU64, PI64, PI64]),
    "aln_segfile_encode": (C.c_int, [P, P, U64, C.c_char_p, U64, PU64]),
    "aln_scan_filter": (C.c_int, [P, C.POINTER(Col), C.c_int, U32, C.POINTER(Insn), C.c_int,
                                  C.c_char_p, P, U64, PU64]),
    "aln_scan_filter_noalloc": (C.c_int, [P, C.POINTER(Col), C.c_int, U32, C.POINTER(Insn), C.c_int,
                                          C.c_char_p, P, U64, PU64]),
    "aln_row_segments": (C.c_int, [P, P, U64, P, P, U32, U32, P]),
    "aln_f64_order_keys": (C.c_int, [P, P, U64, P]),
    "aln_f64_unorder_keys": (C.c_int, [P, P, U64, P]),
    "aln_gather": (C.c_int, [P, C.POINTER(Col), U32, P, U64, P]),
    "aln_eval": (C.c_int, [P, C.POINTER(Col), C.c_int, U32, C.POINTER(Insn), C.c_int, P, U64, P]),
    "aln_apply_perm": (C.c_int, [P, P, P, U64, P]),
    "aln_hash_join": (C.c_int, [P, P, U64, P, U64, P, P, U64, PU64]),
    "aln_join_rows": (C.c_int, [P, P, U64, U64, C.c_int, P, U64, PU64]),
    "aln_dict_lookup": (C.c_int, [P, P, U64, P, U64, P]),
    "aln_like_dict": (C.c_int, [P, P, U64, C.c_uint32, C.c_char_p, P]),
    "aln_groupby": (C.c_int, [P, C.POINTER(P), C.c_int, C.POINTER(P), PI32, C.c_int, U64,
                              C.c_int, U32, C.POINTER(P), C.POINTER(P), U64, PU64]),
    "aln_sort_keys": (C.c_int, [P, C.POINTER(P), PI32, C.c_int, U64, P]),
    "aln_topk": (C.c_int, [P, C.POINTER(P), PI32, C.c_int, U64, U32, P, PU32]),
    "aln_radix_sort_pairs": (C.c_int, [P, P, P, U64, U32]),
    "aln_murmur64a": (C.c_uint64, [C.c_char_p, C.c_int, U32]),
    "aln_hash_rows": (C.c_int, [P, C.POINTER(P), C.c_int, U64, U32, P]),
    "aln_star_agg": (C.c_int, [P, C.POINTER(StarArgs), C.POINTER(StarRow), PU32, P, U64,
                               C.POINTER(StarStats)]),
    "aln_batch_create": (C.c_int, [P, C.POINTER(Col), C.c_int, U32, C.POINTER(BatchQuery), C.c_int,
                                   C.POINTER(BatchTuning), C.POINTER(P)]),
    "aln_batch_set_tuning": (C.c_int, [P, C.POINTER(BatchTuning)]),
    "aln_batch_run": (C.c_int, [P]),
    "aln_batch_results": (C.c_int, [P, C.c_int, P, P, P, U64, PU64]),
    "aln_batch_results_host": (C.c_int, [P, C.c_int, PI64, PI64, PI64, U64, PU64]),
    "aln_batch_autotune": (C.c_int, [P, C.c_int, C.POINTER(BatchTuning), C.POINTER(C.c_double)]),
    "aln_batch_destroy": (None, [P]),
    "aln_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfig), C.POINTER(P)]),
    "aln_server_add_table": (C.c_int, [P, C.POINTER(Col), C.c_int, U32, PI32]),
    "aln_server_submit": (C.c_int, [P, I32, C.POINTER(BatchQuery), PU64]),
    "aln_server_submit_star": (C.c_int, [P, C.POINTER(StarArgs), PU64]),
    "aln_server_wait": (C.c_int, [P, U64, C.POINTER(ServerResult)]),
    "aln_server_fetch": (C.c_int, [P, U64, PI64, PI64, PI64, U64]),
    "aln_server_fetch_star": (C.c_int, [P, U64, C.POINTER(StarRow), U32, PU32]),
    "aln_server_release": (C.c_int, [P, U64]),
    "aln_server_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "aln_server_destroy": (None, [P]),
    "aln_star_agg_batch": (C.c_int, [P, C.POINTER(StarArgs), C.c_int, C.POINTER(C.POINTER(StarRow)), PU32,
                                     C.POINTER(StarStats)]),
    "aln_gen_flat": (C.c_int, [P, U64, U64, P, P, P, P]),
    "aln_gen_dim": (C.c_int, [P, U64, U64, P, P]),
    "aln_gen_orders": (C.c_int, [P, U64, U64, U64, P, P, P, P]),
    "aln_gen_fact_rows": (C.c_int, [P, U64, U64, PU64]),
    "aln_gen_fact": (C.c_int, [P, U64, U64, P, P, P, P]),
    "aln_gen_zipf": (C.c_int, [P, U64, U64, U64, C.c_double, P]),
    "aln_gen_sort_input": (C.c_int, [P, U64, U64, P, P]),
    "aln_dict_first_seen": (C.c_int, [P, P, U64, I64, PI64, P]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the library and bind every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `python -m alenka_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != ALN_OK:
        msg = _lib.aln_last_error().decode() if _lib is not None else ""
        raise AlnError(rc, f"{what}: {msg}" if what else msg)
    return rc


def rpn_array(ops):
    """ops: list of (type, arg, val) tuples -> ctypes array."""
    arr = (Rpn * len(ops))()
    for i, (t, a, v) in enumerate(ops):
        arr[i].type, arr[i].arg, arr[i].val = t, a, v
    return arr


def compile_rpn(ops, col_dec, mode=0):
    """Compile reference RPN (list of (type, arg, val)) into device instructions (host call)."""
    lib = load()
    rp = rpn_array(ops)
    dec = (C.c_int32 * max(1, len(col_dec)))(*col_dec)
    out = (Insn * MAX_INSN)()
    n = C.c_int(0)
    prec = C.c_int32(0)
    check(lib.aln_compile_rpn(rp, len(ops), dec, len(col_dec), mode, out, C.byref(n), C.byref(prec)),
          "aln_compile_rpn")
    return [(out[i].op, out[i].a, out[i].k) for i in range(n.value)], prec.value


def zone_map_check(ops, mins, maxs, col_dec, not_compressed=False):
    lib = load()
    rp = rpn_array(ops)
    n = max(1, len(col_dec))
    mn = (C.c_int64 * n)(*mins)
    mx = (C.c_int64 * n)(*maxs)
    dec = (C.c_int32 * n)(*col_dec)
    out = C.create_string_buffer(1)
    check(lib.aln_zone_map_check(rp, len(ops), mn, mx, dec, len(col_dec), int(not_compressed), out),
          "aln_zone_map_check")
    return out.raw.decode()


def murmur64a(data: bytes, seed: int) -> int:
    return load().aln_murmur64a(data, len(data), seed)


def insn_array(prog):
    arr = (Insn * len(prog))()
    for i, (op, a, k) in enumerate(prog):
        arr[i].op, arr[i].a, arr[i].k = op, a, k
    return arr
