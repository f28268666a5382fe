"""ctypes binding of the operator layer (include/alenka_ops.h): the reference's emit_*
entry points over HBM-resident tables.  Statements are issued as the reference's parser
issues them: expression and select-list emitters first, then one statement call."""
import ctypes as C

import numpy as np

from . import _lib as L

_SIG = {
    "alenka_init": (C.c_int, [C.c_int]),
    "alenka_close": (C.c_int, []),
    "alenka_status": (C.c_int, []),
    "alenka_last_error": (C.c_char_p, []),
    "alenka_set_hash_seed": (None, [C.c_uint32]),
    "alenka_table_create": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32]),
    "alenka_table_add_int": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.c_int32, C.c_int]),
    "alenka_table_add_int_device": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.c_int32, C.c_int]),
    "alenka_table_add_char": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32]),
    "alenka_rows": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64)]),
    "alenka_column_count": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
    "alenka_column_info": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "alenka_fetch_int": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint64]),
    "alenka_fetch_char": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint32]),
    "alenka_drop": (C.c_int, [C.c_char_p]),
    "alenka_last_path": (C.c_char_p, []),
    "alenka_table_add_float": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p]),
    "alenka_fetch_f64": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint64]),
    "emit_float": (None, [C.c_double]),
    "emit_name": (None, [C.c_char_p]),
    "emit_number": (None, [C.c_int64]),
    "emit_decimal": (None, [C.c_char_p]),
    "emit_string": (None, [C.c_char_p]),
    "emit_cmp": (None, [C.c_int]),
    "emit_and": (None, []), "emit_or": (None, []), "emit_add": (None, []), "emit_minus": (None, []),
    "emit_mul": (None, []), "emit_div": (None, []), "emit_eq": (None, []), "emit_count": (None, []),
    "emit_sum": (None, []), "emit_average": (None, []), "emit_min": (None, []), "emit_max": (None, []),
    "emit_sel_name": (None, [C.c_char_p]),
    "emit_var_asc": (None, [C.c_char_p]),
    "emit_var_desc": (None, [C.c_char_p]),
    "emit_join_tab": (None, [C.c_char_p, C.c_char]),
    "emit_limit": (None, [C.c_int]),
    "clean_queues": (None, []),
    "process_error": (None, [C.c_int, C.c_char_p]),
    "emit_filter": (None, [C.c_char_p, C.c_char_p]),
    "emit_select": (None, [C.c_char_p, C.c_char_p, C.c_int]),
    "emit_join": (None, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int]),
    "emit_order": (None, [C.c_char_p, C.c_char_p, C.c_int, C.c_int]),
    "emit_store": (None, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "emit_display": (None, [C.c_char_p, C.c_char_p]),
    "emit_store_binary": (None, [C.c_char_p, C.c_char_p, C.c_int]),
    "emit_load_binary": (None, [C.c_char_p, C.c_char_p, C.c_int]),
    "alenka_set_process_count": (None, [C.c_uint64]),
    "alenka_load_col_data": (C.c_int, [C.c_char_p]),
    "alenka_save_col_data": (C.c_int, [C.c_char_p]),
    "alenkaExecute": (C.c_int, [C.c_char_p]),
}


class Ops:
    """Thin wrapper: attribute access returns the bound C function; strings are encoded."""

    def __init__(self, device=0):
        self.lib = L.load()
        for name, (res, args) in _SIG.items():
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = res, args
        self.check(self.lib.alenka_init(device), "alenka_init")

    def check(self, rc=None, what=""):
        rc = self.lib.alenka_status() if rc is None else rc
        if rc != 0:
            raise L.AlnError(rc, f"{what}: {self.lib.alenka_last_error().decode()}")

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*args):
            a = [x.encode() if isinstance(x, str) else x for x in args]
            return fn(*a)

        return call

    def table(self, name, nrows, seg_rows, columns):
        """columns: list of (name, array, decimals) for ints, or (name, list[str], width) with
        a str list for char columns."""
        self.check(self.lib.alenka_table_create(name.encode(), nrows, seg_rows), "table")
        for col, vals, extra in columns:
            if isinstance(vals, (list, tuple)) and vals and isinstance(vals[0], str):
                buf = b"".join(s.encode().ljust(extra, b"\0")[:extra] for s in vals)
                self.check(self.lib.alenka_table_add_char(name.encode(), col.encode(), buf, extra), col)
            elif extra == "f64":
                a = np.ascontiguousarray(vals, np.float64)
                self.check(self.lib.alenka_table_add_float(name.encode(), col.encode(),
                                                           a.ctypes.data_as(C.c_void_p)), col)
            else:
                a = np.ascontiguousarray(vals, np.int64)
                self.check(self.lib.alenka_table_add_int(name.encode(), col.encode(),
                                                         a.ctypes.data_as(C.c_void_p), extra, 1), col)

    def rows(self, var):
        n = C.c_uint64()
        self.check(self.lib.alenka_rows(var.encode(), C.byref(n)), "rows")
        return n.value

    def fetch(self, var, col):
        n = self.rows(var)
        out = np.empty(max(1, n), np.int64)
        self.check(self.lib.alenka_fetch_int(var.encode(), col.encode(), out.ctypes.data_as(C.c_void_p), n),
                   "fetch")
        return out[:n]

    def fetch_f64(self, var, col):
        n = self.rows(var)
        out = np.empty(max(1, n), np.float64)
        self.check(self.lib.alenka_fetch_f64(var.encode(), col.encode(), out.ctypes.data_as(C.c_void_p), n),
                   "fetch_f64")
        return out[:n]

    def columns(self, var):
        n = C.c_int()
        self.check(self.lib.alenka_column_count(var.encode(), C.byref(n)))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(128)
            t, d = C.c_int(), C.c_int()
            self.check(self.lib.alenka_column_info(var.encode(), i, buf, 128, C.byref(t), C.byref(d)))
            out.append((buf.value.decode(), t.value, d.value))
        return out

    def close(self):
        self.lib.alenka_close()
