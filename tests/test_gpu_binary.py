"""Binary tables in the reference's on-disk format (data.dictionary, .header, segment files,
char dictionaries; emit_load_binary / emit_store_binary, operators.cu:2446-2562): files laid
out as the reference's loader writes them load with their own segment boundaries, and the
files emit_store_binary writes read back through the oracle and through emit_load_binary."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops(ctx):
    from alenka_amd.ops import Ops

    o = Ops(0)
    yield o
    o.close()


def fetch_chars(ops, var, col, width):
    n = ops.rows(var)
    buf = C.create_string_buffer(max(1, n * width))
    ops.check(ops.lib.alenka_fetch_char(var.encode(), col.encode(), buf, n, width), "fetch_char")
    raw = buf.raw[: n * width]
    return [raw[i * width:(i + 1) * width].rstrip(b"\0").decode() for i in range(n)]


SEGS = [1000, 700, 1300]  # uneven, as text loads cut segments by bytes


def fixture_table(tmp_path):
    r = np.random.default_rng(31)
    n = sum(SEGS)
    key = r.integers(0, 400, n).astype(np.int64)
    date = np.sort(r.integers(19920101, 19981231, n)).astype(np.int64)  # sorted: FOR-DELTA segments
    price = r.integers(100, 10000000, n).astype(np.int64)
    seg = [["BUILDING", "AUTOMOBILE", "MACHINERY", "HOUSEHOLD", "FURNITURE"][i] for i in r.integers(0, 5, n)]
    prefix = str(tmp_path / "lineitem")
    entry = O.write_binary_table(prefix, [("l_key", "int", key, False), ("l_date", "int", date, True),
                                          ("l_price", "decimal", price, 2), ("l_seg", "char", seg, 10)], SEGS)
    dd = str(tmp_path / "data.dictionary")
    open(dd, "wb").write(O.data_dictionary_bytes({prefix: entry}))
    return prefix, dd, dict(key=key, date=date, price=price, seg=seg)


def test_load_reference_files(ops, tmp_path):
    prefix, dd, t = fixture_table(tmp_path)
    assert ops.lib.alenka_load_col_data(dd.encode()) == 0, ops.lib.alenka_last_error()
    ops.emit_load_binary("L", prefix, 0)
    ops.check(None, "emit_load_binary")
    assert ops.rows("L") == sum(SEGS)
    for c in ("key", "date", "price"):
        assert np.array_equal(ops.fetch("L", "l_" + c), t[c])
    assert fetch_chars(ops, "L", "l_seg", 10) == t["seg"]
    cols = {c[0]: c for c in ops.columns("L")}
    assert cols["l_price"][2] == 2 and cols["l_seg"][1] == 2
    # a filter on the char column and a group-by: three file segments, so the groups come out
    # in the hash-merge order of a multi-segment set (merge.cu:138-203)
    rc = ops.alenkaExecute(f'F := FILTER L BY l_seg == "MACHINERY" AND l_date < 19960101;\n'
                           f'G := SELECT l_key AS k, SUM(l_price) AS s FROM F GROUP BY l_key;')
    assert rc == 0, ops.lib.alenka_last_error()
    sel = (np.array(t["seg"]) == "MACHINERY") & (t["date"] < 19960101)
    rows = np.nonzero(sel)[0]
    bounds = np.concatenate([[0], np.cumsum(SEGS)])
    seg_of = np.searchsorted(bounds, rows, side="right") - 1
    fb = [int(np.searchsorted(seg_of, s)) for s in range(len(SEGS))] + [len(rows)]
    wk, wv = O.groupby([t["key"][sel]], [t["price"][sel]], [L.AGG_SUM], seg_bounds=fb, seed=100)
    assert np.array_equal(ops.fetch("G", "k"), wk[0])
    assert np.array_equal(ops.fetch("G", "s"), wv[0])


def test_join_pair_order_follows_file_segments(ops, tmp_path):
    prefix, dd, t = fixture_table(tmp_path)
    assert ops.lib.alenka_load_col_data(dd.encode()) == 0
    ops.emit_load_binary("L", prefix, 0)
    ops.check(None, "emit_load_binary")
    dim = np.arange(0, 400, 3, dtype=np.int64)
    ops.table("D", len(dim), 1 << 20, [("d_key", dim, 0), ("d_val", dim * 10, 0)])
    rc = ops.alenkaExecute("J := SELECT l_date AS dt, d_val AS v FROM L JOIN D ON l_key = d_key;")
    assert rc == 0, ops.lib.alenka_last_error()
    bounds = np.concatenate([[0], np.cumsum(SEGS)])
    ol, orr = O.join_sortmerge(t["key"], dim, bounds)
    assert np.array_equal(ops.fetch("J", "dt"), t["date"][ol])
    assert np.array_equal(ops.fetch("J", "v"), (dim * 10)[orr])


def test_store_binary_round_trip(ops, tmp_path, monkeypatch):
    """STORE ... BINARY writes the reference's files (checked with the oracle's decoder and
    murmur hash), the data dictionary saves and reloads, and the stored table loads by name."""
    monkeypatch.chdir(tmp_path)
    r = np.random.default_rng(5)
    n = 7000
    ids = np.arange(n, dtype=np.int64)
    amt = r.integers(-500000, 500000, n).astype(np.int64)
    words = ["PROMO", "AUTOMOBILE", "BUILDING", "", "MACHINERY", "A"]
    name = [words[i] for i in r.integers(0, len(words), n)]
    ops.table("S", n, 2048, [("id", ids, 0), ("amt", amt, 2), ("name", name, 12)])
    ops.lib.alenka_set_process_count(3000)
    rc = ops.alenkaExecute('F := FILTER S BY amt > 0;\nSTORE F INTO \'btab\' BINARY;')
    assert rc == 0, ops.lib.alenka_last_error()
    ops.lib.alenka_set_process_count(1000000000)
    keep = amt > 0
    k = int(keep.sum())
    segs = [min(3000, k - g * 3000) for g in range((k + 2999) // 3000)]
    bounds = np.concatenate([[0], np.cumsum(segs)])
    # the files, read with the oracle
    for col, want in (("id", ids[keep]), ("amt", amt[keep])):
        got = np.concatenate([O.pfor_decode(open(f"btab.{col}.{g}", "rb").read(), segs[g]) for g in range(len(segs))])
        assert np.array_equal(got, want)
        h = struct.unpack("<QIII", open(f"btab.{col}.header", "rb").read())
        assert h[:3] == (k, len(segs), 3000)
    kept_names = [name[i] for i in np.nonzero(keep)[0]]
    first_seen = list(dict.fromkeys(kept_names))
    dict_file = open("btab.name", "rb").read()
    assert [dict_file[i * 12:(i + 1) * 12].rstrip(b"\0").decode() for i in range(len(dict_file) // 12)] == first_seen
    codes = np.concatenate([O.pfor_decode(open(f"btab.name.{g}.idx", "rb").read(), segs[g]) for g in range(len(segs))])
    assert [first_seen[c] for c in codes] == kept_names
    hb = open("btab.name.1.hash", "rb").read()
    assert struct.unpack_from("<I", hb)[0] == segs[1]
    want_h = [O.murmur64a(kept_names[bounds[1] + i].encode().ljust(12, b"\0"), 100) // 2 for i in range(segs[1])]
    assert list(struct.unpack_from(f"<{segs[1]}Q", hb, 4)) == want_h
    # the data dictionary, saved and loaded again; the table then loads by its name
    assert ops.lib.alenka_save_col_data(b"data.dictionary") == 0
    dd = O.parse_data_dictionary(open("data.dictionary", "rb").read())
    assert dd["btab"] == {"id": (0, 0), "amt": (0, 2), "name": (2, 12)}
    assert ops.lib.alenka_load_col_data(b"data.dictionary") == 0
    rc = ops.alenkaExecute('B := FILTER btab BY name == "PROMO";')
    assert rc == 0, ops.lib.alenka_last_error()
    promo = np.array([v == "PROMO" for v in kept_names])
    assert np.array_equal(ops.fetch("B", "id"), ids[keep][promo])
    assert fetch_chars(ops, "btab", "name", 12) == kept_names
    assert np.array_equal(ops.fetch("btab", "amt"), amt[keep])


def test_binary_errors(ops, tmp_path):
    prefix, dd, t = fixture_table(tmp_path)
    assert ops.lib.alenka_load_col_data(dd.encode()) == 0
    open(prefix + ".sort", "wb").write(struct.pack("<II", 2, 5) + b"l_k")  # truncated list
    ops.emit_load_binary("X", prefix, 0)
    assert ops.lib.alenka_status() == L.ALN_E_ARG
    os.remove(prefix + ".sort")
    open(prefix + ".sort", "wb").write(struct.pack("<II", 1, 5) + b"l_key")
    ops.emit_load_binary("X", prefix, 0)
    assert ops.lib.alenka_status() == 0
    os.remove(prefix + ".l_key.1")
    ops.emit_load_binary("Y", prefix, 0)
    assert ops.lib.alenka_status() == L.ALN_E_ARG
    ops.emit_load_binary("Y", str(tmp_path / "nothing"), 0)
    assert ops.lib.alenka_status() == L.ALN_E_ARG
    assert ops.lib.alenka_load_col_data(str(tmp_path / "missing").encode()) == L.ALN_E_ARG
    ops.emit_store_binary("X", str(tmp_path / "z"), 1)
    assert ops.lib.alenka_status() == L.ALN_E_UNSUPPORTED


def test_empty_table_store_and_load(ops, tmp_path):
    """An empty result stores as zero segments and loads back as an empty table."""
    ops.table("E0", 4, 4, [("id", np.arange(4, dtype=np.int64), 0), ("name", ["a", "b", "c", "d"], 3)])
    prefix = str(tmp_path / "empty")
    assert ops.alenkaExecute(f"E1 := FILTER E0 BY id > 10;\nSTORE E1 INTO '{prefix}' BINARY;") == 0, \
        ops.lib.alenka_last_error()
    assert struct.unpack("<QIII", open(prefix + ".id.header", "rb").read()) == (0, 0, 0, 0)
    ops.emit_load_binary("E2", prefix, 0)
    ops.check(None, "emit_load_binary")
    assert ops.rows("E2") == 0
    assert ops.alenkaExecute("E3 := FILTER E2 BY name LIKE \"%\";\nE4 := ORDER E2 BY name ASC;") == 0, \
        ops.lib.alenka_last_error()
    assert ops.rows("E3") == 0 and ops.rows("E4") == 0
