"""aln_scan_filter_noalloc (the scan whose temporaries live in the library's fixed device
array) against aln_scan_filter through the C ABI: identical row ids and counts on calls in
a row, the later ones smaller, with and without zone-map verdicts, on both the interval
fast path and the interpreted path; tables beyond the array's limit are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NAME, NUM, CMP, OR = L.RPN_NAME, L.RPN_NUMBER, L.RPN_CMP, L.RPN_OR
DATE_LT = [(NAME, 1, 0), (NUM, 0, 19950315), (CMP, 1, 0)]                       # intervals
DISC_KEY = [(NAME, 3, 0), (NUM, 0, 5), (CMP, 6, 0), (NAME, 0, 0), (NUM, 0, 1000), (CMP, 1, 0), (OR, 0, 0)]  # VM


def table(ctx, seed, n, seg_rows, sort_dates):
    from alenka_amd import device as D

    cols = O.gen_flat(seed, n)
    if sort_dates:
        cols[1] = np.sort(cols[1])
    dec = [0, 0, 2, 2]
    dcols = [D.Column.encode(ctx, torch.as_tensor(np.ascontiguousarray(c, np.int64)).to(ctx.device), seg_rows,
                             decimals=d) for c, d in zip(cols, dec)]
    return dcols, dec


def both(ctx, dcols, dec, ops, zone_maps):
    """(noalloc row ids, scan_filter row ids) of one predicate over one table."""
    from alenka_amd import device as D

    prog, _ = L.compile_rpn(ops, dec)
    modes = None
    if zone_maps:
        modes = [L.zone_map_check(ops, [c.hsegs[s].vmin for c in dcols], [c.hsegs[s].vmax for c in dcols], dec)
                 for s in range(dcols[0].nsegs)]
    want = D.scan_filter(ctx, dcols, prog, modes).cpu().numpy()
    n = dcols[0].nrows
    out = torch.full((n + 1,), -7, dtype=torch.int32, device=ctx.device)
    cnt = C.c_uint64(12345)
    arr = D._cols_array(dcols)
    ins = L.insn_array(prog)
    m = None if modes is None else "".join(modes).encode()
    L.check(ctx.lib.aln_scan_filter_noalloc(ctx.h, arr, len(dcols), dcols[0].seg_rows, ins, len(prog), m,
                                            D._p(out), n, C.byref(cnt)), "aln_scan_filter_noalloc")
    got = out.cpu().numpy()
    assert (got[cnt.value:] == -7).all(), "wrote past its count"
    return got[: cnt.value], want


@pytest.mark.parametrize("zone_maps", [False, True])
def test_noalloc_matches_scan_filter_twice(ctx, zone_maps):
    # first call: 3.5M rows in 1M-row segments (980 tiles); second: 50,000 rows in 8,192-row
    # segments (14 tiles); third: the first table again, with the interpreted predicate
    big, dec = table(ctx, 21, 3_500_000, 1_000_000, True)
    small, _ = table(ctx, 22, 50_000, 8192, True)
    for dcols, ops in ((big, DATE_LT), (small, DATE_LT), (small, DISC_KEY), (big, DISC_KEY)):
        got, want = both(ctx, dcols, dec, ops, zone_maps)
        assert got.size > 0
        assert np.array_equal(got, want)


def test_noalloc_empty_and_capacity(ctx):
    from alenka_amd import device as D

    dcols, dec = table(ctx, 23, 20_000, 4096, False)
    prog, _ = L.compile_rpn([(NAME, 1, 0), (NUM, 0, 0), (CMP, 1, 0)], dec)  # date < 0: no row
    arr, ins = D._cols_array(dcols), L.insn_array(prog)
    cnt = C.c_uint64(99)
    out = torch.zeros(4, dtype=torch.int32, device=ctx.device)
    L.check(ctx.lib.aln_scan_filter_noalloc(ctx.h, arr, 4, 4096, ins, len(prog), None, D._p(out), 4,
                                            C.byref(cnt)), "aln_scan_filter_noalloc")
    assert cnt.value == 0
    prog, _ = L.compile_rpn(DATE_LT, dec)
    ins = L.insn_array(prog)
    rc = ctx.lib.aln_scan_filter_noalloc(ctx.h, arr, 4, 4096, ins, len(prog), None, D._p(out), 4, C.byref(cnt))
    assert rc == L.ALN_E_CAPACITY and cnt.value > 4


def test_noalloc_refuses_tables_beyond_its_scratch(ctx):
    """262,145 one-tile segments exceed the fixed array: refused before any launch.  The
    descriptors say every segment is empty, so no launch could read a word either way."""
    from alenka_amd import device as D

    nsegs = L.SCAN_NOALLOC_MAX_TILES + 1
    segs = torch.zeros((nsegs * C.sizeof(L.Seg) + 7) // 8, dtype=torch.int64, device=ctx.device)
    words = torch.zeros(16, dtype=torch.int32, device=ctx.device)
    col = L.Col()
    col.words = C.c_void_p(words.data_ptr())
    col.segs = C.c_void_p(segs.data_ptr())
    col.nsegs = nsegs
    col.decimals = 0
    arr = (L.Col * 1)(col)
    prog, _ = L.compile_rpn([(NAME, 0, 0), (NUM, 0, 5), (CMP, 1, 0)], [0])
    ins = L.insn_array(prog)
    cnt = C.c_uint64(0)
    out = torch.zeros(4, dtype=torch.int32, device=ctx.device)
    rc = ctx.lib.aln_scan_filter_noalloc(ctx.h, arr, 1, 4096, ins, len(prog), None, D._p(out), 4, C.byref(cnt))
    assert rc == L.ALN_E_ARG
    assert b"scratch" in ctx.lib.aln_last_error()
