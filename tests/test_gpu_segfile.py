"""The reference's segment files (FOR / FOR-DELTA, compress.cu:180-527) through the C ABI:
files the oracle writes in the reference's layout decode on the device to the original
values; files the device writes are byte-identical to the oracle's and decode back."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(ctx, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.int64)).to(ctx.device)


def cases():
    rng = np.random.default_rng(8)
    yield "const", np.full(1000, 42, np.int64)
    yield "u8", rng.integers(-50, 200, 5000)
    yield "u16", rng.integers(19920101, 19920101 + 60000, 100003)
    yield "u32", rng.integers(0, 1 << 31, 7777)
    yield "u64", rng.integers(-(1 << 62), 1 << 62, 4096)
    yield "one", np.array([-7], np.int64)


@pytest.mark.parametrize("name,v", list(cases()), ids=[n for n, _ in cases()])
def test_segfile_for_roundtrip(ctx, name, v):
    from alenka_amd import device as D

    v = np.ascontiguousarray(v, np.int64)
    ref = O.pfor_encode(v)                      # the reference's FOR layout
    got, lo, hi = D.segfile_decode(ctx, ref)
    assert np.array_equal(got.cpu().numpy(), v)
    assert (lo, hi) == (int(v.min()), int(v.max()))
    mine = D.segfile_encode(ctx, dev(ctx, v))
    assert mine == ref                          # byte-identical file
    assert np.array_equal(O.pfor_decode(mine, len(v)), v)


@pytest.mark.parametrize("n", [1, 2, 63, 1000, 65537])
def test_segfile_delta_decode(ctx, n):
    from alenka_amd import device as D

    rng = np.random.default_rng(n)
    v = np.cumsum(rng.integers(0, 1000, n)).astype(np.int64) + 19000000  # sorted: FOR-DELTA
    ref = O.pfor_encode(v, delta=True)
    got, lo, hi = D.segfile_decode(ctx, ref)
    assert np.array_equal(got.cpu().numpy(), v)
    assert (lo, hi) == (int(v[0]), int(v[-1]))


def test_segfile_errors(ctx):
    from alenka_amd import _lib as L
    from alenka_amd import device as D

    ref = O.pfor_encode(np.arange(100, dtype=np.int64))
    with pytest.raises(L.AlnError):
        D.segfile_decode(ctx, ref[:40])          # truncated
    bad = bytearray(ref)
    bad[20] = 7                                  # unknown compression type
    with pytest.raises(L.AlnError):
        D.segfile_decode(ctx, bytes(bad))
