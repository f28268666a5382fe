"""Batched filter + GROUP BY (aln_batch_*) against the oracle: every query of a batch must
return exactly what its own FILTER + SELECT ... GROUP BY returns (one-segment key order)."""
import numpy as np
import pytest

import oracle_lib as O
from alenka_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REV = dict(a0=0, a1=1, b0=100, b1=-1)  # price * (1 - discount) at scale 4: price * (100 - disc)


def dev(ctx, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.int64)).to(ctx.device)


def host(t):
    return t.cpu().numpy()


def table(ctx, n, seg_rows, seed=11):
    from alenka_amd import device as D

    key, date, price, disc = O.gen_flat(seed, n)
    neg = 5000 - key * 3  # negative keys: floor division below zero
    h = [key, date, price, disc, neg]
    cols = [D.Column.encode(ctx, dev(ctx, c), seg_rows) for c in h[:4]]
    cols.append(D.Column.plain(ctx, dev(ctx, neg), seg_rows))  # a 64-bit column
    return h, cols


def want(h, q):
    p, k, x, y = h[q["pred_col"]], h[q["key_col"]], h[q["x_col"]], h[q["y_col"]]
    s = (p >= q["lo"]) & (p <= q["hi"])
    v = (q["a0"] + q["a1"] * x[s]) * (q["b0"] + q["b1"] * y[s])
    g = np.floor_divide(k[s], q["key_div"])
    wk, wv = O.groupby([g], [v, v], [L.AGG_SUM, L.AGG_COUNT])
    return wk[0], wv[0], wv[1]


def queries():
    qs = []
    for i, (lo, hi) in enumerate([(19920101, 19981231), (19950101, 19950131), (19930615, 19961231),
                                  (19990101, 19920101)]):  # the last keeps no row
        qs.append(dict(pred_col=1, key_col=0, x_col=2, y_col=3, lo=lo, hi=hi, key_div=1 + 7 * i, **REV))
    # GROUP BY month of date, filtered on discount (groups inside the LDS hot set)
    qs.append(dict(pred_col=3, key_col=1, x_col=2, y_col=3, lo=2, hi=6, key_div=100, **REV))
    qs.append(dict(pred_col=3, key_col=3, x_col=2, y_col=0, lo=0, hi=10, key_div=1, a0=0, a1=1, b0=3, b1=1))
    qs.append(dict(pred_col=0, key_col=4, x_col=2, y_col=3, lo=1000, hi=200000, key_div=13, **REV))
    return qs


def check_all(b, h, qs):
    for i, q in enumerate(qs):
        wk, ws, wc = want(h, q)
        k, s, c = b.results(i, max(1, len(wk)))
        assert len(k) == len(wk), f"query {i}"
        assert np.array_equal(host(k), wk), f"query {i} keys"
        assert np.array_equal(host(s), ws), f"query {i} sums"
        assert np.array_equal(host(c), wc), f"query {i} counts"


@pytest.mark.parametrize("seg_rows,hot,bpc", [(65536, -1, 0), (100000, 0, 1), (1 << 20, 8, 4),
                                                 (65536, 4096, 2), (100000, 1024, 1)])
def test_batch_matches_oracle(ctx, seg_rows, hot, bpc):
    from alenka_amd import device as D

    h, cols = table(ctx, 250000, seg_rows)
    qs = queries()
    b = D.Batch(ctx, cols, qs, hot_groups=hot, blocks_per_cu=bpc)
    try:
        for _ in range(2):  # tables are cleared by every run
            b.run()
            check_all(b, h, qs)
    finally:
        b.close()


def test_batch_graph_replays_and_autotune(ctx):
    from alenka_amd import device as D

    h, cols = table(ctx, 120000, 32768, seed=3)
    qs = queries()
    b = D.Batch(ctx, cols, qs, use_graph=True)
    try:
        for r in range(3):  # replays of the captured graph, with eager work between them
            b.run()
            check_all(b, h, qs)
            torch.cuda.synchronize()
            _ = (dev(ctx, h[0]) * 2).sum().item()
        shape, ms = b.autotune(reps=2, use_graph=True)
        assert ms > 0 and shape[1] in (1, 2, 4)
        b.run()
        check_all(b, h, qs)
    finally:
        b.close()


def test_batch_capacity_and_arguments(ctx):
    from alenka_amd import device as D

    h, cols = table(ctx, 20000, 8192)
    qs = queries()[:1]
    b = D.Batch(ctx, cols, qs)
    try:
        b.run()
        with pytest.raises(L.AlnError):
            b.results(0, 3)
        assert ctx.lib.aln_last_error().startswith(b"aln_batch_results")
    finally:
        b.close()
    bad = dict(qs[0], key_div=0)
    with pytest.raises(L.AlnError):
        D.Batch(ctx, cols, [bad])
    with pytest.raises(L.AlnError):
        D.Batch(ctx, cols, qs * 65)


def test_zipf_generator(ctx):
    """Ranks in [0, nkeys), deterministic, frequencies following r^-s (rank 0 most frequent)."""
    from alenka_amd import device as D

    n, nkeys, s = 2_000_000, 1000, 1.1
    z = host(D.gen_zipf(ctx, 9, n, nkeys, s))
    assert np.array_equal(z, host(D.gen_zipf(ctx, 9, n, nkeys, s)))
    assert z.min() >= 0 and z.max() < nkeys
    cnt = np.bincount(z, minlength=nkeys)
    p = np.arange(1, nkeys + 1, dtype=np.float64) ** -s
    p /= p.sum()
    for r in range(10):  # 2M draws: the top ranks are within a few percent of r^-s
        assert abs(cnt[r] / n - p[r]) < 0.03 * p[r] + 3e-4, r
    assert abs(cnt[100:].sum() / n - p[100:].sum()) < 0.01


def test_batch_zipf_keys(ctx):
    """Skewed group keys (most rows in a few groups): LDS hot set and global tables agree."""
    from alenka_amd import device as D

    n = 300000
    key, date, price, disc = O.gen_flat(21, n)
    z = D.gen_zipf(ctx, 5, n, 100000, 1.1)
    h = [key, date, price, disc, host(z)]
    cols = [D.Column.encode(ctx, dev(ctx, c), 65536) for c in h[:4]] + [D.Column.encode(ctx, z, 65536)]
    qs = [dict(pred_col=2, key_col=4, x_col=2, y_col=3, lo=90000 + i * 100000, hi=90000 + i * 100000 + 3000000,
               key_div=div, **REV) for i, div in enumerate([1, 10, 1000, 20000, 100000])]
    # hot 1024 and 4096: LDS tables for 4 and 1 queries per block (query groups of 4 + 1
    # and 5 x 1), every query summing the same expression (computed once per row)
    for hot in (0, 32, 1024, 4096):
        b = D.Batch(ctx, cols, qs, hot_groups=hot)
        try:
            b.run()
            check_all(b, h, qs)
        finally:
            b.close()
    # four shared columns, expressions that differ per query (evaluated per query)
    qv = [dict(q, a0=i, a1=1 + i, b0=100 - i, b1=-1 - i) for i, q in enumerate(qs)]
    b = D.Batch(ctx, cols, qv, hot_groups=1024)
    try:
        b.run()
        check_all(b, h, qv)
    finally:
        b.close()


def test_order_by_sort_input(ctx):
    """ORDER BY of uniform int64 keys with their row numbers as payload: stable and sorted."""
    from alenka_amd import device as D

    keys, pay = D.gen_sort_input(ctx, 4, 1 << 20)
    assert np.array_equal(host(pay), np.arange(1 << 20))
    perm = host(D.sort_keys(ctx, [keys]))
    hk = host(keys)
    assert np.array_equal(perm, np.argsort(hk, kind="stable"))
