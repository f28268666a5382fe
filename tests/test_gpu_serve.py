"""The served form (aln_server_*): requests submitted concurrently from several client
threads are co-batched and overlap on the device; every result must equal the oracle's and
the result of the same request run alone (bit-exact: keys, sums, counts, top rows)."""
import threading

import numpy as np
import pytest

import oracle_lib as O
import q3_reference as Q3
from alenka_amd import _lib as L

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REV = dict(a0=0, a1=1, b0=100, b1=-1)


def dev(ctx, a):
    return torch.as_tensor(np.ascontiguousarray(a, np.int64)).to(ctx.device)


def want(h, q):
    p, k, x, y = h[q["pred_col"]], h[q["key_col"]], h[q["x_col"]], h[q["y_col"]]
    s = (p >= q["lo"]) & (p <= q["hi"])
    v = (q["a0"] + q["a1"] * x[s]) * (q["b0"] + q["b1"] * y[s])
    wk, wv = O.groupby([np.floor_divide(k[s], q["key_div"])], [v, v], [L.AGG_SUM, L.AGG_COUNT])
    return wk[0], wv[0], wv[1]


@pytest.fixture(scope="module")
def tables(ctx):
    from alenka_amd import device as D

    out = {}
    n = 200000
    key, date, price, disc = O.gen_flat(17, n)
    z = D.gen_zipf(ctx, 8, n, 50000, 1.1)
    h = [key, date, price, disc, z.cpu().numpy()]
    cols = [D.Column.encode(ctx, dev(ctx, c), 65536) for c in h[:4]] + [D.Column.encode(ctx, z, 65536)]
    out["flat"] = (h, cols)
    n2 = 70000
    k2, d2, p2, s2 = O.gen_flat(18, n2)
    neg = 1000 - 3 * k2
    h2 = [k2, d2, p2, s2, neg]
    out["small"] = (h2, [D.Column.encode(ctx, dev(ctx, c), 16384) for c in h2])
    t = Q3.tables(99, 3000, 30000)
    enc = lambda a, d=0: D.Column.encode(ctx, dev(ctx, a), 8192, decimals=d)
    dim = [enc(t["c_key"]), enc(t["c_seg"])]
    mid = [enc(t["o_key"]), enc(t["o_cust"]), enc(t["o_date"]), enc(t["o_prio"])]
    fact = [enc(t["l_okey"]), enc(t["l_price"], 2), enc(t["l_disc"], 2), enc(t["l_ship"])]
    b = int(t["code_of"][Q3.SEGMENTS.index("BUILDING")])
    star = (dict(cols=dim, ranges=[(1, b, b)]), dict(cols=mid, ranges=[(2, -(1 << 63), Q3.DATE - 1)]),
            dict(cols=fact, ranges=[(3, Q3.DATE + 1, (1 << 63) - 1)]),
            dict(d_key=0, m_key=0, m_fk=1, m_g1=2, m_g2=3, f_fk=0, f_x=1, f_y=2, a0=0, a1=1, b0=100, b1=-1))
    out["star"] = (star, Q3.run(t)[0])
    return out


def workload(seed, count):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        kind = rng.integers(0, 5)
        if kind == 4:
            reqs.append(("star", None))
            continue
        tab = "flat" if kind < 3 else "small"
        pred = [1, 2, 3][rng.integers(0, 3)]
        lo, hi = {1: (19920101, 19981231), 2: (90000, 10500000), 3: (0, 10)}[pred]
        a = int(rng.integers(lo, hi))
        b = int(rng.integers(a, hi + 1))
        keycol = 4 if rng.random() < 0.6 else 0
        div = int([1, 3, 100, 7919][rng.integers(0, 4)])
        reqs.append((tab, dict(pred_col=pred, key_col=keycol, x_col=2, y_col=3, lo=a, hi=b, key_div=div, **REV)))
    return reqs


def run_mix(srv, tids, tables, reqs, nthreads):
    ids = [None] * len(reqs)

    def client(t):
        for i in range(t, len(reqs), nthreads):
            tab, q = reqs[i]
            if tab == "star":
                D_, M_, F_, spec = tables["star"][0]
                ids[i] = srv.submit_star(D_, M_, F_, spec, limit=10)
            else:
                ids[i] = srv.submit(tids[tab], q)

    th = [threading.Thread(target=client, args=(t,)) for t in range(nthreads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    out = []
    for i, (tab, q) in enumerate(reqs):
        r = srv.wait(ids[i])
        out.append((r.batch_size, srv.fetch_star(ids[i]) if tab == "star" else srv.fetch(ids[i])))
        srv.release(ids[i])
    return out


@pytest.mark.parametrize("plain,streams", [(False, 3), (False, 1), (True, 2)])
def test_served_results_alone_and_in_mix(ctx, tables, plain, streams):
    from alenka_amd import device as D

    reqs = workload(7 + streams, 36)
    srv = D.Server(streams=streams, max_batch=16, plain=plain, budget_bytes=8 << 30)
    try:
        tids = {k: srv.add_table(tables[k][1]) for k in ("flat", "small")}
        mixed = run_mix(srv, tids, tables, reqs, nthreads=3)
        # the same requests alone, one at a time
        alone = [run_mix(srv, tids, tables, [r], nthreads=1)[0] for r in reqs[:12]]
        st = srv.stats()
    finally:
        srv.close()
    for i, ((tab, q), (bs, got)) in enumerate(zip(reqs, mixed)):
        if tab == "star":
            assert got == tables["star"][1], f"request {i}"
        else:
            wk, ws, wc = want(tables[tab][0], q)
            assert np.array_equal(got[0], wk) and np.array_equal(got[1], ws), f"request {i}"
            assert np.array_equal(got[2], wc), f"request {i}"
    for i, (bs, got) in enumerate(alone):
        ref = mixed[i][1]
        if reqs[i][0] == "star":
            assert got == ref
        else:
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), f"request {i} alone"
    assert st["completed"] == len(reqs) + 12
    if not plain and streams == 1:  # a burst of submissions behind one worker is co-batched
        assert max(bs for bs, _ in mixed) > 1 and st["batched_queries"] > 0
    if plain:
        assert max(bs for bs, _ in mixed) == 1


def test_served_budget_admission(ctx, tables):
    from alenka_amd import device as D

    h, cols = tables["flat"]
    big = dict(pred_col=1, key_col=4, x_col=2, y_col=3, lo=19920101, hi=19981231, key_div=1, **REV)
    small = dict(big, key_div=5000)  # 10 groups
    # the budget holds the small query (a few KiB) but not 50000 dense groups (~2.4 MB)
    srv = D.Server(streams=2, budget_bytes=1 << 20)
    try:
        t = srv.add_table(cols)
        ids = [srv.submit(t, big), srv.submit(t, small), srv.submit(t, small)]
        r = srv.wait(ids[0], check=False)
        assert r.status == L.ALN_E_NOMEM
        for i in ids[1:]:
            k, s, c = srv.fetch(i)
            wk, ws, wc = want(h, small)
            assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc)
        assert srv.stats()["mem_in_use"] == 0
    finally:
        srv.close()
