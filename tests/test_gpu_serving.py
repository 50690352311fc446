"""Serving form: continuous batching over several resident indexes of uneven
sizes and dims. Every request's result must be the same alone (a direct
IndexBF.search) and in the mix (bit-exact), admission must respect the HBM
budget, and a request that can never fit is rejected with BUDGET."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # (n, d, precision)
    (12_000, 64, "fp32"),
    (150_000, 96, "bf16"),
    (300_000, 128, "fp8"),
    (50_000, 256, "fp32"),
    (31_000, 96, "fp32"),
    (80_000, 64, "bf16"),
]


@pytest.fixture(scope="module")
def indexes(ca, res):
    out = []
    for i, (n, d, prec) in enumerate(SPECS):
        x = ca.gen_uniform(n, d, 1000 + i, -1, 1)
        idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision=prec))
        idx.add(x)
        out.append(idx)
    return out


def make_requests(ca, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(count):
        i = int(rng.integers(len(SPECS)))
        nq = int(np.clip(np.floor(np.exp(rng.uniform(0, np.log(513)))), 1, 512))
        k = int(rng.choice([1, 10, 100]))
        q = ca.gen_uniform(nq, SPECS[i][1], 50_000 + seed * 1000 + r, -1, 1)
        reqs.append((i, q, k))
    return reqs


def test_alone_equals_in_mix(ca, indexes):
    reqs = make_requests(ca, 150, 1)
    alone = [indexes[i].search(q, k) for i, q, k in reqs]
    srv = ca.Server(slots=4)
    ids = [srv.add_index(ix) for ix in indexes]
    rng = np.random.default_rng(2)
    handles = []
    for i, q, k in reqs:  # Poisson arrivals, mean gap 0.5 ms
        handles.append(srv.submit(ids[i], q, k))
        time.sleep(float(rng.exponential(0.0005)))
    outs = [srv.wait(h, timeout_s=100.0) for h in handles]
    st = srv.stats()
    print("server stats:", st)
    batched = sum(1 for _, _, info in outs if info["batch_requests"] > 1)
    print(f"requests that shared a batch: {batched}")
    for (I, D, info), (Ia, Da), (i, q, k) in zip(outs, alone, reqs):
        assert info["status"] == 0
        assert np.array_equal(I, Ia), f"index {i} nq {q.shape[0]} k {k}"
        assert np.array_equal(D.view(np.uint32), Da.view(np.uint32))
    assert st["completed"] == len(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["peak_inflight"] >= 1
    srv.close()


def test_budget_admission(ca, indexes):
    need = max(ca.batch_bytes(indexes[2], 512, 100), ca.batch_bytes(indexes[1], 512, 100))
    srv = ca.Server(slots=4, hbm_budget_bytes=int(need * 1.5), max_batch_queries=512)
    ids = [srv.add_index(ix) for ix in indexes]
    q = ca.gen_uniform(512, 128, 7, -1, 1)
    q96 = ca.gen_uniform(512, 96, 8, -1, 1)
    hs = [srv.submit(ids[2], q, 100) if j % 2 else srv.submit(ids[1], q96, 100) for j in range(8)]
    for h in hs:
        srv.wait(h, timeout_s=100.0)
    st = srv.stats()
    print("budget stats:", st, "need", need)
    assert st["peak_reserved_bytes"] <= int(need * 1.5)
    assert st["peak_inflight"] == 1  # two such batches never fit together
    srv.close()


def test_request_larger_than_budget_is_rejected(ca, indexes):
    srv = ca.Server(slots=2, hbm_budget_bytes=1 << 20)
    ids = [srv.add_index(ix) for ix in indexes]
    q = ca.gen_uniform(256, 64, 9, -1, 1)
    h = srv.submit(ids[0], q, 10)
    with pytest.raises(ca.RplshError) as e:
        srv.wait(h, timeout_s=100.0)
    assert e.value.status == 6
    assert srv.stats()["rejected"] == 1
    small = srv.submit(ids[0], q[:1], 1)
    I, D, info = srv.wait(small, timeout_s=100.0)
    Ia, Da = indexes[0].search(q[:1], 1)
    assert np.array_equal(I, Ia)
    srv.close()


def test_graph_replays_match_eager(ca, indexes):
    """The same (index, k, bucket) graph replayed many times, with eager
    searches and other graphs in between, always equals the eager result."""
    srv = ca.Server(slots=2, persistent_max_queries=0)
    ids = [srv.add_index(ix) for ix in indexes]
    q = ca.gen_uniform(37, 96, 77, -1, 1)
    Ia, Da = indexes[1].search(q, 10)
    other = ca.gen_uniform(5, 64, 78, -1, 1)
    for rep in range(6):
        I, D, info = srv.wait(srv.submit(ids[1], q[: 20 + rep], 10), timeout_s=100.0)
        assert np.array_equal(I, Ia[: 20 + rep])
        assert np.array_equal(D.view(np.uint32), Da[: 20 + rep].view(np.uint32))
        srv.wait(srv.submit(ids[0], other, 1), timeout_s=100.0)
        indexes[2].search(ca.gen_uniform(3, 128, 79 + rep, -1, 1), 5)  # eager work between
    st = srv.stats()
    print("graph stats:", st)
    # a shape's first batch runs eagerly, later ones replay its graph
    assert st["graph_launches"] >= 10
    assert st["graph_captures"] <= 4  # one per (index, k, bucket) and slot
    srv.close()


def test_eager_server_matches(ca, indexes):
    srv = ca.Server(slots=3, use_graphs=False)
    ids = [srv.add_index(ix) for ix in indexes]
    for i, q, k in make_requests(ca, 20, 5):
        I, D, _ = srv.wait(srv.submit(ids[i], q, k), timeout_s=100.0)
        Ia, Da = indexes[i].search(q, k)
        assert np.array_equal(I, Ia) and np.array_equal(D.view(np.uint32), Da.view(np.uint32))
    assert srv.stats()["graph_launches"] == 0
    srv.close()


def test_registered_index_refuses_changes(ca, res):
    """While a server holds an index, changing it (rows, encoding, tuning)
    fails with STATE instead of leaving cached graphs and in-flight batches
    pointing at freed buffers; after the server is gone it works again."""
    x = ca.gen_uniform(20_000, 64, 601, -1, 1)
    idx = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    srv = ca.Server(slots=2)
    sid = srv.add_index(idx)
    q = ca.gen_uniform(16, 64, 602, -1, 1)
    I0, D0, _ = srv.wait(srv.submit(sid, q, 10), timeout_s=100.0)
    for op in (lambda: idx.add(x[:100]), idx.reset, lambda: idx.set_precision("fp32"),
               idx.clear_tuning, lambda: idx.autotune(q, 10, 64, 0.0)):
        with pytest.raises(ca.RplshError) as e:
            op()
        assert e.value.status == 4
    I1, D1, _ = srv.wait(srv.submit(sid, q, 10), timeout_s=100.0)
    assert np.array_equal(I0, I1) and np.array_equal(D0.view(np.uint32), D1.view(np.uint32))
    srv.close()
    idx.add(x[:100])  # allowed again
    assert idx.num_data == 20_100


def test_graph_key_follows_tuning_entries(ca, res):
    """An index with two tuning entries for one k: batches of different sizes
    resolve to different search parameters, so they must not share a cached
    graph. Every served result equals IndexBF.search on the same queries (which
    resolves the entry the same way)."""
    x = ca.gen_uniform(200_000, 96, 611, -1, 1)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    base = dict(k=10, kprime_factor=4, target_wgs=4096, c_factor=3, variant=0, rho=0,
                recall=1.0, ms_default=0.0, ms_tuned=0.0)
    idx.set_tuning(dict(base, batch_nq=16, kprime_factor=2, c_factor=2, rho=32))
    idx.set_tuning(dict(base, batch_nq=24, kprime_factor=8, c_factor=4, rho=16, variant=1))
    srv = ca.Server(slots=1, persistent_max_queries=0)
    sid = srv.add_index(idx)
    qs = [ca.gen_uniform(nq, 96, 620 + j, -1, 1) for j, nq in enumerate([17, 23, 18, 22, 17, 23])]
    for q in qs:  # buckets: all 32; tuning entries: nearest of 16 / 24
        I, D, info = srv.wait(srv.submit(sid, q, 10), timeout_s=100.0)
        Ia, Da = idx.search(q, 10)
        st = idx.last_stats()
        assert st["kprime"] == 10 * (2 if abs(np.log(q.shape[0] / 16)) < abs(np.log(q.shape[0] / 24)) else 8)
        assert np.array_equal(I, Ia) and np.array_equal(D.view(np.uint32), Da.view(np.uint32))
    s = srv.stats()
    print("two-entry stats:", s)
    assert s["graph_captures"] == 2 and s["graph_launches"] >= 4
    srv.close()


C5_PRECISIONS = ["bf16", "bf16", "fp32", "fp8"]


def test_c5_full_scale(ca, orc, res):
    """The serving configuration at its stated scale: 64 resident indexes with
    n log-uniform in [1e4, 2e6] and d from {64, 96, 128, 256}, U[-1,1) rows;
    1,024 requests with Poisson arrivals, each to a random index with nq
    log-uniform in [1, 2048] and k from {1, 10, 100}. Every request equals its
    alone IndexBF.search bit for bit; requests on FP32 indexes are checked
    against the canonical oracle on their first queries, BF16 / FP8 ones for
    recall against the FP64 oracle."""
    rng = np.random.default_rng(4242)
    specs = []
    for i in range(64):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        d = int(rng.choice([64, 96, 128, 256]))
        specs.append((n, d, C5_PRECISIONS[i % 4]))
    idxs = []
    for i, (n, d, prec) in enumerate(specs):
        x = ca.gen_uniform(n, d, 80_000 + i, -1.0, 1.0)
        ix = ca.IndexBF(res, d, ca.IndexBFConfig(precision=prec))
        ix.add(x)
        idxs.append(ix)
        del x
    print(f"[c5] 64 indexes, {sum(s[0] for s in specs)} rows", flush=True)
    reqs = []
    for r in range(1024):
        i = int(rng.integers(64))
        nq = int(np.clip(np.floor(np.exp(rng.uniform(0.0, np.log(2049.0)))), 1, 2048))
        k = int(rng.choice([1, 10, 100]))
        reqs.append((i, ca.gen_uniform(nq, specs[i][1], 300_000 + r, -1.0, 1.0), k))
    alone = [idxs[i].search(q, k) for i, q, k in reqs]
    print(f"[c5] {len(reqs)} requests, {sum(q.shape[0] for _, q, _ in reqs)} queries; alone done", flush=True)
    srv = ca.Server(slots=4)
    ids = [srv.add_index(ix) for ix in idxs]
    gaps = rng.exponential(1.0 / 2000.0, size=len(reqs))
    handles = []
    for (i, q, k), g in zip(reqs, gaps):
        handles.append(srv.submit(ids[i], q, k))
        time.sleep(float(g))
    outs = [srv.wait(h, timeout_s=300.0) for h in handles]
    st = srv.stats()
    print("[c5] server stats:", st, flush=True)
    for j, ((I, D, info), (Ia, Da), (i, q, k)) in enumerate(zip(outs, alone, reqs)):
        assert info["status"] == 0
        assert np.array_equal(I, Ia), f"request {j}: index {i} nq {q.shape[0]} k {k}"
        assert np.array_equal(D.view(np.uint32), Da.view(np.uint32)), f"request {j}"
    assert st["completed"] == len(reqs)
    assert st["peak_inflight"] >= 2
    assert sum(1 for _, _, info in outs if info["batch_requests"] > 1) > 0
    # oracle on a sample: the first 8 queries of up to 12 requests per precision
    seen = {"fp32": 0, "bf16": 0, "fp8": 0}
    for (I, D, info), (i, q, k) in zip(outs, reqs):
        n, d, prec = specs[i]
        if seen[prec] >= 12:
            continue
        seen[prec] += 1
        x = ca.gen_uniform(n, d, 80_000 + i, -1.0, 1.0)
        qq = q[:8]
        if prec == "fp32":
            Io, Do = orc.search_canonical(x, qq, k)
            assert np.array_equal(I[:8], Io), f"index {i} k {k}"
            assert np.array_equal(D[:8].view(np.uint32), Do.view(np.uint32))
        else:
            I64, _ = orc.search_fp64(x, qq, k)
            r = orc.recall(I[:8], I64)
            assert r >= 0.99, f"{prec} index {i} (n {n}, d {d}) k {k}: recall {r}"
    srv.close()
