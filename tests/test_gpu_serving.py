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
    srv = ca.Server(slots=2)
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
