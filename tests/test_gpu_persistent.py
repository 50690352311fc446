"""Persistent serving kernel: one launch runs whole searches (query prep,
every filter level and its select, the exact re-rank) for several batches of
different indexes and k at once. Every request it serves must equal
IndexBF.search on the same queries bit for bit; the launch must group
queued batches; FP32 indexes stay on the launched pipeline."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # (n, d, precision): sample copies from 65,536 rows; below, strided levels
    (3_000, 64, "bf16"),        # one level (dense select of every row)
    (40_000, 96, "bf16"),       # strided sample levels, no sample copy
    (150_000, 128, "bf16"),
    (500_000, 256, "bf16"),
    (90_000, 128, "fp8"),
    (260_000, 64, "fp8"),
    (70_000, 100, "bf16"),      # dpad 128 with padding
    (50_000, 128, "fp32"),      # launched pipeline
]


@pytest.fixture(scope="module")
def pindexes(ca, res):
    out = []
    for i, (n, d, prec) in enumerate(SPECS):
        x = ca.gen_uniform(n, d, 2000 + i, -1, 1)
        idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision=prec))
        idx.add(x)
        out.append(idx)
    return out


def requests(ca, count, seed, max_nq=256):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(count):
        i = int(rng.integers(len(SPECS)))
        nq = int(np.clip(np.floor(np.exp(rng.uniform(0, np.log(max_nq + 1)))), 1, max_nq))
        k = int(rng.choice([1, 10, 100]))
        reqs.append((i, ca.gen_uniform(nq, SPECS[i][1], 60_000 + seed * 1000 + r, -1, 1), k))
    return reqs


def test_persistent_one_at_a_time(ca, pindexes):
    srv = ca.Server(slots=1, use_graphs=False)
    ids = [srv.add_index(ix) for ix in pindexes]
    for j, (i, q, k) in enumerate(requests(ca, 48, 3)):
        print(f"[persistent] request {j}: index {i} {SPECS[i]} nq {q.shape[0]} k {k}", flush=True)
        I, D, info = srv.wait(srv.submit(ids[i], q, k), timeout_s=60.0)
        Ia, Da = pindexes[i].search(q, k)
        assert np.array_equal(I, Ia), f"index {i} nq {q.shape[0]} k {k}"
        assert np.array_equal(D.view(np.uint32), Da.view(np.uint32))
        assert (info["kernel_batches"] > 0) == (SPECS[i][2] != "fp32")
    st = srv.stats()
    print("stats:", st)
    assert st["persistent_launches"] > 0
    srv.close()


def test_persistent_groups_batches(ca, pindexes):
    """A burst of requests to different indexes and k queued while a slot is
    busy: the next launch takes several batches at once."""
    reqs = requests(ca, 120, 4, max_nq=128)
    alone = [pindexes[i].search(q, k) for i, q, k in reqs]
    srv = ca.Server(slots=1, use_graphs=False)
    ids = [srv.add_index(ix) for ix in pindexes]
    handles = [srv.submit(ids[i], q, k) for i, q, k in reqs]
    outs = [srv.wait(h, timeout_s=120.0) for h in handles]
    for (I, D, info), (Ia, Da), (i, q, k) in zip(outs, alone, reqs):
        assert info["status"] == 0
        assert np.array_equal(I, Ia), f"index {i} nq {q.shape[0]} k {k}"
        assert np.array_equal(D.view(np.uint32), Da.view(np.uint32))
    st = srv.stats()
    grouped = max(info["kernel_batches"] for _, _, info in outs)
    print("stats:", st, "largest group", grouped)
    assert grouped >= 2
    assert st["persistent_batches"] > st["persistent_launches"]
    srv.close()


def test_persistent_retry_path(ca, pindexes):
    """A tuning entry with a survivor multiple of 1 (about k' rows kept per
    query) makes many queries keep fewer than k' rows: the persistent run's
    failed queries are re-run on the launched pipeline with wider margins,
    and the results still equal IndexBF.search's (which retries the same
    way)."""
    idx = pindexes[2]  # 150,000 x 128 BF16
    idx.set_tuning(dict(k=100, batch_nq=64, kprime_factor=4, target_wgs=4096, c_factor=1, variant=0,
                        rho=8, recall=1.0, ms_default=0.0, ms_tuned=0.0))
    try:
        q = ca.gen_uniform(64, 128, 777, -1, 1)
        Ia, Da = idx.search(q, 100)
        retried = idx.last_stats()["retried_queries"]
        print("retried alone:", retried)
        assert retried > 0
        srv = ca.Server(slots=1, use_graphs=False)
        sid = srv.add_index(idx)
        I, D, info = srv.wait(srv.submit(sid, q, 100), timeout_s=60.0)
        assert info["kernel_batches"] == 1
        assert np.array_equal(I, Ia) and np.array_equal(D.view(np.uint32), Da.view(np.uint32))
        srv.close()
    finally:
        idx.clear_tuning()


@pytest.mark.parametrize("which", [0, 1, 3, 5, 6])
def test_indexbf_persistent_equals_launched(ca, pindexes, which):
    """IndexBF.search on the persistent kernel (small batches, one launch per
    step) and on the launched pipeline return the same indices and distance
    bits, at one query tile, a ragged one and several."""
    idx = pindexes[which]
    n, d, prec = SPECS[which]
    try:
        for nq, k in ((1, 10), (7, 100), (32, 10), (33, 1), (100, 100)):
            q = ca.gen_uniform(nq, d, 90_000 + 31 * which + nq, -1, 1)
            print(f"[persistent] IndexBF {SPECS[which]} nq {nq} k {k}", flush=True)
            idx.set_persistent(0)
            Il, Dl = idx.search(q, k)
            assert idx.last_stats()["persistent"] == 0
            idx.set_persistent(128)
            Ip, Dp = idx.search(q, k)
            assert idx.last_stats()["persistent"] == 1
            assert np.array_equal(Ip, Il), f"{SPECS[which]} nq {nq} k {k}"
            assert np.array_equal(Dp.view(np.uint32), Dl.view(np.uint32)), f"{SPECS[which]} nq {nq} k {k}"
    finally:
        idx.set_persistent(32)
