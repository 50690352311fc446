"""Plain path, small batches: batches of at most 16 queries take the
streaming VALU filter (filter_small_kernel) instead of the FP32 MFMA one; its
approximate key sums in another order, so the results must still equal the
canonical restatement bit for bit through the certificate -- at every batch
size it serves (1..16, and 17 on the MFMA kernel beside it), with strided
sample levels, ragged and high dimensions, duplicate rows (ties by lower
index) and the graph-capture path's live-count bound (the serving loop)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("n,d", [(150_000, 96), (70_001, 100), (40_000, 33), (9_000, 300), (3_000, 2048)])
def test_small_batches_bit_exact(ca, orc, res, n, d):
    x = ca.gen_uniform(n, d, 900 + d, -1, 1)
    x[n // 2:n // 2 + 40] = x[7]  # exact duplicates of one row: ties broken by the lower index
    idx = ca.IndexBF(res, d)
    idx.add(x)
    for nq in (1, 2, 3, 8, 16, 17):
        q = ca.gen_uniform(nq, d, 901 + d + nq, -1, 1)
        q[0] = x[7]
        for k in (1, 10, 100):
            print(f"[small] n {n} d {d} nq {nq} k {k}", flush=True)
            I, D = idx.search(q, k)
            Io, Do = orc.search_canonical(x, q, k)
            assert np.array_equal(I, Io), f"n {n} d {d} nq {nq} k {k}"
            assert np.array_equal(bits(D), bits(Do)), f"n {n} d {d} nq {nq} k {k}"


def test_small_batches_through_the_server(ca, orc, res):
    """FP32 requests of 1..16 queries served (graph replays after each shape's
    first batch, the live count read on the device) equal the alone search."""
    d = 64
    x = ca.gen_uniform(60_000, d, 950, -1, 1)
    idx = ca.IndexBF(res, d)
    idx.add(x)
    srv = ca.Server(slots=2)
    sid = srv.add_index(idx)
    try:
        hs = []
        for r, nq in enumerate((1, 5, 16, 1, 9, 16, 3, 12)):
            q = ca.gen_uniform(nq, d, 960 + r, -1, 1)
            hs.append((q, srv.submit(sid, q, 10)))
        for q, h in hs:
            I, D, _ = srv.wait(h, timeout_s=60.0)
            Io, Do = orc.search_canonical(x, q, 10)
            assert np.array_equal(I, Io) and np.array_equal(bits(D), bits(Do))
    finally:
        srv.close()
