"""The five workload shapes at their real sizes (the configurations the
benchmark and the design quote), checked where the CPU oracle is affordable
(query subsets) and through size-independent properties on every query:
indices in [0, n) and unique per row, distances ascending with ties broken by
the lower index, returned distances within the canonical FP32 error bound of
the FP64 distance of the same pair.

C3: 10,000,000 x 96 unit Gaussian (seed 4), 10,000 queries (seed 5), k 10/100.
C4: 20,000,000 x 128 mixture, 1,024 centres, sigma 0.05, 5 % exact duplicate
    rows (seed 6), 4,096 queries (seed 7), k 1/100/1024 (k' up to 4,096).
    Measured on 32 queries: plain FP32 recall 1.0; BF16 0.84 (k=1) / 0.989
    (k=100) at k' = 4k and 1.0 at 16k; FP8 e4m3 <= 0.26 even at 16-64k: its
    3-bit mantissa is coarser than the gaps between neighbours inside a
    ~19,500-point cluster."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def check_rows(I, D, x):
    """Range, uniqueness, ascending distances; where neighbours in a row are
    identical vectors (an exact tie of the selection key) the lower index
    comes first. (Equal final distances of different vectors are not ties of
    the key: ||x||^2 - 2 q.x is selected on, + ||q||^2 is rounded after.)"""
    n = x.shape[0]
    assert ((I >= 0) & (I < n)).all()
    s = np.sort(I, axis=1)
    assert (s[:, 1:] != s[:, :-1]).all(), "duplicate index in a row"
    dd = np.diff(D, axis=1)
    assert (dd >= 0).all(), "distances not ascending"
    r, c = np.nonzero(dd == 0)
    same = (x[I[r, c]] == x[I[r, c + 1]]).all(axis=1)
    assert (I[r, c][same] < I[r, c + 1][same]).all(), "exact tie not broken by the lower index"
    return int(same.sum())


def check_error_bound(orc, x, q, I, D):
    d = x.shape[1]
    pd = orc.pair_dist_fp64(x, q, I)
    qn = orc.norms_canonical(q).astype(np.float64)
    xn_max = float(orc.norms_canonical(x).max())
    bound = 4 * (d + 2) * 2.0 ** -24 * (qn[:, None] + 2 * xn_max)
    assert (np.abs(D.astype(np.float64) - pd) <= bound).all()


@pytest.fixture(scope="module")
def c3(ca, res):
    x = ca.gen_gaussian(10_000_000, 96, 4, True)
    q = ca.gen_gaussian(10_000, 96, 5, True)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision="fp32"))
    idx.add(x)
    yield x, q, idx
    idx.close() if hasattr(idx, "close") else None


@pytest.mark.parametrize("k", [10, 100])
def test_c3_plain(ca, orc, c3, k):
    x, q, idx = c3
    idx.set_precision("fp32")
    I, D = idx.search(q, k)
    check_rows(I, D, x)
    Io, Do = orc.search_canonical(x, q[:64], k)
    assert np.array_equal(I[:64], Io) and np.array_equal(bits(D[:64]), bits(Do))
    Io2, Do2 = orc.search_canonical(x, q[-64:], k)
    assert np.array_equal(I[-64:], Io2) and np.array_equal(bits(D[-64:]), bits(Do2))


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
@pytest.mark.parametrize("k", [10, 100])
def test_c3_precision(ca, orc, c3, precision, k):
    x, q, idx = c3
    idx.set_precision(precision, 4)
    I, D = idx.search(q, k)
    check_rows(I, D, x)
    sub = slice(0, 200)
    I64, _ = orc.search_fp64(x, q[sub], k)
    r = orc.recall(I[sub], I64)
    print(f"C3 {precision} k={k}: recall@{k} {r:.5f} on 200 queries; stats {idx.last_stats()}")
    assert r >= 0.999
    check_error_bound(orc, x, q[sub], I[sub], D[sub])
    idx.set_precision("fp32")


@pytest.fixture(scope="module")
def c4(ca, res):
    x = ca.gen_mixture(20_000_000, 128, 1024, 0.05, 6, 60, 0.05)
    q = ca.gen_mixture(4096, 128, 1024, 0.05, 6, 7)
    idx = ca.IndexBF(res, 128, ca.IndexBFConfig(precision="fp32"))
    idx.add(x)
    yield x, q, idx


@pytest.mark.parametrize("k", [1, 100, 1024])
def test_c4_plain(ca, orc, c4, k):
    x, q, idx = c4
    idx.set_precision("fp32")
    I, D = idx.search(q, k)
    check_rows(I, D, x)
    Io, Do = orc.search_canonical(x, q[:24], k)
    assert np.array_equal(I[:24], Io) and np.array_equal(bits(D[:24]), bits(Do))


@pytest.mark.parametrize("k", [1, 100])
def test_c4_bf16_overfetch(ca, orc, c4, k):
    """BF16 candidates on the tight clusters: with k' = 16k (what the
    autotuner's recall target selects here) recall@k reaches 0.999."""
    x, q, idx = c4
    idx.set_precision("bf16", 16)
    I, D = idx.search(q, k)
    check_rows(I, D, x)
    sub = slice(0, 32)
    I64, _ = orc.search_fp64(x, q[sub], k)
    r = orc.recall(I[sub], I64)
    print(f"C4 bf16 k={k} k'={16 * k}: recall@{k} {r:.4f} on 32 queries")
    assert r >= 0.999
    check_error_bound(orc, x, q[sub], I[sub], D[sub])
    idx.set_precision("fp32")


@pytest.mark.parametrize("k", [1, 100, 1024])
def test_c4_fp8(ca, orc, c4, k):
    """FP8 with k' = 4k (4,096 at k = 1,024) and the FP64 boundary re-rank.
    On these tight clusters the e4m3 candidates lose recall to cancellation
    (see test_gpu_precision: equal to a CPU emulation of the arithmetic at
    200k rows); here the recall is reported and the result must be well formed
    with exact distances."""
    x, q, idx = c4
    idx.set_precision("fp8", 4)
    I, D = idx.search(q, k)
    st = idx.last_stats()
    check_rows(I, D, x)
    sub = slice(0, 32)
    I64, _ = orc.search_fp64(x, q[sub], k)
    r = orc.recall(I[sub], I64)
    print(f"C4 fp8 k={k} (k'={st['kprime']}): recall@{k} {r:.4f} on 32 queries; stats {st}")
    assert st["kprime"] == 4 * k
    check_error_bound(orc, x, q[sub], I[sub], D[sub])
    idx.set_precision("fp32")
