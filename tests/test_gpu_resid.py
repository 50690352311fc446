"""Cluster-residual candidate copy (IndexBFConfig.residual_clusters): on
tight Gaussian-mixture clusters, where BF16 / FP8 copies of the raw rows
cannot order the neighbours inside a cluster, the full filter level over
(q - c).(x - c) reaches recall@k >= 0.999 against the FP64 oracle at the
default over-fetch k' = 4k. Results keep the structure of every search
(range, uniqueness, ascending distances, exact-tie order), and distances
are the exact re-rank's (within the canonical FP32 bound of the FP64
distance)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, D, CENTERS, SIGMA, DUP = 400_000, 128, 64, 0.05, 0.05
NQ = 256


@pytest.fixture(scope="module")
def mix(ca, orc):
    x = ca.gen_mixture(N, D, CENTERS, SIGMA, 11, 12, DUP)
    q = ca.gen_mixture(NQ, D, CENTERS, SIGMA, 11, 13)
    I64, _ = orc.search_fp64(x, q, 100)
    return x, q, I64


def check_rows(I, D, x):
    n = x.shape[0]
    assert ((I >= 0) & (I < n)).all()
    s = np.sort(I, axis=1)
    assert (s[:, 1:] != s[:, :-1]).all(), "duplicate index in a row"
    assert (np.diff(D, axis=1) >= 0).all(), "distances not ascending"
    r, c = np.nonzero(np.diff(D, axis=1) == 0)
    same = (x[I[r, c]] == x[I[r, c + 1]]).all(axis=1)
    assert (I[r, c][same] < I[r, c + 1][same]).all(), "exact tie not broken by the lower index"


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_residual_recall_on_tight_clusters(ca, orc, res, mix, prec):
    x, q, I64 = mix
    plain = ca.IndexBF(res, D, ca.IndexBFConfig(precision=prec))
    plain.add(x)
    idx = ca.IndexBF(res, D, ca.IndexBFConfig(precision=prec, residual_clusters=128))
    idx.add(x)
    for k in (1, 10, 100):
        print(f"[resid] {prec} k={k}", flush=True)
        Ip, _ = plain.search(q, k)
        I, Dd = idx.search(q, k)
        st = idx.last_stats()
        check_rows(I, Dd, x)
        rp = orc.recall(Ip, I64[:, :k])
        r = orc.recall(I, I64[:, :k])
        print(f"[resid] {prec} k={k} k'={st['kprime']}: recall {r:.5f} (raw copy {rp:.5f}); "
              f"rows scored {st['rows_filter_full']}, retried {st['retried_queries']}")
        assert st["rows_filter_full"] >= N
        assert r >= 0.999, f"{prec} k={k}: recall {r}"
        # distances: the exact re-rank's, within the canonical FP32 bound of the FP64 distance
        pd = orc.pair_dist_fp64(x, q, I)
        qn = orc.norms_canonical(q).astype(np.float64)
        bound = 4 * (D + 2) * 2.0 ** -24 * (qn[:, None] + 2 * float(orc.norms_canonical(x).max()))
        assert (np.abs(Dd.astype(np.float64) - pd) <= bound).all()


def test_residual_precision_switch_and_ragged_dim(ca, orc, res):
    """d = 100 (padded to 128 / 112 of the tile), precision switched back
    and forth on one index: the same results as an index built directly in
    each precision."""
    d = 100
    x = ca.gen_mixture(150_000, d, 32, SIGMA, 21, 22, DUP)
    q = ca.gen_mixture(64, d, 32, SIGMA, 21, 23)
    I64, _ = orc.search_fp64(x, q, 10)
    idx = ca.IndexBF(res, d, ca.IndexBFConfig(precision="bf16", residual_clusters=64))
    idx.add(x)
    out = {}
    for prec in ("bf16", "fp8", "bf16"):
        idx.set_precision(prec, 4)
        I, Dd = idx.search(q, 10)
        check_rows(I, Dd, x)
        r = orc.recall(I, I64)
        print(f"[resid] d={d} {prec}: recall {r:.5f}", flush=True)
        assert r >= 0.999
        if prec in out:
            assert np.array_equal(out[prec][0], I) and np.array_equal(out[prec][1].view(np.uint32), Dd.view(np.uint32))
        out[prec] = (I, Dd)
    fresh = ca.IndexBF(res, d, ca.IndexBFConfig(precision="fp8", residual_clusters=64))
    fresh.add(x)
    If, Df = fresh.search(q, 10)
    assert np.array_equal(If, out["fp8"][0]) and np.array_equal(Df.view(np.uint32), out["fp8"][1].view(np.uint32))


def test_residual_incremental_add_and_small_index(ca, orc, res):
    """Rows added in two calls rebuild the layout over all rows (the same
    results as one add); an index of at most 2,048 rows has one dense level
    and no residual copy."""
    x = ca.gen_mixture(120_000, 64, 16, SIGMA, 31, 32, DUP)
    q = ca.gen_mixture(48, 64, 16, SIGMA, 31, 33)
    a = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="fp8", residual_clusters=32))
    a.add(x[:70_000])
    a.add(x[70_000:])
    b = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="fp8", residual_clusters=32))
    b.add(x)
    Ia, Da = a.search(q, 20)
    Ib, Db = b.search(q, 20)
    assert np.array_equal(Ia, Ib) and np.array_equal(Da.view(np.uint32), Db.view(np.uint32))
    I64, _ = orc.search_fp64(x, q, 20)
    assert orc.recall(Ia, I64) >= 0.999
    s = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="bf16", residual_clusters=32))
    s.add(x[:2000])
    Is, Ds = s.search(q, 5)
    I64s, _ = orc.search_fp64(x[:2000], q, 5)
    assert orc.recall(Is, I64s) >= 0.999
