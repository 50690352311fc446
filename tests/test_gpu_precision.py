"""Precision forms (BF16 / FP8 candidates, over-fetch k' = 4k, exact FP32
re-rank, FP64 boundary) against the FP64 oracle: recall@k >= 0.999 and the
reported distances are the canonical FP32 distances of the returned rows.
Configurations are the benchmark's shapes scaled to oracle-friendly sizes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_recall(orc, x, q, I, D, k, min_recall=0.999):
    I64, _ = orc.search_fp64(x, q, k)
    r = orc.recall(I, I64)
    pd = orc.pair_dist_fp64(x, q, I)
    valid = I >= 0
    err = float(np.max(np.abs(D[valid] - pd[valid]) / np.maximum(pd[valid], 1e-30)))
    print(f"recall@{k} = {r:.5f}, max rel distance error {err:.3e}")
    assert r >= min_recall
    assert err < 1e-4
    return r


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
@pytest.mark.parametrize("k", [10, 100])
def test_unit_gaussian_96(ca, orc, res, precision, k):
    """10M x 96 unit-Gaussian workload shape at 1M rows, 1,000 queries."""
    x = ca.gen_gaussian(1_000_000, 96, 4, True)
    q = ca.gen_gaussian(1000, 96, 5, True)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision=precision, kprime_factor=4))
    idx.add(x)
    I, D = idx.search(q, k)
    check_recall(orc, x, q, I, D, k)
    st = idx.last_stats()
    assert st["kprime"] == 4 * k


@pytest.mark.parametrize("k", [1, 100])
def test_fp8_mixture_duplicates(ca, orc, res, k):
    """20M x 128 FP8 workload shape at 2M rows: 1,024 centres, sigma 0.05,
    5% exact duplicates (ties), 512 queries."""
    x = ca.gen_mixture(2_000_000, 128, 1024, 0.05, 6, 7, 0.05)
    q = ca.gen_mixture(512, 128, 1024, 0.05, 6, 70)
    idx = ca.IndexBF(res, 128, ca.IndexBFConfig(precision="fp8", kprime_factor=4))
    idx.add(x)
    I, D = idx.search(q, k)
    check_recall(orc, x, q, I, D, k)


def test_fp64_boundary_orders_like_oracle(ca, orc, res):
    """With the FP64 boundary re-rank, the k-th result agrees with the FP64
    oracle's set even where FP32 distances tie."""
    x = ca.gen_mixture(300_000, 64, 64, 0.02, 1, 2, 0.05)
    q = ca.gen_mixture(300, 64, 64, 0.02, 1, 20)
    idx = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    I, D = idx.search(q, 10)
    I64, _ = orc.search_fp64(x, q, 10)
    assert orc.recall(I, I64) >= 0.999


def test_precision_switch_keeps_plain_exact(ca, orc, res):
    x = ca.gen_uniform(50000, 128, 1)
    q = ca.gen_uniform(100, 128, 2)
    idx = ca.IndexBF(res, 128, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    idx.search(q, 10)
    idx.set_precision("fp32")
    I, D = idx.search(q, 10)
    Io, Do = orc.search_canonical(x, q, 10)
    assert np.array_equal(I, Io) and np.array_equal(D.view(np.uint32), Do.view(np.uint32))
