"""Dimensions above 256 on the plain FP32 path (the reference handles any d:
chooseTileSize's d > 1024 branch, src/internal/search.cu:72, and the
VeryHighDim norm kernel, src/internal/norm.cu:15,50; its GIST1M dataset is
960-d, include/dataset.h:200). The K-looped MFMA filter streams both the rows
and the queries through LDS in 32-element panels; results must equal the
canonical restatement bit for bit, as at every other d."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("n,d,nq,k", [
    (20_000, 960, 64, 10),
    (20_000, 960, 300, 100),
    (9_000, 257, 33, 7),
    (12_345, 300, 129, 50),
    (5_000, 1001, 17, 1),
    (3_000, 4096, 20, 16),
])
def test_indexbf_high_dim_bit_exact(ca, orc, res, n, d, nq, k):
    x = ca.gen_uniform(n, d, 700 + d, -1, 1)
    q = ca.gen_uniform(nq, d, 701 + d, -1, 1)
    idx = ca.IndexBF(res, d)
    idx.add(x)
    I, D = idx.search(q, k)
    Io, Do = orc.search_canonical(x, q, k)
    assert np.array_equal(I, Io)
    assert np.array_equal(bits(D), bits(Do))


def test_gist_shape_clustered(ca, orc, res):
    """960-d Gaussian mixture (close neighbours, heavy cancellation in the
    expansion): still bit-exact, with the plain path's certificate doing the
    work."""
    x = ca.gen_mixture(50_000, 960, 64, 0.05, 11, 12)
    q = ca.gen_mixture(100, 960, 64, 0.05, 11, 13)
    idx = ca.IndexBF(res, 960)
    idx.add(x)
    I, D = idx.search(q, 100)
    Io, Do = orc.search_canonical(x, q, 100)
    assert np.array_equal(I, Io) and np.array_equal(bits(D), bits(Do))
    I64, _ = orc.search_fp64(x, q, 100)
    print("960-d mixture: recall@100 vs FP64", orc.recall(I, I64), "stats", idx.last_stats())


def test_search_l2_distance_high_dim(ca, orc):
    """searchL2Distance (include/internal/search.h:8-11) on device buffers at
    d = 960, exact and relative distances."""
    n, d, nq, k = 15_000, 960, 40, 20
    x = ca.gen_uniform(n, d, 801, 0, 1)
    q = ca.gen_uniform(nq, d, 802, 0, 1)
    dx, dq = ca.DeviceArray.from_host(x), ca.DeviceArray.from_host(q)
    dI = ca.DeviceArray((nq, k), np.int32)
    dD = ca.DeviceArray((nq, k), np.float32)
    for exact in (True, False):
        ca.search_l2_distance(dx, None, dq, k, dI, dD, exact=exact)
        Io, Do = orc.search_canonical(x, q, k, exact=exact)
        assert np.array_equal(dI.to_host(), Io)
        assert np.array_equal(bits(dD.to_host()), bits(Do))


def test_precision_forms_refuse_high_dim(ca, res):
    with pytest.raises(ca.RplshError) as e:
        ca.IndexBF(res, 960, ca.IndexBFConfig(precision="bf16"))
    assert e.value.status == 1
    idx = ca.IndexBF(res, 960)
    with pytest.raises(ca.RplshError):
        idx.set_precision("fp8")
