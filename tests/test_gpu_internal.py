"""The reference's internal selection / broadcast functions (l2Select,
blockSelect x2, broadcastAlongRowsSum; include/internal/select.h:8-34,
broadcast.h:7-8) against numpy: one FP32 add then the k smallest (largest) by
(value, label), padded when k exceeds the row; bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _values(rows, n, seed, dup=True):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-4, 4, (rows, n)).astype(np.float32)
    v[v == 0] = 1.0  # keep +-0 out: their relative order is a convention
    if dup and n > 8:  # exact ties: copy some values along each row
        v[:, 1::7] = v[:, 0:1]
    return v


def _ref_select(keys, labels, k, largest=False):
    rows, n = keys.shape
    K = np.full((rows, k), -np.finfo(np.float32).max if largest else np.finfo(np.float32).max, np.float32)
    V = np.full((rows, k), -1, np.int32)
    for r in range(rows):
        order = np.lexsort((labels[r], -keys[r] if largest else keys[r]))[:k]
        K[r, :order.size] = keys[r][order]
        V[r, :order.size] = labels[r][order]
    return K, V


@pytest.mark.parametrize("rows,n,k", [(7, 100, 10), (33, 5000, 100), (5, 20000, 1024), (4, 12, 40), (3, 1, 1)])
def test_l2_select(ca, res, rows, n, k):
    p = _values(rows, n, rows + n)
    norms = np.random.default_rng(n).uniform(0, 3, n).astype(np.float32)
    dp, dn = ca.DeviceArray.from_host(p), ca.DeviceArray.from_host(norms)
    dD, dI = ca.DeviceArray((rows, k), np.float32), ca.DeviceArray((rows, k), np.int32)
    ca.l2_select(dp, dn, k, dD, dI)
    keys = (norms[None, :] + p).astype(np.float32)
    Kr, Vr = _ref_select(keys, np.broadcast_to(np.arange(n, dtype=np.int32), (rows, n)), k)
    assert np.array_equal(dI.to_host(), Vr)
    assert np.array_equal(bits(dD.to_host()), bits(Kr))


@pytest.mark.parametrize("largest", [False, True])
@pytest.mark.parametrize("with_labels", [False, True])
@pytest.mark.parametrize("rows,n,k", [(9, 700, 25), (6, 9000, 600), (2, 5, 9)])
def test_block_select(ca, res, rows, n, k, largest, with_labels):
    v = _values(rows, n, 3 * rows + n)
    labels = (np.random.default_rng(k).permutation(rows * n).reshape(rows, n).astype(np.int32) if with_labels
              else np.broadcast_to(np.arange(n, dtype=np.int32), (rows, n)).copy())
    dv = ca.DeviceArray.from_host(v)
    dl = ca.DeviceArray.from_host(labels) if with_labels else None
    dK, dV = ca.DeviceArray((rows, k), np.float32), ca.DeviceArray((rows, k), np.int32)
    ca.block_select(dv, k, dK, dV, labels=dl, select_max=largest)
    Kr, Vr = _ref_select(v, labels, k, largest)
    assert np.array_equal(dV.to_host(), Vr)
    assert np.array_equal(bits(dK.to_host()), bits(Kr))


@pytest.mark.parametrize("rows,cols", [(1, 1), (17, 129), (1000, 96)])
def test_broadcast_rows_sum(ca, res, rows, cols):
    m = _values(rows, cols, 5, dup=False)
    v = np.random.default_rng(6).uniform(-2, 2, rows).astype(np.float32)
    dm, dv = ca.DeviceArray.from_host(m), ca.DeviceArray.from_host(v)
    ca.broadcast_rows_sum(dv, dm)
    assert np.array_equal(bits(dm.to_host()), bits((m + v[:, None]).astype(np.float32)))


def test_search_decomposes_into_the_internal_functions(ca, orc, res):
    """searchL2Distance == l2Select over -2 Q X^T (canonical dots) + broadcast of
    ||q||^2, on the canonical restatement's own products."""
    x = ca.gen_uniform(3000, 32, 1, -1, 1)
    q = ca.gen_uniform(20, 32, 2, -1, 1)
    xn = orc.norms_canonical(x)
    qn = orc.norms_canonical(q)
    prod = np.empty((20, 3000), np.float32)
    for i in range(20):
        for j in range(3000):
            prod[i, j] = np.float32(-2.0) * np.float32(orc.load().oracle_dot_canonical(
                q[i].ctypes.data, x[j].ctypes.data, 32))
    dp, dn = ca.DeviceArray.from_host(prod), ca.DeviceArray.from_host(xn)
    dD, dI = ca.DeviceArray((20, 10), np.float32), ca.DeviceArray((20, 10), np.int32)
    ca.l2_select(dp, dn, 10, dD, dI)
    ca.broadcast_rows_sum(ca.DeviceArray.from_host(qn), dD)
    Io, Do = orc.search_canonical(x, q, 10)
    assert np.array_equal(dI.to_host(), Io)
    assert np.array_equal(bits(dD.to_host()), bits(Do))
