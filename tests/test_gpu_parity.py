"""Plain FP32 path on the GPU against the canonical CPU restatement: indices
and distances bit-exact (the canonical order is the parity contract; the
reference's own FP32 order is unobservable), plus the reference's fixtures
and edge cases (empty index, k > n, ties, odd shapes)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

FMAX = np.finfo(np.float32).max


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bitexact(I, D, Io, Do):
    bad = np.nonzero((I != Io).any(1) | (bits(D) != bits(Do)).any(1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: {I[bad[0]][:8]} vs {Io[bad[0]][:8]}"


@pytest.mark.parametrize("n,d", [(10000, 128), (5000, 96), (3001, 100), (777, 256), (1, 7), (65, 3)])
def test_norms_bitexact(ca, orc, res, n, d):
    x = ca.gen_uniform(n, d, 7, -1, 1)
    dx = ca.DeviceArray.from_host(x)
    dn = ca.DeviceArray((n,), np.float32)
    ca.compute_l2_norm(dx, dn)
    assert np.array_equal(bits(dn.to_host()), bits(orc.norms_canonical(x)))


@pytest.mark.parametrize("d", [32, 64, 128, 100, 1500])
def test_norm_fixture_all_ones(ca, res, d):
    x = np.ones((1001, d), np.float32)
    dx = ca.DeviceArray.from_host(x)
    dn = ca.DeviceArray((1001,), np.float32)
    ca.compute_l2_norm(dx, dn)
    assert (dn.to_host() == d).all()


def test_encode_bf16_bitexact(ca, orc, res):
    x = ca.gen_gaussian(300, 96, 1) * np.float32(7.5)
    dx = ca.DeviceArray.from_host(x)
    out = ca.DeviceArray((300, 112), np.uint16)
    ca.encode_bf16(dx, 112, out)
    got = out.to_host()
    assert (got[:, 96:] == 0).all()
    ref = np.array([orc.bf16_rne(float(v)) for v in x.ravel()], np.uint16).reshape(300, 96)
    assert np.array_equal(got[:, :96], ref)


def test_encode_e4m3_bitexact(ca, orc, res):
    x = ca.gen_mixture(200, 100, 16, 0.3, 1, 2)
    x[3] = 0.0
    dx = ca.DeviceArray.from_host(x)
    out = ca.DeviceArray((200, 128), np.uint8)
    sc = ca.DeviceArray((200,), np.float32)
    ca.encode_e4m3(dx, 128, out, sc)
    got, gs = out.to_host(), sc.to_host()
    assert (got[:, 100:] == 0).all()
    for i in range(200):
        mx = np.float32(np.abs(x[i]).max())
        if mx == 0:
            s = np.float32(1.0)
        else:
            s0 = np.float32(mx / np.float32(448.0))
            m, e = np.frexp(s0)
            s = s0 if m == 0.5 else np.float32(np.ldexp(np.float32(1.0), int(e)))
        assert gs[i] == s
        ref = [orc.e4m3_rne_sat(float(np.float32(v / s))) for v in x[i]]
        assert list(got[i, :100]) == ref


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_plain_matches_golden(ca, res, name):
    g = dict(np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False))
    gen, k = make_golden.CASES[name]
    x, q = gen()
    idx = ca.IndexBF(res, x.shape[1])
    idx.add(x)
    I, D = idx.search(q, k)
    assert_bitexact(I, D, g["I"], g["D"])


@pytest.mark.parametrize("d", [32, 64, 96, 100, 128, 200, 256])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_plain_shapes(ca, orc, res, d, k):
    n, nq = 30011, 97
    x = ca.gen_uniform(n, d, 100 + d, -1, 1)
    q = ca.gen_uniform(nq, d, 200 + d, -1, 1)
    dx, dq = ca.DeviceArray.from_host(x), ca.DeviceArray.from_host(q)
    dI, dD = ca.DeviceArray((nq, k), np.int32), ca.DeviceArray((nq, k), np.float32)
    ca.search_l2_distance(dx, None, dq, k, dI, dD, exact=True)
    Io, Do = orc.search_canonical(x, q, k)
    assert_bitexact(dI.to_host(), dD.to_host(), Io, Do)
    # relative distances (needRealDistances = false)
    ca.search_l2_distance(dx, None, dq, k, dI, dD, exact=False)
    Io, Do = orc.search_canonical(x, q, k, exact=False)
    assert_bitexact(dI.to_host(), dD.to_host(), Io, Do)


def test_plain_c1_full(ca, orc, res):
    """100,000 x 128 U[0,1) seed 0, 1,000 queries seed 1, k = 10."""
    x = ca.gen_uniform(100000, 128, 0)
    q = ca.gen_uniform(1000, 128, 1)
    idx = ca.IndexBF(res, 128)
    idx.add(x)
    I, D = idx.search(q, 10)
    Io, Do = orc.search_canonical(x, q, 10)
    assert_bitexact(I, D, Io, Do)


def test_plain_c2_subset(ca, orc, res):
    """1,000,000 x 128 mixture (256 centres U[-1,1), sigma 0.1), k = 100, on
    1,000 of the 10,000 queries: bit-exact vs canonical; distance error vs
    FP64 reported (the FP32 expansion itself is ~1e-5 off FP64 here)."""
    x = ca.gen_mixture(1_000_000, 128, 256, 0.1, 2, 3)
    q = ca.gen_mixture(1000, 128, 256, 0.1, 2, 30)
    idx = ca.IndexBF(res, 128)
    idx.add(x)
    I, D = idx.search(q, 100)
    Io, Do = orc.search_canonical(x, q, 100)
    assert_bitexact(I, D, Io, Do)
    pd = orc.pair_dist_fp64(x, q[:100], I[:100])
    err = np.max(np.abs(D[:100] - pd) / pd)
    print(f"max rel error vs FP64 of the canonical FP32 distances: {err:.3e}")
    assert err < 1e-4


def fixture_search():
    dim, nb, nq = 128, 1000, 100
    q = np.ones((nq, dim), np.float32)
    b = np.minimum(np.arange(nb), 10).astype(np.float32)[:, None] * np.ones((1, dim), np.float32)
    return b, q


def test_search_fixture_internal_api(ca, res):
    """test/test_search.cpp:17-36 through searchL2Distance, relative distances."""
    b, q = fixture_search()
    db, dq = ca.DeviceArray.from_host(b), ca.DeviceArray.from_host(q)
    dI, dD = ca.DeviceArray((100, 10), np.int32), ca.DeviceArray((100, 10), np.float32)
    ca.search_l2_distance(db, None, dq, 10, dI, dD, exact=False)
    assert (dI.to_host() == np.array([1, 0, 2, 3, 4, 5, 6, 7, 8, 9])).all()
    assert (dD.to_host() == np.array([-128, 0, 0, 384, 1024, 1920, 3072, 4480, 6144, 8064])).all()


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_search_fixture_ties(ca, res, precision):
    """990 identical rows: the lowest indices win, for every precision."""
    b, q = fixture_search()
    idx = ca.IndexBF(res, 128, ca.IndexBFConfig(precision=precision))
    idx.add(b)
    I, D = idx.search(q, 100)
    assert (I[:, :10] == np.array([1, 0, 2, 3, 4, 5, 6, 7, 8, 9])).all()
    assert (I[:, 10:] == np.arange(10, 100)).all()
    assert (D[:, 10:] == 128 * 81).all()
    assert (D[0, :10] == np.array([0, 128, 128, 512, 1152, 2048, 3200, 4608, 6272, 8192])).all()


def test_empty_index_and_padding(ca, orc, res):
    idx = ca.IndexBF(res, 16)
    q = ca.gen_uniform(5, 16, 1)
    I, D = idx.search(q, 4)
    assert (I == -1).all() and (D == FMAX).all()
    x = ca.gen_uniform(3, 16, 2)
    idx.add(x)
    I, D = idx.search(q, 6)
    Io, Do = orc.search_canonical(x, q, 6)
    assert_bitexact(I, D, Io, Do)
    assert (I[:, 3:] == -1).all() and (D[:, 3:] == FMAX).all()


def test_single_query_and_large_k(ca, orc, res):
    x = ca.gen_uniform(200000, 64, 3, -1, 1)
    q = ca.gen_uniform(3, 64, 4, -1, 1)
    idx = ca.IndexBF(res, 64)
    idx.add(x)
    for k in (1, 1024):
        I, D = idx.search(q[:1], k)
        Io, Do = orc.search_canonical(x, q[:1], k)
        assert_bitexact(I, D, Io, Do)


def test_api_pointers_add_reset(ca, orc, res):
    x = ca.gen_uniform(40000, 96, 5)
    q = ca.gen_uniform(200, 96, 6)
    Io, Do = orc.search_canonical(x, q, 25)
    idx = ca.IndexBF(res, 96)
    idx.add(x[:12345])
    idx.add(ca.DeviceArray.from_host(x[12345:]))  # device-pointer add, appended
    assert idx.num_data == 40000
    I, D = idx.search(q, 25)
    assert_bitexact(I, D, Io, Do)
    dq = ca.DeviceArray.from_host(q)
    dI, dD = ca.DeviceArray((200, 25), np.int32), ca.DeviceArray((200, 25), np.float32)
    idx.search(dq, 25, dI, dD)
    assert_bitexact(dI.to_host(), dD.to_host(), Io, Do)
    idx.reset()
    assert idx.num_data == 0
    I, _ = idx.search(q, 5)
    assert (I == -1).all()


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_batch_invariance(ca, res, precision):
    """A query's result does not depend on the other queries of its batch."""
    x = ca.gen_gaussian(150000, 96, 8, True)
    q = ca.gen_gaussian(700, 96, 9, True)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision=precision))
    idx.add(x)
    I, D = idx.search(q, 50)
    parts = [idx.search(q[a:b], 50) for a, b in [(0, 1), (1, 300), (300, 700)]]
    assert np.array_equal(np.concatenate([p[0] for p in parts]), I)
    assert np.array_equal(bits(np.concatenate([p[1] for p in parts])), bits(D))
