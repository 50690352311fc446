"""Pins the LSH restatement in the oracle (oracle_lsh_hash / oracle_lsh_search)
against formulations that do not share its code:

* codes: signs of float64 projections minus the offsets (numpy matmul), on
  every bit whose float64 value is far from its offset; margins within the
  FP32 forward-error bound;
* probing every bucket (1 bit, 1 extra probe) with an unlimited budget must
  reproduce the canonical brute-force search exactly (itself pinned to the
  golden vectors);
* unlimited budget: the candidate set is {rows sharing a probed bucket},
  computed by set membership over numpy codes, and the result is the exact
  search restricted to that set;
* limited budget: the first B entries of the bucket concatenation, built in
  numpy from the codes, then de-duplicated;
* BF16 pre-rank: keeps everything when f*k covers the candidates or the
  budget; on integer data (exact in BF16) changes nothing; elsewhere keeps the
  f*k smallest BF16 distances recomputed in float64 from the pinned encoder.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def _data(n, nq, d, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d)).astype(np.float32)
    q = rng.standard_normal((nq, d)).astype(np.float32)
    return x, q


def _w(L, K, d, seed):
    return np.random.default_rng(seed).standard_normal((L * K, d)).astype(np.float32)


def _probe_codes(codes_q, marg_q, L, K, P):
    """[nq][P+1][L] probe codes from numpy: bit order by (|dot|, bit)."""
    nq = codes_q.shape[1]
    out = np.empty((nq, P + 1, L), np.uint64)
    for i in range(nq):
        for t in range(L):
            c = int(codes_q[t, i])
            out[i, 0, t] = c
            m = np.abs(marg_q[i, t * K:(t + 1) * K].astype(np.float64))
            order = np.lexsort((np.arange(K), m))  # by |m|, then bit
            for p in range(1, P + 1):
                out[i, p, t] = c ^ (1 << int(order[p - 1]))
    return out


@pytest.mark.parametrize("centered", [False, True])
def test_codes_match_float64_signs(centered):
    L, K, d = 3, 7, 33
    x, _ = _data(2000, 1, d, 1)
    w = _w(L, K, d, 2)
    off = np.random.default_rng(3).standard_normal(L * K).astype(np.float32) if centered else None
    codes, marg = orc.lsh_hash(x, w, L, K, off)
    o64 = np.zeros(L * K) if off is None else off.astype(np.float64)
    m64 = x.astype(np.float64) @ w.astype(np.float64).T - o64  # [n][L*K]
    scale = np.abs(x).astype(np.float64) @ np.abs(w).astype(np.float64).T + np.abs(o64)
    bound = (d + 3) * 2.0 ** -24 * scale
    assert np.all(np.abs(marg - m64) <= bound)
    for t in range(L):
        for b in range(K):
            col = t * K + b
            sure = np.abs(m64[:, col]) > bound[:, col]
            bit = (codes[t] >> np.uint32(b)) & np.uint32(1)
            assert np.array_equal(bit[sure].astype(bool), m64[sure, col] >= 0), (t, b)


def test_all_buckets_is_brute_force():
    x, q = _data(700, 40, 24, 3)
    w = _w(1, 1, 24, 4)
    for k in (1, 10, 50):
        I, D, raw, distinct = orc.lsh_search(x, w, 1, 1, 1, 4096, q, k)
        Io, Do = orc.search_canonical(x, q, k)
        assert np.array_equal(I, Io) and np.array_equal(D.view(np.uint32), Do.view(np.uint32))
        assert raw == 40 * 700 and distinct == 40 * 700
    # three tables, every bucket of each: each row appears three times before de-duplication
    w3 = _w(3, 1, 24, 5)
    I, D, raw, distinct = orc.lsh_search(x[:500], w3, 3, 1, 1, 4096, q, 10)
    Io, Do = orc.search_canonical(x[:500], q, 10)
    assert np.array_equal(I, Io) and np.array_equal(D.view(np.uint32), Do.view(np.uint32))
    assert raw == 40 * 1500 and distinct == 40 * 500


def _medians(x, w, L, K):
    """Offsets at the projections' medians over x (float64 ranks of the FP32 dots)."""
    _, m = orc.lsh_hash(x, w, L, K)
    return np.sort(m, axis=0)[x.shape[0] // 2].astype(np.float32)


@pytest.mark.parametrize("L,K,P,centered", [(4, 6, 0, False), (4, 6, 2, False), (8, 3, 3, False),
                                            (2, 10, 5, False), (4, 6, 2, True), (2, 9, 4, True)])
def test_unlimited_budget_is_bucket_union(L, K, P, centered):
    x, q = _data(400, 30, 16, 10 + L + K + P)
    x = x + np.float32(2.0)  # away from the origin
    q = q + np.float32(2.0)
    w = _w(L, K, 16, 20 + P)
    off = _medians(x, w, L, K) if centered else None
    codes, _ = orc.lsh_hash(x, w, L, K, off)
    cq, mq = orc.lsh_hash(q, w, L, K, off)
    probes = _probe_codes(cq, mq, L, K, P)
    k = 8
    I, D, raw, distinct = orc.lsh_search(x, w, L, K, P, 4096, q, k, off)
    tot_distinct = 0
    for i in range(q.shape[0]):
        member = np.zeros(x.shape[0], bool)
        for t in range(L):
            member |= np.isin(codes[t].astype(np.uint64), probes[i, :, t])
        rows = np.nonzero(member)[0]
        entries = sum(int(np.isin(codes[t].astype(np.uint64), probes[i, :, t]).sum()) for t in range(L))
        assert entries <= 4096  # nothing was cut
        tot_distinct += rows.size
        if rows.size == 0:
            assert (I[i] == -1).all()
            continue
        Io, Do = orc.search_canonical(x[rows], q[i:i + 1], min(k, rows.size))
        m = min(k, rows.size)
        assert np.array_equal(I[i, :m], rows[Io[0]]), i
        assert np.array_equal(D[i, :m].view(np.uint32), Do[0].view(np.uint32)), i
        assert (I[i, m:] == -1).all() and (D[i, m:] == np.finfo(np.float32).max).all()
    assert distinct == tot_distinct


@pytest.mark.parametrize("budget", [1, 7, 64, 300])
def test_budget_cuts_the_concatenation(budget):
    L, K, P = 5, 4, 2
    x, q = _data(800, 12, 20, 30)
    w = _w(L, K, 20, 31)
    off = _medians(x, w, L, K)
    codes, _ = orc.lsh_hash(x, w, L, K, off)
    cq, mq = orc.lsh_hash(q, w, L, K, off)
    probes = _probe_codes(cq, mq, L, K, P)
    k = 5
    I, D, raw, distinct = orc.lsh_search(x, w, L, K, P, budget, q, k, off)
    raw_ref = dist_ref = 0
    for i in range(q.shape[0]):
        seq = []
        for p in range(P + 1):
            for t in range(L):
                seq.extend(np.nonzero(codes[t].astype(np.uint64) == probes[i, p, t])[0].tolist())
        seq = seq[:budget]
        raw_ref += len(seq)
        rows = np.unique(np.array(seq, np.int64))
        dist_ref += rows.size
        m = min(k, rows.size)
        if m:
            Io, Do = orc.search_canonical(x[rows], q[i:i + 1], m)
            assert np.array_equal(I[i, :m], rows[Io[0]]), i
            assert np.array_equal(D[i, :m].view(np.uint32), Do[0].view(np.uint32)), i
        assert (I[i, m:] == -1).all()
    assert raw == raw_ref and distinct == dist_ref


def test_median_offsets_balance_every_bit():
    """With median offsets each bit splits the rows in half (ties aside)."""
    L, K = 4, 8
    x, _ = _data(1001, 1, 12, 50)
    x = x * np.float32(0.1) + np.float32(3.0)  # a tight cluster far from the origin
    w = _w(L, K, 12, 51)
    c0, _ = orc.lsh_hash(x, w, L, K)
    c1, _ = orc.lsh_hash(x, w, L, K, _medians(x, w, L, K))
    for t in range(L):
        for b in range(K):
            ones = int(((c1[t] >> np.uint32(b)) & np.uint32(1)).sum())
            assert ones == 1001 - 1001 // 2, (t, b, ones)  # rows >= the median
    # through the origin the cluster mostly lands in few buckets
    assert len(np.unique(c0[0])) < len(np.unique(c1[0]))


def test_empty_index_pads():
    _, q = _data(1, 3, 8, 40)
    I, D, raw, distinct = orc.lsh_search(np.zeros((0, 8), np.float32), _w(2, 3, 8, 41), 2, 3, 1, 10, q, 4)
    assert (I == -1).all() and (D == np.finfo(np.float32).max).all() and raw == 0 and distinct == 0


def test_prerank_covering_all_candidates_changes_nothing():
    """prerank * k >= every query's distinct candidates: the pre-rank keeps them all."""
    L, K, P, k = 4, 6, 1, 5
    x, q = _data(300, 20, 16, 60)
    w = _w(L, K, 16, 61)
    I0, D0, r0, d0 = orc.lsh_search(x, w, L, K, P, 4096, q, k)
    # 64 * 5 = 320 >= 300 rows: every query's distinct set survives the pre-rank
    I1, D1, r1, d1 = orc.lsh_search(x, w, L, K, P, 4096, q, k, prerank=64)
    assert np.array_equal(I1, I0) and np.array_equal(D1.view(np.uint32), D0.view(np.uint32))
    assert (r1, d1) == (r0, d0)
    # f * k >= the budget: no pre-rank stage at all
    I2, D2, _, _ = orc.lsh_search(x, w, L, K, P, 200, q, k, prerank=40)
    I3, D3, _, _ = orc.lsh_search(x, w, L, K, P, 200, q, k)
    assert np.array_equal(I2, I3) and np.array_equal(D2.view(np.uint32), D3.view(np.uint32))
    # a pre-rank that cuts (f * k = 10 < distinct) still keeps most exact neighbours
    I4, _, _, _ = orc.lsh_search(x, w, L, K, P, 4096, q, k, prerank=2)
    assert (I4 == I0).mean() > 0.5


def test_prerank_on_integer_data_is_exact():
    """Small integers are exact in BF16 and their dot products exact in FP32,
    so the BF16 distances equal the exact ones: with f >= 1 the pre-rank keeps
    the f*k nearest candidates and the result equals the unranked one."""
    rng = np.random.default_rng(70)
    x = rng.integers(-6, 7, (900, 12)).astype(np.float32)
    q = rng.integers(-6, 7, (25, 12)).astype(np.float32)
    L, K, P = 3, 5, 2
    w = _w(L, K, 12, 71)
    I0, D0, _, _ = orc.lsh_search(x, w, L, K, P, 2048, q, 8)
    for f in (1, 2, 4):
        I1, D1, _, _ = orc.lsh_search(x, w, L, K, P, 2048, q, 8, prerank=f)
        assert np.array_equal(I1, I0) and np.array_equal(D1.view(np.uint32), D0.view(np.uint32)), f


def test_prerank_keeps_the_bf16_nearest():
    """The rows kept are the f*k smallest BF16 distances, rebuilt in numpy
    (bf16 via the pinned encoder, distances in float64: ties and near-ties of
    the FP32 chain excluded by a margin) -- the final result is the exact
    search over that set."""
    L, K, P, k, f = 4, 5, 1, 6, 2
    x, q = _data(700, 15, 24, 80)
    w = _w(L, K, 24, 81)
    codes, _ = orc.lsh_hash(x, w, L, K)
    cq, mq = orc.lsh_hash(q, w, L, K)
    probes = _probe_codes(cq, mq, L, K, P)
    I, D, _, _ = orc.lsh_search(x, w, L, K, P, 4096, q, k, prerank=f)
    bf = np.vectorize(lambda v: orc.bf16_to_float(orc.bf16_rne(float(v))), otypes=[np.float32])
    xb, qb = bf(x).astype(np.float64), bf(q).astype(np.float64)
    xn = orc.norms_canonical(x).astype(np.float64)
    checked = 0
    for i in range(q.shape[0]):
        member = np.zeros(x.shape[0], bool)
        for t in range(L):
            member |= np.isin(codes[t].astype(np.uint64), probes[i, :, t])
        rows = np.nonzero(member)[0]
        if rows.size <= f * k:
            continue
        relb = xn[rows] - 2.0 * (xb[rows] @ qb[i])
        order = np.argsort(relb, kind="stable")
        gap = relb[order[f * k]] - relb[order[f * k - 1]]
        if gap < 1e-3 * (1 + abs(relb[order[f * k]])):
            continue  # the FP32 chain may order these two differently
        keep = np.sort(rows[order[:f * k]])
        Io, Do = orc.search_canonical(x[keep], q[i:i + 1], k)
        assert np.array_equal(I[i], keep[Io[0]]), i
        assert np.array_equal(D[i].view(np.uint32), Do[0].view(np.uint32)), i
        checked += 1
    assert checked >= 10


def test_bf16_tree_dot_order_and_exactness():
    """The piece/tree order: with 4 pieces holding 2^24, 1, -2^24, 1 the tree
    (0+2, 1+3, then 0+1) gives 2 exactly, where a sequential chain gives 1
    (2^24 + 1 rounds to 2^24) and adjacent pairing gives 1 as well."""
    a = np.zeros(32, np.float32)
    b = np.zeros(32, np.float32)
    a[[0, 8, 16, 24]] = 1.0
    b[[0, 8, 16, 24]] = [2.0 ** 24, 1.0, -(2.0 ** 24), 1.0]
    assert orc.bf16_tree_dot(a, b) == np.float32(2.0)
    assert orc.bf16_tree_dot(a[:25], b[:25]) == np.float32(2.0)  # ragged last piece, same pieces
    # 16 pieces (G = 16): the first pairing is p with p + 8
    a = np.zeros(128, np.float32)
    b = np.zeros(128, np.float32)
    a[[0, 64, 8, 72]] = 1.0
    b[[0, 64, 8, 72]] = [2.0 ** 24, -(2.0 ** 24), 1.0, 1.0]  # pieces 0, 8, 1, 9
    assert orc.bf16_tree_dot(a, b) == np.float32(2.0)
    # integer data: every order is exact, the result is the integer dot product
    rng = np.random.default_rng(3)
    for d in (1, 7, 8, 9, 96, 100, 128, 200, 513, 1030):
        u = rng.integers(-9, 10, d).astype(np.float32)
        v = rng.integers(-9, 10, d).astype(np.float32)
        assert orc.bf16_tree_dot(u, v) == np.float32(int(u.astype(np.int64) @ v.astype(np.int64))), d
    # BF16 rounding is applied to both operands: 1 + 2^-9 rounds to 1
    u = np.full(4, 1.0 + 2.0 ** -9, np.float32)
    assert orc.bf16_tree_dot(u, np.ones(4, np.float32)) == np.float32(4.0)
    # random data: within the FP32 forward-error bound of the float64 BF16 product
    bf = np.vectorize(lambda t: orc.bf16_to_float(orc.bf16_rne(float(t))), otypes=[np.float64])
    for d in (33, 128, 700):
        u = rng.standard_normal(d).astype(np.float32)
        v = rng.standard_normal(d).astype(np.float32)
        ub, vb = bf(u), bf(v)
        ex = float(ub @ vb)
        bound = (d + 8) * 2.0 ** -24 * float(np.abs(ub) @ np.abs(vb))
        assert abs(float(orc.bf16_tree_dot(u, v)) - ex) <= bound, d
