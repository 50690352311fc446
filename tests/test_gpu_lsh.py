"""IndexRPLSH (random-projection LSH front end + exact re-rank) on the GPU
against the oracle's LSH restatement: projections, median offsets, codes,
margins and bucket tables bit-exact; search results (indices and distance bits) and candidate
counts equal to the oracle's on the same seeded data, across table / bit /
probe / budget shapes, ragged dimensions, small batches and padding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def make(ca, res, d, **kw):
    return ca.IndexRPLSH(res, d, ca.IndexRPLSHConfig(**kw))


def test_projections_are_seeded_gaussians(ca, res):
    idx = make(ca, res, 40, tables=3, bits=5, seed=77)
    assert np.array_equal(idx.projections(), ca.gen_gaussian(15, 40, 77))


@pytest.mark.parametrize("center", [True, False])
def test_codes_tables_and_query_dots_bitexact(ca, orc, res, center):
    n, d, L, K = 5000, 33, 4, 9
    x = ca.gen_uniform(n, d, 1, -1, 1) + np.float32(0.5)
    q = ca.gen_uniform(50, d, 2, -1, 1) + np.float32(0.5)
    idx = make(ca, res, d, tables=L, bits=K, seed=5, center=center)
    idx.add(x)
    w = idx.projections()
    off = idx.offsets()
    if center:  # element n // 2 of each projection's sorted canonical dots
        _, dots = orc.lsh_hash(x, w, L, K)
        assert np.array_equal(off, np.sort(dots, axis=0)[n // 2])
    else:
        assert (off == 0).all()
    codes, srt, perm = idx.tables()
    c_ref, _ = orc.lsh_hash(x, w, L, K, off)
    assert np.array_equal(codes, c_ref)
    for t in range(L):
        order = np.lexsort((np.arange(n), c_ref[t]))  # by code, then row
        assert np.array_equal(perm[t], order)
        assert np.array_equal(srt[t], c_ref[t][order])
    qc, qm = idx.hash_queries(q)
    qc_ref, qm_ref = orc.lsh_hash(q, w, L, K, off)
    assert np.array_equal(qc, qc_ref)
    assert np.array_equal(bits(qm), bits(qm_ref))


CASES = [  # n, d, nq, k, tables, bits, probes, budget, center
    (20000, 64, 200, 10, 8, 12, 0, 1024, 1),
    (20000, 64, 200, 10, 8, 12, 3, 1024, 1),
    (20000, 64, 200, 10, 8, 12, 3, 1024, 0),
    (30000, 100, 97, 100, 16, 10, 2, 4096, 1),
    (5000, 33, 40, 10, 4, 6, 1, 64, 1),      # budget cuts the bucket concatenation
    (3000, 200, 20, 20, 6, 8, 4, 512, 1),    # d > 128
    (4000, 96, 7, 10, 8, 10, 2, 300, 1),     # batch under 16 queries
    (2000, 17, 30, 50, 2, 14, 0, 100, 1),    # many queries with < k candidates: padding
    (70000, 32, 64, 10, 4, 20, 2, 512, 1),   # median over a 65536-row sample
    (500, 16, 10, 10, 64, 1, 1, 4096, 1),    # each row ~64 times in the cut: de-duplication
]


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}d{c[1]}q{c[2]}k{c[3]}L{c[4]}K{c[5]}P{c[6]}B{c[7]}"
                                              f"c{c[8]}" for c in CASES])
def test_search_bitexact_vs_oracle(ca, orc, res, case):
    n, d, nq, k, L, K, P, B, center = case
    x = ca.gen_mixture(n, d, 32, 0.2, 3, 4)
    q = ca.gen_mixture(nq, d, 32, 0.2, 3, 5)
    idx = make(ca, res, d, tables=L, bits=K, probes=P, max_candidates=B, seed=11, center=bool(center))
    idx.add(x)
    off = idx.offsets()
    if center and n > 65536:
        _, dots = orc.lsh_hash(x[:65536], idx.projections(), L, K)
        assert np.array_equal(off, np.sort(dots, axis=0)[65536 // 2])
    I, D = idx.search(q, k)
    Io, Do, raw, distinct = orc.lsh_search(x, idx.projections(), L, K, P, B, q, k, off)
    bad = np.nonzero((I != Io).any(1) | (bits(D) != bits(Do)).any(1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    st = idx.last_stats()
    assert st["raw_candidates"] == raw and st["distinct_candidates"] == distinct
    assert st["queries"] == nq


PRE_CASES = [  # n, d, nq, k, tables, bits, probes, budget, prerank
    (20000, 64, 200, 10, 8, 12, 3, 1024, 4),
    (30000, 100, 97, 100, 16, 10, 2, 4096, 2),
    (3000, 200, 20, 20, 6, 8, 4, 512, 3),     # d > 128, ragged BF16 rows
    (4000, 33, 7, 10, 8, 10, 2, 300, 1),      # batch under 16 queries, ld padding
    (2000, 17, 30, 50, 2, 14, 0, 1000, 4),    # fewer distinct candidates than f*k: padding
    (3000, 5, 20, 5, 4, 6, 1, 200, 2),        # one piece per row: 64 rows per load
    (3000, 60, 20, 10, 4, 8, 1, 400, 2),      # 8 pieces
    (1500, 256, 12, 10, 4, 8, 1, 256, 2),     # 32 pieces: two rows per load
    (30000, 64, 40, 10, 64, 10, 4, 6000, 4),  # budget above the exact re-rank's 4096
    (40000, 32, 30, 20, 64, 9, 6, 8192, 8),   # the largest budget
    (500, 16, 10, 10, 64, 1, 1, 8192, 4),     # 8192 entries of 500 rows: heavy de-duplication
]


@pytest.mark.parametrize("case", PRE_CASES, ids=[f"n{c[0]}d{c[1]}q{c[2]}k{c[3]}P{c[6]}B{c[7]}f{c[8]}"
                                                  for c in PRE_CASES])
def test_prerank_bitexact_vs_oracle(ca, orc, res, case):
    n, d, nq, k, L, K, P, B, f = case
    x = ca.gen_mixture(n, d, 32, 0.2, 3, 14)
    q = ca.gen_mixture(nq, d, 32, 0.2, 3, 15)
    idx = make(ca, res, d, tables=L, bits=K, probes=P, max_candidates=B, seed=12, prerank=f)
    idx.add(x[: n // 3])
    idx.add(x[n // 3:])  # the BF16 copy grows with the index
    off = idx.offsets()
    I, D = idx.search(q, k)
    Io, Do, raw, distinct = orc.lsh_search(x, idx.projections(), L, K, P, B, q, k, off, prerank=f)
    bad = np.nonzero((I != Io).any(1) | (bits(D) != bits(Do)).any(1))[0]
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    st = idx.last_stats()
    assert st["raw_candidates"] == raw and st["distinct_candidates"] == distinct


def test_prerank_integer_data_equals_full_rerank(ca, res):
    """Integer data is exact in BF16: the pre-rank keeps the true f*k nearest."""
    rng = np.random.default_rng(5)
    x = rng.integers(-8, 9, (20000, 48)).astype(np.float32)
    q = rng.integers(-8, 9, (64, 48)).astype(np.float32)
    cfg = dict(tables=8, bits=10, probes=2, max_candidates=2048, seed=4)
    full = make(ca, res, 48, **cfg)
    full.add(x)
    I0, D0 = full.search(q, 10)
    pre = make(ca, res, 48, prerank=2, **cfg)
    pre.add(x)
    I1, D1 = pre.search(q, 10)
    assert np.array_equal(I1, I0) and np.array_equal(bits(D1), bits(D0))


def test_all_buckets_equals_plain_brute_force(ca, res):
    x = ca.gen_uniform(3000, 48, 7, -1, 1)
    q = ca.gen_uniform(64, 48, 8, -1, 1)
    idx = make(ca, res, 48, tables=1, bits=1, probes=1, max_candidates=4096, seed=3)
    idx.add(x)
    I, D = idx.search(q, 25)
    bf = ca.IndexBF(res, 48)
    bf.add(x)
    Ib, Db = bf.search(q, 25)
    assert np.array_equal(I, Ib) and np.array_equal(bits(D), bits(Db))
    assert idx.last_stats()["distinct_candidates"] == 64 * 3000


def test_incremental_add_device_pointers_and_reset(ca, res):
    d, k = 64, 10
    x = ca.gen_mixture(12000, d, 16, 0.3, 1, 2)
    q = ca.gen_mixture(100, d, 16, 0.3, 1, 3)
    cfg = dict(tables=6, bits=10, probes=2, max_candidates=800, seed=9)
    one = make(ca, res, d, center=False, **cfg)
    one.add(x)
    I1, D1 = one.search(q, k)
    two = make(ca, res, d, center=False, **cfg)
    two.add(x[:5000])
    two.add(ca.DeviceArray.from_host(x[5000:]))
    assert two.num_data == 12000
    dq = ca.DeviceArray.from_host(q)
    dI, dD = ca.DeviceArray((100, k), np.int32), ca.DeviceArray((100, k), np.float32)
    two.search(dq, k, dI, dD)
    assert np.array_equal(dI.to_host(), I1) and np.array_equal(bits(dD.to_host()), bits(D1))
    two.reset()
    assert two.num_data == 0
    I0, D0 = two.search(q, k)
    assert (I0 == -1).all() and (D0 == np.finfo(np.float32).max).all()
    two.add(x)
    I2, _ = two.search(q, k)
    assert np.array_equal(I2, I1)
    # centred: offsets come from the first add (after a reset, the next one)
    c1 = make(ca, res, d, center=True, **cfg)
    c1.add(x[:5000])
    off = c1.offsets()
    c1.add(x[5000:])
    assert np.array_equal(c1.offsets(), off)
    c1.reset()
    assert (c1.offsets() == 0).all()
    c1.add(x[5000:])
    assert not np.array_equal(c1.offsets(), off)


def test_recall_on_clustered_data(ca, orc, res):
    """Not a parity check: the front end must find most true neighbours on
    clustered data (256 centres, sigma 0.1, d = 128), and more with probes."""
    x = ca.gen_mixture(200_000, 128, 256, 0.1, 2, 3)
    q = ca.gen_mixture(500, 128, 256, 0.1, 2, 30)
    Io, _ = orc.search_canonical(x, q, 10)
    rec = {}
    for P in (0, 3):
        idx = make(ca, res, 128, tables=16, bits=14, probes=P, max_candidates=2048, seed=1)
        idx.add(x)
        I, _ = idx.search(q, 10)
        rec[P] = orc.recall(I, Io)
        print("probes", P, "recall@10", rec[P], idx.last_stats())
    assert rec[3] >= rec[0] and rec[3] >= 0.8


def test_budget_above_exact_limit_needs_prerank(ca, res):
    with pytest.raises(ca.RplshError):
        make(ca, res, 16, max_candidates=5000)               # no pre-rank: 5000 exact
    with pytest.raises(ca.RplshError):
        make(ca, res, 16, max_candidates=8193, prerank=4)    # past the gather's limit
    idx = make(ca, res, 16, tables=8, bits=6, probes=2, max_candidates=8192, prerank=64, seed=2)
    x = ca.gen_uniform(5000, 16, 3, -1, 1)
    idx.add(x)
    I, _ = idx.search(x[:4], 64)                             # 64 * 64 = 4096 exact: fine
    assert I.shape == (4, 64)
    with pytest.raises(ca.RplshError):
        idx.search(x[:4], 65)                                # 65 * 64 > 4096


def test_invalid_configs_rejected(ca, res):
    for kw in (dict(bits=0), dict(bits=25), dict(probes=17, bits=16), dict(max_candidates=9000),
               dict(tables=100, probes=10, bits=12), dict(tables=0), dict(prerank=-1),
               dict(prerank=65)):
        with pytest.raises(ca.RplshError):
            make(ca, res, 16, **kw)
