"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

  * closed-form fixtures of the reference's own drivers
    (test/test_search.cpp:17-36, test/test_norm.cpp:15-22);
  * the canonical FP32 restatement against exact rational arithmetic with
    correctly rounded fmaf (so a dropped FMA, a reordered sum or a wrong sign
    fails), on tiny inputs;
  * the FP64 brute force against an independent numpy float64 brute force;
  * recall against hand-computed cases (include/dataset.h:135-152);
  * the bf16 / e4m3 encoders against torch's bfloat16 / float8_e4m3fn
    conversions and the exhaustive e4m3 code table.
"""
from fractions import Fraction

import numpy as np
import pytest


# ------------------------------------------------------------------ helpers
def rn32(v: Fraction) -> np.float32:
    """Round an exact rational to the nearest float32, ties to even."""
    if v == 0:
        return np.float32(0.0)
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    qe = max(e, -126) - 23
    s = a / Fraction(2) ** qe
    n = s.numerator // s.denominator
    rem = s - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(sign * float(Fraction(n) * Fraction(2) ** qe))


def fma32(a, b, c) -> np.float32:
    return rn32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def canonical_exact(x, q, k, exact=True):
    """Canonical order evaluated with exact rationals + correct rounding."""
    n, d = x.shape
    xn = []
    for i in range(n):
        s = np.float32(0)
        for j in range(d):
            s = fma32(x[i, j], x[i, j], s)
        xn.append(s)
    I = np.full((q.shape[0], k), -1, np.int32)
    D = np.full((q.shape[0], k), np.finfo(np.float32).max, np.float32)
    for r in range(q.shape[0]):
        qn = np.float32(0)
        for j in range(d):
            qn = fma32(q[r, j], q[r, j], qn)
        rel = []
        for i in range(n):
            s = np.float32(0)
            for j in range(d):
                s = fma32(q[r, j], x[i, j], s)
            rel.append((fma32(-2.0, s, xn[i]), i))
        rel.sort()
        for t, (v, i) in enumerate(rel[:k]):
            I[r, t] = i
            D[r, t] = rn32(Fraction(float(v)) + Fraction(float(qn))) if exact else v
    return I, D


def fixture_search():
    dim, nb, nq = 128, 1000, 100
    q = np.ones((nq, dim), np.float32)
    b = np.minimum(np.arange(nb), 10).astype(np.float32)[:, None] * np.ones((1, dim), np.float32)
    return b, q


# -------------------------------------------------------------- fixtures
def test_search_fixture_relative(orc):
    b, q = fixture_search()
    I, D = orc.search_canonical(b, q, 10, exact=False)
    assert (I == np.array([1, 0, 2, 3, 4, 5, 6, 7, 8, 9])).all()
    assert (D == np.array([-128, 0, 0, 384, 1024, 1920, 3072, 4480, 6144, 8064], np.float32)).all()


def test_search_fixture_exact_and_ties(orc):
    b, q = fixture_search()
    I, D = orc.search_canonical(b, q, 100, exact=True)
    assert (I[:, :10] == np.array([1, 0, 2, 3, 4, 5, 6, 7, 8, 9])).all()
    assert (D[0, :10] == np.array([0, 128, 128, 512, 1152, 2048, 3200, 4608, 6272, 8192])).all()
    # rows 10..999 are identical (distance 128*81): lowest indices first
    assert (I[:, 10:] == np.arange(10, 100)).all()
    assert (D[:, 10:] == 128 * 81).all()
    I64, D64 = orc.search_fp64(b, q, 100)
    assert (I64 == I).all() and (D64 == D).all()


@pytest.mark.parametrize("dim", [1, 3, 32, 64, 96, 100, 128, 256, 1500])
def test_norm_fixture_all_ones(orc, dim):
    x = np.ones((37, dim), np.float32)
    assert (orc.norms_canonical(x) == dim).all()


# ------------------------------------------------- canonical vs exact arithmetic
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_canonical_matches_exact_rational(orc, seed):
    rng = np.random.default_rng(seed)
    n, d, nq, k = 40, 7, 3, 6
    x = (rng.standard_normal((n, d)) * rng.choice([1e-3, 1.0, 30.0], size=(n, 1))).astype(np.float32)
    q = rng.standard_normal((nq, d)).astype(np.float32)
    for exact in (False, True):
        Ie, De = canonical_exact(x, q, k, exact)
        I, D = orc.search_canonical(x, q, k, exact=exact, threads=1)
        assert (I == Ie).all()
        assert (D.view(np.uint32) == De.view(np.uint32)).all()
    xn = orc.norms_canonical(x)
    for i in range(n):
        s = np.float32(0)
        for j in range(d):
            s = fma32(x[i, j], x[i, j], s)
        assert xn[i] == s


def test_exact_check_detects_unfused_order(orc):
    """A restatement without FMA (round(a*b) then add) differs from the
    canonical chain on such data, so the rational check above would catch it."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((200, 16)).astype(np.float32)
    diffs = 0
    for i in range(200):
        s_fma = np.float32(0)
        s_mul = np.float32(0)
        for j in range(16):
            s_fma = fma32(x[i, j], x[i, j], s_fma)
            s_mul = np.float32(s_mul + np.float32(x[i, j] * x[i, j]))
        diffs += int(s_fma != s_mul)
    assert diffs > 0
    xn = orc.norms_canonical(x)
    recomputed = [np.float32(0)] * 200
    for i in range(200):
        s = np.float32(0)
        for j in range(16):
            s = fma32(x[i, j], x[i, j], s)
        recomputed[i] = s
    assert (xn == np.array(recomputed, np.float32)).all()


def test_canonical_padding_and_empty(orc):
    x = np.arange(12, dtype=np.float32).reshape(4, 3)
    q = np.zeros((2, 3), np.float32)
    I, D = orc.search_canonical(x, q, 6)
    assert (I[:, 4:] == -1).all() and (D[:, 4:] == np.finfo(np.float32).max).all()
    assert (I[:, :4] == np.arange(4)).all()
    I, D = orc.search_canonical(np.zeros((0, 3), np.float32), q, 3)
    assert (I == -1).all() and (D == np.finfo(np.float32).max).all()


# ------------------------------------------------------------ FP64 oracle
def test_fp64_matches_numpy(orc):
    rng = np.random.default_rng(7)
    x = rng.standard_normal((3000, 24)).astype(np.float32)
    q = rng.standard_normal((20, 24)).astype(np.float32)
    k = 15
    I, D = orc.search_fp64(x, q, k)
    dd = ((q[:, None, :].astype(np.float64) - x[None, :, :].astype(np.float64)) ** 2).sum(-1)
    for r in range(q.shape[0]):
        order = np.lexsort((np.arange(x.shape[0]), dd[r]))[:k]
        assert (I[r] == order).all()
        np.testing.assert_allclose(D[r], dd[r, order], rtol=1e-13)
    pd = orc.pair_dist_fp64(x, q, I)
    np.testing.assert_allclose(pd, D, rtol=1e-13)


def test_canonical_close_to_fp64(orc, ca):
    x = ca.gen_uniform(20000, 64, 3)
    q = ca.gen_uniform(50, 64, 4)
    I, D = orc.search_canonical(x, q, 10)
    I64, D64 = orc.search_fp64(x, q, 10)
    assert orc.recall(I, I64) > 0.99
    pd = orc.pair_dist_fp64(x, q, I)
    assert np.max(np.abs(D - pd) / pd) < 1e-4


# ------------------------------------------------------------------ recall
def test_recall_definition(orc):
    truth = np.array([[1, 2, 3, 4], [5, 6, 7, 8]], np.int32)
    res = np.array([[4, 3, 9, 1], [-1, -1, -1, 5]], np.int32)
    assert orc.recall(res, truth) == pytest.approx((3 / 4 + 1 / 4) / 2)
    assert orc.recall(truth, truth) == 1.0


# ---------------------------------------------------------------- encoders
def test_bf16_matches_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4000),
        np.array([0.0, -0.0, 1.0, np.inf, -np.inf, 3.4e38, 1e-40, -1e-45], np.float32),
        # exact ties: low 16 bits 0x8000
        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
    ])
    ref = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    ours = np.array([orc.bf16_rne(float(v)) for v in vals], np.uint16)
    assert (ours == ref).all()
    assert orc.bf16_rne(float("nan")) & 0x7FC0 == 0x7FC0


def test_e4m3_code_table(orc):
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    ref = torch.from_numpy(codes).view(torch.float8_e4m3fn).to(torch.float32).numpy()
    ours = np.array([orc.e4m3_to_float(int(c)) for c in codes], np.float32)
    finite = ~np.isnan(ref)
    assert (np.isnan(ours) == ~finite).all()
    assert (ours[finite] == ref[finite]).all()
    # every finite code round-trips through the encoder
    for c in codes[finite]:
        v = orc.e4m3_to_float(int(c))
        if v == 0.0 and c == 0x80:
            continue
        assert orc.e4m3_rne_sat(v) == c


def test_e4m3_matches_torch_in_range(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.uniform(-447.9, 447.9, 5000).astype(np.float32),
        (rng.standard_normal(5000) * 2.0 ** rng.integers(-12, 6, 5000)).astype(np.float32),
        # midpoints between neighbouring codes (ties to even)
        np.array([(orc.e4m3_to_float(c) + orc.e4m3_to_float(c + 1)) / 2 for c in range(0, 0x7e)], np.float32),
    ])
    ref = torch.from_numpy(vals).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    ours = np.array([orc.e4m3_rne_sat(float(v)) for v in vals], np.uint8)
    assert (ours == ref).all()
    assert orc.e4m3_rne_sat(1e6) == 0x7E and orc.e4m3_rne_sat(-1e6) == 0xFE
    assert orc.e4m3_rne_sat(448.0) == 0x7E and orc.e4m3_rne_sat(460.0) == 0x7E
