"""TEXMEX vector files and recall (the reference's include/dataset.h):
host-side reading / writing against numpy's reading of the same bytes, the
format's failure modes, and recall against the oracle's restatement. No GPU."""
import numpy as np
import pytest


def np_write(path, a):
    """The .fvecs / .ivecs / .bvecs layout written by numpy alone."""
    a = np.ascontiguousarray(a)
    n, d = a.shape
    with open(path, "wb") as fh:
        for i in range(n):
            fh.write(np.int32(d).tobytes())
            fh.write(a[i].tobytes())


def np_read(path, dtype):
    raw = np.fromfile(path, np.uint8)
    d = int(raw[:4].view(np.int32)[0])
    rec = 4 + d * np.dtype(dtype).itemsize
    rows = raw.reshape(-1, rec)
    assert (rows[:, :4].copy().view(np.int32)[:, 0] == d).all()
    return rows[:, 4:].copy().view(dtype).reshape(-1, d)


@pytest.mark.parametrize("dtype,d", [(np.float32, 128), (np.int32, 100), (np.uint8, 128), (np.float32, 1)])
def test_read_matches_numpy(ca, tmp_path, dtype, d):
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((777, d)) * 50).astype(dtype) if dtype != np.uint8 else rng.integers(0, 256, (777, d), dtype=np.uint8)
    p = str(tmp_path / "x.vecs")
    np_write(p, a)
    assert ca.xvecs_info(p, dtype) == (777, d)
    got = ca.read_xvecs(p, dtype)
    assert got.dtype == np.dtype(dtype) and np.array_equal(got, a)
    assert np.array_equal(ca.read_xvecs(p, dtype, first=100, count=250), a[100:350])
    assert np.array_equal(ca.read_xvecs(p, dtype, first=776), a[776:])
    assert ca.read_xvecs(p, dtype, first=10, count=0).shape == (0, d)


def test_write_matches_numpy(ca, tmp_path):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((500, 96)).astype(np.float32)
    g = rng.integers(0, 10**6, (500, 100)).astype(np.int32)
    pf, pi = str(tmp_path / "a.fvecs"), str(tmp_path / "g.ivecs")
    ca.write_xvecs(pf, a[:300])
    ca.write_xvecs(pf, a[300:], append=True)
    ca.write_xvecs(pi, g)
    assert np.array_equal(np_read(pf, np.float32), a)
    assert np.array_equal(np_read(pi, np.int32), g)
    with open(pf, "rb") as fh:
        assert len(fh.read()) == 500 * (4 + 96 * 4)


def test_large_file_is_read_in_chunks(ca, tmp_path):
    """> 64 MB: the host path reads and compacts several chunks."""
    a = ca.gen_uniform(140_000, 128, 9)  # 140k x 516 B = 72 MB
    p = str(tmp_path / "big.fvecs")
    ca.write_xvecs(p, a)
    assert np.array_equal(ca.read_fvecs(p), a)


def test_malformed_files_fail_loudly(ca, tmp_path):
    p = str(tmp_path / "bad.fvecs")
    np_write(p, np.ones((10, 8), np.float32))
    with open(p, "ab") as fh:
        fh.write(b"\x00\x00")  # not a whole record
    with pytest.raises(ca.RplshError, match="whole number"):
        ca.read_fvecs(p)
    q = str(tmp_path / "mixed.fvecs")
    np_write(q, np.ones((4, 8), np.float32))
    raw = bytearray(open(q, "rb").read())
    raw[2 * 36: 2 * 36 + 4] = np.int32(9).tobytes()  # record 2 claims dimension 9
    open(q, "wb").write(bytes(raw))
    with pytest.raises(ca.RplshError, match="record 2 has dimension 9"):
        ca.read_fvecs(q)
    with pytest.raises(ca.RplshError, match="cannot open"):
        ca.read_fvecs(str(tmp_path / "missing.fvecs"))
    with pytest.raises(ca.RplshError, match="outside"):
        ca.read_xvecs(str(tmp_path / "bad.fvecs").replace("bad", "mixed"), np.float32, first=3, count=2)
    e = str(tmp_path / "empty.fvecs")
    open(e, "wb").close()
    assert ca.xvecs_info(e) == (0, 0)


def test_recall_matches_oracle_and_leaves_inputs(ca, orc):
    rng = np.random.default_rng(5)
    truth = np.stack([rng.permutation(1000)[:100] for _ in range(300)]).astype(np.int32)
    res = truth.copy()
    for i in range(300):  # replace a random number of entries, shuffle
        m = rng.integers(0, 101)
        res[i, :m] = rng.integers(1000, 2000, m)
        rng.shuffle(res[i])
    r0, t0 = res.copy(), truth.copy()
    rec, missed = ca.recall_at(res, truth, return_missed=True)
    assert rec == pytest.approx(orc.recall(res, truth), abs=1e-12)
    hits = np.array([len(set(res[i]) & set(truth[i])) for i in range(300)])
    assert np.array_equal(missed, 100 - hits)
    assert np.array_equal(res, r0) and np.array_equal(truth, t0)  # not reordered (dataset.h:137 does)
    # -1 pads never match; recall@k looks at the first k columns only
    pad = truth.copy()
    pad[:, 50:] = -1
    assert ca.recall_at(pad, truth) == pytest.approx(0.5)
    assert ca.recall_at(pad, truth, k=50) == pytest.approx(1.0)


def test_dataset_irisa_layout(ca, tmp_path):
    rng = np.random.default_rng(6)
    base = rng.standard_normal((2000, 32)).astype(np.float32)
    query = rng.standard_normal((20, 32)).astype(np.float32)
    gt = rng.integers(0, 2000, (20, 10)).astype(np.int32)
    for nm, a in (("base.fvecs", base), ("query.fvecs", query), ("groundtruth.ivecs", gt)):
        ca.write_xvecs(str(tmp_path / f"toy_{nm}"), a)
    ds = ca.Dataset("toy", str(tmp_path))
    assert (ds.num_base, ds.num_query, ds.dimension, ds.gt_k) == (2000, 20, 32, 10)
    assert np.array_equal(ds.base, base) and np.array_equal(ds.query, query)
    assert ds.evaluate(gt) == 1.0


def test_index_file_rejects_foreign_files(ca, tmp_path):
    """rplsh_bf_load checks the header before touching a device."""
    p = str(tmp_path / "not_an_index.bin")
    ca.write_xvecs(p, np.ones((3, 4), np.float32))
    with pytest.raises(ca.RplshError, match="not an index file"):
        ca.IndexBF.load(_NoResources(), p)


class _NoResources:
    handle = 1  # never dereferenced: the header check fails first
