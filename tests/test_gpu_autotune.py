"""Autotuner: the chosen k' reaches the recall target on the sample, tuned
parameters never change plain-path results, and the cache round-trips."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_precision_form(ca, orc, res, tmp_path):
    from curplsh_amd import autotune as at

    x = ca.gen_gaussian(200_000, 96, 21, True)
    sample = ca.gen_gaussian(256, 96, 22, True)
    q = ca.gen_gaussian(300, 96, 23, True)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    path = str(tmp_path / "tune.json")
    t = at.autotune(idx, sample, 10, 1024, 0.999, cache_path=path)
    print("tuned:", t)
    assert 1 <= t["kprime_factor"] <= 16
    assert t["recall"] >= 0.999 or t["kprime_factor"] == 16
    # the tuner's grid: target workgroups {4096, 8192, 16384}, survivor multiple
    # {2, 3, 4, 8}, filter variant LDS wide / narrow / ring wide, and for the
    # LDS ones possibly the other LDS kernel on the sample levels (bits 8-15:
    # 1 + its variant), and possibly the full level in L2-sized slices (bit 16)
    assert t["target_wgs"] in (4096, 8192, 16384) and t["c_factor"] in (2, 3, 4, 8)
    full, sample = t["variant"] & 0xFF, ((t["variant"] >> 8) & 0xFF) - 1
    assert t["variant"] >> 16 in (0, 1)
    assert full in (0, 1, 3)
    assert sample == -1 or (full in (0, 1) and sample == 1 - full)
    assert t["ms_tuned"] <= t["ms_default"] * 1.5
    I, _ = idx.search(q, 10)
    assert idx.last_stats()["kprime"] == 10 * t["kprime_factor"]
    I64, _ = orc.search_fp64(x, q, 10)
    print("recall on fresh queries:", orc.recall(I, I64))
    assert orc.recall(I, I64) >= 0.99
    # cache round trip onto a new index
    idx2 = ca.IndexBF(res, 96, ca.IndexBFConfig(precision="bf16"))
    idx2.add(x)
    t2 = at.autotune(idx2, sample, 10, 1024, 0.999, cache_path=path)
    assert t2 == pytest.approx(t)
    I2, _ = idx2.search(q, 10)
    assert np.array_equal(I2, I)


def test_autotune_plain_keeps_exactness(ca, orc, res):
    x = ca.gen_uniform(100_000, 128, 31)
    q = ca.gen_uniform(200, 128, 32)
    idx = ca.IndexBF(res, 128)
    idx.add(x)
    t = idx.autotune(q[:64], 10, 512)
    assert t["kprime_factor"] == 0
    I, D = idx.search(q, 10)
    Io, Do = orc.search_canonical(x, q, 10)
    assert np.array_equal(I, Io) and np.array_equal(D.view(np.uint32), Do.view(np.uint32))
