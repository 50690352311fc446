"""Index files: a saved and reloaded index (rows, configuration, tuning)
answers every search bit-identically to the original."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_save_load_roundtrip(ca, orc, res, tmp_path, precision):
    x = ca.gen_gaussian(120_000, 96, 51, True)
    q = ca.gen_gaussian(500, 96, 52, True)
    idx = ca.IndexBF(res, 96, ca.IndexBFConfig(precision=precision))
    idx.add(x[:70_000])
    idx.add(x[70_000:])
    tune = dict(k=10, batch_nq=500, kprime_factor=4 if precision != "fp32" else 0, target_wgs=8192,
                c_factor=3, variant=0, rho=32, recall=0.0, ms_default=0.0, ms_tuned=0.0)
    idx.set_tuning(tune)
    I, D = idx.search(q, 10)
    st = idx.last_stats()
    p = str(tmp_path / "index.rplsh")
    idx.save(p)
    back = ca.IndexBF.load(res, p)
    assert back.dim == 96
    I2, D2 = back.search(q, 10)
    st2 = back.last_stats()
    assert np.array_equal(I, I2) and np.array_equal(D.view(np.uint32), D2.view(np.uint32))
    assert (st2["target_wgs"], st2["kprime"]) == (st["target_wgs"], st["kprime"])
    if precision == "fp32":
        Io, Do = orc.search_canonical(x, q, 10)
        assert np.array_equal(I2, Io)
    back.close()
    # truncated file
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(ca.RplshError, match="truncated"):
        ca.IndexBF.load(res, p)
