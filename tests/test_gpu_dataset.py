"""Vector files streamed into device memory (pinned staging, several chunks)
equal the host read; a TEXMEX-layout dataset with FP64 ground truth searched
from device memory gives the expected recall."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_device_read_streams_in_chunks(ca, tmp_path):
    a = ca.gen_uniform(300_000, 128, 17)  # 155 MB of records: three 64 MB chunks
    p = str(tmp_path / "big.fvecs")
    ca.write_xvecs(p, a)
    dev = ca.read_fvecs(p, device=0)
    assert dev.shape == a.shape
    assert np.array_equal(dev.to_host(), a)
    part = ca.read_fvecs(p, device=0, first=123_457, count=99_999)
    assert np.array_equal(part.to_host(), a[123_457:123_457 + 99_999])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_dataset_search_recall(ca, orc, res, tmp_path, precision):
    x = ca.gen_gaussian(80_000, 96, 31, True)
    q = ca.gen_gaussian(200, 96, 32, True)
    gt, _ = orc.search_fp64(x, q, 100)
    for nm, arr in (("base.fvecs", x), ("query.fvecs", q), ("groundtruth.ivecs", gt)):
        ca.write_xvecs(str(tmp_path / f"syn_{nm}"), arr)
    ds = ca.Dataset("syn", str(tmp_path), device=0)
    idx = ca.IndexBF(res, ds.dimension, ca.IndexBFConfig(precision=precision))
    idx.add(ds.base)
    I = ca.DeviceArray((ds.num_query, 100), np.int32)
    D = ca.DeviceArray((ds.num_query, 100), np.float32)
    idx.search(ds.query, 100, I, D)
    r, missed = ds.evaluate(I, return_missed=True)
    print(precision, "recall@100", r)
    assert r >= 0.999
    assert missed.sum() == round((1 - r) * ds.num_query * 100)
