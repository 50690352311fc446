"""The HBM budget behind DeviceMemory (include/device_memory.h:12-89):
reservations within the budget, blocks reused after release without a new
allocation, waits and refusals at the limit, and a serving loop admitting its
batches against a budget shared with other reservations."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MB = 1 << 20


def test_budget_reservations(ca):
    b = ca.HbmBudget(0, 64 * MB)
    p1 = b.reserve(10 * MB)
    p2 = b.reserve(10 * MB)
    assert p1 and p2 and p1 != p2
    st = b.stats()
    assert st["reserved_bytes"] == 20 * MB and st["mallocs"] == 2
    b.release(p1)
    p3 = b.reserve(9 * MB)  # fits the released 10 MB block: no allocation
    assert p3 == p1 and b.stats()["mallocs"] == 2
    with pytest.raises(ca.RplshError) as e:
        b.reserve(60 * MB, timeout_s=0.0)  # 20 MB held: does not fit now
    assert e.value.status == 6
    with pytest.raises(ca.RplshError) as e:
        b.reserve(65 * MB)  # never fits
    assert e.value.status == 6
    t0 = time.time()
    with pytest.raises(ca.RplshError):
        b.reserve(50 * MB, timeout_s=0.2)
    assert time.time() - t0 >= 0.15
    b.release(p2)
    b.release(p3)
    st = b.stats()
    assert st["reserved_bytes"] == 0 and st["peak_reserved_bytes"] == 20 * MB
    # a new large block first returns cached ones to the device
    p4 = b.reserve(60 * MB)
    st = b.stats()
    assert st["reserved_bytes"] + st["cached_bytes"] <= 64 * MB
    b.release(p4)
    b.close()


def test_server_admits_against_shared_budget(ca, res):
    x = ca.gen_uniform(100_000, 64, 951, -1, 1)
    idx = ca.IndexBF(res, 64, ca.IndexBFConfig(precision="bf16"))
    idx.add(x)
    q = ca.gen_uniform(256, 64, 952, -1, 1)
    need = ca.batch_bytes(idx, 256, 10)
    budget = ca.HbmBudget(0, 3 * need)
    srv = ca.Server(slots=2, memory=budget, use_graphs=False)
    sid = srv.add_index(idx)
    Ia, Da = idx.search(q, 10)
    hold = budget.reserve(int(2.5 * need))  # leaves half a batch
    h = srv.submit(sid, q, 10)
    time.sleep(0.1)
    assert not srv.done(h), "admitted past the shared budget"
    budget.release(hold)
    I, D, info = srv.wait(h, timeout_s=30.0)
    assert np.array_equal(I, Ia) and np.array_equal(D.view(np.uint32), Da.view(np.uint32))
    assert info["queue_ms"] >= 80.0
    st = budget.stats()
    assert st["peak_reserved_bytes"] <= 3 * need
    assert srv.memory().stats()["budget_bytes"] == 3 * need
    srv.close()
    assert budget.stats()["reserved_bytes"] == 0
    budget.close()
