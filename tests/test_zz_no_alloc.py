"""rplsh_search_l2_dev_noalloc: the searchL2Distance entry point that takes
every temporary from the library's fixed per-device scratch array. It must
return exactly what rplsh_search_l2_dev returns, on two calls in a row (the
second smaller, so stale scratch contents from the first must not leak into
it), and reject sizes whose temporaries do not fit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def run_both(ca, x, q, k, exact, with_norms):
    n, d = x.shape
    nq = q.shape[0]
    dx, dq = ca.DeviceArray.from_host(x), ca.DeviceArray.from_host(q)
    dn = None
    if with_norms:
        import oracle
        dn = ca.DeviceArray.from_host(oracle.norms_canonical(x))
    outs = []
    for fn in (ca.search_l2_distance, ca.search_l2_distance_noalloc):
        dI = ca.DeviceArray((nq, k), np.int32)
        dD = ca.DeviceArray((nq, k), np.float32)
        fn(dx, dn, dq, k, dI, dD, exact=exact)
        outs.append((dI.to_host(), dD.to_host()))
    return outs


def test_noalloc_matches_allocating_entry_point(ca, orc):
    print("[no-alloc] call 1: 200000x128, 512 queries, k=10, norms computed", flush=True)
    x = ca.gen_uniform(200_000, 128, 901, -1, 1)
    q = ca.gen_uniform(512, 128, 902, -1, 1)
    (Ia, Da), (Ib, Db) = run_both(ca, x, q, 10, True, False)
    assert np.array_equal(Ia, Ib) and np.array_equal(bits(Da), bits(Db))
    Io, Do = orc.search_canonical(x, q[:32], 10)
    assert np.array_equal(Ib[:32], Io) and np.array_equal(bits(Db[:32]), bits(Do))
    print("[no-alloc] call 2: 50000x100 (padded rows), 64 queries, k=5, given norms, relative", flush=True)
    x2 = ca.gen_uniform(50_000, 100, 903, 0, 1)
    q2 = ca.gen_uniform(64, 100, 904, 0, 1)
    (Ia, Da), (Ib, Db) = run_both(ca, x2, q2, 5, False, True)
    assert np.array_equal(Ia, Ib) and np.array_equal(bits(Da), bits(Db))
    Io, Do = orc.search_canonical(x2, q2, 5, exact=False)
    assert np.array_equal(Ib, Io) and np.array_equal(bits(Db), bits(Do))


def test_noalloc_rejects_oversize(ca):
    x = ca.gen_uniform(20_000, 64, 905, -1, 1)
    q = ca.gen_uniform(4096, 64, 906, -1, 1)
    dx, dq = ca.DeviceArray.from_host(x), ca.DeviceArray.from_host(q)
    dI = ca.DeviceArray((4096, 100), np.int32)
    dD = ca.DeviceArray((4096, 100), np.float32)
    assert ca.noalloc_scratch_bytes() == 64 << 20
    with pytest.raises(ca.RplshError) as e:
        ca.search_l2_distance_noalloc(dx, None, dq, 100, dI, dD)
    assert e.value.status == 1 and "scratch" in str(e.value)
    # a fitting request right after still works
    outs = []
    for fn in (ca.search_l2_distance_noalloc, ca.search_l2_distance):
        dq16 = ca.DeviceArray.from_host(q[:16])
        dI2 = ca.DeviceArray((16, 10), np.int32)
        dD2 = ca.DeviceArray((16, 10), np.float32)
        fn(dx, None, dq16, 10, dI2, dD2)
        outs.append((dI2.to_host(), dD2.to_host()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(bits(outs[0][1]), bits(outs[1][1]))
