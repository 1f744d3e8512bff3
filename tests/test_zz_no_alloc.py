"""GPU: aj_join_dev_noalloc (the device join with every workspace buffer in the
library's fixed scratch array) through the C ABI, against aj_join_dev on the same
device relations: count, checksum and the pair list are identical, on two calls
in a row, the second with smaller relations, for PRO and NPO; the library's
device-allocation counter does not move across any of its calls; relations above the
limit are AJ_ERR_CAPACITY and SHJ is AJ_ERR_ARG."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(aj, dR, nR, dS, nS, algo, pairs_ptr=0, pairs_capacity=0, records=0, bits=0,
             passes=0):
    """One call through the C ABI; the library's device-allocation counter must not move."""
    L = aj._lib
    o = L.aj_join_opts_t(algo, bits, passes, 0, 0, pairs_ptr or None, pairs_capacity, records,
                         0, None)
    out = L.aj_join_out_t()
    used = aj.get_option("device_used")
    rc = aj.lib.aj_join_dev_noalloc(dR or None, nR, dS or None, nS, ctypes.byref(o),
                                    ctypes.byref(out), None)
    assert aj.get_option("device_used") == used
    return rc, out


def _inputs(aj, kind, nR, nS):
    if kind == "unique_uniform":
        return aj.gen_unique(nR, 12345), aj.gen_uniform(nS, 1, nR, 54321)
    return aj.gen_uniform(nR, 1, max(nR // 8, 1), 7), aj.gen_uniform(nS, 1, max(nR // 8, 1), 8)


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_noalloc_matches_join_dev(aj, oracle, algo):
    a = {"PRO": aj.AJ_ALGO_PRO, "NPO": aj.AJ_ALGO_NPO}[algo]
    with aj.Context(0) as ctx:
        # the largest relations it takes, then smaller ones with duplicates
        for kind, nR, nS in (("unique_uniform", 1 << 20, 1 << 20), ("dups", 3000, 5000),
                             ("unique_uniform", 70000, 1 << 20)):
            R, S = _inputs(aj, kind, nR, nS)
            dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
            try:
                want = ctx.join(dR.ptr, nR, dS.ptr, nS, a)
                print(f"case noalloc {algo} {kind} {nR}x{nS}", flush=True)
                rc, got = _noalloc(aj, dR.ptr, nR, dS.ptr, nS, a)
                assert rc == aj.AJ_OK, aj.last_error()
                assert (got.count, got.checksum) == (want.count, want.checksum)
                assert (got.count, got.checksum) == oracle.aggregate(R, S)
                assert got.device_ms == 0  # no events
                # pairs into a caller-owned buffer, against aj_join_dev's
                cnt = int(want.count)
                b1, b2 = aj.DeviceArray(8 * cnt), aj.DeviceArray(8 * cnt)
                try:
                    w2 = ctx.join(dR.ptr, nR, dS.ptr, nS, a, pairs_ptr=b1.ptr,
                                  pairs_capacity=cnt)
                    rc, g2 = _noalloc(aj, dR.ptr, nR, dS.ptr, nS, a, b2.ptr, cnt)
                    assert rc == aj.AJ_OK, aj.last_error()
                    assert g2.npairs == w2.npairs == cnt
                    assert np.array_equal(np.sort(b2.to_host(np.uint64, cnt)),
                                          np.sort(b1.to_host(np.uint64, cnt)))
                finally:
                    b1.free()
                    b2.free()
            finally:
                dR.free()
                dS.free()


def test_noalloc_limits(aj):
    n = (1 << 20) + 1
    R = aj.gen_unique(n, 1)
    dR = aj.DeviceArray.from_host(R)
    try:
        rc, _ = _noalloc(aj, dR.ptr, n, dR.ptr, 1000, aj.AJ_ALGO_PRO)
        assert rc == aj.AJ_ERR_CAPACITY and "at most" in aj.last_error()
        rc, _ = _noalloc(aj, dR.ptr, 1000, dR.ptr, 1000, aj.AJ_ALGO_SHJ)
        assert rc == aj.AJ_ERR_ARG
        # a geometry whose workspace does not fit the array: 24 radix bits in 2
        # passes give 2^24 over-allocated partitions
        rc, _ = _noalloc(aj, dR.ptr, 1 << 20, dR.ptr, 1 << 20, aj.AJ_ALGO_PRO, bits=24, passes=2)
        assert rc == aj.AJ_ERR_CAPACITY, aj.last_error()
        # and the next call in range works again
        rc, out = _noalloc(aj, dR.ptr, 1 << 20, dR.ptr, 1 << 20, aj.AJ_ALGO_NPO)
        assert rc == aj.AJ_OK and out.count == 1 << 20
    finally:
        dR.free()
