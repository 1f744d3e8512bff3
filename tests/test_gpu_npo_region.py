"""GPU parity of NPO's region build (npo.hip) against the oracle.

Counting NPO joins of at least `npo_region_min` R tuples build the hash table
region by region: R is partitioned by hashed key into the table's regions (the
staged passes, keys replaced by fmix32(key) in pass 1 and recovered by its
inverse), one work-group builds each region in LDS and writes it out.  A
partition that overflows its capacity makes the flat atomic build run instead,
on the device.  Results must equal the oracle's bit for bit on either path:
unique, duplicate, negative and sentinel keys, ragged sizes down to the
smallest table that splits into regions, the key whose hash is all ones, heavy
duplicates (the fallback), and the 8192-slot regions of a 2^29-slot table.
"""
import contextlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEY_HASH_ALL_ONES = 0x331DA083  # fmix32 of this key is 0xFFFFFFFF


@contextlib.contextmanager
def options(aj, **kw):
    old = {k: aj.get_option(k) for k in kw}
    try:
        for k, v in kw.items():
            aj.set_option(k, v)
        yield
    finally:
        for k, v in old.items():
            aj.set_option(k, v)


def _npo(aj, R, S, **kw):
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    try:
        with aj.Context(0) as ctx:
            return ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_NPO, **kw)
    finally:
        dR.free()
        dS.free()


def _inputs(aj, kind, nR, nS):
    rng = np.random.default_rng(nR ^ nS)
    if kind == "unique_uniform":
        return aj.gen_unique(nR, 11), aj.gen_uniform(nS, 1, nR, 12)
    if kind == "dups":  # about 4 copies of every key on each side
        k = max(nR // 4, 1)
        return aj.gen_uniform(nR, 1, k, 13), aj.gen_uniform(nS, 1, k, 14)
    if kind == "negative":
        keys = rng.integers(-2**31, 2**31, max(nR // 2, 1), dtype=np.int64).astype(np.int32)
        pay = lambda m: rng.integers(-2**31, 2**31, m, dtype=np.int64).astype(np.int32)
        return (aj.as_tuples(rng.choice(keys, nR), pay(nR)),
                aj.as_tuples(rng.choice(keys, nS), pay(nS)))
    if kind == "sentinels":  # all-ones R tuples, and the key hashing to all ones
        # (~200 copies each: every copy of a key lands in one partition, whose
        # capacity leaves room for a few hundred beyond its expected size)
        R = aj.gen_uniform(nR, 1, nR, 15)
        R["key"][::20000] = -1
        R["payloadID"][::20000] = -1
        R["key"][3::21000] = np.int32(KEY_HASH_ALL_ONES)
        R["payloadID"][3::21000] = -1
        S = aj.gen_uniform(nS, 1, nR, 16)
        S["key"][::5] = -1
        S["key"][2::13] = np.int32(KEY_HASH_ALL_ONES)
        return R, S
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["unique_uniform", "dups", "negative", "sentinels"])
@pytest.mark.parametrize("nR,nS", [(1 << 22, 1 << 22), ((1 << 22) + 4321, (1 << 23) - 77)])
def test_region_build_matches_oracle(aj, oracle, kind, nR, nS):
    R, S = _inputs(aj, kind, nR, nS)
    print(f"case npo region {kind} {nR}x{nS}", flush=True)
    got = _npo(aj, R, S)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)
    assert got.flags == aj.AJ_OUT_STAGED
    with options(aj, npo_region_min=0):  # the flat build: the same result
        flat = _npo(aj, R, S)
    assert (flat.count, flat.checksum, flat.flags) == (got.count, got.checksum, 0)


@pytest.mark.parametrize("nR", [4097, 6000, 65536, 100003, 1 << 20])
def test_region_build_small_tables(aj, oracle, nR):
    """Down to a table of 2^14 slots (4 regions of 4096)."""
    nS = 3 * nR + 5
    R, S = _inputs(aj, "dups", nR, nS)
    print(f"case npo region small {nR}", flush=True)
    with options(aj, npo_region_min=1):
        got = _npo(aj, R, S)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)
    assert got.flags == aj.AJ_OUT_STAGED


@pytest.mark.parametrize("kind", ["zipf", "one_key"])
def test_region_overflow_runs_the_flat_build(aj, oracle, kind):
    n = 1 << 22
    if kind == "zipf":  # the top keys overflow their partitions
        R, S = aj.gen_zipf(n, n, 1.2, 21), aj.gen_uniform(n, 1, n, 22)
    else:
        R = aj.as_tuples(np.full(n, 99), np.arange(n))
        S = aj.gen_uniform(n, 90, 110, 23)
    print(f"case npo region overflow {kind}", flush=True)
    got = _npo(aj, R, S)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)
    assert got.flags == aj.AJ_OUT_FALLBACK
    with options(aj, kernel_timing=1):  # again on the reused workspace, with timing
        again = _npo(aj, R, S)
    assert (again.count, again.checksum, again.flags) == (got.count, got.checksum, got.flags)
    assert again.phases()["npo_build"] > 0 and again.phases()["npo_probe"] > 0


def test_region_build_pairs_use_the_flat_build(aj, oracle):
    n = 1 << 22
    R, S = _inputs(aj, "dups", n, n)
    cnt, chk = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    buf = aj.DeviceArray(8 * cnt)
    print(f"case npo region pairs {cnt}", flush=True)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, n, dS.ptr, n, aj.AJ_ALGO_NPO, pairs_ptr=buf.ptr,
                           pairs_capacity=cnt)
            got = np.sort(buf.to_host(np.uint64, cnt))
            agg = ctx.join(dR.ptr, n, dS.ptr, n, aj.AJ_ALGO_NPO)
    finally:
        dR.free()
        dS.free()
        buf.free()
    assert (out.count, out.checksum, out.npairs, out.flags) == (cnt, chk, cnt, 0)
    assert (agg.count, agg.checksum, agg.flags) == (cnt, chk, aj.AJ_OUT_STAGED)
    assert np.array_equal(got, oracle.pairs_sortmerge(R, S))


def test_region_build_8192_slot_regions(aj):
    """|R| = 2^27 + 1000: a table of 2^29 slots, 2^16 regions of 8192 slots.
    R keys 1..|R| (shuffled) with payload 3 * key; S keys over [1, |R| + 2^20]:
    count and checksum in closed form."""
    nR, nS = (1 << 27) + 1000, 1 << 24
    rng = np.random.default_rng(77)
    keys = rng.permutation(np.arange(1, nR + 1, dtype=np.uint32))
    R = aj.as_tuples(keys.view(np.int32), (keys * np.uint32(3)).view(np.int32))
    del keys
    skeys = rng.integers(1, nR + (1 << 20) + 1, nS, dtype=np.uint32)
    spay = rng.integers(0, 2**32, nS, dtype=np.uint64).astype(np.uint32)
    S = aj.as_tuples(skeys.view(np.int32), spay.view(np.int32))
    hit = skeys <= nR
    want_cnt = int(hit.sum())
    rpay = (skeys[hit] * np.uint32(3)).astype(np.uint64)
    want_chk = int(((rpay << np.uint64(32)) + spay[hit].astype(np.uint64)).sum(dtype=np.uint64))
    print("case npo region 8192-slot regions", flush=True)
    got = _npo(aj, R, S)
    assert (got.count, got.checksum) == (want_cnt, want_chk)
    assert got.flags == aj.AJ_OUT_STAGED
