"""GPU parity of the staged two-pass partitioning (staged.hip) against the oracle.

The staged path partitions super-chunk by super-chunk into over-allocated
partitions whose room is reserved by atomics; a partition that overflows sends
the whole join to the histogram passes.  Bit-exact bar as everywhere: count,
64-bit checksum and the sorted pair list equal the oracle's, whichever path ran
(aj_join_out_t.flags says which), over pass bit splits, cursor sets, super-chunk
sizes (down to one tile: every super-chunk boundary is a tile boundary), work-group
sizes, ragged sizes and inputs built to overflow.
"""
import contextlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _inputs(aj, kind, nR, nS):
    if kind == "unique_uniform":
        return aj.gen_unique(nR, 12345), aj.gen_uniform(nS, 1, nR, 54321)
    if kind == "dups":
        k = max(nR // 8, 1)
        return aj.gen_uniform(nR, 1, k, 1), aj.gen_uniform(nS, 1, k, 2)
    if kind == "negative":
        rng = np.random.default_rng(3)
        keys = rng.integers(-2**31, 2**31, max(nR // 4, 1), dtype=np.int64).astype(np.int32)
        return aj.as_tuples(rng.choice(keys, nR)), aj.as_tuples(rng.choice(keys, nS))
    if kind == "all_ones":  # most tuples are the bit pattern all ones: one hot partition
        R = aj.as_tuples(np.full(nR, -1), np.full(nR, -1))
        R["key"][::3] = 5
        S = aj.as_tuples(np.full(nS, -1))
        S["key"][1::2] = 5
        return R, S
    if kind == "low_bits_zero":  # every key a multiple of 2^16: all in partition 0
        return (aj.as_tuples(np.arange(nR, dtype=np.int64) << 16),
                aj.as_tuples((np.arange(nS, dtype=np.int64) % nR) << 16))
    if kind == "zipf":
        return aj.gen_unique(nR, 5), aj.gen_zipf(nS, nR, 1.2, 6)
    raise ValueError(kind)


def _join(aj, R, S, bits, **kw):
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    try:
        with aj.Context(0) as ctx:
            return ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO, radix_bits=bits,
                            passes=2, **kw)
    finally:
        dR.free()
        dS.free()


@pytest.mark.parametrize("kind", ["unique_uniform", "dups", "negative", "all_ones"])
@pytest.mark.parametrize("bits", [2, 9, 12, 14, 16])
def test_staged_matches_oracle(aj, oracle, kind, bits):
    nR, nS = 1 << 20, (1 << 20) + 777
    R, S = _inputs(aj, kind, nR, nS)
    print(f"case staged {kind} bits={bits}", flush=True)
    out = _join(aj, R, S, bits)
    assert (out.count, out.checksum) == oracle.aggregate(R, S)
    if kind != "all_ones":  # the low bits of these keys spread evenly
        assert out.flags == aj.AJ_OUT_STAGED
    else:
        assert out.flags == aj.AJ_OUT_FALLBACK


@pytest.mark.parametrize("sets", [1, 2, 8])
@pytest.mark.parametrize("chunk", [16384, 1 << 20, 0])  # 16384: one 1024-lane tile
@pytest.mark.parametrize("block", [0, 512, 1024])  # 0: 512-lane pass 1, 1024-lane pass 2
def test_staged_geometry_invariance(aj, oracle, sets, chunk, block):
    nR, nS = (1 << 21) + 12345, (1 << 21) - 999
    R, S = _inputs(aj, "unique_uniform", nR, nS)
    want = oracle.aggregate(R, S)
    print(f"case staged sets={sets} chunk={chunk} block={block}", flush=True)
    with options(aj, stage_sets=sets, stage_chunk=chunk, stage_block=block):
        for bits in (10, 14, 15):
            out = _join(aj, R, S, bits)
            assert (out.count, out.checksum) == want, bits
            assert out.flags == aj.AJ_OUT_STAGED


@pytest.mark.parametrize("kind", ["low_bits_zero", "zipf"])
def test_staged_overflow_falls_back_exactly(aj, oracle, kind):
    nR, nS = 1 << 20, 1 << 21
    R, S = _inputs(aj, kind, nR, nS)
    want = oracle.aggregate(R, S)
    print(f"case staged overflow {kind}", flush=True)
    out = _join(aj, R, S, 14)
    assert out.flags == aj.AJ_OUT_FALLBACK
    assert (out.count, out.checksum) == want
    # the same join again (workspace reused after a fallback) and with timing on
    with options(aj, kernel_timing=1):
        out = _join(aj, R, S, 14)
    assert out.flags == aj.AJ_OUT_FALLBACK and (out.count, out.checksum) == want


@pytest.mark.parametrize("records", [0, 1])
def test_staged_pairs_and_records(aj, oracle, records):
    nR, nS = 1 << 20, 1 << 20
    R, S = _inputs(aj, "dups", nR, nS)
    cnt, _ = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    buf = aj.DeviceArray(8 * cnt)
    print(f"case staged pairs records={records} matches={cnt}", flush=True)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, nR, dS.ptr, nS, aj.AJ_ALGO_PRO, radix_bits=14, passes=2,
                           pairs_ptr=buf.ptr, pairs_capacity=cnt, records=records)
            got = np.sort(buf.to_host(np.uint64, cnt))
    finally:
        dR.free()
        dS.free()
        buf.free()
    assert out.flags == aj.AJ_OUT_STAGED and out.npairs == cnt
    if records:  # the matching S tuple per match: (key << 32 | payloadID)
        want = oracle.shj_records_by_side(R, S, nR + nS)[1]
        assert np.array_equal(got, want)
    else:
        assert np.array_equal(got, oracle.pairs_sortmerge(R, S))


def test_staged_timing_and_switch(aj, oracle):
    nR = nS = 1 << 22
    R, S = _inputs(aj, "unique_uniform", nR, nS)
    want = oracle.aggregate(R, S)
    with options(aj, kernel_timing=1):
        out = _join(aj, R, S, 14)
    assert out.flags == aj.AJ_OUT_STAGED and (out.count, out.checksum) == want
    ph = out.phases()
    assert ph["scatter1"] > 0 and ph["scatter2"] > 0 and ph["join"] > 0
    assert ph["hist1"] < 0.01 and ph["hist2"] < 0.01  # no histogram kernels
    with options(aj, staged=0):
        out = _join(aj, R, S, 14)
    assert out.flags == 0 and (out.count, out.checksum) == want


def _join1(aj, ctx, R, S, bits, **kw):
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    try:
        return ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO, radix_bits=bits,
                        passes=1, **kw)
    finally:
        dR.free()
        dS.free()


@pytest.mark.parametrize("kind", ["unique_uniform", "dups", "negative", "zipf", "all_ones",
                                  "low_bits_zero"])
@pytest.mark.parametrize("bits", [6, 11])
def test_one_pass_staged_matches_oracle(aj, oracle, kind, bits):
    """One pass of up to 11 bits runs as a one-segment window batch: no
    histogram, S runs past a partition's capacity spilled into its extension
    range (Zipf S stays on this path), an R partition that overflows (one hot
    key, keys with equal low bits) reruns the pass from a histogram on the
    device.  Exact either way; flags say which ran."""
    nR, nS = 1 << 20, (1 << 21) + 777
    R, S = _inputs(aj, kind, nR, nS)
    want = oracle.aggregate(R, S)
    print(f"case one-pass staged {kind} bits={bits}", flush=True)
    with aj.Context(0) as ctx:
        out = _join1(aj, ctx, R, S, bits)
        assert (out.count, out.checksum) == want
        hot_r = kind in ("all_ones", "low_bits_zero")
        assert out.flags == (aj.AJ_OUT_FALLBACK if hot_r else aj.AJ_OUT_STAGED)
        # again on the same context (its plan reused), with per-kernel timing
        with options(aj, kernel_timing=1):
            out = _join1(aj, ctx, R, S, bits)
        assert (out.count, out.checksum) == want
        if not hot_r:
            ph = out.phases()
            assert ph["scatter1"] > 0 and ph["join"] > 0 and ph["scatter2"] < 0.01


def test_one_pass_staged_plan_follows_sizes(aj, oracle):
    """The context's one-segment plan is rebuilt for larger inputs, other radix
    bits and round sizes, and after the option "staged" changes."""
    cases = [(1 << 20, 1 << 20, 10, 8704), ((1 << 21) + 5, (1 << 22) - 3, 11, 8704),
             (1 << 20, 1 << 21, 11, 12288), (1 << 19, 1 << 20, 9, 8192)]
    with aj.Context(0) as ctx:
        for nR, nS, bits, cap in cases:
            R, S = _inputs(aj, "unique_uniform", nR, nS)
            want = oracle.aggregate(R, S)
            print(f"case plan nR={nR} nS={nS} bits={bits} cap={cap}", flush=True)
            out = _join1(aj, ctx, R, S, bits, cap=cap)
            assert (out.count, out.checksum) == want and out.flags == aj.AJ_OUT_STAGED
        with options(aj, staged=0):
            out = _join1(aj, ctx, R, S, bits, cap=cap)
        assert (out.count, out.checksum) == want and out.flags == 0


@pytest.mark.parametrize("kind", ["unique_uniform", "dups", "negative", "zipf"])
def test_one_pass_histogram_at_11_bits(aj, oracle, kind):
    """One histogram pass of 11 bits (option "staged" off) scatters 2048 digits
    per tile with the block-wide digit scan; exact against the oracle."""
    nR, nS = 1 << 20, (1 << 21) + 777
    R, S = _inputs(aj, kind, nR, nS)
    want = oracle.aggregate(R, S)
    print(f"case one-pass histogram 11 bits {kind}", flush=True)
    with options(aj, staged=0), aj.Context(0) as ctx:
        out = _join1(aj, ctx, R, S, 11)
    assert (out.count, out.checksum) == want and out.flags == 0


@pytest.mark.parametrize("nR,nS", [(0, 1 << 20), ((1 << 20) + 3, 0), (1 << 10, (1 << 20) - 1),
                                   ((1 << 20) - 5, 9), ((1 << 19) + 1, (1 << 19) + 1)])
def test_one_pass_staged_edge_sizes(aj, oracle, nR, nS):
    """Empty and lopsided sides through the one-pass staged path (11 bits):
    the result equals the oracle's (0 matches when a side is empty)."""
    R = aj.gen_unique(nR, 12345) if nR else aj.as_tuples(np.zeros(0, np.int32))
    S = aj.gen_uniform(nS, 1, max(nR, 1), 54321) if nS else aj.as_tuples(np.zeros(0, np.int32))
    want = oracle.aggregate(R, S) if nR and nS else (0, 0)
    print(f"case one-pass staged edge nR={nR} nS={nS}", flush=True)
    with aj.Context(0) as ctx:
        dR = aj.DeviceArray.from_host(R) if nR else aj.DeviceArray(64)
        dS = aj.DeviceArray.from_host(S) if nS else aj.DeviceArray(64)
        try:
            out = ctx.join(dR.ptr, nR, dS.ptr, nS, aj.AJ_ALGO_PRO, radix_bits=11, passes=1)
        finally:
            dR.free()
            dS.free()
    assert (out.count, out.checksum) == want
