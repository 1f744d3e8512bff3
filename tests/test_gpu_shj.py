"""GPU: the eager symmetric hash join (SHJ_JM_NP / SHJ_JBCR_NP / SHJ_st and
algo AJ_ALGO_SHJ of the device API) against the oracle, counts, progress curves,
pairs and the reference's per-match records.

Bit-exact bar: count and checksum equal the equi-join's (closed form by key
aggregation, and the reference's own single-thread SHJ count), and the
per-step progress curve equals the one the pair list fixes (a pair is found in
the step of its later tuple)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _kinds(aj, n):
    rng = np.random.default_rng(n)
    yield "unique_uniform", aj.gen_unique(n, 1), aj.gen_uniform(n, 1, max(n, 1), 2)
    yield "dups", aj.gen_uniform(n, 1, max(n // 8, 1), 3), aj.gen_uniform(n, 1, max(n // 8, 1), 4)
    keys = rng.integers(-2**31, 2**31, max(n // 4, 1), dtype=np.int64).astype(np.int32)
    yield "negative", aj.as_tuples(rng.choice(keys, n)), aj.as_tuples(rng.choice(keys, n))
    R = aj.as_tuples(np.full(n, -1), np.full(n, -1))  # all-ones tuples on both sides
    R["key"][::3] = 5
    S = aj.as_tuples(np.full(n, -1), np.full(n, -1))
    S["key"][1::2] = 5
    yield "all_ones", R, S
    yield "zipf", aj.gen_unique(n, 5), aj.gen_zipf(2 * n, n, 1.2, 6)


@pytest.mark.parametrize("fn", ["SHJ_JM_NP", "SHJ_JBCR_NP", "SHJ_st"])
@pytest.mark.parametrize("n", [2, 1000, 1 << 17])
def test_shj_dropin(aj, oracle, fn, n):
    for kind, R, S in _kinds(aj, n):
        print(f"case {fn} {kind} n={n}", flush=True)
        got = getattr(aj, fn)(R, S, 8)
        want = oracle.aggregate(R, S)
        assert (got.count, got.checksum) == want, kind
        if oracle.ref_available() and len(R) >= 2 and len(S) >= 2 and n <= 1000:
            assert got.count == oracle.ref_shj(R, S), kind


def test_shj_golden_fixtures(aj):
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    exp = json.load(open(os.path.join(here, "golden", "joins.json")))
    arr = np.load(os.path.join(here, "golden", "joins.npz"))
    for name, e in sorted(exp.items()):
        R = arr[name + "_R"].view(aj.TUPLE_DTYPE)
        S = arr[name + "_S"].view(aj.TUPLE_DTYPE)
        print(f"case SHJ {name}", flush=True)
        got = aj.SHJ_JM_NP(R, S, 4)
        assert (got.count, got.checksum) == (e["count"], int(e["checksum"])), name


# steps of more than 2^19 tuples (more than one per lane of the step grid) use
# two separate tables instead of interleaved slot pairs
@pytest.mark.parametrize("nR,nS,batch", [(300, 300, 1), (5000, 7000, 7), (100000, 60000, 1000),
                                         (1 << 20, 1 << 20, 65536),
                                         ((1 << 22) + 999, 1 << 22, 1 << 20),
                                         (3 << 20, (1 << 20) + 5, 1 << 21)])
def test_shj_progress(aj, oracle, nR, nS, batch):
    R = aj.gen_uniform(nR, 1, max(nR // 4, 1), 21)
    S = aj.gen_uniform(nS, 1, max(nR // 4, 1), 22)
    want = oracle.shj_progress(R, S, batch)
    print(f"case SHJ progress {nR}x{nS} batch={batch}", flush=True)
    with aj.Context(0) as ctx:
        dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
        prog = np.zeros(len(want), np.uint64)
        try:
            out = ctx.join(dR.ptr, nR, dS.ptr, nS, algo=aj.AJ_ALGO_SHJ, batch=batch,
                           progress_ptr=prog.ctypes.data)
        finally:
            dR.free()
            dS.free()
    assert out.count == int(want[-1]) and (out.count, out.checksum) == oracle.aggregate(R, S)
    assert np.array_equal(prog, want)


def _shj_inputs(aj, kind):
    if kind == "dups":
        return aj.gen_uniform(3000, 1, 300, 31), aj.gen_uniform(2500, 1, 300, 32)
    R = aj.as_tuples(np.full(1200, -1), np.full(1200, -1))  # all-ones tuples on both sides
    R["key"][::3] = 5
    S = aj.as_tuples(np.full(900, -1), np.full(900, -1))
    S["key"][1::2] = 5
    return R, S


@pytest.mark.parametrize("kind", ["dups", "all_ones"])
@pytest.mark.parametrize("batch", [1, 7, 1000])
def test_shj_records_step_rule(aj, oracle, kind, batch):
    """Device API with records: per match the probing (later) tuple, as the step rule
    fixes it; at batch 1 that is the reference's own shj records."""
    R, S = _shj_inputs(aj, kind)
    want = oracle.shj_records(R, S, batch)
    if batch == 1 and oracle.ref_available():
        c, ref = oracle.ref_shj_records(R, S)
        assert c == len(want) and np.array_equal(ref, want)
    cnt = len(want)
    print(f"case SHJ records {kind} batch={batch} matches={cnt}", flush=True)
    with aj.Context(0) as ctx:
        dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
        buf = aj.DeviceArray(8 * cnt)
        try:
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), algo=aj.AJ_ALGO_SHJ, batch=batch,
                           pairs_ptr=buf.ptr, pairs_capacity=cnt, records=1)
            got = np.sort(buf.to_host(np.uint64, cnt))
        finally:
            dR.free()
            dS.free()
            buf.free()
    assert (out.count, out.checksum) == oracle.aggregate(R, S) and out.npairs == cnt
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", ["dups", "all_ones"])
def test_shj_pairs_and_dropin_results(aj, oracle, kind):
    """Pairs through the device API and the drop-ins' materialize 1 (pairs) and 2 (the
    reference's record lists); a too small pair buffer is an error."""
    R, S = _shj_inputs(aj, kind)
    pairs = oracle.pairs_sortmerge(R, S)
    cnt = len(pairs)
    with aj.Context(0) as ctx:
        dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
        buf = aj.DeviceArray(8 * cnt)
        try:
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), algo=aj.AJ_ALGO_SHJ, batch=64,
                           pairs_ptr=buf.ptr, pairs_capacity=cnt)
            assert out.npairs == cnt
            assert np.array_equal(np.sort(buf.to_host(np.uint64, cnt)), pairs)
            with pytest.raises(aj.AjError) as e:
                ctx.join(dR.ptr, len(R), dS.ptr, len(S), algo=aj.AJ_ALGO_SHJ, batch=64,
                         pairs_ptr=buf.ptr, pairs_capacity=cnt // 2)
            assert e.value.status == aj.AJ_ERR_CAPACITY
        finally:
            dR.free()
            dS.free()
            buf.free()
    aj.set_option("materialize", 1)
    try:
        got = aj.SHJ_JM_NP(R, S, 4)
    finally:
        aj.set_option("materialize", 0)
    assert got.count == cnt and np.array_equal(np.sort(got.pairs), pairs)
    aj.set_option("materialize", 2)
    try:
        got = aj.SHJ_st(R, S, 1)
    finally:
        aj.set_option("materialize", 0)
    # the drop-in's step: max(65536, n / 256) tuples per side, one step here
    assert got.count == cnt and got.pairs is None
    assert np.array_equal(np.sort(got.records[0]), oracle.shj_records(R, S, 65536))


@pytest.mark.parametrize("kind", ["dups", "all_ones"])
def test_shj_dropin_step_one_gives_reference_records(aj, oracle, kind):
    """Option shj_batch = 1: the drop-in SHJ_st alternates R[i], S[i] as the reference's
    shj does (localjoiner.cpp:43-52), so its record lists (materialize 2) are the
    reference's own records, and the timed records split into R and S probes as the
    reference's recordR / recordS lists do."""
    R, S = _shj_inputs(aj, kind)
    want = oracle.shj_records(R, S, 1)
    if oracle.ref_available():
        c, ref = oracle.ref_shj_records(R, S)
        assert c == len(want) and np.array_equal(ref, want)
    by_side = oracle.shj_records_by_side(R, S, 1)
    print(f"case SHJ drop-in shj_batch=1 {kind} matches={len(want)}", flush=True)
    aj.set_option("shj_batch", 1)
    aj.set_option("materialize", 2)
    try:
        got = aj.SHJ_st(R, S, 1)
        aj.set_option("materialize", 0)
        aj.set_option("timing", 1)
        timed = aj.SHJ_JM_NP(R, S, 2)
    finally:
        aj.set_option("materialize", 0)
        aj.set_option("timing", 0)
        aj.set_option("shj_batch", 0)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)
    assert np.array_equal(np.sort(got.records[0]), want)
    assert timed.count == len(want) and timed.timed_r == len(by_side[0])


def test_shj_empty_side(aj):
    R = aj.gen_unique(100, 1)
    E = np.empty(0, dtype=aj.TUPLE_DTYPE)
    assert aj.SHJ_JM_NP(R, E, 1).count == 0


def test_shj_timed_records_both_sides(aj, oracle, tmp_path):
    """Option timing = 1 on the eager join: one timed record per match, the probing
    tuple's payloadID; the last timed_r of them are R tuples' probes.  Under the
    drop-in's step (65536 tuples per side, four steps here) both sides probe, and each
    side's payloadIDs are the ones the step rule assigns to it.  The driver writes the
    progressive files under the algorithm's name."""
    import os
    import subprocess
    R = aj.gen_unique(200000, 12345)
    S = aj.gen_uniform(220000, 1, 200000, 54321)
    aj.set_option("timing", 1)
    try:
        print("case SHJ timed records", flush=True)
        got = aj.SHJ_st(R, S, 1)
    finally:
        aj.set_option("timing", 0)
    r_rec, s_rec = oracle.shj_records_by_side(R, S, 65536)
    lo = lambda v: np.sort(v & np.uint64(0xFFFFFFFF))  # noqa: E731
    assert got.count == len(r_rec) + len(s_rec) == len(got.timed)
    assert 0 < got.timed_r == len(r_rec) < got.count
    assert np.array_equal(lo(got.timed[:got.count - got.timed_r]), lo(s_rec))
    assert np.array_equal(lo(got.timed[got.count - got.timed_r:]), lo(r_rec))
    ms = (got.timed >> np.uint64(32)).astype(np.float64) / got.clock_hz * 1e3
    assert ms.min() >= 0 and ms.max() < 1000
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    d = tmp_path / "exp"
    r = subprocess.run([exe, "-a", "SHJ_JM_NP", "-r", "200000", "-s", "220000", "--timing",
                        str(d), "--exp-id", "5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = d / "results"
    assert (res / "records" / "SHJ_JM_NP_5.txt").read_text().split() == ["420000",
                                                                       str(got.count)]
    for sub in ("timestamps", "latency", "gaps"):
        v = np.array((res / sub / "SHJ_JM_NP_5.txt").read_text().split(), float)
        assert len(v) == got.count and np.all(np.diff(v) >= 0), sub


@pytest.mark.parametrize("batch", [1 << 19, (1 << 19) + 1])  # interleaved, then separate tables
def test_shj_pairs_both_table_layouts(aj, oracle, batch):
    n = 3 << 19
    R, S = aj.gen_uniform(n, 1, n // 8, 41), aj.gen_uniform(n, 1, n // 8, 42)
    cnt, chk = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    buf = aj.DeviceArray(8 * cnt)
    print(f"case SHJ pairs batch={batch} matches={cnt}", flush=True)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, n, dS.ptr, n, algo=aj.AJ_ALGO_SHJ, batch=batch,
                           pairs_ptr=buf.ptr, pairs_capacity=cnt)
            got = np.sort(buf.to_host(np.uint64, cnt))
    finally:
        dR.free()
        dS.free()
        buf.free()
    assert (out.count, out.checksum, out.npairs) == (cnt, chk, cnt)
    assert np.array_equal(got, oracle.pairs_sortmerge(R, S))
