"""GPU parity: PRO and NPO through the C ABI vs the CPU oracle.

Bit-exact bar: count, 64-bit wrapping checksum of (R.payload << 32 | S.payload),
and the sorted pair list.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED_R, SEED_S = 12345, 54321


def _inputs(aj, kind, n):
    if kind == "unique_uniform":  # the benchmark's key distribution (R a permutation, S uniform)
        return aj.gen_unique(n, SEED_R), aj.gen_uniform(n, 1, max(n, 1), SEED_S)
    if kind == "dups":  # duplicate factor ~8 on both sides
        k = max(n // 8, 1)
        return aj.gen_uniform(n, 1, k, 1), aj.gen_uniform(n, 1, k, 2)
    if kind == "one_key":
        return aj.as_tuples(np.full(n, 7)), aj.as_tuples(np.full(n, 7))
    if kind == "disjoint":
        return aj.as_tuples(np.arange(n) * 2), aj.as_tuples(np.arange(n) * 2 + 1)
    if kind == "negative":  # full 32-bit range incl. negative keys
        rng = np.random.default_rng(3)
        keys = rng.integers(-2**31, 2**31, max(n // 4, 1), dtype=np.int64).astype(np.int32)
        return (aj.as_tuples(rng.choice(keys, n)), aj.as_tuples(rng.choice(keys, n)))
    if kind == "all_ones":  # tuples whose bit pattern is all ones (key -1, payload -1)
        R = aj.as_tuples(np.full(n, -1), np.full(n, -1))
        R["key"][::3] = 5
        S = aj.as_tuples(np.full(n, -1))
        S["key"][1::2] = 5
        return R, S
    if kind == "zipf":
        return aj.gen_unique(n, SEED_R), aj.gen_zipf(4 * n, n, 1.0, SEED_S)
    raise ValueError(kind)


KINDS = ["unique_uniform", "dups", "one_key", "disjoint", "negative", "all_ones", "zipf"]


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [1, 1000, 1 << 16])
def test_dropin_count_checksum(aj, oracle, algo, kind, n):
    if kind == "one_key" and n > 1000:
        n = 1000  # |R|*|S| pairs
    R, S = _inputs(aj, kind, n)
    print(f"case {algo} {kind} n={n}", flush=True)
    got = getattr(aj, algo)(R, S, 8)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_empty_relations(aj, algo):
    R = aj.gen_unique(100, 1)
    E = np.empty(0, dtype=aj.TUPLE_DTYPE)
    for a, b in ((R, E), (E, R), (E, E)):
        got = getattr(aj, algo)(a, b, 4)
        assert (got.count, got.checksum) == (0, 0)


@pytest.mark.parametrize("bits,passes", [(0, 1), (4, 1), (8, 1), (13, 1), (8, 2), (13, 2),
                                         (14, 2), (15, 2), (16, 2), (18, 2)])
@pytest.mark.parametrize("cap", [4096, 8192, 12288])
def test_pro_geometry_invariance(aj, oracle, bits, passes, cap):
    """Results do not depend on radix bits, pass count or LDS round size."""
    n = 1 << 18
    R, S = aj.gen_uniform(n, 1, n // 2, 11), aj.gen_uniform(n, 1, n // 2, 12)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    print(f"case bits={bits} passes={passes} cap={cap}", flush=True)
    with aj.Context(0) as ctx:
        out = ctx.join(dR.ptr, n, dS.ptr, n, aj.AJ_ALGO_PRO, radix_bits=bits,
                       passes=passes, cap=cap)
    assert (out.count, out.checksum) == want


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
@pytest.mark.parametrize("kind", ["unique_uniform", "dups", "all_ones"])
def test_materialized_pairs(aj, oracle, algo, kind):
    R, S = _inputs(aj, kind, 5000)
    aj.set_option("materialize", 1)
    try:
        got = getattr(aj, algo)(R, S, 8)
    finally:
        aj.set_option("materialize", 0)
    ref = oracle.pairs_sortmerge(R, S)
    assert got.pairs is not None and len(got.pairs) == got.count == len(ref)
    assert np.array_equal(np.sort(got.pairs), ref)


def test_pro_split_heavy_partition(aj, oracle):
    """A hot key puts most of S in one partition: S is split over many tasks."""
    n = 1 << 18
    R = aj.gen_unique(n, 5)
    S = aj.gen_zipf(1 << 21, n, 1.5, 6)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    with aj.Context(0) as ctx:
        for split in (1024, 65536):
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO, split=split)
            assert (out.count, out.checksum) == want


def test_device_pairs_capacity_error(aj, oracle):
    R, S = _inputs(aj, "dups", 4096)
    cnt, _ = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    buf = aj.DeviceArray(8 * (cnt // 2))
    with aj.Context(0) as ctx:
        with pytest.raises(aj.AjError) as e:
            ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO,
                     pairs_ptr=buf.ptr, pairs_capacity=cnt // 2)
        assert e.value.status == aj.AJ_ERR_CAPACITY


def test_baseline_shape_closed_form(aj):
    """|R|=|S|=2^24 of the benchmark's key distribution: every S tuple matches the one R
    tuple with its key, so count = |S| and the checksum follows from R's inverse."""
    n = 1 << 24
    R, S = aj.gen_unique(n, SEED_R), aj.gen_uniform(n, 1, n, SEED_S)
    inv = np.empty(n + 1, np.uint64)
    inv[R["key"]] = R["payloadID"].astype(np.uint64)
    want_sum = int((inv[S["key"]] << np.uint64(32)).sum(dtype=np.uint64)
                   + S["payloadID"].astype(np.uint64).sum(dtype=np.uint64)) % (1 << 64)
    for fn in (aj.PRO, aj.NPO):
        got = fn(R, S, 8)
        assert got.count == n
        assert got.checksum == want_sum


def _fixtures():
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "joins.json")) as f:
        exp = json.load(f)
    arr = np.load(os.path.join(here, "golden", "joins.npz"))
    dt = np.dtype([("payloadID", "<i4"), ("key", "<i4")])
    return [(n, arr[n + "_R"].view(dt), arr[n + "_S"].view(dt), e) for n, e in sorted(exp.items())]


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_golden_fixtures(aj, algo):
    """HIP path vs the committed fixtures (counts from the reference's own RJ_st/NPO_st)."""
    import xxhash
    aj.set_option("materialize", 1)
    try:
        for name, R, S, e in _fixtures():
            print(f"case {algo} {name}", flush=True)
            got = getattr(aj, algo)(R, S, 8)
            assert got.count == e["count"], name
            assert got.checksum == int(e["checksum"]), name
            pairs = np.sort(got.pairs) if got.pairs is not None else np.zeros(0, np.uint64)
            assert xxhash.xxh64(pairs.tobytes()).hexdigest() == e["pairs_digest"], name
    finally:
        aj.set_option("materialize", 0)


def _u64(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _parse_tbl(path):
    """Out.tbl lines "key payloadID" -> uint64 records (key << 32 | payloadID)."""
    v = np.array(open(path).read().split(), dtype=np.int64).reshape(-1, 2)
    k = v[:, 0].astype(np.int32).view(np.uint32).astype(np.uint64)
    p = v[:, 1].astype(np.int32).view(np.uint32).astype(np.uint64)
    return (k << np.uint64(32)) | p


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_golden_records(aj, algo, tmp_path):
    """materialize = 2: per-match <S.key, S.payloadID> records in the reference's
    chained result lists vs the reference's own materialised records (s_digest,
    from its RJ_st / NPO_st), and the Out.tbl text written from them."""
    import xxhash
    aj.set_option("materialize", 2)
    try:
        for name, R, S, e in _fixtures():
            print(f"case {algo} {name}", flush=True)
            f = tmp_path / f"{algo}_{name}.tbl"
            got = getattr(aj, algo)(R, S, 4, result_file=str(f))
            assert got.count == e["count"] and got.checksum == int(e["checksum"]), name
            assert got.pairs is None
            assert [c[0] for c in got.chains] == [0, 1, 2, 3], name
            assert all(len(r) == 0 for r in got.records[1:]), name
            rec = got.records[0]
            assert len(rec) == e["count"], name
            assert xxhash.xxh64(np.sort(rec).tobytes()).hexdigest() == e["s_digest"], name
            assert np.array_equal(_parse_tbl(f), rec[::-1]), name
    finally:
        aj.set_option("materialize", 0)


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_records_multi_buffer(aj, oracle, algo, tmp_path):
    """A result of 2.6M records spans three 1M-tuple buffers: the chain layout,
    the records (against S rows picked by the oracle's pair list) and the Out.tbl
    bytes against the reference's own write_result_relation on the same result."""
    import ctypes
    from alliancedb_amd import _lib
    n = 1 << 21
    R = aj.gen_unique(n, 11)
    S = aj.gen_uniform(2600000, 1, n, 12)  # payloadID = row index
    pairs = oracle.pairs_sortmerge(R, S)
    want = np.sort(_u64(S)[(pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)])
    aj.set_option("materialize", 2)
    rr = _lib.relation_t(R.ctypes.data, len(R), None)
    rs = _lib.relation_t(S.ctypes.data, len(S), None)
    print(f"case {algo} records n={n}", flush=True)
    try:
        res = getattr(_lib.lib, algo)(ctypes.byref(rr), ctypes.byref(rs), 2)
    finally:
        aj.set_option("materialize", 0)
    assert res, aj.last_error()
    try:
        assert res.contents.totalresults == len(pairs) and res.contents.nthreads == 2
        tl = ctypes.cast(res.contents.resultlist, ctypes.POINTER(_lib.threadresult_t))
        cb = ctypes.cast(tl[0].results, ctypes.POINTER(_lib.aj_chainedtuplebuffer_t)).contents
        per = _lib.CHAINEDBUFF_NUMTUPLESPERBUF
        assert cb.numbufs == 3 and cb.writepos == len(pairs) - 2 * per
        assert tl[1].nresults == 0 and tl[1].threadid == 1
        rec = _lib._read_chain(cb, len(pairs))
        assert np.array_equal(np.sort(rec), want)
        ours = tmp_path / "ours.tbl"
        _lib.check(_lib.lib.aj_write_result_file(res, str(ours).encode()))
        assert np.array_equal(_parse_tbl(ours), rec[::-1])
        if oracle.ref_available():
            theirs = tmp_path / "ref.tbl"
            oracle.ref_write_result(res, str(theirs))
            assert ours.read_bytes() == theirs.read_bytes()
    finally:
        _lib.lib.aj_free_result(res)


def test_driver_persists_result(aj, oracle, tmp_path):
    """ajbench --out writes the reference's Out.tbl: one "key payloadID" line per
    match, the S tuple of the match."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    out = tmp_path / "Out.tbl"
    r = subprocess.run([exe, "-a", "PRO", "-r", "100000", "-s", "150000", "-n", "3",
                        "--out", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    R = aj.gen_unique(100000, 12345)
    S = aj.gen_uniform(150000, 1, 100000, 54321)
    pairs = oracle.pairs_sortmerge(R, S)
    want = np.sort(_u64(S)[(pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)])
    assert np.array_equal(np.sort(_parse_tbl(out)), want)


def test_driver_persists_shj_records(aj, oracle, tmp_path):
    """ajbench -a SHJ_st --out writes the eager join's records: per match the probing
    (later-arriving) tuple, under the drop-in's step of max(65536, n / 256) tuples."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    out = tmp_path / "Out.tbl"
    r = subprocess.run([exe, "-a", "SHJ_st", "-r", "100000", "-s", "150000", "--out", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    R = aj.gen_unique(100000, 12345)
    S = aj.gen_uniform(150000, 1, 100000, 54321)
    assert np.array_equal(np.sort(_parse_tbl(out)), oracle.shj_records(R, S, 65536))


@pytest.mark.parametrize("mode", ["zipf_both", "fk_zipf"])
@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_driver_runs_join(aj, oracle, algo, mode):
    """The C++ driver joins generated relations and prints the reference's result line.

    -d 2 draws both relations' keys Zipf over their own sizes, as the reference's
    createRelation does for R and for S (benchmark.cpp:103, 188, 204); --fk-zipf
    keeps R unique and draws S over R's keys."""
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    extra = ["--fk-zipf"] if mode == "fk_zipf" else []
    print(f"case driver {algo} {mode}", flush=True)
    r = subprocess.run([exe, "-a", algo, "-r", "200000", "-s", "300000", "-d", "2", "-z", "1.0",
                        "--repeats", "2", *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = re.findall(r"Results = (\d+)\. checksum = (\d+)", r.stdout)
    assert len(got) == 2
    if mode == "fk_zipf":
        R = aj.gen_unique(200000, 12345)
        S = aj.gen_zipf(300000, 200000, 1.0, 54321)
    else:
        R = aj.gen_zipf(200000, 200000, 1.0, 12345)
        S = aj.gen_zipf(300000, 300000, 1.0, 54321)
    want = oracle.aggregate(R, S)
    for c, x in got:
        assert (int(c), int(x)) == want


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0, 1.5])
@pytest.mark.parametrize("algo", [0, 1])
def test_skewed_foreign_key_join(aj, oracle, theta, algo):
    """Skewed workload shape at 1/16 size: |R| = 2^20 unique keys, |S| = 2^24 Zipf over R's keys."""
    nR, nS = 1 << 20, 1 << 24
    R = aj.gen_unique(nR, 77)
    S = aj.gen_zipf(nS, nR, theta, 78)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    print(f"case skew theta={theta} algo={algo}", flush=True)
    with aj.Context(0) as ctx:
        out = ctx.join(dR.ptr, nR, dS.ptr, nS, algo)
    assert (out.count, out.checksum) == want


def test_autotuner_picks_a_geometry_with_identical_results(aj, oracle):
    """aj_tune_pro times candidates on device data; the kept geometry joins with
    the same count and checksum as the oracle, and open-geometry calls use it."""
    n = 1 << 20
    R = aj.gen_unique(n, 5)
    S = aj.gen_uniform(n, 1, n, 6)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    aj.tune_clear()
    try:
        with aj.Context(0) as ctx:
            print("case autotune 2^20", flush=True)
            best = aj.tune_pro(ctx, dR.ptr, n, dS.ptr, n, reps=2)
            assert best["log2_r"] == 20 and best["log2_s"] == 20
            assert best["cap"] in (8192, 8704, 12288) and best["device_ms"] > 0
            assert aj.tune_get(n, n) == best
            out = ctx.join(dR.ptr, n, dS.ptr, n)  # geometry from the table
            assert (out.count, out.checksum) == want
            out = ctx.join(dR.ptr, n, dS.ptr, n, radix_bits=best["radix_bits"],
                           passes=best["passes"], cap=best["cap"])
            assert (out.count, out.checksum) == want
    finally:
        aj.tune_clear()


@pytest.mark.parametrize("bits,passes", [(8, 1), (14, 2), (15, 2), (18, 2)])
@pytest.mark.parametrize("block,chunk", [(256, 0), (512, 8192), (512, 65536), (256, 16384)])
def test_pro_block_and_chunk_invariance(aj, oracle, bits, passes, block, chunk):
    """Results do not depend on the partition work-group size (256 or 512 lanes: the
    256-lane scatters, 2-D histogram and pass 2) or the pass-1 chunk (8192-tuple chunks
    are half a 16384-tuple tile, so every tile is partly filled)."""
    n = 1 << 18
    R, S = aj.gen_uniform(n, 1, n // 2, 13), aj.gen_uniform(n + 777, 1, n // 2, 14)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    print(f"case bits={bits} passes={passes} block={block} chunk={chunk}", flush=True)
    aj.set_option("part_block", block)
    aj.set_option("part_chunk", chunk)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO, radix_bits=bits,
                           passes=passes)
    finally:
        aj.set_option("part_block", 0)
        aj.set_option("part_chunk", 0)
        dR.free()
        dS.free()
    assert (out.count, out.checksum) == want


@pytest.mark.parametrize("bits,passes", [(14, 2), (15, 2)])
@pytest.mark.parametrize("option,value", [("part_hist2d", 0), ("kernel_timing", 1), ("tuned", 0)])
def test_pro_option_invariance(aj, oracle, bits, passes, option, value):
    """Results do not depend on the remaining switches: the generic pass 2 instead of the
    2-D-histogram one, per-phase kernel timing, and ignoring the tuned-geometry table."""
    n = 1 << 18
    R, S = aj.gen_unique(n, 15), aj.gen_zipf(n, n, 1.0, 16)
    want = oracle.aggregate(R, S)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    print(f"case bits={bits} passes={passes} {option}={value}", flush=True)
    old = aj.get_option(option)
    aj.set_option(option, value)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, n, dS.ptr, n, aj.AJ_ALGO_PRO, radix_bits=bits, passes=passes)
    finally:
        aj.set_option(option, old)
        dR.free()
        dS.free()
    assert (out.count, out.checksum) == want


def test_device_budget_failure_returns_null_and_nomem(aj):
    """A join whose device memory does not fit the option device_budget fails cleanly:
    the drop-in returns NULL with AJ_ERR_NOMEM (the reference exit()s when an
    allocation fails, parallel_radix_join.cpp:58-65, common_functions.cpp:83-88), the
    device API returns the status, and after the budget is lifted the same process
    joins again.  Runs in a fresh process: the default context keeps the buffers
    earlier tests grew."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = r"""
import sys
sys.path.insert(0, %r)
import alliancedb_amd as aj
n = 1 << 22
R = aj.gen_unique(n, 1)
S = aj.gen_uniform(n, 1, n, 2)
d = aj.DeviceArray.from_host(R)  # 32 MiB held before the budget is set
aj.set_option("device_budget", 16 << 20)
for fn in (aj.PRO, aj.NPO):
    try:
        fn(R, S, 1)
        print("NOFAIL", flush=True)
    except aj.AjError as e:
        print("STATUS", e.status, flush=True)
with aj.Context(0) as ctx:
    try:
        ctx.join(d.ptr, n, d.ptr, n, aj.AJ_ALGO_NPO)
        print("NOFAIL", flush=True)
    except aj.AjError as e:
        print("STATUS", e.status, flush=True)
aj.set_option("device_budget", 0)
d.free()
r = aj.PRO(R, S, 1)
print("OK", r.count, flush=True)
""" % root
    print("case device_budget 16 MiB, 2^22 tuples per side", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    assert lines[:3] == [f"STATUS {aj.AJ_ERR_NOMEM}"] * 3, r.stdout
    assert lines[3] == f"OK {1 << 22}"


@pytest.mark.parametrize("algo", ["PRO", "NPO", "SHJ_st"])
def test_zipf_keys_on_both_sides(aj, oracle, algo):
    """Duplicate-heavy R and S: both relations' keys Zipf(1.0) over their own sizes, as
    the reference's -d 2 draws them (create_relation_zipf, generator.cpp:851-869); the
    top key has ~73K copies on each side."""
    n = 1 << 20
    R = aj.gen_zipf(n, n, 1.0, 31)
    S = aj.gen_zipf(n, n, 1.0, 32)
    print(f"case zipf both sides {algo}", flush=True)
    got = getattr(aj, algo)(R, S, 8)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)


@pytest.mark.parametrize("algo", ["NPO", "SHJ_st"])
@pytest.mark.parametrize("kind", ["zipf15", "one_key", "wide_payloads"])
def test_hash_tables_with_hot_keys(aj, oracle, algo, kind):
    """The hash-table joins with millions of copies of one key: an insert costs one
    atomic once the key's slot is found and a counting probe reads one slot, so the
    time does not grow with the copies (count up to ~10^13 matches).  With payloads
    over the full 32-bit range the per-key payload sums wrap many times: the S sums
    feed the low word of the checksum with their carries."""
    n = 1 << 24
    if kind == "zipf15":  # top key ~38% of each side
        R = aj.gen_zipf(n, n, 1.5, 41)
        S = aj.gen_zipf(n, n, 1.5, 42)
    elif kind == "wide_payloads":  # 64 keys, ~2^18 copies each per side
        rng = np.random.default_rng(44)
        pay = lambda m: rng.integers(-2**31, 2**31, m, dtype=np.int64).astype(np.int32)
        R = aj.as_tuples(rng.integers(0, 64, n >> 2).astype(np.int32), pay(n >> 2))
        S = aj.as_tuples(rng.integers(0, 64, n >> 2).astype(np.int32), pay(n >> 2))
    else:  # every R tuple has key 7; 1/16 of S matches
        R = aj.as_tuples(np.full(n >> 2, 7), np.arange(n >> 2))
        S = aj.gen_uniform(n >> 2, 1, 16, 43)
        S["key"][S["key"] != 7] += 100
    want = oracle.aggregate(R, S)
    print(f"case hot keys {algo} {kind} count={want[0]}", flush=True)
    got = getattr(aj, algo)(R, S, 8)
    assert (got.count, got.checksum) == want


@pytest.mark.parametrize("copies", [255, 256, 257])
def test_join_round_table_duplicate_counts(aj, oracle, copies):
    """A round's direct table counts the further copies of a key in 8 bits: R a
    permutation of 1..2^18 plus copies of three of its keys (one of them
    `copies` times in all, around the 8-bit limit, past which the round goes to
    the bucket table), S uniform over R's keys.  Counts and checksums equal the
    oracle's."""
    n = 1 << 18
    rng = np.random.default_rng(copies)
    base = rng.permutation(np.arange(1, n + 1))
    extra = np.concatenate([np.full(copies - 1, 1000), np.full(7, 2000), np.full(3, 3000)])
    keys = np.concatenate([base, extra]).astype(np.int64)
    rng.shuffle(keys)
    R = aj.as_tuples(keys.astype(np.int32), rng.integers(-2**31, 2**31, len(keys)).astype(np.int32))
    S = aj.as_tuples(rng.integers(1, n + 1, 2 * n).astype(np.int32),
                     rng.integers(-2**31, 2**31, 2 * n).astype(np.int32))
    print(f"case copies={copies} nR={len(R)} nS={len(S)}", flush=True)
    got = aj.PRO(R, S, 8)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)


def _heavy_inputs(aj, kind):
    nR, nS = 1 << 20, 1 << 22
    rng = np.random.default_rng(404)
    if kind in ("zipf10", "zipf15"):
        return aj.gen_unique(nR, 7), aj.gen_zipf(nS, nR, 1.0 if kind == "zipf10" else 1.5, 8)
    if kind == "one_key":  # every S tuple the same key, held three times by R
        keys = rng.permutation(np.arange(1, nR + 1))
        keys[:3] = 77
        R = aj.as_tuples(keys.astype(np.int32), rng.integers(-2**31, 2**31, nR).astype(np.int32))
        return R, aj.as_tuples(np.full(nS, 77, np.int32),
                               rng.integers(-2**31, 2**31, nS).astype(np.int32))
    if kind == "mixed":  # hot keys with R duplicates, hot keys missing from R, negative keys
        rk = rng.permutation(np.arange(1, nR + 1)).astype(np.int64)
        rk[:40] = np.repeat([5, 6, -3, 2**31 - 1], 10)
        hot = np.array([5, 6, -3, 2**31 - 1, 99999999, -77777])
        sk = rng.integers(1, nR + 1, nS)
        pick = rng.random(nS) < 0.5
        sk[pick] = rng.choice(hot, pick.sum())
        R = aj.as_tuples(rk.astype(np.int32), rng.integers(-2**31, 2**31, nR).astype(np.int32))
        S = aj.as_tuples(sk.astype(np.int32), rng.integers(-2**31, 2**31, nS).astype(np.int32))
        return R, S
    raise ValueError(kind)


@pytest.mark.parametrize("kind,heavy", [("zipf10", False), ("zipf15", True), ("one_key", True),
                                        ("mixed", True)])
def test_heavy_keys_joined_by_aggregates(aj, oracle, kind, heavy):
    """A skewed S (|S| >= 4|R|): when a few keys hold most of it (Zipf 1.5, one
    key, six keys holding half), they are joined by their per-key counts and
    payload sums and the rest by the partitioned join; at Zipf 1.0 the screen
    leaves the join to the partitioned path.  Count and checksum equal the
    oracle's either way, and with option "heavy" 0."""
    R, S = _heavy_inputs(aj, kind)
    want = oracle.aggregate(R, S)
    print(f"case heavy {kind} nR={len(R)} nS={len(S)}", flush=True)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO)
            assert (out.count, out.checksum) == want
            assert bool(out.flags & aj.AJ_OUT_HEAVY) == heavy
            if heavy:
                # the hottest key is always among those taken: its tuples are not
                # left to the partitioned join
                keys, counts = np.unique(S["key"], return_counts=True)
                assert out.light <= len(S) - counts.max()
                assert out.heavy_ms > 0.0
            else:
                assert out.light == 0 and out.heavy_ms == 0.0
            for bits, passes in ((11, 1), (10, 2)):  # other geometries of the light join
                out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO,
                               radix_bits=bits, passes=passes)
                assert (out.count, out.checksum) == want
                assert bool(out.flags & aj.AJ_OUT_HEAVY) == heavy
            aj.set_option("heavy", 0)
            try:
                out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO)
            finally:
                aj.set_option("heavy", 1)
            assert (out.count, out.checksum) == want and not out.flags & aj.AJ_OUT_HEAVY
    finally:
        dR.free()
        dS.free()


def test_heavy_keys_not_taken_without_skew(aj, oracle):
    """Uniform S of the same shape: the sample finds no heavy keys and the
    partitioned join runs alone."""
    R, S = aj.gen_unique(1 << 20, 7), aj.gen_uniform(1 << 22, 1, 1 << 20, 9)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    try:
        with aj.Context(0) as ctx:
            out = ctx.join(dR.ptr, len(R), dS.ptr, len(S), aj.AJ_ALGO_PRO)
    finally:
        dR.free()
        dS.free()
    assert (out.count, out.checksum) == oracle.aggregate(R, S)
    assert not out.flags & aj.AJ_OUT_HEAVY
