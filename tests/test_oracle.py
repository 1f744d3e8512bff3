"""CPU: pin the oracle before trusting it.

Each oracle function is checked against something it does not compute itself:
the reference's own RJ_st/NPO_st (oracle/_ref, built from the reference
sources) and its materialised S records, the committed golden fixtures, closed
forms, and brute force.  Planted-mistake checks show the checksum catches the
plausible slips (swapped halves, dropped term, off-by-one payload).
"""
import json
import os

import numpy as np
import pytest
import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))
TUPLE = np.dtype([("payloadID", "<i4"), ("key", "<i4")])


def tuples(keys, payloads=None):
    t = np.empty(len(keys), TUPLE)
    t["key"] = keys
    t["payloadID"] = np.arange(len(keys)) if payloads is None else payloads
    return t


def digest(a):
    return xxhash.xxh64(np.ascontiguousarray(a, np.uint64).tobytes()).hexdigest()


def load_fixtures():
    with open(os.path.join(HERE, "golden", "joins.json")) as f:
        exp = json.load(f)
    arr = np.load(os.path.join(HERE, "golden", "joins.npz"))
    for name, e in sorted(exp.items()):
        R = arr[name + "_R"].view(TUPLE)
        S = arr[name + "_S"].view(TUPLE)
        yield name, R, S, e


FIXTURES = list(load_fixtures())


@pytest.mark.parametrize("name,R,S,e", FIXTURES, ids=[f[0] for f in FIXTURES])
def test_oracle_matches_golden(oracle, name, R, S, e):
    assert oracle.aggregate(R, S) == (e["count"], int(e["checksum"]))
    for bits, passes in ((8, 2), (4, 1), (12, 2)):
        r = oracle.pro(R, S, nthreads=4, radix_bits=bits, passes=passes, pairs=True)
        assert (r.count, r.checksum) == (e["count"], int(e["checksum"]))
        assert digest(oracle.sort_u64(r.pairs)) == e["pairs_digest"]
    r = oracle.npo(R, S, nthreads=3, pairs=True)
    assert (r.count, r.checksum) == (e["count"], int(e["checksum"]))
    assert digest(oracle.sort_u64(r.pairs)) == e["pairs_digest"]
    assert digest(oracle.pairs_sortmerge(R, S)) == e["pairs_digest"]


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("shape", ["unique", "dups", "wide"])
def test_oracle_vs_reference_build(oracle, seed, shape):
    """Counts of oj_pro/oj_npo equal the reference's own RJ_st and NPO_st; the S half
    of every pair equals the reference's materialised <S.key, S.payloadID>."""
    rng = np.random.default_rng(seed)
    n = 1 << 15
    if shape == "unique":
        R, S = tuples(rng.permutation(n) + 1), tuples(rng.integers(1, n + 1, n))
    elif shape == "dups":
        R, S = tuples(rng.integers(0, 900, n)), tuples(rng.integers(0, 900, n))
    else:
        R = tuples(rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32) >> 12)
        S = tuples(rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32) >> 12)
    c_rj, s_rj = oracle.ref_join("RJ_st", R, S, s_list=True)
    c_np, s_np = oracle.ref_join("NPO_st", R, S, s_list=True)
    assert c_rj == c_np
    p = oracle.pro(R, S, nthreads=8, pairs=True)
    q = oracle.npo(R, S, nthreads=8, pairs=True)
    assert p.count == q.count == c_rj
    assert p.checksum == q.checksum == oracle.aggregate(R, S)[1]
    # S half of the pairs, joined back to S's key, equals the reference's records
    skey = S["key"].astype(np.int64) & 0xFFFFFFFF
    s_pay = (p.pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)
    mine = np.sort((skey[s_pay].astype(np.uint64) << np.uint64(32)) | s_pay.astype(np.uint64))
    assert np.array_equal(mine, s_rj)


def test_aggregate_closed_forms(oracle):
    rng = np.random.default_rng(7)
    n = 5000
    # two permutations, payload = row id: count n, checksum (2^32+1) n(n-1)/2
    R, S = tuples(rng.permutation(n) + 1), tuples(rng.permutation(n) + 1)
    assert oracle.aggregate(R, S) == (n, ((1 << 32) + 1) * n * (n - 1) // 2 % (1 << 64))
    # one key on both sides: |R||S| pairs
    nr, ns = 300, 700
    R, S = tuples(np.full(nr, 9)), tuples(np.full(ns, 9))
    want = (ns * (nr * (nr - 1) // 2) << 32) + nr * (ns * (ns - 1) // 2)
    assert oracle.aggregate(R, S) == (nr * ns, want % (1 << 64))
    # disjoint keys
    assert oracle.aggregate(tuples(np.arange(100) * 2), tuples(np.arange(100) * 2 + 1)) == (0, 0)


@pytest.mark.parametrize("seed", range(5))
def test_aggregate_and_joins_vs_bruteforce(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    nr, ns = rng.integers(0, 300, 2)
    R = tuples(rng.integers(-5, 40, nr), rng.integers(-2**31, 2**31, nr, dtype=np.int64))
    S = tuples(rng.integers(-5, 40, ns), rng.integers(-2**31, 2**31, ns, dtype=np.int64))
    bf = oracle.pairs_bruteforce(R, S)
    cnt, chk = oracle.aggregate(R, S)
    assert cnt == len(bf)
    assert chk == int(bf.sum(dtype=np.uint64)) if len(bf) else chk == 0
    assert np.array_equal(oracle.pairs_sortmerge(R, S), bf)
    for bits, passes in ((1, 1), (6, 2), (16, 2)):
        r = oracle.pro(R, S, nthreads=1 + seed, radix_bits=bits, passes=passes, pairs=True)
        assert np.array_equal(oracle.sort_u64(r.pairs), bf)
    r = oracle.npo(R, S, nthreads=1 + seed, pairs=True)
    assert np.array_equal(oracle.sort_u64(r.pairs), bf)


def test_checksum_catches_planted_mistakes(oracle):
    """Each plausible slip changes the checksum on asymmetric data."""
    rng = np.random.default_rng(11)
    n = 2000
    R, S = tuples(rng.integers(1, 500, n)), tuples(rng.integers(1, 500, n))
    cnt, chk = oracle.aggregate(R, S)
    pairs = oracle.pairs_bruteforce(R, S)
    r_half, s_half = pairs >> np.uint64(32), pairs & np.uint64(0xFFFFFFFF)
    swapped = int(((s_half << np.uint64(32)) | r_half).sum(dtype=np.uint64))
    dropped_r = int(s_half.sum(dtype=np.uint64))
    off_by_one = int(((r_half << np.uint64(32)) | (s_half + np.uint64(1))).sum(dtype=np.uint64))
    assert chk == int(pairs.sum(dtype=np.uint64))
    assert len({chk, swapped, dropped_r, off_by_one}) == 4
    # aggregate itself distinguishes the orientation of the join
    assert oracle.aggregate(S, R)[1] != chk


def test_sort_u64(oracle):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**63, 10000, dtype=np.int64).astype(np.uint64) * np.uint64(2)
    assert np.array_equal(oracle.sort_u64(a), np.sort(a))


def test_reference_shj_counts_the_equi_join(aj, oracle):
    """The reference's eager single-thread SHJ produces exactly the equi-join
    matches (the closed-form count), so the lazy joins' oracle pins SHJ too."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref/libref.so not built")
    rng = np.random.default_rng(11)
    for nR, nS, dom in ((2, 2, 1), (500, 700, 50), (4096, 3000, 4096), (1000, 5000, 7)):
        R = aj.as_tuples(rng.integers(-5, dom, nR))
        S = aj.as_tuples(rng.integers(-5, dom, nS))
        assert oracle.ref_shj(R, S) == oracle.aggregate(R, S)[0]


def test_shj_progress_closed_form(aj, oracle):
    """Progress of a one-key join: with |R| = |S| = n all on one key, pairs with
    max(i, j) < (k + 1) b number ((k + 1) b)^2 after step k."""
    n, b = 64, 8
    R = aj.as_tuples(np.full(n, 3))
    S = aj.as_tuples(np.full(n, 3))
    want = np.array([((k + 1) * b) ** 2 for k in range(n // b)], np.uint64)
    assert np.array_equal(oracle.shj_progress(R, S, b), want)


def test_shj_records_pinned(aj, oracle):
    """The SHJ record restatement: at batch 1 the reference's own shj records (when the
    reference build is present); in one step (batch >= |R|, |S|) every match is found by
    the S tuple's probe, so each S tuple appears once per R tuple with its key."""
    R = aj.gen_uniform(700, 1, 90, 41)
    S = aj.gen_uniform(650, 1, 90, 42)
    if oracle.ref_available():
        c, ref = oracle.ref_shj_records(R, S)
        assert np.array_equal(oracle.shj_records(R, S, 1), ref) and c == len(ref)
    one = oracle.shj_records(R, S, 1 << 20)
    cnt = np.bincount(R["key"], minlength=91)
    sv = (S["key"].astype(np.uint64) << np.uint64(32)) | S["payloadID"].astype(np.uint64)
    assert np.array_equal(one, np.sort(np.repeat(sv, cnt[S["key"]])))
    # every record is one of the pair's two tuples and the count is the join's
    assert len(oracle.shj_records(R, S, 7)) == oracle.aggregate(R, S)[0]
