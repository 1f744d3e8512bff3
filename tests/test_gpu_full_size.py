"""GPU: the headline workload at full size, |R| = |S| = 2^27 8-byte tuples, and the
skewed foreign-key workload at full size (|R| = 2^24, |S| = 2^28 Zipf).

R keys are a seeded permutation of 1..2^27 (payload = row id) and S keys are
uniform over 1..2^27 (payload = row id), so every S tuple matches exactly the R
tuple holding its key.  The pair list is therefore fixed without any join:
{(R payload of s.key) << 32 | s.payload : s in S}.

Bit-exact bar for the drop-in PRO and NPO: the count and checksum equal the CPU
PRO oracle's on the host cores (which equal the closed form), and the pair list
equals the closed form's.  Because every S payload occurs in exactly one pair,
list equality is checked in O(n) instead of by sorting 2^27 pairs: the low words
of the returned pairs are each S payload exactly once, and each high word is the
R payload that S row's key fixes -- the same multiset, so the same sorted list."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 27
SEED_R, SEED_S = 12345, 54321


@pytest.fixture(scope="module")
def full(aj, oracle):
    print("case full size: inputs, CPU PRO oracle (16 threads), closed form", flush=True)
    R = aj.gen_unique(N, SEED_R)
    S = aj.gen_uniform(N, 1, N, SEED_S)
    cpu = oracle.pro(R, S, nthreads=16)
    # closed form: the R payload matched by S row p, indexed by p = S payload
    inv = np.zeros(N + 1, np.uint32)
    inv[R["key"]] = R["payloadID"].view(np.uint32)
    sp = S["payloadID"].view(np.uint32)
    want_hi = np.empty(N, np.uint32)
    want_hi[sp] = inv[S["key"]]
    del inv
    assert np.array_equal(np.sort(sp), np.arange(N, dtype=np.uint32))  # S payloads unique
    closed_sum = ((int(want_hi.sum(dtype=np.uint64)) << 32) + int(sp.sum(dtype=np.uint64))) % (1 << 64)
    assert (cpu.count, cpu.checksum) == (N, closed_sum)
    return R, S, cpu.count, cpu.checksum, want_hi


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_full_size_pair_list(aj, full, algo):
    R, S, count, checksum, want_hi = full
    print(f"case full size {algo} 2^27 x 2^27 with pairs", flush=True)
    aj.set_option("materialize", 1)
    try:
        got = getattr(aj, algo)(R, S, 8)
    finally:
        aj.set_option("materialize", 0)
    assert (got.count, got.checksum) == (count, checksum)
    assert got.pairs is not None and len(got.pairs) == N
    lo = (got.pairs & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    hi = (got.pairs >> np.uint64(32)).astype(np.uint32)
    assert (np.bincount(lo, minlength=N) == 1).all()  # every S payload exactly once
    assert np.array_equal(hi, want_hi[lo])


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.0, 1.5])
def test_full_size_skewed_foreign_key(aj, theta):
    """|R| = 2^24 unique keys, |S| = 2^28 keys Zipf(theta) over R's keys: every S tuple
    matches exactly the R tuple holding its key, so count = |S| and the checksum is
    sum over S of (R payload of s.key) << 32 | s.payload (mod 2^64)."""
    nR, nS = 1 << 24, 1 << 28
    print(f"case full size skewed foreign key theta={theta}", flush=True)
    R = aj.gen_unique(nR, SEED_R)
    S = aj.gen_zipf(nS, nR, theta, SEED_S)
    inv = np.zeros(nR + 1, np.uint64)
    inv[R["key"]] = R["payloadID"].astype(np.uint64)
    hi_sum = int(inv[S["key"]].sum(dtype=np.uint64))
    del inv
    lo_sum = int(S["payloadID"].astype(np.uint64).sum(dtype=np.uint64))
    want = (nS, ((hi_sum << 32) + lo_sum) % (1 << 64))
    for algo in ("PRO", "NPO"):
        got = getattr(aj, algo)(R, S, 8)
        assert (got.count, got.checksum) == want, algo
