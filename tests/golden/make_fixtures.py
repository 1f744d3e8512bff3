"""Writes the golden join fixtures of tests/golden/ (run in the dev container).

Inputs are drawn with numpy's seeded generator (independent of the library's
own generators).  Expected values:
  count          from the reference's own RJ_st and NPO_st, compiled from its
                 sources (oracle/_ref/libref.so) -- both must agree
  s_digest       xxh64 of the reference's materialised <S.key, S.payloadID>
                 records (JOIN_RESULT_MATERIALIZE), sorted
  checksum       closed form by key aggregation (no join algorithm)
  pairs_digest   xxh64 of the sorted (R.payload << 32 | S.payload) list from a
                 textbook sort-merge
Usage: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np
import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

TUPLE = np.dtype([("payloadID", "<i4"), ("key", "<i4")])


def tuples(keys, payloads=None):
    t = np.empty(len(keys), TUPLE)
    t["key"] = keys
    t["payloadID"] = np.arange(len(keys)) if payloads is None else payloads
    return t


def cases():
    rng = np.random.default_rng(20240601)
    n = 4096
    yield "perm_perm_4096", tuples(rng.permutation(n) + 1), tuples(rng.permutation(n) + 1)
    n = 8192
    yield "unique_uniform_8192", tuples(rng.permutation(n) + 1), tuples(rng.integers(1, n + 1, n))
    yield "dups_4096", tuples(rng.integers(1, 257, 4096)), tuples(rng.integers(1, 257, 4096))
    z = rng.zipf(1.3, 20000)
    z = z[z <= 2048][:8192]
    yield "zipf_2048x8192", tuples(rng.permutation(2048) + 1), tuples(z)
    keys = rng.integers(-2**31, 2**31, 1024, dtype=np.int64).astype(np.int32)
    yield "negative_keys_4096", tuples(rng.choice(keys, 4096)), tuples(rng.choice(keys, 4096))
    yield "one_key_512", tuples(np.full(512, 42)), tuples(np.full(300, 42))
    yield "disjoint_2048", tuples(np.arange(2048) * 2), tuples(np.arange(2048) * 2 + 1)
    yield "empty_R", tuples(np.zeros(0, np.int32)), tuples(rng.integers(1, 10, 100))
    yield "single_pair", tuples(np.array([7])), tuples(np.array([7]))
    R = tuples(np.full(600, -1), np.full(600, -1))
    R["key"][::3] = 5
    S = tuples(np.full(700, -1))
    S["key"][1::2] = 5
    yield "all_ones_tuples", R, S


def digest(a: np.ndarray) -> str:
    return xxhash.xxh64(np.ascontiguousarray(a, np.uint64).tobytes()).hexdigest()


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    expected = {}
    arrays = {}
    for name, R, S in cases():
        c_rj, s_rj = oracle.ref_join("RJ_st", R, S, s_list=True)
        if len(R) >= 2:  # the reference's NPO_st cannot size a table for |R| < 2
            c_np, s_np = oracle.ref_join("NPO_st", R, S, s_list=True)
            assert c_rj == c_np and np.array_equal(s_rj, s_np), name
        cnt, chk = oracle.aggregate(R, S)
        assert cnt == c_rj, (name, cnt, c_rj)
        pairs = oracle.pairs_sortmerge(R, S)
        expected[name] = {"nR": len(R), "nS": len(S), "count": c_rj, "checksum": str(chk),
                          "s_digest": digest(s_rj), "pairs_digest": digest(pairs)}
        arrays[name + "_R"] = R.view(np.uint64) if len(R) else np.zeros(0, np.uint64)
        arrays[name + "_S"] = S.view(np.uint64) if len(S) else np.zeros(0, np.uint64)
        print(name, expected[name])
    np.savez_compressed(os.path.join(HERE, "joins.npz"), **arrays)
    with open(os.path.join(HERE, "joins.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
