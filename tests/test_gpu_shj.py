"""GPU: the eager symmetric hash join (SHJ_JM_NP / SHJ_JBCR_NP / SHJ_st and
algo AJ_ALGO_SHJ of the device API) against the oracle.

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


@pytest.mark.parametrize("nR,nS,batch", [(300, 300, 1), (5000, 7000, 7), (100000, 60000, 1000),
                                         (1 << 20, 1 << 20, 65536)])
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


def test_shj_rejects_materialisation(aj):
    R = aj.gen_unique(100, 1)
    aj.set_option("materialize", 1)
    try:
        with pytest.raises(aj.AjError):
            aj.SHJ_st(R, R, 1)
    finally:
        aj.set_option("materialize", 0)
    E = np.empty(0, dtype=aj.TUPLE_DTYPE)
    assert aj.SHJ_JM_NP(R, E, 1).count == 0
