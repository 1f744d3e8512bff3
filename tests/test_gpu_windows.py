"""GPU parity of batched tumbling-window joins: every window's count and checksum
equals the CPU oracle's PRO on that window's tuples, alone or in the batch, with
or without graph capture, over repeated replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _streams(aj, ctx, nwin, lo, hi, seed, dup=4, rate=1.6):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), nwin)).astype(np.uint32)
    cap = int(sizes.sum() * 1.01 + 8 * np.sqrt(sizes.sum()) + 2048)
    out = {}
    for side, s in (("R", seed * 2 + 1), ("S", seed * 2 + 2)):
        d = aj.DeviceArray(12 * cap)
        b = aj.DeviceArray(4 * (nwin + 1))
        n = aj.gen_window_streams(ctx, s, sizes, rate, dup, d.ptr, cap, b.ptr)
        out[side] = (d, n, b)
    assert np.array_equal(out["R"][2].to_host(np.uint32, nwin + 1),
                          out["S"][2].to_host(np.uint32, nwin + 1))
    return sizes, out


def _expected(aj, oracle, out, nwin):
    bounds = out["R"][2].to_host(np.uint32, nwin + 1)
    host = {}
    for side in ("R", "S"):
        d, n, _ = out[side]
        host[side] = d.to_host(aj.STREAM_DTYPE, n)
    want = []
    for w in range(nwin):
        win = []
        for side in ("R", "S"):
            t = host[side]
            lo, hi = np.searchsorted(t["ts"], [bounds[w], bounds[w + 1]], side="left")
            win.append(aj.as_tuples(t["key"][lo:hi].view(np.int32),
                                    t["payload"][lo:hi].view(np.int32)))
        want.append(oracle.aggregate(*win))
    return bounds, host, want


@pytest.mark.parametrize("graph", [True, False])
def test_batch_windows_match_oracle(aj, oracle, graph):
    nwin = 96
    with aj.Context(0) as ctx:
        sizes, out = _streams(aj, ctx, nwin, 1 << 10, 1 << 17, seed=3)
        bounds, host, want = _expected(aj, oracle, out, nwin)
        assert np.all(np.diff(host["R"]["ts"].astype(np.int64)) >= 0)
        cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
        print(f"case graph={graph} nR={nR} nS={nS}", flush=True)
        with aj.BatchPlan(ctx, nR, nS, nwin, int(sizes.max() * 1.1), graph=graph) as plan:
            for rep in range(3):  # replays, with host work and syncs between them
                plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
                aj.device_sync()
                got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                               chk.to_host(np.uint64, nwin).tolist()))
                assert got == [(int(c), int(x)) for c, x in want], rep
                # a plain join on the same context between replays
                one = ctx.join(dR.ptr, 0, dS.ptr, 0)
                assert one.count == 0


def test_batch_graph_update_on_new_inputs(aj, oracle):
    """A second input set of other sizes updates the captured graph in place."""
    with aj.Context(0) as ctx:
        runs = []
        for seed, nwin in ((5, 40), (6, 25)):
            sizes, out = _streams(aj, ctx, nwin, 1 << 11, 1 << 16, seed=seed)
            runs.append((nwin, sizes, out, _expected(aj, oracle, out, nwin)[2]))
        max_R = max(r[2]["R"][1] for r in runs)
        max_S = max(r[2]["S"][1] for r in runs)
        with aj.BatchPlan(ctx, max_R, max_S, 64, 1 << 17) as plan:
            for rep in range(2):
                for nwin, sizes, out, want in runs:
                    cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
                    (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
                    plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
                    aj.device_sync()
                    got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                                   chk.to_host(np.uint64, nwin).tolist()))
                    assert got == [(int(c), int(x)) for c, x in want], (rep, nwin)


def test_batch_window_equals_single_join(aj):
    """A window's batched result equals a plain device PRO join of its tuples."""
    nwin = 12
    with aj.Context(0) as ctx:
        sizes, out = _streams(aj, ctx, nwin, 1 << 12, 1 << 15, seed=9)
        bounds = out["R"][2].to_host(np.uint32, nwin + 1)
        (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
        cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        with aj.BatchPlan(ctx, nR, nS, nwin) as plan:
            plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
            aj.device_sync()
        got_c, got_x = cnt.to_host(np.uint64, nwin), chk.to_host(np.uint64, nwin)
        hR, hS = dR.to_host(aj.STREAM_DTYPE, nR), dS.to_host(aj.STREAM_DTYPE, nS)
        for w in range(nwin):
            parts = []
            for t in (hR, hS):
                lo, hi = np.searchsorted(t["ts"], [bounds[w], bounds[w + 1]])
                parts.append(aj.as_tuples(t["key"][lo:hi].view(np.int32),
                                          t["payload"][lo:hi].view(np.int32)))
            single = aj.PRO(parts[0], parts[1], 1)
            assert (single.count, single.checksum) == (int(got_c[w]), int(got_x[w]))


@pytest.mark.parametrize("cap,target,chunk", [(4096, 2048, 16384), (12288, 8192, 65536),
                                              (8704, 0, 16384), (8192, 4096, 0)])
def test_batch_geometry_invariance(aj, oracle, cap, target, chunk):
    """Per-window results do not depend on the plan's LDS round size, its per-partition
    target (which sets each window's radix bits) or its chunk size (16K-tuple chunks split
    larger windows over several work-groups and leave partly filled tiles)."""
    nwin = 48
    with aj.Context(0) as ctx:
        sizes, out = _streams(aj, ctx, nwin, 1 << 9, 1 << 17, seed=7)
        bounds, host, want = _expected(aj, oracle, out, nwin)
        cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
        print(f"case cap={cap} target={target} chunk={chunk} nR={nR} nS={nS}", flush=True)
        with aj.BatchPlan(ctx, nR, nS, nwin, int(sizes.max() * 1.1), cap=cap, target=target,
                          chunk=chunk, graph=False) as plan:
            plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
            aj.device_sync()
            got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                           chk.to_host(np.uint64, nwin).tolist()))
        assert got == [(int(c), int(x)) for c, x in want]


def test_batch_autotuner_picks_a_geometry_with_identical_results(aj, oracle):
    """aj_tune_batch times 18 plan geometries on the resident batch (all must agree);
    a plan with the chosen geometry returns the oracle's per-window results."""
    nwin = 64
    with aj.Context(0) as ctx:
        sizes, out = _streams(aj, ctx, nwin, 1 << 12, 1 << 18, seed=11)
        bounds, host, want = _expected(aj, oracle, out, nwin)
        (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
        maxw = int(sizes.max() * 1.1)
        print(f"case batch autotuner nR={nR} nS={nS}", flush=True)
        opts, ms = aj.tune_batch(ctx, dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, maxw, reps=1)
        assert opts["cap"] in (8192, 8704, 12288) and ms > 0
        assert opts["target"] in (4096, 8192, 2048) and opts["chunk"] in (0, 65536, 262144)
        cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        with aj.BatchPlan(ctx, nR, nS, nwin, maxw, **opts) as plan:
            plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
            aj.device_sync()
            got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                           chk.to_host(np.uint64, nwin).tolist()))
        assert got == [(int(c), int(x)) for c, x in want]
        for d in (cnt, chk, out["R"][0], out["R"][2], out["S"][0], out["S"][2]):
            d.free()


@pytest.mark.parametrize("staged,case", [(1, "RS"), (1, "S"), (1, "zipf"), (1, "S_all"),
                                         (0, "RS"), (0, "zipf")])
def test_batch_windows_overflow_falls_back_exactly(aj, oracle, staged, case):
    """Skewed windows overflow the staged pass's fixed partition capacities.
    RS: windows 2 and 6 have all keys equal on both sides; an R partition that
    overflows sends the batch to the histogram pass on the device.  S: only their
    S keys are equal; the S runs past the capacity are spilled and gathered
    into the partition's extension range.  zipf: every window's S keys are Zipf
    distributed (many partitions spill).  S_all: every window's S keys are equal
    (nearly all of S spills).  Every
    window stays exact, with the staged pass on (option "staged" 1) or off, and
    so do an empty side and the ordinary windows of the same batch."""
    rng = np.random.default_rng(77)
    nwin = 10
    hot = (2, 6)
    sizes = {"R": [], "S": []}
    for w in range(nwin):
        for side in ("R", "S"):
            n = 1 << 14 if w in hot else int(rng.integers(1 << 11, 1 << 15))
            sizes[side].append(0 if (w == 4 and side == "S") else n)
    bounds = np.arange(nwin + 1, dtype=np.uint32) * 100000
    host = {}
    for side in ("R", "S"):
        parts = []
        for w in range(nwin):
            n = sizes[side][w]
            t = np.zeros(n, dtype=aj.STREAM_DTYPE)
            t["ts"] = bounds[w] + np.sort(rng.integers(0, 100000, n)).astype(np.uint32)
            dom = max(2, sizes["R"][w] // 4)
            keys = rng.integers(1, dom, n)
            if side == "S" and case == "zipf":
                keys = np.minimum(rng.zipf(1.3, n), dom - 1)
            if (w in hot and (case == "RS" or (case == "S" and side == "S"))) or \
                    (case == "S_all" and side == "S"):
                keys = np.full(n, 42)
            t["key"] = keys
            t["payload"] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
            parts.append(t)
        host[side] = np.concatenate(parts)
    want = []
    for w in range(nwin):
        win = []
        for side in ("R", "S"):
            t = host[side]
            lo, hi = np.searchsorted(t["ts"], [bounds[w], bounds[w + 1]], side="left")
            win.append(aj.as_tuples(t["key"][lo:hi].view(np.int32),
                                    t["payload"][lo:hi].view(np.int32)))
        want.append(oracle.aggregate(*win))
    assert want[4] == (0, 0)
    if case == "RS":
        assert want[2][0] == (1 << 28)
    aj.set_option("staged", staged)
    try:
        with aj.Context(0) as ctx:
            dR, dS = aj.DeviceArray.from_host(host["R"]), aj.DeviceArray.from_host(host["S"])
            db = aj.DeviceArray.from_host(bounds)
            cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
            nR, nS = len(host["R"]), len(host["S"])
            print(f"case {case} staged={staged} nR={nR} nS={nS}", flush=True)
            with aj.BatchPlan(ctx, nR, nS, nwin, 1 << 15) as plan:
                for rep in range(2):
                    plan.run(dR.ptr, nR, dS.ptr, nS, db.ptr, nwin, cnt.ptr, chk.ptr)
                    aj.device_sync()
                    got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                                   chk.to_host(np.uint64, nwin).tolist()))
                    assert got == [(int(c), int(x)) for c, x in want], rep
    finally:
        aj.set_option("staged", 1)


@pytest.mark.parametrize("staged", [1, 0])
def test_batch_plan_of_11_bits(aj, oracle, staged):
    """A plan whose largest windows take 11 radix bits (windows of up to 2^18
    tuples at 128 R tuples per partition): its scatters scan 2^11 digits
    block-wide, smaller windows take fewer bits, and every window is exact,
    staged pass on or off."""
    nwin = 24
    aj.set_option("staged", staged)
    try:
        with aj.Context(0) as ctx:
            sizes, out = _streams(aj, ctx, nwin, 1 << 12, 1 << 18, seed=21)
            bounds, host, want = _expected(aj, oracle, out, nwin)
            cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
            (dR, nR, bR), (dS, nS, _) = out["R"], out["S"]
            print(f"case 11-bit plan staged={staged} nR={nR} nS={nS}", flush=True)
            with aj.BatchPlan(ctx, nR, nS, nwin, 1 << 18, target=128, graph=False) as plan:
                plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
                aj.device_sync()
                got = list(zip(cnt.to_host(np.uint64, nwin).tolist(),
                               chk.to_host(np.uint64, nwin).tolist()))
            assert got == [(int(c), int(x)) for c, x in want]
    finally:
        aj.set_option("staged", 1)
