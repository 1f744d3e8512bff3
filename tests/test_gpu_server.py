"""GPU: the serving loop.  Every window's result equals the CPU oracle on that
window's tuples and a plain single join of the same tuples, whether the
window ran alone or batched with others; admission queues windows (does not
fail them) while the HBM budget is taken; misuse returns status codes."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _window_tuples(aj, ctx, n, seed, domain, theta):
    """The tuples aj_server_window_fill writes, regenerated on the device."""
    if n == 0:
        return np.zeros(0, dtype=aj.TUPLE_DTYPE)
    d = aj.DeviceArray(12 * n)
    aj.gen_stream_tuples(ctx, d.ptr, n, seed, domain, theta)
    t = d.to_host(aj.STREAM_DTYPE, n)
    d.free()
    return aj.as_tuples(t["key"].view(np.int32), t["payload"].view(np.int32))


def _expected(aj, oracle, ctx, spec):
    nR, nS, seed, domain, theta = spec
    R = _window_tuples(aj, ctx, nR, seed, domain, 0.0)
    S = _window_tuples(aj, ctx, nS, seed ^ 0x5DEECE66D, domain, theta)
    return R, S, oracle.aggregate(R, S)


def _specs(n, seed, lo=1 << 12, hi=1 << 17):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nR = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        nS = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        domain = int(rng.choice([1 << 10, 1 << 14, 1 << 16]))
        theta = 1.0 if i % 2 else 0.0
        out.append((nR, nS, 1000 + i, domain, theta))
    return out


def _run(aj, srv, specs):
    wins = []
    for i, (nR, nS, seed, domain, theta) in enumerate(specs):
        w = srv.alloc(nR, nS)
        srv.fill(w, seed, domain, theta)
        srv.submit(w, tag=i)
        wins.append(w)
    srv.drain()
    res = {}
    while True:
        got = srv.poll(256, 0)
        if not got:
            break
        for r in got:
            res[r.tag] = r
    return res


def test_server_results_match_oracle_and_plain_join(aj, oracle):
    specs = _specs(40, 3)
    specs.append((0, 5000, 77, 1 << 10, 0.0))   # empty R side
    specs.append((3000, 0, 78, 1 << 10, 0.0))   # empty S side
    with aj.Context(0) as ctx:
        with aj.Server(ctx, hbm_budget=4 << 30, slots=3, max_batch_windows=16,
                       max_batch_tuples=1 << 20, max_window_tuples=1 << 18) as srv:
            print(f"case server windows={len(specs)}", flush=True)
            res = _run(aj, srv, specs)
            st = srv.stats()
        assert len(res) == len(specs)
        assert st["windows_done"] == len(specs) and st["batches"] >= 3
        assert st["window_bytes_used"] == 0
        batched = 0
        for i, spec in enumerate(specs):
            r = res[i]
            assert r.status == 0
            R, S, want = _expected(aj, oracle, ctx, spec)
            assert (r.count, r.checksum) == want, (i, spec)
            assert r.close_us <= r.start_us + 1e-3 <= r.done_us + 1e-3
            batched += r.batch_windows > 1
            if len(R) and len(S):
                dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
                alone = ctx.join(dR.ptr, len(R), dS.ptr, len(S))
                assert (alone.count, alone.checksum) == want, i
        assert batched > 0  # some windows shared a batch


def test_server_admission_queues_under_budget(aj, oracle):
    """A budget that holds only a few windows: producers wait in window_alloc
    and every window still completes with the right result."""
    specs = [(60000, 60000, 500 + i, 1 << 12, 0.5 if i % 3 == 0 else 0.0) for i in range(24)]
    with aj.Context(0) as ctx:
        # workspaces for 2 slots of 2^18 tuples, plus room for about 4 windows
        probe = aj.Server(ctx, hbm_budget=8 << 30, slots=2, max_batch_windows=8,
                          max_batch_tuples=1 << 18, max_window_tuples=1 << 17)
        ws = probe.stats()["workspace_bytes"]
        probe.close()
        budget = ws + 4 * 120000 * 12 + 1000
        with aj.Server(ctx, hbm_budget=budget, slots=2, max_batch_windows=8,
                       max_batch_tuples=1 << 18, max_window_tuples=1 << 17) as srv:
            assert srv.stats()["window_bytes"] < 5 * 120000 * 12
            results, errors = {}, []

            def producer(part):
                try:
                    for i in part:
                        nR, nS, seed, domain, theta = specs[i]
                        w = srv.alloc(nR, nS, timeout_us=-1)
                        srv.fill(w, seed, domain, theta)
                        srv.submit(w, tag=i)
                except Exception as e:  # pragma: no cover - reported below
                    errors.append(e)

            threads = [threading.Thread(target=producer, args=(range(k, 24, 3),)) for k in range(3)]
            print("case admission 24 windows, budget for 4", flush=True)
            for t in threads:
                t.start()
            while len(results) < 24 and not errors:
                for r in srv.poll(64, 200000):
                    results[r.tag] = r
            for t in threads:
                t.join()
            assert not errors, errors
            st = srv.stats()
        assert st["alloc_waits"] > 0
        assert st["window_bytes_peak"] <= st["window_bytes"]
        for i, spec in enumerate(specs):
            _, _, want = _expected(aj, oracle, ctx, spec)
            assert (results[i].count, results[i].checksum) == want, i


def test_server_misuse_returns_status(aj):
    with aj.Context(0) as ctx:
        with aj.Server(ctx, hbm_budget=2 << 30, slots=1, max_batch_windows=4,
                       max_batch_tuples=1 << 16, max_window_tuples=1 << 15) as srv:
            with pytest.raises(aj.AjError) as e:
                srv.alloc(1 << 16, 10)            # above max_window_tuples
            assert e.value.status == aj.AJ_ERR_ARG
            w = srv.alloc(1000, 1000)
            srv.fill(w, 1, 100)
            srv.submit(w, tag=1)
            with pytest.raises(aj.AjError) as e:
                srv.submit(w, tag=2)              # twice
            assert e.value.status in (aj.AJ_ERR_STATE, aj.AJ_ERR_ARG)
            w2 = srv.alloc(500, 500)
            srv.release(w2)                       # cancelled before submit
            with pytest.raises(aj.AjError) as e:
                srv.release(w2)                   # unknown now
            assert e.value.status == aj.AJ_ERR_ARG
            srv.drain()
            got = srv.poll(8, 0)
            assert [r.tag for r in got] == [1]
            assert srv.stats()["window_bytes_used"] == 0
        with pytest.raises(aj.AjError) as e:
            aj.Server(ctx, hbm_budget=1 << 20)    # budget below the workspaces
        assert e.value.status == aj.AJ_ERR_CAPACITY


def test_server_fill_order_across_release_refill_and_upload(aj, oracle):
    """Window fills are queued and launched together; every later use of the
    same memory still sees them in the order they were made: a filled window
    released unsubmitted gives its memory to the next window (whose own fill
    must win), a window filled twice keeps the second fill, a filled window
    overwritten by an upload keeps the uploaded tuples, and a window with an
    empty side is published without a batch while its other side's fill is
    queued."""
    with aj.Context(0) as ctx:
        with aj.Server(ctx, hbm_budget=3 << 30, slots=2, max_batch_windows=8,
                       max_batch_tuples=1 << 20, max_window_tuples=1 << 18) as srv:
            n = 1 << 15
            specs = {}
            # released after a fill: the memory goes to the next window
            w0 = srv.alloc(n, n)
            srv.fill(w0, 11, 1 << 12, 1.0)
            srv.release(w0)
            w1 = srv.alloc(n, n)
            srv.fill(w1, 12, 1 << 14, 0.0)
            srv.submit(w1, tag=1)
            specs[1] = (n, n, 12, 1 << 14, 0.0)
            # filled twice: the second fill is the window's content
            w2 = srv.alloc(n, 2 * n)
            srv.fill(w2, 21, 1 << 10, 0.0)
            srv.fill(w2, 22, 1 << 13, 1.0)
            srv.submit(w2, tag=2)
            specs[2] = (n, 2 * n, 22, 1 << 13, 1.0)
            # an empty side: published at once, its R fill still queued
            w3 = srv.alloc(n, 0)
            srv.fill(w3, 31, 1 << 12, 0.0)
            srv.submit(w3, tag=3)
            w4 = srv.alloc(n, n)
            srv.fill(w4, 41, 1 << 12, 0.0)
            srv.submit(w4, tag=4)
            specs[4] = (n, n, 41, 1 << 12, 0.0)
            # filled, then uploaded: the upload is the window's content
            rng = np.random.default_rng(5)
            up = []
            for _ in range(2):
                t = np.zeros(n, dtype=aj.STREAM_DTYPE)
                t["key"] = rng.integers(1, 1 << 11, n)
                t["payload"] = np.arange(n)
                up.append(t)
            R5, S5 = (aj.as_tuples(t["key"].view(np.int32), t["payload"].view(np.int32))
                      for t in up)
            w5 = srv.alloc(n, n)
            srv.fill(w5, 52, 1 << 15, 1.0)
            srv.upload(w5, up[0], up[1])
            srv.submit(w5, tag=5)
            srv.drain()
            res = {}
            while True:
                got = srv.poll(64, 0)
                if not got:
                    break
                for r in got:
                    res[r.tag] = r
            assert sorted(res) == [1, 2, 3, 4, 5]
            assert (res[3].count, res[3].checksum) == (0, 0)
            for tag, spec in specs.items():
                _, _, want = _expected(aj, oracle, ctx, spec)
                assert (res[tag].count, res[tag].checksum) == want, tag
            assert (res[5].count, res[5].checksum) == oracle.aggregate(R5, S5)
            assert all(r.status == aj.AJ_OK for r in res.values())
            assert srv.stats()["window_bytes_used"] == 0
