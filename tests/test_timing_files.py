"""CPU: progressive-result files (aj_write_timing_files) on results built here
in host memory.  The breakdown file is compared byte for byte with the
reference's own breakdown_global (timer/t_timer.cpp:113-140, compiled into
oracle/_ref/libref.so).  The record files are compared with the oracle's
restatement of merge / sortRecords / dump_timing (t_timer.cpp:31-99, 401-452:
the reference compiles those bodies out unless its scripts edit
common_functions.h:21 in place) and with closed-form values."""
import ctypes
import os

import numpy as np
import pytest

from alliancedb_amd import _lib

def _result(keep, count, timed, clock_hz, algo, phase_ms, device_ms):
    t = np.ascontiguousarray(timed, np.uint64)
    keep.append(t)
    r = _lib.aj_result_t()
    r.base = _lib.result_t(count, None, 1)
    r.ntimed = len(t)
    r.timed = t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)) if len(t) else None
    r.clock_hz = clock_hz
    r.algo = algo
    r.device_ms = device_ms
    for i, v in enumerate(phase_ms):
        r.phase_ms[i] = v
    keep.append(r)
    return r


def _cyc(ms):
    """Cycles of a float32 device time as the writer converts it."""
    return int(float(np.float32(ms)) * (2.1 * 1e6))


@pytest.mark.parametrize("algo,gap,last_ts", [(0, 1, 0), (0, 3, 1000), (1, 1, 1000), (1, 7, 0)])
def test_timing_files(aj, oracle, tmp_path, algo, gap, last_ts):
    rng = np.random.default_rng(algo * 10 + gap)
    nR, nS, count = 5000, 7000, 20000
    ticks = np.sort(rng.integers(0, 3_000_000, count)).astype(np.uint64)
    ticks[100:140] = ticks[100]  # ties: ordered by S.payloadID
    sids = rng.integers(0, nS, count).astype(np.uint64)
    s_ts = np.sort(rng.integers(0, int(last_ts * 2.1e6) + 1, nS)).astype(np.uint64)
    timed = (ticks << np.uint64(32)) | sids
    phases = [0.0] * 12
    if algo == 0:
        phases[0:8] = [0.01, 0.2, 0.03, 0.4, 0.01, 0.0, 0.02, 0.35]
        phases[8:10] = [0.05, 0.6]
    else:
        phases[10:12] = [0.3, 0.9]
    device_ms = 1.9
    keep = []
    res = _result(keep, count, timed, 2.1e9, algo, phases, device_ms)
    relR = _lib.relation_t(None, nR, None)
    pl = _lib.relation_payload_t(s_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nS)
    relS = _lib.relation_t(None, nS, ctypes.addressof(pl))
    ours = tmp_path / "ours"
    summ = _lib.aj_timing_summary_t()
    name = "PRJ" if algo == 0 else "NPJ"
    _lib.check(_lib.lib.aj_write_timing_files(
        ctypes.cast(ctypes.pointer(res), ctypes.POINTER(_lib.result_t)), ctypes.byref(relR),
        ctypes.byref(relS), str(ours).encode(), name.encode(), 7, last_ts, gap,
        ctypes.byref(summ)))
    # the same samples, in production order, through the restatement
    order = np.sort(timed)[gap - 1::gap]
    cyc = order >> np.uint64(32)  # 2.1e9 ticks/s: one tick per cycle
    sid = (order & np.uint64(0xFFFFFFFF)).astype(np.int64)
    times, lat, gaps = oracle.timing_records(cyc, sid, s_ts, last_ts)
    res_dir = ours / "results"
    rd = lambda d: (res_dir / d / f"{name}_7.txt").read_text()  # noqa: E731
    assert rd("records") == f"{nR + nS}\n{count}\n"
    assert rd("timestamps") == oracle.timing_text(times, last_ts)
    assert rd("latency") == oracle.timing_text(lat, last_ts)
    assert rd("gaps") == oracle.timing_text(gaps, last_ts)
    if algo == 0:
        part, build = _cyc(sum(np.float32(p) for p in phases[0:8])), 0
        join = _cyc(np.float32(phases[8]) + np.float32(phases[9]))
    else:
        part, build, join = 0, _cyc(phases[10]), _cyc(phases[11])
    overall = _cyc(device_ms)
    if oracle.ref_available():
        theirs = tmp_path / "ref_breakdown.txt"
        oracle.ref_breakdown_file(str(theirs), nR, nS, last_ts, overall, part, build, join)
        assert rd("breakdown") == theirs.read_text()
    n = len(order) - 1
    times = np.array(open(ours / "results" / "timestamps" / f"{name}_7.txt").read().split(),
                     float)
    assert summ.samples == len(order)
    assert summ.progress_ms[4] == pytest.approx(times[int(np.ceil(n * 0.5))], abs=1e-6)


def test_timing_files_need_timed_records(aj, tmp_path):
    keep = []
    res = _result(keep, 10, np.zeros(0, np.uint64), 0.0, 0, [0.0] * 12, 1.0)
    rel = _lib.relation_t(None, 5, None)
    rc = _lib.lib.aj_write_timing_files(
        ctypes.cast(ctypes.pointer(res), ctypes.POINTER(_lib.result_t)), ctypes.byref(rel),
        ctypes.byref(rel), str(tmp_path).encode(), b"PRJ", 1, 0, 1, None)
    assert rc == aj.AJ_ERR_STATE
    assert _lib.lib.aj_write_timing_files(None, None, None, None, None, 0, 0, 1, None) == \
        aj.AJ_ERR_ARG


def test_timing_files_closed_form(aj, tmp_path):
    """Three matches at 0, 1 and 2 ms of a 100 MHz clock (2.1e6 cycles per ms),
    S timestamps 0, 1 ms and 0.5 ms, window 10 ms: times 0, 1, 2 (+10); latency
    t - ts + 10 = 10, 10, 11.5 (+10); gaps sid - i = 0, 0, 0 (+10)."""
    keep = []
    ticks = np.array([0, 100000, 200000], np.uint64)
    sids = np.array([0, 1, 2], np.uint64)
    res = _result(keep, 3, (ticks << np.uint64(32)) | sids, 1e8, 0, [0.0] * 12, 0.5)
    s_ts = np.array([0, 2100000, 1050000], np.uint64)
    pl = _lib.relation_payload_t(s_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 3)
    relS = _lib.relation_t(None, 3, ctypes.addressof(pl))
    relR = _lib.relation_t(None, 2, None)
    summ = _lib.aj_timing_summary_t()
    _lib.check(_lib.lib.aj_write_timing_files(
        ctypes.cast(ctypes.pointer(res), ctypes.POINTER(_lib.result_t)), ctypes.byref(relR),
        ctypes.byref(relS), str(tmp_path).encode(), b"PRJ", 0, 10, 1, ctypes.byref(summ)))
    rd = lambda d: (tmp_path / "results" / d / "PRJ_0.txt").read_text().split()  # noqa: E731
    assert rd("records") == ["5", "3"]
    assert [float(x) for x in rd("timestamps")] == [10.0, 11.0, 12.0]
    assert [float(x) for x in rd("latency")] == [20.0, 20.0, 21.5]
    assert [float(x) for x in rd("gaps")] == [10.0, 10.0, 10.0]
    bd = [float(x) for x in rd("breakdown")]
    assert bd[0] == pytest.approx(10 * 2.1e6 / 5) and bd[1:6] == [0, 0, 0, 0, 0]
    assert summ.samples == 3 and summ.progress_ms[4] == 11.0 and summ.latency95_ms == 10.0


@pytest.mark.parametrize("gap,last_ts", [(1, 0), (4, 500)])
def test_timing_files_both_probe_sides(aj, oracle, tmp_path, gap, last_ts):
    """Eager-join results: the last ntimed_r records are R tuples' probes.  Samples are
    taken in production order over both sides, then each is measured against its own
    side's timestamps and counted in its own side's gap index, as merge does for the
    reference's recordR / recordS lists (t_timer.cpp:412-434)."""
    rng = np.random.default_rng(gap + 11)
    nR, nS, ns, nr = 4000, 6000, 9000, 5000
    s_ticks = rng.integers(0, 2_000_000, ns).astype(np.uint64)
    r_ticks = rng.integers(0, 2_000_000, nr).astype(np.uint64)
    r_ticks[:30] = s_ticks[0]  # ties across the sides: S first
    sids = rng.integers(0, nS, ns).astype(np.uint64)
    rids = rng.integers(0, nR, nr).astype(np.uint64)
    s_ts = np.sort(rng.integers(0, int(last_ts * 2.1e6) + 1, nS)).astype(np.uint64)
    r_ts = np.sort(rng.integers(0, int(last_ts * 2.1e6) + 1, nR)).astype(np.uint64)
    timed = np.concatenate([(s_ticks << np.uint64(32)) | sids, (r_ticks << np.uint64(32)) | rids])
    keep = []
    res = _result(keep, ns + nr, timed, 2.1e9, 2, [0.0] * 12, 1.5)
    res.ntimed_r = nr
    pr = _lib.relation_payload_t(r_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nR)
    ps = _lib.relation_payload_t(s_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nS)
    relR = _lib.relation_t(None, nR, ctypes.addressof(pr))
    relS = _lib.relation_t(None, nS, ctypes.addressof(ps))
    ours = tmp_path / "ours"
    _lib.check(_lib.lib.aj_write_timing_files(
        ctypes.cast(ctypes.pointer(res), ctypes.POINTER(_lib.result_t)), ctypes.byref(relR),
        ctypes.byref(relS), str(ours).encode(), b"SHJ", 1, last_ts, gap, None))
    # production order: ticks, then side (S before R), then payloadID
    t = np.concatenate([s_ticks, r_ticks])
    side = np.concatenate([np.zeros(ns, np.int64), np.ones(nr, np.int64)])
    pid = np.concatenate([sids, rids]).astype(np.int64)
    order = np.lexsort((pid, side, t))[gap - 1::gap]
    times, lat, gaps = oracle.timing_records(t[order], pid[order], s_ts, last_ts,
                                             sides=side[order], r_ts=r_ts)
    rd = lambda d: (ours / "results" / d / "SHJ_1.txt").read_text()  # noqa: E731
    assert rd("timestamps") == oracle.timing_text(times, last_ts)
    assert rd("latency") == oracle.timing_text(lat, last_ts)
    assert rd("gaps") == oracle.timing_text(gaps, last_ts)
    # a plain PRO-style result (no R-side records) is unchanged by the side handling
    assert rd("records") == f"{nR + nS}\n{ns + nr}\n"


def test_timing_files_both_sides_by_hand(aj, tmp_path):
    """Four matches worked out by hand: S probes (payloadIDs 2, 0) at 2.1e6 and 6.3e6
    cycles, R probes (payloadIDs 1, 0) at 4.2e6 and 8.4e6; S timestamps [0, 0, 1.05e6],
    R timestamps [2.1e6, 0]; lastTS 0, every match sampled.  Times are cycles / 2.1e6 ms;
    latency subtracts the probing tuple's own side's timestamp; the gap is its
    payloadID minus its index among its side's samples."""
    c = 2_100_000  # cycles per ms; clock_hz 2.1e9 gives one cycle per tick
    s = [(1 * c, 2), (3 * c, 0)]
    r = [(2 * c, 1), (4 * c, 0)]
    timed = np.array([(t << 32) | i for t, i in s + r], np.uint64)
    keep = []
    res = _result(keep, 4, timed, 2.1e9, 2, [0.0] * 12, 1.0)
    res.ntimed_r = 2
    s_ts = np.array([0, 0, c // 2], np.uint64)
    r_ts = np.array([c, 0], np.uint64)
    pr = _lib.relation_payload_t(r_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 2)
    ps = _lib.relation_payload_t(s_ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 3)
    relR = _lib.relation_t(None, 2, ctypes.addressof(pr))
    relS = _lib.relation_t(None, 3, ctypes.addressof(ps))
    _lib.check(_lib.lib.aj_write_timing_files(
        ctypes.cast(ctypes.pointer(res), ctypes.POINTER(_lib.result_t)), ctypes.byref(relR),
        ctypes.byref(relS), str(tmp_path).encode(), b"SHJ_st", 0, 0, 1, None))
    rd = lambda d: [float(x) for x in (tmp_path / "results" / d / "SHJ_st_0.txt").read_text().split()]  # noqa: E731,E501
    assert rd("timestamps") == [1.0, 2.0, 3.0, 4.0]
    # S: 1 - 0.5 (ts of S row 2), 3 - 0 ; R: 2 - 0 (R row 1), 4 - 1 (R row 0)
    assert rd("latency") == [0.5, 2.0, 3.0, 3.0]
    # S: 2 - 0, 0 - 1 ; R: 1 - 0, 0 - 1
    assert rd("gaps") == [-1.0, -1.0, 1.0, 2.0]
    assert rd("records") == [5.0, 4.0]
