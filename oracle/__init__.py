"""CPU oracle for the lazy equi-join path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  Wraps:
  liboracle.so     clean-room C restatement (oracle.c): PRO, NPO, aggregate,
                   sort-merge and brute-force pair lists
  _ref/libref.so   the reference's own RJ_st / NPO_st compiled from its sources
                   (present where the reference tree was available at build time)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AJ_ORACLE_LIB names another build of the oracle (the ASan build of the CPU suite)
ORACLE_PATH = os.environ.get("AJ_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

TUPLE_DTYPE = np.dtype([("payloadID", "<i4"), ("key", "<i4")])


class _Out(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("checksum", ctypes.c_uint64),
                ("npairs", ctypes.c_uint64), ("part_ms", ctypes.c_double),
                ("join_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


@dataclass
class OracleResult:
    count: int
    checksum: int
    pairs: Optional[np.ndarray]
    part_ms: float = 0.0
    join_ms: float = 0.0
    total_ms: float = 0.0


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_PATH)
        P, U64 = ctypes.c_void_p, ctypes.c_uint64
        L.oj_pro.argtypes = [P, U64, P, U64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.POINTER(_Out), P, U64]
        L.oj_npo.argtypes = [P, U64, P, U64, ctypes.c_int, ctypes.POINTER(_Out), P, U64]
        L.oj_aggregate.argtypes = [P, U64, P, U64, ctypes.POINTER(U64), ctypes.POINTER(U64)]
        L.oj_pairs_sortmerge.argtypes = [P, U64, P, U64, P, U64]
        L.oj_pairs_sortmerge.restype = ctypes.c_int64
        L.oj_pairs_bruteforce.argtypes = [P, U64, P, U64, P, U64]
        L.oj_pairs_bruteforce.restype = ctypes.c_int64
        L.oj_sort_u64.argtypes = [P, U64]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_PATH)
        P, U64 = ctypes.c_void_p, ctypes.c_uint64
        L.ref_join.argtypes = [ctypes.c_int, P, U64, P, U64, ctypes.POINTER(ctypes.c_int64), P,
                               U64, ctypes.POINTER(U64)]
        _ref = L
    return _ref


def _c(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=TUPLE_DTYPE)


def _p(a: np.ndarray):
    return a.ctypes.data if a.size else None


def pro(R, S, nthreads=8, radix_bits=8, passes=2, pairs=False) -> OracleResult:
    R, S = _c(R), _c(S)
    out = _Out()
    buf = None
    cap = 0
    if pairs:
        cnt, _ = aggregate(R, S)
        cap = cnt
        buf = np.empty(max(cap, 1), np.uint64)
    rc = lib().oj_pro(_p(R), len(R), _p(S), len(S), nthreads, radix_bits, passes,
                      ctypes.byref(out), None if buf is None else buf.ctypes.data, cap)
    if rc:
        raise RuntimeError(f"oj_pro failed: {rc}")
    return OracleResult(out.count, out.checksum, None if buf is None else buf[:out.npairs],
                        out.part_ms, out.join_ms, out.total_ms)


def npo(R, S, nthreads=8, pairs=False) -> OracleResult:
    R, S = _c(R), _c(S)
    out = _Out()
    buf = None
    cap = 0
    if pairs:
        cnt, _ = aggregate(R, S)
        cap = cnt
        buf = np.empty(max(cap, 1), np.uint64)
    rc = lib().oj_npo(_p(R), len(R), _p(S), len(S), nthreads, ctypes.byref(out),
                      None if buf is None else buf.ctypes.data, cap)
    if rc:
        raise RuntimeError(f"oj_npo failed: {rc}")
    return OracleResult(out.count, out.checksum, None if buf is None else buf[:out.npairs],
                        out.part_ms, out.join_ms, out.total_ms)


def aggregate(R, S):
    """(count, checksum) from key groups alone -- no join algorithm involved."""
    R, S = _c(R), _c(S)
    c, x = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().oj_aggregate(_p(R), len(R), _p(S), len(S), ctypes.byref(c), ctypes.byref(x))
    if rc:
        raise RuntimeError(f"oj_aggregate failed: {rc}")
    return int(c.value), int(x.value)


def pairs_sortmerge(R, S) -> np.ndarray:
    R, S = _c(R), _c(S)
    cnt, _ = aggregate(R, S)
    buf = np.empty(max(cnt, 1), np.uint64)
    n = lib().oj_pairs_sortmerge(_p(R), len(R), _p(S), len(S), buf.ctypes.data, cnt)
    if n < 0:
        raise RuntimeError("oj_pairs_sortmerge failed")
    return buf[:n]


def pairs_bruteforce(R, S) -> np.ndarray:
    R, S = _c(R), _c(S)
    cap = len(R) * len(S)
    buf = np.empty(max(cap, 1), np.uint64)
    n = lib().oj_pairs_bruteforce(_p(R), len(R), _p(S), len(S), buf.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("oj_pairs_bruteforce failed")
    return buf[:n]


def sort_u64(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().oj_sort_u64(a.ctypes.data if a.size else None, len(a))
    return a


def ref_join(algo: str, R, S, s_list: bool = False):
    """The reference's own RJ_st ("RJ_st") or NPO_st ("NPO_st") on copies of R, S.
    Returns (count, sorted list of (S.key << 32 | S.payloadID) per match or None)."""
    R, S = _c(R), _c(S)
    a = {"RJ_st": 0, "NPO_st": 1}[algo]
    cnt = ctypes.c_int64()
    n = ctypes.c_uint64()
    buf = None
    cap = 0
    if s_list:
        cap, _ = aggregate(R, S)
        buf = np.empty(max(cap, 1), np.uint64)
    rc = ref().ref_join(a, _p(R), len(R), _p(S), len(S), ctypes.byref(cnt),
                        None if buf is None else buf.ctypes.data, cap, ctypes.byref(n))
    if rc == -5:
        raise ValueError("the reference's NPO_st is undefined for |R| < 2")
    if rc:
        raise RuntimeError(f"ref_join failed: {rc}")
    lst = None if buf is None else np.sort(buf[:n.value])
    return int(cnt.value), lst


def ref_write_result(res_ptr, path: str) -> None:
    """The reference's own write_result_relation (Out.tbl writer) on a result_t*
    (ctypes pointer or address) whose resultlist holds chained tuple buffers."""
    f = ref().ref_write_result
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    f.restype = ctypes.c_int
    if f(ctypes.cast(res_ptr, ctypes.c_void_p), os.fsencode(path)):
        raise RuntimeError("ref_write_result failed")


def ref_load_relation(path: str, keyby: int, ts_key: int, n: int, partitions: int):
    """The reference's own load_relation on a text file -> (tuples, ts)."""
    f = ref().ref_load_relation
    f.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    tup = np.zeros(max(n, 1), np.uint64)
    ts = np.zeros(max(n, 1), np.uint64)
    rc = f(os.fsencode(path), keyby, ts_key, n, partitions, tup.ctypes.data, ts.ctypes.data)
    if rc:
        raise RuntimeError(f"ref_load_relation failed: {rc}")
    return tup[:n].view(TUPLE_DTYPE), ts[:n]


def load_relation_text(text: str, keyby: int, ts_key: int, n: int, partitions: int):
    """Pure-Python restatement of the reference's read_relation + smooth
    (hashing/utils/generator.cpp:899-1027, 1050-1067) for small inputs.
    Returns (keys int32 array, payloadIDs, ts uint64 array) in placement order."""
    sep = next((c for c in text if c in ",|"), None)
    if sep is None:
        raise ValueError("no separator")
    lines = text.split("\n")
    if text.endswith("\n"):
        lines = lines[:-1]
    if len(lines) < n:
        raise ValueError("too few lines")
    keys, ts = [], []
    for ln in lines[:n]:
        f = ln.split(sep)
        keys.append(int(_c_long_prefix(f[keyby])))
        ts.append(int(float(_c_long_prefix(f[ts_key])) * 2.1 * 1e6) if ts_key else 0)
    if n:  # smooth (generator.cpp:1050-1067)
        run, cur, head = 0, 0, ts[0]
        for k in range(n):
            if ts[k] == head:
                run += 1
                continue
            inner = float((ts[k] - ts[cur]) % 2**64) / float(run)
            for j in range(run):
                ts[cur + j] = int(float(ts[cur + j]) + j * inner)
            head, cur, run = ts[k], k, 1
    per = n // partitions
    fill = [0] * partitions
    out_k = np.zeros(n, np.int32)
    out_p = np.zeros(n, np.int32)
    out_t = np.zeros(n, np.uint64)
    for k in range(n):  # placement (generator.cpp:1011-1024)
        p = k % partitions
        end = n if p == partitions - 1 else per * (p + 1)
        if per * p + fill[p] == end:
            p = partitions - 1
        idx = per * p + fill[p]
        fill[p] += 1
        out_k[idx], out_p[idx], out_t[idx] = keys[k], idx, ts[k]
    return out_k, out_p, out_t


def _c_long_prefix(tok: str) -> str:
    """The prefix strtol(.., 10) parses: white space, sign, decimal digits."""
    s = tok.lstrip(" \t\n\v\f\r")
    i = 1 if s[:1] in "+-" else 0
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    if j == i:
        raise ValueError(f"not a number: {tok!r}")
    return s[:j]


def ref_breakdown_file(path: str, nR: int, nS: int, last_ts: int, overall: int, part: int,
                       build: int, join: int) -> None:
    """The reference's own breakdown_global (lazy joins) written to path."""
    f = ref().ref_breakdown_file
    f.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_long,
                  ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
    f.restype = ctypes.c_int
    if f(os.fsencode(path), nR, nS, last_ts, overall, part, build, join):
        raise RuntimeError("ref_breakdown_file failed")


def timing_records(cycles, sids, s_ts, last_ts: int, sides=None, r_ts=None):
    """Restatement of the reference's merge + sortRecords (timer/t_timer.cpp:
    401-452): per sample, the match time in ms, the latency (unsigned cycle
    arithmetic, then signed) in ms against the sample's own side's timestamps,
    and the disorder gap payloadID - index of the sample among its side's samples;
    each list sorted ascending.  sides[i] = 1 marks an R-side sample (an R tuple's
    probe in the eager joins), default all S."""
    wait = int(last_ts * 2.1 * 1e6)
    times, lat, gaps = [], [], []
    idx = [0, 0]
    for i, (c, s) in enumerate(zip(cycles, sids)):
        c, s = int(c), int(s)
        r = int(sides[i]) if sides is not None else 0
        ts = r_ts if r else s_ts
        times.append(c / (2.1 * 1e6))
        u = (c - (int(ts[s]) if ts is not None else 0) + wait) % 2**64
        lat.append((u - 2**64 if u >= 2**63 else u) / (2.1 * 1e6))
        gaps.append(float(s - idx[r]))
        idx[r] += 1
    return sorted(times), sorted(lat), sorted(gaps)


def timing_text(values, last_ts: int) -> str:
    """dump_timing's lines: std::to_string(element + lastTS) (t_timer.cpp:58-61)."""
    return "".join("%f\n" % (v + last_ts) for v in values)


def ref_shj(R, S) -> int:
    """The reference's own single-thread symmetric hash join (shj,
    helper/localjoiner.cpp:33-64) -> match count.  Needs |R|, |S| >= 2."""
    R, S = _c(R), _c(S)
    f = ref().ref_shj
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_int64)]
    f.restype = ctypes.c_int
    c = ctypes.c_int64()
    if f(_p(R), len(R), _p(S), len(S), ctypes.byref(c)):
        raise ValueError("the reference's SHJ needs |R|, |S| >= 2")
    return int(c.value)


def ref_shj_records(R, S):
    """The reference's own single-thread shj with its materialised records: per match
    the probing (later-arriving) tuple's key << 32 | payloadID, sorted; and the count."""
    R, S = _c(R), _c(S)
    cnt, _ = aggregate(R, S)
    buf = np.empty(max(cnt, 1), np.uint64)
    f = ref().ref_shj_records
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_uint64,
                  ctypes.POINTER(ctypes.c_uint64)]
    f.restype = ctypes.c_int
    c, n = ctypes.c_int64(), ctypes.c_uint64()
    if f(_p(R), len(R), _p(S), len(S), ctypes.byref(c), buf.ctypes.data, cnt, ctypes.byref(n)):
        raise ValueError("the reference's SHJ needs |R|, |S| >= 2")
    return int(c.value), np.sort(buf[:n.value])


def shj_records_by_side(R, S, batch: int):
    """Sorted SHJ records under the step rule, split by the probing side: arrivals in
    steps of `batch` tuples per side, R's part of a step before S's; a pair (r_i, s_j)
    of equal keys is found by the probe of its later tuple, which is r_i when
    floor(i/batch) > floor(j/batch) and s_j otherwise, and the record is that tuple
    (key << 32 | payloadID).  For batch = 1 this is the reference's alternating
    R[0], S[0], R[1], ... order (localjoiner.cpp:44-53).  Returns (records of R
    tuples' probes, records of S tuples' probes); every pair is expanded, so the
    inputs must be small enough for their pair count."""
    R, S = _c(R), _c(S)
    rk, sk = R["key"].astype(np.int64), S["key"].astype(np.int64)
    rv = (R["key"].astype(np.uint32).astype(np.uint64) << np.uint64(32)) | \
        R["payloadID"].astype(np.uint32).astype(np.uint64)
    sv = (S["key"].astype(np.uint32).astype(np.uint64) << np.uint64(32)) | \
        S["payloadID"].astype(np.uint32).astype(np.uint64)
    ro = np.argsort(rk, kind="stable")
    rks = rk[ro]
    lo = np.searchsorted(rks, sk, side="left")
    cnt = np.searchsorted(rks, sk, side="right") - lo
    total = int(cnt.sum())
    sj = np.repeat(np.arange(len(S), dtype=np.int64), cnt)
    offs = np.arange(total, dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ri = ro[np.repeat(lo, cnt) + offs]
    r_later = (ri // batch) > (sj // batch)
    return np.sort(rv[ri[r_later]]), np.sort(sv[sj[~r_later]])


def shj_records(R, S, batch: int) -> np.ndarray:
    """All SHJ records under the step rule (shj_records_by_side), sorted."""
    r, s = shj_records_by_side(R, S, batch)
    return np.sort(np.concatenate([r, s]))


def shj_progress(R, S, batch: int) -> np.ndarray:
    """Matches found after each eager step of `batch` tuples per side (R part
    first): a pair (r_i, s_j) belongs to step max(i // batch, j // batch) --
    the step in which its later tuple arrives.  From the sort-merge pair list;
    needs payloadID = row index on both sides."""
    R, S = _c(R), _c(S)
    assert np.array_equal(R["payloadID"], np.arange(len(R)))
    assert np.array_equal(S["payloadID"], np.arange(len(S)))
    n = max(len(R), len(S))
    steps = (n + batch - 1) // batch
    pairs = pairs_sortmerge(R, S)
    i = (pairs >> np.uint64(32)).astype(np.int64)
    j = (pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)
    k = np.maximum(i // batch, j // batch)
    return np.cumsum(np.bincount(k, minlength=steps)[:steps]).astype(np.uint64)
