"""ctypes binding of libaj.so, the C ABI declared in include/alliancedb/aj.h.

The library is the product: every call here goes straight to the HIP
implementation.  If libaj.so is missing the import fails loudly; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AJ_LIB names another build of the same library (the host-ASan build of the
# CPU suite); by default the in-tree libaj.so
LIB_PATH = os.environ.get("AJ_LIB") or os.path.join(_HERE, "libaj.so")

# tuple_t of hashing/utils/types.h:67-71: {int32 payloadID @0, int32 key @4}
TUPLE_DTYPE = np.dtype([("payloadID", "<i4"), ("key", "<i4")])
# aj_stream_tuple_t: {uint32 key, uint32 payload, uint32 ts}
STREAM_DTYPE = np.dtype([("key", "<u4"), ("payload", "<u4"), ("ts", "<u4")])

AJ_OK = 0
AJ_ERR_ARG = -1
AJ_ERR_NOMEM = -2
AJ_ERR_HIP = -3
AJ_ERR_CAPACITY = -4
AJ_ERR_STATE = -5
AJ_ERR_NODEV = -6
AJ_ALGO_PRO = 0
# aj_join_out_t.flags (enum aj_out_flags)
AJ_OUT_STAGED = 1
AJ_OUT_FALLBACK = 2
AJ_OUT_HEAVY = 4
AJ_ACCESS_LOAD, AJ_ACCESS_CAS, AJ_ACCESS_LOAD_CAS = 0, 1, 2
AJ_ALGO_NPO = 1
AJ_ALGO_SHJ = 2


class relation_payload_t(ctypes.Structure):
    _fields_ = [("ts", ctypes.POINTER(ctypes.c_uint64)), ("num_tuples", ctypes.c_uint64)]


class relation_t(ctypes.Structure):
    _fields_ = [("tuples", ctypes.c_void_p), ("num_tuples", ctypes.c_uint64),
                ("payload", ctypes.c_void_p)]


class result_t(ctypes.Structure):
    _fields_ = [("totalresults", ctypes.c_int64), ("resultlist", ctypes.c_void_p),
                ("nthreads", ctypes.c_int)]


class threadresult_t(ctypes.Structure):
    _fields_ = [("nresults", ctypes.c_int64), ("results", ctypes.c_void_p),
                ("threadid", ctypes.c_uint32)]


class aj_tuplebuffer_t(ctypes.Structure):
    pass


aj_tuplebuffer_t._fields_ = [("tuples", ctypes.POINTER(ctypes.c_uint64)),
                             ("next", ctypes.POINTER(aj_tuplebuffer_t))]


class aj_chainedtuplebuffer_t(ctypes.Structure):
    _fields_ = [("buf", ctypes.POINTER(aj_tuplebuffer_t)),
                ("readcursor", ctypes.POINTER(aj_tuplebuffer_t)),
                ("writecursor", ctypes.POINTER(aj_tuplebuffer_t)),
                ("writepos", ctypes.c_uint32), ("readpos", ctypes.c_uint32),
                ("readlen", ctypes.c_uint32), ("numbufs", ctypes.c_uint32)]


CHAINEDBUFF_NUMTUPLESPERBUF = 1024 * 1024


class aj_timing_summary_t(ctypes.Structure):
    _fields_ = [("progress_ms", ctypes.c_double * 5), ("latency95_ms", ctypes.c_double),
                ("latency99_ms", ctypes.c_double), ("samples", ctypes.c_uint64)]


class aj_result_t(ctypes.Structure):
    _fields_ = [("base", result_t), ("checksum", ctypes.c_uint64), ("npairs", ctypes.c_uint64),
                ("pairs", ctypes.POINTER(ctypes.c_uint64)), ("status", ctypes.c_int),
                ("device_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("ntimed", ctypes.c_uint64), ("timed", ctypes.POINTER(ctypes.c_uint64)),
                ("clock_hz", ctypes.c_double), ("algo", ctypes.c_int),
                ("phase_ms", ctypes.c_float * 12), ("ntimed_r", ctypes.c_uint64)]


class aj_join_opts_t(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int), ("radix_bits", ctypes.c_int), ("passes", ctypes.c_int),
                ("cap", ctypes.c_int), ("split", ctypes.c_uint32),
                ("pairs", ctypes.c_void_p), ("pairs_capacity", ctypes.c_uint64),
                ("records", ctypes.c_int), ("batch", ctypes.c_uint64),
                ("progress", ctypes.c_void_p)]


AJ_NPHASES = 12
PHASE_NAMES = ["plan1", "hist1", "scan1", "scatter1", "plan2", "hist2", "scan2", "scatter2",
               "tasks", "join", "npo_build", "npo_probe"]


class aj_join_out_t(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("checksum", ctypes.c_uint64),
                ("npairs", ctypes.c_uint64), ("device_ms", ctypes.c_float),
                ("phase_ms", ctypes.c_float * AJ_NPHASES), ("nrecords_r", ctypes.c_uint64),
                ("flags", ctypes.c_uint32), ("heavy_ms", ctypes.c_float),
                ("light", ctypes.c_uint64)]

    def phases(self) -> dict:
        return {n: float(self.phase_ms[i]) for i, n in enumerate(PHASE_NAMES)}


class aj_batch_opts_t(ctypes.Structure):
    _fields_ = [("cap", ctypes.c_int), ("split", ctypes.c_uint32), ("target", ctypes.c_uint32),
                ("chunk", ctypes.c_uint32), ("no_graph", ctypes.c_int)]


class aj_server_opts_t(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("slots", ctypes.c_int),
                ("max_batch_windows", ctypes.c_int), ("max_batch_tuples", ctypes.c_uint64),
                ("max_window_tuples", ctypes.c_uint64), ("cap", ctypes.c_int)]


class aj_window_t(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("dR", ctypes.c_void_p), ("dS", ctypes.c_void_p),
                ("nR", ctypes.c_uint64), ("nS", ctypes.c_uint64)]


class aj_server_result_t(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("checksum", ctypes.c_uint64), ("status", ctypes.c_int),
                ("batch_windows", ctypes.c_int), ("close_us", ctypes.c_double),
                ("start_us", ctypes.c_double), ("done_us", ctypes.c_double)]


class aj_server_stats_t(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "windows_done", "batches", "tuples_done", "window_bytes", "window_bytes_used",
        "window_bytes_peak", "workspace_bytes", "alloc_waits", "pending", "in_flight")]


class aj_tune_entry_t(ctypes.Structure):
    _fields_ = [("log2_r", ctypes.c_int), ("log2_s", ctypes.c_int), ("radix_bits", ctypes.c_int),
                ("passes", ctypes.c_int), ("cap", ctypes.c_int), ("part_block", ctypes.c_int),
                ("part_chunk", ctypes.c_uint32), ("device_ms", ctypes.c_float),
                ("one_pass_hist", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class AjError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"aj status {status}: {msg}")
        self.status = status


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C alliancedb_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    c = ctypes
    P = c.c_void_p
    sig = {
        "aj_last_error": (c.c_char_p, []),
        "aj_last_status": (c.c_int, []),
        "PRO": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t), c.c_int]),
        "NPO": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t), c.c_int]),
        "SHJ_JM_NP": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t),
                                            c.c_int]),
        "SHJ_JBCR_NP": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t),
                                              c.c_int]),
        "SHJ_st": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t), c.c_int]),
        "PRO_CPU": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t), c.c_int]),
        "NPO_CPU": (c.POINTER(result_t), [c.POINTER(relation_t), c.POINTER(relation_t), c.c_int]),
        "aj_result_ext": (c.POINTER(aj_result_t), [c.POINTER(result_t)]),
        "aj_free_result": (None, [c.POINTER(result_t)]),
        "aj_write_timing_files": (c.c_int, [c.POINTER(result_t), c.POINTER(relation_t),
                                            c.POINTER(relation_t), c.c_char_p, c.c_char_p,
                                            c.c_int, c.c_long, c.c_int,
                                            c.POINTER(aj_timing_summary_t)]),
        "aj_load_relation": (c.c_int, [c.POINTER(relation_t), c.POINTER(relation_payload_t),
                                       c.c_int32, c.c_int32, c.c_char_p, c.c_uint64,
                                       c.c_uint32]),
        "aj_write_result_file": (c.c_int, [c.POINTER(result_t), c.c_char_p]),
        "aj_set_option": (c.c_int, [c.c_char_p, c.c_int64]),
        "aj_get_option": (c.c_int64, [c.c_char_p]),
        "aj_ctx_create": (c.c_int, [c.c_int, c.POINTER(P)]),
        "aj_ctx_destroy": (c.c_int, [P]),
        "aj_join_dev": (c.c_int, [P, P, c.c_uint64, P, c.c_uint64, c.POINTER(aj_join_opts_t),
                                  c.POINTER(aj_join_out_t), P]),
        "aj_join_dev_noalloc": (c.c_int, [P, c.c_uint64, P, c.c_uint64,
                                          c.POINTER(aj_join_opts_t), c.POINTER(aj_join_out_t),
                                          P]),
        "aj_device_alloc": (c.c_int, [c.POINTER(P), c.c_uint64]),
        "aj_device_free": (c.c_int, [P]),
        "aj_copy_h2d": (c.c_int, [P, P, c.c_uint64]),
        "aj_copy_d2h": (c.c_int, [P, P, c.c_uint64]),
        "aj_device_sync": (c.c_int, []),
        "aj_copy_d2d": (c.c_int, [P, P, c.c_uint64]),
        "aj_access_kernel": (c.c_int, [c.c_int, P, c.c_uint64, c.c_uint64,
                                       c.POINTER(c.c_float)]),
        "aj_batch_plan_create": (c.c_int, [P, c.c_uint64, c.c_uint64, c.c_int, c.c_uint64,
                                           c.POINTER(aj_batch_opts_t), c.POINTER(P)]),
        "aj_batch_plan_destroy": (c.c_int, [P]),
        "aj_batch_run": (c.c_int, [P, P, c.c_uint64, P, c.c_uint64, P, c.c_int, P, P, P]),
        "aj_gen_window_streams": (c.c_int, [P, c.c_uint64, c.c_int, P, c.c_double, c.c_uint32,
                                            P, c.c_uint64, P, c.POINTER(c.c_uint64)]),
        "aj_gen_stream_tuples": (c.c_int, [P, P, c.c_uint64, c.c_uint64, c.c_uint32, c.c_double,
                                           c.c_uint32, P]),
        "aj_server_create": (c.c_int, [P, c.POINTER(aj_server_opts_t), c.POINTER(P)]),
        "aj_server_destroy": (c.c_int, [P]),
        "aj_server_window_alloc": (c.c_int, [P, c.c_uint64, c.c_uint64, c.c_int64,
                                             c.POINTER(aj_window_t)]),
        "aj_server_window_release": (c.c_int, [P, c.POINTER(aj_window_t)]),
        "aj_server_window_fill": (c.c_int, [P, c.POINTER(aj_window_t), c.c_uint64, c.c_uint32,
                                            c.c_double]),
        "aj_server_window_upload": (c.c_int, [P, c.POINTER(aj_window_t), P, P]),
        "aj_server_submit": (c.c_int, [P, c.POINTER(aj_window_t), c.c_uint64, c.c_double]),
        "aj_server_poll": (c.c_int, [P, c.POINTER(aj_server_result_t), c.c_int, c.c_int64,
                                     c.POINTER(c.c_int)]),
        "aj_server_drain": (c.c_int, [P]),
        "aj_server_now_us": (c.c_double, [P]),
        "aj_server_stats": (c.c_int, [P, c.POINTER(aj_server_stats_t)]),
        "aj_tune_batch": (c.c_int, [P, P, c.c_uint64, P, c.c_uint64, P, c.c_int, c.c_uint64,
                                    c.c_int, c.POINTER(aj_batch_opts_t), c.POINTER(c.c_float)]),
        "aj_tune_pro": (c.c_int, [P, P, c.c_uint64, P, c.c_uint64, c.c_int,
                                  c.POINTER(aj_tune_entry_t)]),
        "aj_tune_put": (c.c_int, [c.POINTER(aj_tune_entry_t)]),
        "aj_tune_get": (c.c_int, [c.c_uint64, c.c_uint64, c.POINTER(aj_tune_entry_t)]),
        "aj_tune_clear": (c.c_int, []),
        "aj_tune_save": (c.c_int, [c.c_char_p]),
        "aj_tune_load": (c.c_int, [c.c_char_p]),
        "aj_gen_unique": (c.c_int, [P, c.c_uint64, c.c_uint64]),
        "aj_gen_uniform": (c.c_int, [P, c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint64]),
        "aj_gen_zipf": (c.c_int, [P, c.c_uint64, c.c_uint32, c.c_double, c.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()
_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]
_libc.free.restype = None


def last_error() -> str:
    return lib.aj_last_error().decode(errors="replace")


def check(status: int) -> None:
    if status != AJ_OK:
        raise AjError(status, last_error())


def set_option(name: str, value: int) -> None:
    check(lib.aj_set_option(name.encode(), int(value)))


def get_option(name: str) -> int:
    return int(lib.aj_get_option(name.encode()))


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


def as_tuples(keys: np.ndarray, payloads: Optional[np.ndarray] = None) -> np.ndarray:
    t = np.empty(len(keys), dtype=TUPLE_DTYPE)
    t["key"] = keys
    t["payloadID"] = np.arange(len(keys)) if payloads is None else payloads
    return t


@dataclass
class JoinResult:
    count: int
    checksum: int
    pairs: Optional[np.ndarray]
    device_ms: float
    total_ms: float = 0.0
    # option materialize = 2: per resultlist entry, its <S.key, S.payloadID>
    # records as uint64 (key << 32 | payloadID) in append order (oldest buffer
    # first); the reference's backward read gives them in reverse
    records: Optional[list] = None
    # per resultlist entry: (threadid, numbufs, writepos)
    chains: Optional[list] = None
    # option timing = 1: (device clock ticks since the join started << 32 | S.payloadID)
    # per match, the clock's rate and the per-phase device times (enum aj_phase)
    timed: Optional[np.ndarray] = None
    clock_hz: float = 0.0
    phase_ms: Optional[list] = None
    # SHJ: the last timed_r timed records are R tuples' probes (R.payloadID)
    timed_r: int = 0


def _read_chain(cb: aj_chainedtuplebuffer_t, n: int) -> np.ndarray:
    """All n records of a chain in append order: the head buffer (newest, cb.buf)
    holds writepos records, each older buffer a full CHAINEDBUFF_NUMTUPLESPERBUF
    (hashing/utils/tuple_buffer.h:93-109)."""
    out = []
    b = cb.buf
    take = cb.writepos
    left = n
    while left > 0:
        if not b:
            raise AjError(AJ_ERR_STATE, "record chain shorter than its count")
        k = min(take, left)
        if k:
            out.append(np.ctypeslib.as_array(b.contents.tuples, shape=(k,)).copy())
        left -= k
        b = b.contents.next
        take = CHAINEDBUFF_NUMTUPLESPERBUF
    return np.concatenate(out[::-1]) if out else np.zeros(0, np.uint64)


def _dropin(fn, R: np.ndarray, S: np.ndarray, nthreads: int,
            result_file: Optional[str] = None) -> JoinResult:
    R = np.ascontiguousarray(R, dtype=TUPLE_DTYPE)
    S = np.ascontiguousarray(S, dtype=TUPLE_DTYPE)
    rr = relation_t(R.ctypes.data if len(R) else None, len(R), None)
    rs = relation_t(S.ctypes.data if len(S) else None, len(S), None)
    res = fn(ctypes.byref(rr), ctypes.byref(rs), int(nthreads))
    if not res:
        raise AjError(lib.aj_last_status(), last_error())
    try:
        ext = lib.aj_result_ext(res).contents
        pairs = None
        if ext.npairs:
            pairs = np.ctypeslib.as_array(ext.pairs, shape=(ext.npairs,)).copy()
        records = chains = None
        if ext.base.resultlist:
            tl = ctypes.cast(ext.base.resultlist, ctypes.POINTER(threadresult_t))
            records, chains = [], []
            for i in range(ext.base.nthreads):
                cb = ctypes.cast(tl[i].results, ctypes.POINTER(aj_chainedtuplebuffer_t)).contents
                records.append(_read_chain(cb, int(tl[i].nresults)))
                chains.append((int(tl[i].threadid), int(cb.numbufs), int(cb.writepos)))
        if result_file is not None:
            check(lib.aj_write_result_file(res, os.fsencode(result_file)))
        timed = None
        if ext.ntimed:
            timed = np.ctypeslib.as_array(ext.timed, shape=(ext.ntimed,)).copy()
        return JoinResult(int(ext.base.totalresults), int(ext.checksum), pairs,
                          float(ext.device_ms), float(ext.total_ms), records, chains, timed,
                          float(ext.clock_hz), list(ext.phase_ms), int(ext.ntimed_r))
    finally:
        lib.aj_free_result(res)


def load_relation(path: str, num_tuples: int, keyby: int = 0, ts_key: int = 0,
                  partitions: int = 1):
    """A relation from the reference's text format (aj_load_relation; see aj.h).
    Returns (tuples as TUPLE_DTYPE, timestamps as uint64)."""
    rel = relation_t(None, 0, None)
    pl = relation_payload_t(None, 0)
    check(lib.aj_load_relation(ctypes.byref(rel), ctypes.byref(pl), int(keyby), int(ts_key),
                               os.fsencode(path), int(num_tuples), int(partitions)))
    n = int(rel.num_tuples)
    try:
        tup = np.ctypeslib.as_array(ctypes.cast(rel.tuples, ctypes.POINTER(ctypes.c_uint64)),
                                    shape=(max(n, 1),))[:n].copy().view(TUPLE_DTYPE)
        ts = np.ctypeslib.as_array(pl.ts, shape=(max(n, 1),))[:n].copy()
    finally:
        _libc.free(rel.tuples)
        _libc.free(ctypes.cast(pl.ts, ctypes.c_void_p))
    return tup, ts


def PRO(R: np.ndarray, S: np.ndarray, nthreads: int = 1,
        result_file: Optional[str] = None) -> JoinResult:
    """Drop-in PRO (radix join) on host relations; see aj.h.  With result_file
    the result's records are written there as the reference's Out.tbl."""
    return _dropin(lib.PRO, R, S, nthreads, result_file)


def NPO(R: np.ndarray, S: np.ndarray, nthreads: int = 1,
        result_file: Optional[str] = None) -> JoinResult:
    """Drop-in NPO (no-partitioning join) on host relations; see aj.h."""
    return _dropin(lib.NPO, R, S, nthreads, result_file)


def SHJ_JM_NP(R: np.ndarray, S: np.ndarray, nthreads: int = 1) -> JoinResult:
    """Drop-in eager symmetric hash join (join-matrix name); see aj.h."""
    return _dropin(lib.SHJ_JM_NP, R, S, nthreads)


def SHJ_JBCR_NP(R: np.ndarray, S: np.ndarray, nthreads: int = 1) -> JoinResult:
    """Drop-in eager symmetric hash join (join-biclique ContRand name); see aj.h."""
    return _dropin(lib.SHJ_JBCR_NP, R, S, nthreads)


def SHJ_st(R: np.ndarray, S: np.ndarray, nthreads: int = 1) -> JoinResult:
    """Drop-in eager symmetric hash join (single-thread name); see aj.h."""
    return _dropin(lib.SHJ_st, R, S, nthreads)


def PRO_CPU(R: np.ndarray, S: np.ndarray, nthreads: int = 0) -> JoinResult:
    """The multithreaded CPU PRO baseline (host threads only; see aj.h)."""
    return _dropin(lib.PRO_CPU, R, S, nthreads)


def NPO_CPU(R: np.ndarray, S: np.ndarray, nthreads: int = 0) -> JoinResult:
    """The multithreaded CPU NPO baseline (host threads only; see aj.h)."""
    return _dropin(lib.NPO_CPU, R, S, nthreads)


class Context:
    """An aj_ctx_t: device scratch reused across joins on one GPU."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib.aj_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            lib.aj_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def join(self, dR: int, nR: int, dS: int, nS: int, algo: int = AJ_ALGO_PRO,
             radix_bits: int = 0, passes: int = 0, cap: int = 0, split: int = 0,
             pairs_ptr: int = 0, pairs_capacity: int = 0, stream: int = 0, records: int = 0,
             batch: int = 0, progress_ptr: int = 0) -> aj_join_out_t:
        """Join device-resident tuple arrays (device pointers as ints)."""
        o = aj_join_opts_t(algo, radix_bits, passes, cap, split, pairs_ptr or None,
                           pairs_capacity, records, batch, progress_ptr or None)
        out = aj_join_out_t()
        check(lib.aj_join_dev(self.handle, dR or None, nR, dS or None, nS, ctypes.byref(o),
                              ctypes.byref(out), stream or None))
        return out


class BatchPlan:
    """aj_batch_plan_t: batched tumbling-window joins captured in one hipGraph."""

    def __init__(self, ctx: "Context", max_R: int, max_S: int, max_windows: int,
                 max_window_tuples: int = 0, cap: int = 0, split: int = 0, target: int = 0,
                 chunk: int = 0, graph: bool = True):
        self.ctx = ctx
        o = aj_batch_opts_t(cap, split, target, chunk, 0 if graph else 1)
        h = ctypes.c_void_p()
        check(lib.aj_batch_plan_create(ctx.handle, max_R, max_S, max_windows,
                                       max_window_tuples, ctypes.byref(o), ctypes.byref(h)))
        self.handle = h

    def run(self, dR: int, nR: int, dS: int, nS: int, d_bounds: int, nwin: int, d_count: int,
            d_checksum: int, stream: int = 0) -> None:
        check(lib.aj_batch_run(self.handle, dR or None, nR, dS or None, nS, d_bounds, nwin,
                               d_count, d_checksum, stream or None))

    def close(self) -> None:
        if self.handle:
            lib.aj_batch_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def tune_batch(ctx: "Context", dR: int, nR: int, dS: int, nS: int, d_bounds: int, nwin: int,
               max_window_tuples: int = 0, reps: int = 2) -> tuple:
    """aj_tune_batch: the fastest batch geometry for these resident windows, as
    (dict of BatchPlan keyword arguments, mean launch ms)."""
    o = aj_batch_opts_t()
    ms = ctypes.c_float(0)
    check(lib.aj_tune_batch(ctx.handle, dR, nR, dS, nS, d_bounds, nwin, max_window_tuples,
                            reps, ctypes.byref(o), ctypes.byref(ms)))
    return {"cap": o.cap, "target": o.target, "chunk": o.chunk}, float(ms.value)


def gen_window_streams(ctx: "Context", seed: int, sizes: np.ndarray, rate_per_us: float,
                       dup: int, d_out: int, capacity: int, d_bounds: int) -> int:
    """Device stream of 12-byte tuples for tumbling windows of the given sizes."""
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    n = ctypes.c_uint64()
    check(lib.aj_gen_window_streams(ctx.handle, seed, len(sizes), sizes.ctypes.data,
                                    float(rate_per_us), dup, d_out, capacity, d_bounds,
                                    ctypes.byref(n)))
    return int(n.value)


def tune_pro(ctx: "Context", dR: int, nR: int, dS: int, nS: int, reps: int = 3) -> dict:
    """Autotunes the PRO geometry for the size bucket of (nR, nS) on device data."""
    e = aj_tune_entry_t()
    check(lib.aj_tune_pro(ctx.handle, dR, nR, dS, nS, reps, ctypes.byref(e)))
    return e.as_dict()


def tune_get(nR: int, nS: int):
    e = aj_tune_entry_t()
    if lib.aj_tune_get(nR, nS, ctypes.byref(e)) != 0:
        return None
    return e.as_dict()


def tune_put(**fields) -> None:
    e = aj_tune_entry_t(**{k: v for k, v in fields.items()})
    check(lib.aj_tune_put(ctypes.byref(e)))


def tune_save(path: str) -> None:
    check(lib.aj_tune_save(path.encode()))


def tune_load(path: str) -> None:
    check(lib.aj_tune_load(path.encode()))


def tune_clear() -> None:
    check(lib.aj_tune_clear())


def gen_stream_tuples(ctx: "Context", d_out: int, n: int, seed: int, key_domain: int,
                      theta: float = 0.0, payload0: int = 0) -> None:
    """Device generator of n 12-byte window tuples (blocks until done)."""
    check(lib.aj_gen_stream_tuples(ctx.handle, d_out, n, seed, key_domain, float(theta),
                                   payload0, None))


class Server:
    """aj_server_t: admission by HBM budget, continuous batching over HIP streams."""

    def __init__(self, ctx: "Context", hbm_budget: int = 0, slots: int = 0,
                 max_batch_windows: int = 0, max_batch_tuples: int = 0,
                 max_window_tuples: int = 0, cap: int = 0):
        o = aj_server_opts_t(hbm_budget, slots, max_batch_windows, max_batch_tuples,
                             max_window_tuples, cap)
        h = ctypes.c_void_p()
        check(lib.aj_server_create(ctx.handle, ctypes.byref(o), ctypes.byref(h)))
        self.handle = h

    def alloc(self, nR: int, nS: int, timeout_us: int = -1) -> aj_window_t:
        w = aj_window_t()
        check(lib.aj_server_window_alloc(self.handle, nR, nS, timeout_us, ctypes.byref(w)))
        return w

    def release(self, w: aj_window_t) -> None:
        check(lib.aj_server_window_release(self.handle, ctypes.byref(w)))

    def fill(self, w: aj_window_t, seed: int, key_domain: int, theta_s: float = 0.0) -> None:
        check(lib.aj_server_window_fill(self.handle, ctypes.byref(w), seed, key_domain,
                                        float(theta_s)))

    def upload(self, w: aj_window_t, R: np.ndarray, S: np.ndarray) -> None:
        R = np.ascontiguousarray(R, dtype=STREAM_DTYPE)
        S = np.ascontiguousarray(S, dtype=STREAM_DTYPE)
        assert len(R) == w.nR and len(S) == w.nS
        check(lib.aj_server_window_upload(self.handle, ctypes.byref(w), R.ctypes.data,
                                          S.ctypes.data))

    def submit(self, w: aj_window_t, tag: int = 0, close_us: float = -1.0) -> None:
        check(lib.aj_server_submit(self.handle, ctypes.byref(w), tag, float(close_us)))

    def poll(self, max_results: int = 256, timeout_us: int = 0) -> list:
        buf = (aj_server_result_t * max_results)()
        n = ctypes.c_int()
        check(lib.aj_server_poll(self.handle, buf, max_results, timeout_us, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def drain(self) -> None:
        check(lib.aj_server_drain(self.handle))

    def now_us(self) -> float:
        return lib.aj_server_now_us(self.handle)

    def stats(self) -> dict:
        st = aj_server_stats_t()
        check(lib.aj_server_stats(self.handle, ctypes.byref(st)))
        return {n: int(getattr(st, n)) for n, _ in aj_server_stats_t._fields_}

    def close(self) -> None:
        if self.handle:
            lib.aj_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceArray:
    """A device allocation made through the C ABI (aj_device_alloc)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        check(lib.aj_device_alloc(ctypes.byref(p), int(nbytes)))
        self.ptr = int(p.value or 0)
        self.nbytes = int(nbytes)

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            check(lib.aj_copy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib.aj_copy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib.aj_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync() -> None:
    check(lib.aj_device_sync())


def gen_unique(n: int, seed: int) -> np.ndarray:
    t = np.empty(n, dtype=TUPLE_DTYPE)
    check(lib.aj_gen_unique(_ptr(t), n, seed))
    return t


def gen_uniform(n: int, lo: int, hi: int, seed: int) -> np.ndarray:
    t = np.empty(n, dtype=TUPLE_DTYPE)
    check(lib.aj_gen_uniform(_ptr(t), n, lo, hi, seed))
    return t


def gen_zipf(n: int, nkeys: int, theta: float, seed: int) -> np.ndarray:
    t = np.empty(n, dtype=TUPLE_DTYPE)
    check(lib.aj_gen_zipf(_ptr(t), n, nkeys, float(theta), seed))
    return t


def access_rate(kind: int, table_ptr: int, slots: int, n: int) -> float:
    """G operations/s of one aj_access_kernel launch (see aj.h)."""
    ms = ctypes.c_float(0)
    check(lib.aj_access_kernel(kind, table_ptr, slots, n, ctypes.byref(ms)))
    return n / (ms.value * 1e-3) / 1e9
