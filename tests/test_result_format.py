"""CPU: the reference-format result lists (option materialize = 2) and their
Out.tbl writer, on result_t blocks built here in host memory (no GPU).

aj_write_result_file is checked byte for byte against the reference's own
write_result_relation (hashing/utils/tuple_buffer.h:155-236, compiled into
oracle/_ref/libref.so) on the same chains: several thread entries, an empty
entry, a chain of two 1M-tuple buffers and negative / extreme values."""
import ctypes

import numpy as np
import pytest

from alliancedb_amd import _lib

PER = _lib.CHAINEDBUFF_NUMTUPLESPERBUF


class _Keep:
    """Holds the numpy buffers and ctypes objects a hand-built result points at."""

    def __init__(self):
        self.objs = []

    def add(self, o):
        self.objs.append(o)
        return o


def _chain(keep, records):
    """A chained buffer holding records (uint64 key << 32 | payloadID) in append
    order, laid out as the reference's cb_next_writepos leaves it."""
    nbuf = max(1, -(-len(records) // PER))
    head = ctypes.POINTER(_lib.aj_tuplebuffer_t)()
    for k in range(nbuf):
        t = keep.add(np.zeros(PER, np.uint64))
        part = records[k * PER:(k + 1) * PER]
        t[:len(part)] = part
        b = keep.add(_lib.aj_tuplebuffer_t(t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           head))
        head = ctypes.pointer(b)
    cb = keep.add(_lib.aj_chainedtuplebuffer_t(head, head, head,
                                               len(records) - (nbuf - 1) * PER, 0, 0, nbuf))
    return cb


def _result(keep, lists):
    tl = keep.add((_lib.threadresult_t * len(lists))())
    for i, rec in enumerate(lists):
        tl[i].nresults = len(rec)
        tl[i].results = ctypes.addressof(_chain(keep, rec))
        tl[i].threadid = i
    res = keep.add(_lib.result_t(sum(len(r) for r in lists), ctypes.addressof(tl), len(lists)))
    return res


def _rec(keys, pays):
    k = np.asarray(keys, np.int64).astype(np.int32).view(np.uint32).astype(np.uint64)
    p = np.asarray(pays, np.int64).astype(np.int32).view(np.uint32).astype(np.uint64)
    return (k << np.uint64(32)) | p


def _expected_text(lists):
    """Restated from tuple_buffer.h:214-233: per entry, newest record first."""
    out = []
    for rec in lists:
        for v in rec[::-1]:
            key = np.int64(np.uint64(v) >> np.uint64(32)).astype(np.int32)
            pay = np.int64(np.uint64(v) & np.uint64(0xFFFFFFFF)).astype(np.int32)
            out.append(f"{int(key)} {int(pay)}\n")
    return "".join(out)


def _small_lists():
    return [_rec([5, -1, 2**31 - 1, -2**31, 0], [0, 7, -2**31, 2**31 - 1, -1]),
            _rec([], []),
            _rec([42, 43], [1000000, 3])]


def test_chain_struct_layouts():
    assert ctypes.sizeof(_lib.threadresult_t) == 24
    assert ctypes.sizeof(_lib.aj_tuplebuffer_t) == 16
    assert ctypes.sizeof(_lib.aj_chainedtuplebuffer_t) == 40
    assert _lib.aj_chainedtuplebuffer_t.writepos.offset == 24
    assert _lib.aj_chainedtuplebuffer_t.numbufs.offset == 36


def test_write_result_file_small(tmp_path):
    keep = _Keep()
    lists = _small_lists()
    res = _result(keep, lists)
    path = tmp_path / "Out.tbl"
    _lib.check(_lib.lib.aj_write_result_file(ctypes.byref(res), str(path).encode()))
    assert path.read_text() == _expected_text(lists)


def test_write_result_file_matches_reference_writer(oracle, tmp_path):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref/libref.so not built")
    rng = np.random.default_rng(7)
    n0 = PER + 1234  # two buffers in entry 0
    big = _rec(rng.integers(-2**31, 2**31, n0), rng.integers(-2**31, 2**31, n0))
    lists = [big, _rec([], []), _rec([1, 2, 3], [-4, -5, -6]), _rec(np.arange(PER), np.arange(PER))]
    keep = _Keep()
    res = _result(keep, lists)
    ours, theirs = tmp_path / "ours.tbl", tmp_path / "ref.tbl"
    _lib.check(_lib.lib.aj_write_result_file(ctypes.byref(res), str(ours).encode()))
    oracle.ref_write_result(ctypes.byref(res), str(theirs))
    a, b = ours.read_bytes(), theirs.read_bytes()
    assert len(a) == len(b) and a == b
    assert a.count(b"\n") == sum(len(r) for r in lists)


def test_read_chain_append_order():
    keep = _Keep()
    n = PER + 5
    rec = np.arange(n, dtype=np.uint64) * np.uint64(3)
    cb = _chain(keep, rec)
    assert (cb.numbufs, cb.writepos) == (2, 5)
    assert np.array_equal(_lib._read_chain(cb, n), rec)


def test_write_result_file_errors(aj, tmp_path):
    assert _lib.lib.aj_write_result_file(None, b"x") == aj.AJ_ERR_ARG
    # counts matches but carries no record lists
    res = _lib.result_t(3, None, 1)
    assert _lib.lib.aj_write_result_file(ctypes.byref(res), str(tmp_path / "y").encode()) == \
        aj.AJ_ERR_STATE
    # an entry whose chain holds fewer records than it counts
    keep = _Keep()
    lists = [_rec([1, 2], [3, 4])]
    r2 = _result(keep, lists)
    ctypes.cast(r2.resultlist, ctypes.POINTER(_lib.threadresult_t))[0].nresults = 3
    assert _lib.lib.aj_write_result_file(ctypes.byref(r2), str(tmp_path / "z").encode()) == \
        aj.AJ_ERR_STATE
    _lib.lib.aj_free_result(None)  # no-op


def test_materialize_option_range(aj):
    old = aj.get_option("materialize")
    try:
        aj.set_option("materialize", 2)
        assert aj.get_option("materialize") == 2
        with pytest.raises(aj.AjError):
            aj.set_option("materialize", 3)
    finally:
        aj.set_option("materialize", old)
