"""CPU: relations from the reference's text format (aj_load_relation).

Checked against the reference's own load_relation (hashing/utils/generator.cpp:
645-666, 899-1027, 1050-1067, compiled into oracle/_ref/libref.so) on the same
files -- both separators, several columns, runs of equal timestamps (the
smoothing), partition counts that do and do not divide |R|, surplus lines,
CRLF line ends and padded fields -- and against the oracle's pure-Python
restatement where the reference is undefined (tsKey = 0) or absent."""
import numpy as np
import pytest


def _write(tmp_path, name, rows, sep, crlf=False, pad=False, extra=0):
    end = "\r\n" if crlf else "\n"
    lines = []
    for r in rows:
        f = [(" " + str(x)) if pad else str(x) for x in r]
        lines.append(sep.join(f))
    for k in range(extra):
        lines.append(sep.join(["1"] * len(rows[0])))
    p = tmp_path / name
    p.write_text(end.join(lines) + end)
    return str(p)


def _rows(rng, n, ncols, key_col, ts_col, run_len=4, back=False):
    rows = []
    t = int(rng.integers(1000, 2000))
    for i in range(n):
        if i % run_len == 0:  # runs of equal timestamps
            t += int(rng.integers(-40 if back else 1, 50))
        r = [int(rng.integers(-1000, 1000)) for _ in range(ncols)]
        r[key_col] = int(rng.integers(-2**31, 2**31))
        r[ts_col] = t
        rows.append(r)
    return rows


CASES = [
    # n, ncols, keyby, tsKey, partitions, sep, crlf, pad, extra lines, run length,
    # timestamps that also step back (the smoothing's unsigned difference wraps)
    (1000, 2, 0, 1, 1, ",", False, False, 0, 4, False),
    (1000, 4, 0, 3, 8, "|", False, False, 7, 3, False),
    (997, 3, 2, 1, 8, ",", True, False, 0, 5, False),
    (513, 3, 1, 2, 3, "|", False, True, 2, 1, False),
    (5, 2, 0, 1, 8, ",", False, False, 0, 2, False),  # more partitions than tuples
    (4096, 2, 1, 0 + 1, 16, ",", False, False, 0, 64, False),
    (600, 2, 0, 1, 4, ",", False, False, 0, 3, True),
]


@pytest.mark.parametrize("case", CASES)
def test_loader_matches_reference(aj, oracle, tmp_path, case):
    n, ncols, keyby, ts_key, parts, sep, crlf, pad, extra, run, back = case
    rng = np.random.default_rng(n * 31 + ncols)
    rows = _rows(rng, n, ncols, keyby, ts_key, run, back)
    path = _write(tmp_path, "rel.txt", rows, sep, crlf, pad, extra)
    tup, ts = aj.load_relation(path, n, keyby, ts_key, parts)
    # the oracle's restatement
    k, p, t = oracle.load_relation_text(open(path, newline="").read(), keyby, ts_key, n, parts)
    assert np.array_equal(tup["key"], k) and np.array_equal(tup["payloadID"], p)
    assert np.array_equal(ts, t)
    assert np.array_equal(tup["payloadID"], np.arange(n))
    if not oracle.ref_available():
        pytest.skip("oracle/_ref/libref.so not built")
    rt, rts = oracle.ref_load_relation(path, keyby, ts_key, n, parts)
    assert np.array_equal(tup.view(np.uint64), rt.view(np.uint64))
    assert np.array_equal(ts, rts)


def test_smoothing_known_answer(aj, tmp_path):
    """ts column 1,1,1,4 -> cycles 2.1e6 * (1, 2, 3, 4): the run of three equal
    values is spread evenly towards the next one; the trailing run stays."""
    path = _write(tmp_path, "s.txt", [[7, 1], [8, 1], [9, 1], [10, 4], [11, 4]], ",")
    tup, ts = aj.load_relation(path, 5, 0, 1, 1)
    assert list(tup["key"]) == [7, 8, 9, 10, 11]
    assert list(ts) == [2100000, 4200000, 6300000, 8400000, 8400000]


def test_round_robin_placement(aj, tmp_path):
    """7 lines over 3 partitions of 2, 2 and 3 slots: lines 0,3 -> [0,2), 1,4 -> [2,4),
    2,5 -> [4,7) and line 6 (partition 0, full) -> the last partition."""
    path = _write(tmp_path, "p.txt", [[100 + i, i + 1] for i in range(7)], "|")
    tup, _ = aj.load_relation(path, 7, 0, 1, 3)
    assert list(tup["key"]) == [100, 103, 101, 104, 102, 105, 106]
    assert list(tup["payloadID"]) == list(range(7))


def test_no_timestamp_column(aj, oracle, tmp_path):
    path = _write(tmp_path, "k.txt", [[5, 1], [6, 2], [-7, 3]], ",")
    tup, ts = aj.load_relation(path, 3, 0, 0, 2)
    k, p, t = oracle.load_relation_text(open(path).read(), 0, 0, 3, 2)
    assert np.array_equal(tup["key"], k) and not ts.any()


def test_loader_errors(aj, tmp_path):
    def bad(text, n=2, keyby=0, ts_key=1, parts=1):
        p = tmp_path / "bad.txt"
        p.write_text(text)
        with pytest.raises(aj.AjError) as e:
            aj.load_relation(str(p), n, keyby, ts_key, parts)
        assert e.value.status == aj.AJ_ERR_ARG
    bad("1 2\n3 4\n")                   # no separator
    bad("1,2\n", n=2)                   # too few lines
    bad("1,2\nx,3\n")                   # key not a number
    bad("1,2\n3000000000,3\n")          # key outside int32
    bad("1,2\n3\n")                     # missing ts field
    bad("1,2\n3,-4\n")                  # negative timestamp
    # a step back after a run of huge equal timestamps: the smoothed value passes 2^64
    bad("1,8000000000000\n2,8000000000000\n3,8000000000000\n4,5000000000000\n", n=4)
    bad("1,2\n3,4\n", parts=0)
    with pytest.raises(aj.AjError):
        aj.load_relation(str(tmp_path / "missing.txt"), 1)


def test_loader_empty_and_zero(aj, tmp_path):
    path = _write(tmp_path, "e.txt", [[1, 1]], ",")
    tup, ts = aj.load_relation(path, 0, 0, 1, 4)
    assert len(tup) == 0 and len(ts) == 0
