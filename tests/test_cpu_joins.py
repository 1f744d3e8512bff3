"""PRO_CPU / NPO_CPU: the library's multithreaded CPU joins (host threads only).

Checked against the golden fixtures (counts from the reference's own RJ_st /
NPO_st), the aggregation closed form, radix-configuration invariance, and the
driver's reference-style -d 2 workload (Zipf keys on R and S over their own
sizes, benchmark.cpp:103, 188, 204) selected by name as in hashing/main.cpp:
306-338.  None of this needs a GPU.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from test_oracle import FIXTURES


@pytest.mark.parametrize("name,R,S,e", FIXTURES, ids=[f[0] for f in FIXTURES])
@pytest.mark.parametrize("fn", ["PRO_CPU", "NPO_CPU"])
def test_cpu_joins_match_golden(aj, fn, name, R, S, e):
    for nt in (1, 3, 8):
        r = getattr(aj, fn)(R, S, nt)
        assert (r.count, r.checksum) == (e["count"], int(e["checksum"])), (fn, name, nt)


@pytest.mark.parametrize("bits,passes", [(0, 0), (4, 1), (8, 2), (13, 2), (16, 2), (9, 1)])
def test_pro_cpu_radix_configurations(aj, oracle, bits, passes):
    R = aj.gen_uniform(300000, 1, 50000, 11)
    S = aj.gen_zipf(400000, 50000, 0.8, 12)
    want = oracle.aggregate(R, S)
    aj.set_option("radix_bits", bits)
    aj.set_option("passes", passes)
    try:
        r = aj.PRO_CPU(R, S, 8)
    finally:
        aj.set_option("radix_bits", 0)
        aj.set_option("passes", 0)
    assert (r.count, r.checksum) == want


@pytest.mark.parametrize("fn", ["PRO_CPU", "NPO_CPU"])
def test_cpu_joins_closed_forms(aj, fn):
    n = 1 << 18
    R = aj.gen_unique(n, 5)
    S = aj.gen_unique(n, 6)
    r = getattr(aj, fn)(R, S, 0)
    # both permutations of 1..n with payload = row id: every key once on each side
    assert r.count == n
    assert r.checksum == ((1 << 32) + 1) * (n * (n - 1) // 2) % (1 << 64)
    empty = np.zeros(0, aj.TUPLE_DTYPE)
    assert getattr(aj, fn)(empty, S, 2).count == 0
    assert getattr(aj, fn)(R, empty, 2).count == 0


def test_cpu_joins_reject_bad_relations(aj):
    import ctypes
    rel = aj._lib.relation_t(None, 5, None)  # tuples missing
    ok = aj._lib.relation_t(None, 0, None)
    assert not aj.lib.PRO_CPU(ctypes.byref(rel), ctypes.byref(ok), 1)
    assert aj.lib.aj_last_status() == aj.AJ_ERR_ARG
    assert not aj.lib.NPO_CPU(ctypes.byref(ok), ctypes.byref(rel), 1)
    assert aj.lib.aj_last_status() == aj.AJ_ERR_ARG


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _driver(aj, *args):
    exe = os.path.join(ROOT, "alliancedb_amd", "ajbench")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = re.findall(r"Results = (\d+)\. checksum = (\d+)", r.stdout)
    assert got
    return [(int(c), int(x)) for c, x in got]


@pytest.mark.parametrize("algo", ["PRO_CPU", "NPO_CPU"])
def test_driver_zipf_both_sides(aj, oracle, algo):
    got = _driver(aj, "-a", algo, "-r", "200000", "-s", "300000", "-d", "2", "-z", "1.0",
                  "-n", "4", "--repeats", "2")
    R = aj.gen_zipf(200000, 200000, 1.0, 12345)
    S = aj.gen_zipf(300000, 300000, 1.0, 54321)
    want = oracle.aggregate(R, S)
    assert got == [want, want]


def test_driver_fk_zipf_and_default(aj, oracle):
    got = _driver(aj, "-a", "PRO_CPU", "-r", "100000", "-s", "150000", "-d", "2", "-z", "1.5",
                  "--fk-zipf", "-n", "3")
    R = aj.gen_unique(100000, 12345)
    assert got == [oracle.aggregate(R, aj.gen_zipf(150000, 100000, 1.5, 54321))]
    got = _driver(aj, "-a", "NPO_CPU", "-r", "100000", "-s", "150000", "-n", "2")
    assert got == [oracle.aggregate(R, aj.gen_uniform(150000, 1, 100000, 54321))]
