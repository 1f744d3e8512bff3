"""GPU: joins over relations loaded from the reference's text format.

Loaded relations (aj_load_relation) are joined through the drop-in PRO/NPO and
through the driver's -R/-S options, and checked against the oracle.  The
reference's application datasets (stock, rovio, YSB, DEBS) are not shipped;
when a directory holding them is named by AJ_DATASETS, their join counts are
checked against the known answers the reference's scripts record
(hashing/scripts/benchmark.sh:382-436)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_rel(path, keys, ts, sep=","):
    with open(path, "w") as f:
        for k, t in zip(keys, ts):
            f.write(f"{int(k)}{sep}{int(t)}\n")


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_loaded_relations_join(aj, oracle, algo, tmp_path):
    rng = np.random.default_rng(99)
    nR, nS = 60000, 80000
    kr = rng.integers(0, 20000, nR)
    ks = rng.integers(0, 20000, nS)
    ts_r = np.sort(rng.integers(0, 1000, nR))
    ts_s = np.sort(rng.integers(0, 1000, nS))
    pr, ps = str(tmp_path / "r.txt"), str(tmp_path / "s.txt")
    _write_rel(pr, kr, ts_r, ",")
    _write_rel(ps, ks, ts_s, "|")
    print(f"case {algo} loaded {nR}x{nS}", flush=True)
    R, _ = aj.load_relation(pr, nR, 0, 1, 8)
    S, _ = aj.load_relation(ps, nS, 0, 1, 8)
    got = getattr(aj, algo)(R, S, 8)
    assert (got.count, got.checksum) == oracle.aggregate(R, S)
    # the driver on the same files
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    r = subprocess.run([exe, "-a", algo, "-r", str(nR), "-s", str(nS), "-n", "8", "-R", pr,
                        "-J", "0", "-L", "1", "-S", ps, "-K", "0", "-M", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.findall(r"Results = (\d+)\. checksum = (\d+)", r.stdout)
    assert m and (int(m[0][0]), int(m[0][1])) == (got.count, got.checksum)


# name: (R file, S file, |R|, |S|, key columns, ts columns, matches)
DATASETS = {
    "stock": ("stock/cj_1000ms_1t.txt", "stock/sb_1000ms_1t.txt", 60527, 77227, (0, 0), (1, 1),
              15598112),
    "rovio": ("rovio/1000ms_1t.txt", "rovio/1000ms_1t.txt", 2873604, 2873604, (0, 0), (3, 3),
              87856849382),
    "ysb": ("YSB/campaigns_id.txt", "YSB/ad_events.txt", 1000, 10000000, (0, 0), (0, 1),
            10000000),
    "debs": ("DEBS/posts_key32_partitioned.csv", "DEBS/comments_key32_partitioned.csv", 1000000,
             1000000, (0, 0), (0, 0), 251033140),
}


@pytest.mark.parametrize("name", sorted(DATASETS))
def test_dataset_known_answers(aj, name):
    root = os.environ.get("AJ_DATASETS")
    fr, fs, nR, nS, keys, tss, want = DATASETS[name]
    if not root or not os.path.exists(os.path.join(root, fr)):
        pytest.skip("dataset not supplied (set AJ_DATASETS to the reference's datasets folder)")
    R, _ = aj.load_relation(os.path.join(root, fr), nR, keys[0], tss[0], 1)
    S, _ = aj.load_relation(os.path.join(root, fs), nS, keys[1], tss[1], 1)
    for algo in ("PRO", "NPO"):
        assert getattr(aj, algo)(R, S, 1).count == want, algo


@pytest.mark.parametrize("algo", ["PRO", "NPO"])
def test_timed_records(aj, oracle, algo):
    """Option timing = 1: one timed record per match whose S.payloadIDs are the
    matches' S rows (from the oracle's pair list), times inside the join, and
    per-phase device times."""
    R = aj.gen_unique(1 << 18, 5)
    S = aj.gen_uniform(300000, 1, 1 << 18, 6)
    aj.set_option("timing", 1)
    try:
        print(f"case {algo} timed", flush=True)
        got = getattr(aj, algo)(R, S, 4)
    finally:
        aj.set_option("timing", 0)
    pairs = oracle.pairs_sortmerge(R, S)
    assert got.count == len(pairs) and len(got.timed) == got.count
    sids = np.sort(got.timed & np.uint64(0xFFFFFFFF))
    assert np.array_equal(sids, np.sort(pairs & np.uint64(0xFFFFFFFF)))
    assert 1e6 <= got.clock_hz <= 1e10
    ms = (got.timed >> np.uint64(32)).astype(np.float64) / got.clock_hz * 1e3
    assert ms.min() >= 0 and ms.max() < 1000
    ph = got.phase_ms
    if algo == "PRO":
        assert sum(ph[0:8]) > 0 and ph[9] > 0
    else:
        assert ph[10] > 0 and ph[11] > 0


def test_driver_timing_files(aj, tmp_path):
    """ajbench --timing on loaded relations writes the reference's five result
    files: counts, one line per match in timestamps / latency / gaps, and the
    7-line breakdown."""
    rng = np.random.default_rng(3)
    nR, nS = 20000, 30000
    pr, ps = str(tmp_path / "r.txt"), str(tmp_path / "s.txt")
    _write_rel(pr, rng.integers(0, 5000, nR), np.sort(rng.integers(0, 1000, nR)))
    _write_rel(ps, rng.integers(0, 5000, nS), np.sort(rng.integers(0, 1000, nS)))
    exe = os.path.join(os.path.dirname(aj.LIB_PATH), "ajbench")
    d = tmp_path / "exp"
    r = subprocess.run([exe, "-a", "PRO", "-r", str(nR), "-s", str(nS), "-n", "4", "-R", pr,
                        "-L", "1", "-S", ps, "-M", "1", "-w", "1000", "--timing", str(d),
                        "--exp-id", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "95th latency" in r.stdout
    count = int(re.findall(r"Results = (\d+)", r.stdout)[0])
    res = d / "results"
    assert (res / "records" / "PRJ_3.txt").read_text().split() == [str(nR + nS), str(count)]
    for sub in ("timestamps", "latency", "gaps"):
        v = np.array((res / sub / "PRJ_3.txt").read_text().split(), float)
        assert len(v) == count and np.all(np.diff(v) >= 0), sub
    t = np.array((res / "timestamps" / "PRJ_3.txt").read_text().split(), float)
    assert t.min() >= 1000  # lastTS = the window
    assert len((res / "breakdown" / "PRJ_3.txt").read_text().split()) == 7
