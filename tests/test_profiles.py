"""CPU: the committed traffic summary that bench.py reports as roofline.traffic holds
byte counts per kernel launch (not GB), consistent with the algorithmic bytes of the
2^27 x 2^27 PRO join (2^28 tuples: 8 B per tuple for the histogram and the join, 16 B
for each scatter), and bench.py reads them back."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_summary_is_in_bytes():
    with open(os.path.join(ROOT, "profiles", "pro_2e27_traffic.json")) as f:
        d = json.load(f)
    alg = {"k_hist2d (pass 1)": 8, "k_scatter_pipe (pass 1)": 16,
           "k_scatter_p2 (pass 2)": 16, "k_join": 8}
    got = d["hbm_bytes_per_launch"]
    assert set(got) == set(alg)
    for k, per_tuple in alg.items():
        b = got[k]
        assert isinstance(b, int)
        # measured traffic within [algorithmic, 1.5 x algorithmic]
        assert per_tuple * (1 << 28) <= b * 1.01 and b <= 1.5 * per_tuple * (1 << 28), k


def test_bench_reads_traffic_summary():
    import sys
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "pro_2e27_traffic.json")) as f:
        d = json.load(f)
    k = "k_scatter_pipe (pass 1)"
    assert bench.measured_traffic(k, d["workload"]) == d["hbm_bytes_per_launch"][k]
    assert bench.measured_traffic(k, "another workload") is None
