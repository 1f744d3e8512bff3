"""CPU: the committed traffic summary that bench.py reports as roofline.traffic holds
byte counts per kernel launch (not GB), consistent with the algorithmic bytes of the
2^27 x 2^27 PRO join (2^28 tuples: the staged passes launch once per super-chunk of
2^25 tuples and move 16 B per tuple, the join reads 8 B per tuple in one launch),
and bench.py reads them back."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_summary_is_in_bytes():
    with open(os.path.join(ROOT, "profiles", "pro_2e27_traffic.json")) as f:
        d = json.load(f)
    alg = {"k_stage1 (pass 1)": 16 * (1 << 25), "k_stage2 (pass 2)": 16 * (1 << 25),
           "k_join": 8 * (1 << 28)}
    got = d["hbm_bytes_per_launch"]
    assert set(got) == set(alg)
    for k, per_launch in alg.items():
        b = got[k]
        assert isinstance(b, int)
        # measured traffic within [algorithmic, 1.5 x algorithmic]
        assert per_launch <= b * 1.01 and b <= 1.5 * per_launch, k


def test_bench_reads_traffic_summary():
    import sys
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "pro_2e27_traffic.json")) as f:
        d = json.load(f)
    k = "k_stage2 (pass 2)"
    assert bench.measured_traffic(k, d["workload"]) == d["hbm_bytes_per_launch"][k]
    assert bench.measured_traffic(k, "another workload") is None


def test_hash_join_traffic_summaries():
    """NPO's probe moves about one 128-byte line per S tuple (random slot loads)
    plus the tuple itself; its region build reads R and writes every table slot
    (8 + 32 B per R tuple); an SHJ step launch moves at least a line per tuple."""
    n = 1 << 27
    with open(os.path.join(ROOT, "profiles", "npo_2e27_traffic.json")) as f:
        npo = json.load(f)["hbm_bytes_per_launch"]
    assert 128 * n <= npo["k_npo_probe"] <= 160 * n
    assert 40 * n <= npo["k_npo_region"] <= 1.1 * 40 * n
    with open(os.path.join(ROOT, "profiles", "shj_2e27_traffic.json")) as f:
        d = json.load(f)
    per_join = d["hbm_bytes_per_launch"]["k_shj_step (all eager steps)"]
    assert d["launches_per_join"]["k_shj_step (all eager steps)"] == 512
    assert 128 * 2 * n <= per_join <= 400 * 2 * n
