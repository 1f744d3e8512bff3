"""CPU: the autotuner's table (C ABI): entries per size bucket, JSON save/load
round trip, rejection of malformed files and invalid entries.  Tuning itself
(timing candidates) needs a GPU and is in tests/test_gpu_join.py."""
import pytest


@pytest.fixture(autouse=True)
def _clean_table(aj):
    aj.tune_clear()
    yield
    aj.tune_clear()


def _entry(**kw):
    e = dict(log2_r=27, log2_s=27, radix_bits=14, passes=2, cap=8704, part_block=512,
             part_chunk=0, device_ms=3.5)
    e.update(kw)
    return e


def test_bucket_lookup(aj):
    assert aj.tune_get(1 << 27, 1 << 27) is None
    aj.tune_put(**_entry())
    got = aj.tune_get((1 << 27) + 5, (1 << 28) - 1)  # same bucket: floor(log2)
    assert got["radix_bits"] == 14 and got["cap"] == 8704
    assert aj.tune_get(1 << 26, 1 << 27) is None


def test_save_load_round_trip(aj, tmp_path):
    aj.tune_put(**_entry())
    aj.tune_put(**_entry(log2_r=20, log2_s=24, radix_bits=8, passes=1, cap=12288,
                         part_block=256, part_chunk=65536, device_ms=0.25, one_pass_hist=1))
    path = str(tmp_path / "tuning.json")
    aj.tune_save(path)
    import json
    doc = json.load(open(path))  # the file is plain JSON
    assert len(doc["entries"]) == 2
    aj.tune_clear()
    assert aj.tune_get(1 << 20, 1 << 24) is None
    aj.tune_load(path)
    got = aj.tune_get(1 << 20, 1 << 24)
    assert (got["radix_bits"], got["passes"], got["cap"], got["part_block"], got["part_chunk"],
            got["one_pass_hist"]) == (8, 1, 12288, 256, 65536, 1)
    assert abs(got["device_ms"] - 0.25) < 1e-4
    assert aj.tune_get(1 << 27, 1 << 27)["one_pass_hist"] == 0


def test_load_table_without_one_pass_choice(aj, tmp_path):
    """A table saved before entries recorded the one-pass path loads with the
    staged pass allowed (one_pass_hist 0)."""
    old = tmp_path / "old.json"
    old.write_text('{"format": "alliancedb_amd PRO tuning", "version": 1, "entries": [\n'
                   '{"log2_r": 24, "log2_s": 28, "radix_bits": 11, "passes": 1, "cap": 8704, '
                   '"part_block": 512, "part_chunk": 0, "device_ms": 2.5000}\n]}\n')
    aj.tune_load(str(old))
    got = aj.tune_get(1 << 24, 1 << 28)
    assert (got["radix_bits"], got["passes"], got["one_pass_hist"]) == (11, 1, 0)


def test_invalid_entries_and_files(aj, tmp_path):
    with pytest.raises(aj.AjError):
        aj.tune_put(**_entry(cap=1000))
    with pytest.raises(aj.AjError):
        aj.tune_put(**_entry(part_block=300))
    with pytest.raises(aj.AjError):
        aj.tune_put(**_entry(one_pass_hist=2))
    bad = tmp_path / "bad.json"
    bad.write_text('{"format": "alliancedb_amd PRO tuning", "version": 1, "entries": [\n'
                   '{"log2_r": 27, "log2_s": oops}\n]}\n')
    with pytest.raises(aj.AjError):
        aj.tune_load(str(bad))
    other = tmp_path / "other.json"
    other.write_text('{"entries": []}\n')
    with pytest.raises(aj.AjError):
        aj.tune_load(str(other))
    with pytest.raises(aj.AjError):
        aj.tune_load(str(tmp_path / "missing.json"))
