"""CPU: the C-ABI library loads, exports what the headers declare, rejects bad
input with status codes (never exit()), and its generators are deterministic
and have the stated distributions.  No compute call needs a GPU here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest
import xxhash

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "alliancedb", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src,
                             flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "return")):
                names.add(m.group(1))
    return sorted(names)


def test_header_symbols_exported(aj):
    names = declared_functions()
    assert {"PRO", "NPO", "aj_join_dev", "aj_last_error"} <= set(names)
    lib = ctypes.CDLL(aj.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts(aj):
    from alliancedb_amd import _lib
    assert aj.TUPLE_DTYPE.itemsize == 8 and aj.TUPLE_DTYPE.fields["key"][1] == 4
    assert ctypes.sizeof(_lib.relation_t) == 24
    assert ctypes.sizeof(_lib.result_t) == 24
    # aj_result_t starts with result_t so free(result) frees the whole block
    assert _lib.aj_result_t.base.offset == 0
    assert aj.STREAM_DTYPE.itemsize == 12


def test_join_out_mirror_matches_header(tmp_path):
    """The ctypes mirror of aj_join_out_t has the C header's size and offsets."""
    import shutil
    import subprocess
    from alliancedb_amd import _lib
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    T = _lib.aj_join_out_t
    names = [f[0] for f in T._fields_]
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "alliancedb/aj.h"\n'
           'int main(void) { printf("%zu", sizeof(aj_join_out_t));' +
           "".join(f' printf(" %zu", offsetof(aj_join_out_t, {n}));' for n in names) +
           " return 0; }\n")
    (tmp_path / "layout.c").write_text(src)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(tmp_path / "layout.c")], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == [ctypes.sizeof(T)] + [getattr(T, n).offset for n in names]


def test_null_relation_is_an_error_not_exit(aj):
    from alliancedb_amd import _lib
    res = _lib.lib.PRO(None, None, 4)
    assert not res
    assert _lib.lib.aj_last_status() == aj.AJ_ERR_ARG
    assert "NULL" in aj.last_error()
    bad = _lib.relation_t(None, 10, None)  # tuples missing
    ok = _lib.relation_t(None, 0, None)
    assert not _lib.lib.NPO(ctypes.byref(ok), ctypes.byref(bad), 1)
    assert _lib.lib.aj_last_status() == aj.AJ_ERR_ARG


def test_options(aj):
    with pytest.raises(aj.AjError):
        aj.set_option("no_such_option", 1)
    with pytest.raises(aj.AjError):
        aj.set_option("radix_bits", 99)
    aj.set_option("radix_bits", 12)
    assert aj.get_option("radix_bits") == 12
    aj.set_option("radix_bits", 0)


def test_no_gpu_fails_loudly(aj):
    """Without a GPU the product reports AJ_ERR_NODEV instead of computing on the CPU."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    R = aj.gen_unique(64, 1)
    with pytest.raises(aj.AjError) as e:
        aj.PRO(R, R, 1)
    assert e.value.status == aj.AJ_ERR_NODEV




def _gen(aj, kind, n, seed):
    if kind == "unique":
        return aj.gen_unique(n, seed)
    if kind == "uniform":
        return aj.gen_uniform(n, 1, n, seed)
    return aj.gen_zipf(n, 1 << 16, 1.0, seed)


def test_generator_determinism_fixture(aj):
    """Seed -> byte stream is fixed (fixture in tests/golden/generators.json)."""
    import json
    path = os.path.join(ROOT, "tests", "golden", "generators.json")
    with open(path) as f:
        want = json.load(f)
    for key, digest in want.items():
        kind, n, seed = key.split(":")
        got = xxhash.xxh64(_gen(aj, kind, int(n), int(seed)).tobytes()).hexdigest()
        assert got == digest, key


def test_unique_is_a_permutation(aj):
    for n in (1, 2, 3, 1000, 1 << 17):
        t = aj.gen_unique(n, 99)
        assert np.array_equal(np.sort(t["key"]), np.arange(1, n + 1))
        assert np.array_equal(t["payloadID"], np.arange(n))
    a, b = aj.gen_unique(4096, 1), aj.gen_unique(4096, 2)
    assert not np.array_equal(a["key"], b["key"])


def test_uniform_range_and_mean(aj):
    t = aj.gen_uniform(1 << 20, 5, 104, 3)
    k = t["key"]
    assert k.min() == 5 and k.max() == 104
    counts = np.bincount(k - 5, minlength=100)
    # chi-square against uniform, 99 dof: well inside 99.9% quantile (~148)
    exp = len(k) / 100
    assert ((counts - exp) ** 2 / exp).sum() < 148


@pytest.mark.parametrize("theta", [0.5, 1.0, 1.5])
def test_zipf_head_frequency(aj, theta):
    nkeys, n = 1 << 14, 1 << 20
    t = aj.gen_zipf(n, nkeys, theta, 17)
    assert t["key"].min() >= 1 and t["key"].max() <= nkeys
    counts = np.sort(np.bincount(t["key"], minlength=nkeys + 1))[::-1]
    H = (1.0 / np.arange(1, nkeys + 1) ** theta).sum()
    for rank in (1, 2, 10):
        want = n / (rank ** theta * H)
        assert abs(counts[rank - 1] - want) < 6 * np.sqrt(want) + 0.01 * want, (rank, theta)


def test_zipf_theta0_is_uniform(aj):
    t = aj.gen_zipf(1 << 18, 256, 0.0, 1)
    counts = np.bincount(t["key"], minlength=257)[1:]
    exp = (1 << 18) / 256
    assert ((counts - exp) ** 2 / exp).sum() < 350


def test_driver_registry_by_name():
    """ajbench selects joins by name like the reference's -a (main.cpp:521-539)."""
    import subprocess
    exe = os.path.join(ROOT, "alliancedb_amd", "ajbench")
    r = subprocess.run([exe, "-a", "NO_SUCH_JOIN"], capture_output=True, text=True)
    assert r.returncode != 0 and "does not exist" in r.stderr
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "PRO|NPO" in r.stdout
