"""The reference-side binding shown in the integration guide (section 1) builds.

The adapter block and the registry rows are taken verbatim from the guide,
compiled against the reference's own headers (hashing/benchmark.h and
everything it includes: utils/types.h, joins/parallel_radix_join.h, ...) with
this repository's include/ on the path, in both include orders, and linked
against libaj.so.  The adapters must bind the C entry points (unmangled PRO,
NPO, SHJ_JM_NP), not the reference's C++ overloads that take a param_t
(parallel_radix_join.h:37, no_partitioning_join.h:36, onlinejoins.h:33).
Run, the adapter either joins (a GPU is present) or takes the reference's
failure path: a message and exit(EXIT_FAILURE) (parallel_radix_join.cpp:58-65).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/hashing"
LIB = os.path.join(ROOT, "alliancedb_amd", "libaj.so")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(REF) and os.path.exists(LIB)),
    reason="needs the reference tree and a built libaj.so")


def _section1_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text.split("## 1.", 1)[1].split("\n## 2.", 1)[0]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, re.S)
    assert len(blocks) >= 2, "the integration guide's section 1 holds the adapter and the registry rows"
    return blocks[0], blocks[1]


MAIN = r"""
static algo_t gpu_algos[] = {
@ROWS@
    {"NULL", NULL}};

int main(int argc, char **argv) {
    for (int i = 0; gpu_algos[i].joinAlgo; ++i) printf("%s\n", gpu_algos[i].name);
    if (argc < 2) return 0;
    tuple_t rt[4] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}};
    tuple_t st[4] = {{0, 4}, {1, 3}, {2, 2}, {3, 9}};
    relation_t R = {rt, 4, NULL}, S = {st, 4, NULL};
    param_t p;
    p.nthreads = 2;
    for (int i = 0; gpu_algos[i].joinAlgo; ++i) {
        if (strcmp(gpu_algos[i].name, argv[1]) == 0) {
            result_t *res = gpu_algos[i].joinAlgo(&R, &S, p);
            printf("matches %lld\n", (long long)res->totalresults);
            free(res);
            return 0;
        }
    }
    return 2;
}
"""


def _build(tmp_path, order):
    adapter, registry = _section1_blocks()
    rows = "\n".join(l for l in registry.splitlines() if l.strip().startswith("{"))
    lines = adapter.splitlines()
    inc = [l for l in lines if l.startswith("#include")]
    assert inc == ['#include "benchmark.h"', "#include <alliancedb/aj.h>"], inc
    if order == "aj_first":
        body = [l for l in lines if not l.startswith("#include")]
        lines = [inc[1], inc[0]] + body
    src = tmp_path / f"gpu_joins_{order}.cpp"
    src.write_text("\n".join(lines) + "\n" + MAIN.replace("@ROWS@", rows))
    obj = tmp_path / f"gpu_joins_{order}.o"
    exe = tmp_path / f"gpu_joins_{order}"
    flags = ["-std=c++11", "-msse4.2", "-mavx2", "-w", "-I", REF, "-I",
             os.path.join(ROOT, "include")]
    r = subprocess.run(["g++", *flags, "-c", str(src), "-o", str(obj)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run(["g++", str(obj), "-o", str(exe), "-L", os.path.dirname(LIB), "-laj",
                        "-Wl,-rpath," + os.path.dirname(LIB), "-lpthread"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return obj, exe


@pytest.mark.parametrize("order", ["reference_first", "aj_first"])
def test_adapter_compiles_and_links(tmp_path, order):
    obj, exe = _build(tmp_path, order)
    undef = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True,
                           check=True).stdout.split()
    for sym in ("PRO", "NPO", "SHJ_JM_NP", "aj_set_option", "aj_last_error"):
        assert sym in undef, f"adapter does not bind the C symbol {sym}"
    names = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert names == ["PRO_GPU", "NPO_GPU", "SHJ_JM_NP_GPU"]


def test_adapter_runs_or_fails_like_the_reference(tmp_path):
    _, exe = _build(tmp_path, "reference_first")
    r = subprocess.run([str(exe), "PRO_GPU"], capture_output=True, text=True)
    if r.returncode == 0:  # a GPU joined the four tuples: keys 2, 3, 4 match
        assert "matches 3" in r.stdout
    else:  # no GPU: exit(EXIT_FAILURE) after the message, as the reference fails
        assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
        assert "[ERROR] PRO_GPU" in r.stdout
