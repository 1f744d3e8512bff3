"""The CPU-side tests again, against host-instrumented builds.

`make -C alliancedb_amd/csrc asan` builds the library with AddressSanitizer and
UBSan on its host code (device code as usual, -Xarch_host before every
-fsanitize), `make -C oracle asan` the oracle with the same sanitizers and the
same clang, so both share one runtime; the listed test files then run in a
child pytest that loads these builds (AJ_LIB, AJ_ORACLE_LIB) under an
LD_PRELOAD of that runtime.  Any sanitizer report fails the run (UBSan does
not recover).  Covered host code: the drop-in argument checks and status codes,
options, generators, the text loader, the result-file and progressive-file
writers, the autotuner's table, the CPU baseline joins and the oracle.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNTIMES = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
FILES = ["test_cpu_joins.py", "test_loader.py", "test_timing_files.py", "test_result_format.py",
         "test_abi.py", "test_tune.py", "test_oracle.py"]


@pytest.mark.skipif(not RUNTIMES, reason="clang sanitizer runtime not found")
def test_host_code_under_asan_and_ubsan(tmp_path):
    for d in ("alliancedb_amd/csrc", "oracle"):
        r = subprocess.run(["make", "-C", os.path.join(ROOT, d), "-j8", "asan"],
                           capture_output=True, text=True, timeout=900)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = tmp_path / "sanitizer"
    env = dict(os.environ,
               LD_PRELOAD=RUNTIMES[-1],
               ASAN_OPTIONS=f"detect_leaks=0:log_path={log}",
               UBSAN_OPTIONS="print_stacktrace=1",
               AJ_LIB=os.path.join(ROOT, "_build", "asan", "libaj.so"),
               AJ_ORACLE_LIB=os.path.join(ROOT, "_build", "asan", "liboracle.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", *[os.path.join(ROOT, "tests", f) for f in FILES]],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    reports = "".join(open(p).read() for p in glob.glob(str(log) + "*"))
    assert r.returncode == 0 and not reports, (r.stdout[-3000:], reports[-3000:])
