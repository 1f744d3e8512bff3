"""CPU: the serving loop's window memory and admission controller
(alliancedb_amd/csrc/arena.h) against a brute-force model, with requests that
must queue (not fail) while the budget is taken.  Built with g++ and run under
ThreadSanitizer when the toolchain has it."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "arena_test.cpp")
INC = os.path.join(ROOT, "alliancedb_amd", "csrc")


def _build(tmp_path, extra):
    exe = str(tmp_path / ("arena_test" + ("_tsan" if extra else "")))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I", INC, SRC, "-o", exe] + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    return exe, r


def test_arena_and_admission(tmp_path):
    exe, r = _build(tmp_path, [])
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "arena_test OK" in out.stdout


def test_admission_under_thread_sanitizer(tmp_path):
    exe, r = _build(tmp_path, ["-fsanitize=thread"])
    if r.returncode != 0:
        pytest.skip("g++ cannot build with -fsanitize=thread here: " + r.stderr[-200:])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    if out.returncode != 0 and "FATAL: ThreadSanitizer" in out.stderr:
        pytest.skip("ThreadSanitizer cannot run in this environment")
    assert out.returncode == 0, out.stderr[-2000:]
    assert "WARNING: ThreadSanitizer" not in out.stderr
