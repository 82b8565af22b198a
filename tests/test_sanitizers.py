"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs (CPU): the oracle
under tests/sanitize/oracle_driver.cpp, libqsolver's host-only code (error
handling, Mesh set-up and teardown, CSV text I/O) under
tests/sanitize/host_driver.cpp, and the column-job placement of the fused
column kernels under tests/sanitize/place_driver.cpp, all built by
tests/sanitize/Makefile with the sanitizers aborting on their first finding
(leak detection on)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None or not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("sanitizer toolchain absent")
    subprocess.check_call(["make", "-s", "-j4", "-C", SAN])
    return os.path.join(SAN, "_build")


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)


def test_oracle_under_sanitizers(built):
    r = run(os.path.join(built, "oracle_san"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle sanitizer run: ok" in r.stdout


def test_host_code_under_sanitizers(built, tmp_path):
    r = run(os.path.join(built, "host_san"), str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host sanitizer run: ok" in r.stdout


def test_column_job_placement(built):
    """col_job_place maps the jobs of every fused side, tile width and stage
    kind one to one onto (field, tile), with the 1024^2 order of an XCD range
    (r1, vorticity, r2, the partial tile; final stage: vorticity jobs ending
    the first round)."""
    r = run(os.path.join(built, "place_san"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "placement run: ok" in r.stdout

