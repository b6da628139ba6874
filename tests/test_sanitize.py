"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code (a memory and undefined-behaviour
check of the CPU side): the C oracle over its entry points (`make -C oracle sanitize`, driver
oracle/sanitize_check.c) and the library's host protocol tables (tables.cpp, driver
tests/sanitize/tables_check.cpp).  Every finding aborts the driver with a non-zero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _need(tool):
    if not shutil.which(tool):
        pytest.skip(f"{tool} not available")


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    _need("gcc")
    _need("make")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True, env=ENV, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize_check ok" in r.stdout
    assert "runtime error" not in r.stderr


def test_host_tables_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    _need("g++")
    exe = str(tmp_path / "tables_check")
    src = [os.path.join(ROOT, "mrsim-acq_amd", "csrc", "tables.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "tables_check.cpp")]
    b = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-Wall", *src, "-o", exe], capture_output=True, text=True,
                       timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "tables_check ok" in r.stdout
    assert "runtime error" not in r.stderr
