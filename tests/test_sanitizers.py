"""The library's threaded host code (graph generators, weights, file formats,
relabeling) and the oracle's threaded routines, built on the CPU with
AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer
(tests/sanitize/Makefile, g++/gcc only) and run on small graphs by
tests/sanitize/san_driver.cpp, which also checks their results against each
other. Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("g++/gcc not available")
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-j4", "-C", HERE, "OUT=%s" % out], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_under_sanitizers(built, kind, tmp_path):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([os.path.join(str(built), "san_" + kind), str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize ok" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
