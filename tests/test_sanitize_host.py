"""The library's host code (proxies, GCT/GMT/CLS loaders) and the oracle under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/cpp/sanitize_host.cpp,
built with g++/gcc by tests/cpp/sanitize.mk).  CPU only: the
engine entry points that host code calls are answered by the oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "_build", "sanitize_host")


@pytest.fixture(scope="module")
def sanitized():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++/make not available")
    subprocess.run(["make", "-s", "-C", CPP, "-f", "sanitize.mk"], check=True)
    return BIN


def test_host_code_clean_under_asan_ubsan(sanitized, tmp_path):
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sanitized, os.path.join(HERE, "golden"), str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized host checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr


def test_sanitizers_are_linked(sanitized):
    # the binary carries the AddressSanitizer runtime (its flag list prints)
    r = subprocess.run([sanitized], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="help=1"), timeout=60)
    assert "AddressSanitizer" in r.stdout + r.stderr
