"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

The oracle is rebuilt with -fsanitize=address,undefined (oracle/Makefile,
target sanitize) and every pin of tests/test_oracle.py runs against it in a
child process that preloads the sanitizer runtime; the drop-in drivers' own C
code is rebuilt the same way (everythinggraph_amd/Makefile, target sanitize)
and test_abi.py's command-line checks run against those binaries.  Any
sanitizer report fails the child, and so the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


ASAN = _runtime("libasan.so")
pytestmark = pytest.mark.skipif(ASAN is None, reason="no AddressSanitizer runtime")


def _env(**extra):
    env = dict(os.environ)
    env.update(LD_PRELOAD=ASAN, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.update(extra)
    return env


@pytest.mark.timeout(900)
def test_oracle_pins_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    so = os.path.join(ROOT, "oracle", "liboracle_san.so")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(TESTS, "test_oracle.py")],
                       env=_env(EG_ORACLE_SO=so), capture_output=True, text=True, timeout=850)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.timeout(600)
def test_driver_host_code_under_asan_ubsan():
    pkg = os.path.join(ROOT, "everythinggraph_amd")
    subprocess.run(["make", "-s", "-C", pkg, "sanitize"], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-k", "driver_usage", os.path.join(TESTS, "test_abi.py")],
                       env=_env(EG_BIN_DIR=os.path.join(pkg, "bin_san")), capture_output=True,
                       text=True, timeout=550)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
