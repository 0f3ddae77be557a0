"""CPU sanitizer runs (no GPU): the library's host-only C++ -- the per-node
stencil drop-ins, the threaded row (de)interleaving of the operator
transfers, the reference file formats -- under AddressSanitizer +
UndefinedBehaviorSanitizer and under ThreadSanitizer (tests/sanitize), and
the CPU oracle that every parity test trusts under ASan + UBSan, running its
golden-vector and closed-form tests."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SAN = os.path.join(HERE, "sanitize")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", SAN], check=True)


@pytest.mark.parametrize("flavour", ["asan", "tsan"])
def test_host_code_is_clean_under_sanitizers(built, tmp_path, flavour):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(SAN, "host_check_" + flavour), str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "clean" in r.stdout
    assert "Sanitizer" not in r.stderr, r.stderr


def test_oracle_is_clean_under_asan_and_ubsan(built):
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.exists(libasan), libasan
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               IBIM_ORACLE_LIB=os.path.join(REPO, "oracle", "liboracle_asan.so"))
    # the golden vectors and the closed forms (the largest dense solves left out for time)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_oracle_golden.py"), os.path.join(HERE, "test_oracle_math.py"),
                        "-k", "not well_conditioned and not converges and not mie"],
                       capture_output=True, text=True, env=env, timeout=900, cwd=REPO)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr
