"""The oracle's own test files, run once more against its AddressSanitizer +
UndefinedBehaviorSanitizer build (oracle/Makefile `sanitize`), in a child
process that preloads the sanitizer runtimes.  Any out-of-bounds access,
use after free or undefined arithmetic in the oracle fails the run.  CPU only."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN_LIB = ROOT / "oracle" / "libvf_oracle_san.so"


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if p and os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.timeout(900)
def test_oracle_suite_under_asan_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], check=True, capture_output=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan}:{ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "VFO_ORACLE_LIB": str(SAN_LIB),
    })
    files = ["tests/test_oracle_fracture.py", "tests/test_oracle_geometry.py", "tests/test_oracle_subseeds.py",
             "tests/test_oracle_erosion.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *files],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout, tail
