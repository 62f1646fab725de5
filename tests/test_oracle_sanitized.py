"""The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle's CPU tests run in a child process against liboracle_san.so (the
same source built with -fsanitize=address,undefined, oracle/Makefile), with
the ASan and UBSan runtimes preloaded into that Python process; any
out-of-bounds access, use after free or undefined behaviour in the oracle
aborts the child and fails this test.  Host code only.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_SO = os.path.join(ORACLE, "liboracle_san.so")
# the CPU test modules that drive the oracle's C functions
MODULES = ["tests/test_oracle.py", "tests/test_golden.py", "tests/test_force_oracle.py",
           "tests/test_face_reference.py", "tests/test_qphix_blocks.py"]


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    r = subprocess.run(["make", "-B", "-C", ORACLE, "liboracle_san.so"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", QAMD_ORACLE_SO=SAN_SO)
    # the child really runs the sanitized build
    probe = ("import sys; sys.path.insert(0, %r); from oracle import oracle as o; o.lib(); "
             "m = open('/proc/self/maps').read(); print(o._SO, 'liboracle_san.so' in m, 'libasan' in m)" % ROOT)
    r = subprocess.run([sys.executable, "-c", probe], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.split() == [SAN_SO, "True", "True"], r.stdout
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
                       + MODULES, env=env, cwd=ROOT, capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout
