"""The multiply-shift division the kernels use for their index arithmetic
(qphix-codegen_amd/csrc/fastdiv.h, Granlund & Montgomery's round-up method),
compiled from the same header with the host compiler and checked against the
built-in / and % over every divisor 1..2048 and large divisors, at both ends
of the numerator range."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fastdiv_matches_integer_division(tmp_path):
    exe = tmp_path / "fastdiv_check"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "qphix-codegen_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cpp", "fastdiv_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
