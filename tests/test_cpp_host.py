"""The C++ host classes (include/qphix_amd.hpp) without a GPU: the header
compiles as strict C++11 with the host compiler, and the C++ host program's
context-only checks pass (geometry queries, the traversal choice, and the
exceptions thrown for bad arguments before any device work)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def test_header_compiles_as_cxx11(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "qphix_amd.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_host_program_context_checks(qamd):
    exe = os.path.join(CPP, "host_api_check")
    if not os.path.exists(exe):  # normally built with the library
        r = subprocess.run(["make", "-s", "-C", CPP, "host_api_check"], capture_output=True, text=True)
        assert r.returncode == 0, "C++ host program does not build:\n" + r.stderr[-2000:]
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "cpu ok" in r.stdout, r.stdout + r.stderr
