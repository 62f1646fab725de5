"""Build step: compiles every HIP source of the library for gfx950 with the
Makefile in csrc/ (always a full rebuild) into libqphix_amd.so."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs=8, verbose=False):
    cmd = ["make", "-B", f"-j{jobs}", "-C", CSRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"HIP build failed:\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}")
    if verbose:
        print(r.stdout[-2000:])
    return os.path.join(_HERE, "libqphix_amd.so")
