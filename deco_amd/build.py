"""Build step: compiles libdeco_hip.so (hipcc, gfx950) from deco_amd/csrc with its Makefile."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs=8, verbose=False):
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stdout.write(res.stdout[-4000:])
        sys.stderr.write(res.stderr[-8000:])
    if res.returncode != 0:
        raise RuntimeError("building libdeco_hip.so failed")
    return os.path.join(HERE, "libdeco_hip.so")


if __name__ == "__main__":
    print(build(verbose=True))
