"""Build recipe for the oracle's C restatement (test infrastructure only).

``python oracle/build.py`` compiles ``oracle/msoracle.c`` into
``oracle/libmsoracle.so`` with gcc.  -ffp-contract=off keeps every multiply
and add separately rounded, as in the reference's x86-64 build (no -mfma).

The reference's own C++ (Qt5 + FFTW3, qmake build) is not built here: it needs
external libraries that this image does not provide as part of the project, so
``oracle/_ref`` stays empty and the oracle is pinned by closed forms,
invariants and brute force instead.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "msoracle.c")
OUT = os.path.join(HERE, "libmsoracle.so")


def build(verbose=False):
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-std=c11", "-o", OUT + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
