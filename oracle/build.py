"""Build recipe of the oracle (test infrastructure only).

Compiles oracle/ntlp_oracle.c with gcc into oracle/liboracle.so.  The
reference itself (Fortran, needing Intel mpiifort, an MPI library and netCDF)
is not built here.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build_oracle() -> str:
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c99", "-fPIC", "-shared", "-fopenmp",
           "-Wall", "-o", tmp, os.path.join(HERE, "ntlp_oracle.c"), "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + proc.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_oracle())
