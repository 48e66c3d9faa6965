"""Build recipe for the oracle's C restatement (TEST INFRASTRUCTURE ONLY).

``python -m oracle._cbuild`` compiles ``oracle/cpu_xray.c`` into
``oracle/libxray_cpu.so`` with gcc (-O3 -march=x86-64-v3 -fno-fast-math
-ffp-contract=off -fopenmp).  The reference itself is pure Python, so
there is no reference code to compile into ``oracle/_ref``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_xray.c"
LIB = HERE / "libxray_cpu.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fno-fast-math", "-ffp-contract=off", "-fopenmp",
           "-fPIC", "-shared", "-o", str(tmp), str(SRC), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def load() -> ctypes.CDLL:
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    P = ctypes.c_void_p
    d, i = ctypes.c_double, ctypes.c_int
    lib.cpu_xr2d_weights.argtypes = [i, i, i, P, P, d, P, P, P]
    lib.cpu_xr2d_forward.argtypes = [i, i, i, i, P, P, d, P, P, P]
    lib.cpu_xr2d_adjoint.argtypes = [i, i, i, i, P, P, d, P, P, P]
    return lib


if __name__ == "__main__":
    print(build())
