"""Build recipe for the oracle's C restatement (TEST INFRASTRUCTURE ONLY).

``python -m oracle._cbuild`` compiles ``oracle/cpu_xray.c`` into
``oracle/libxray_cpu.so`` with gcc (-O3 -march=x86-64-v3 -fno-fast-math
-ffp-contract=off -fopenmp): the portable build the tests load on any
x86-64 host.  ``build(native=True, lib=path)`` compiles the same source
with -march=native for the machine it runs on (the CPU baseline timed by
bench.py is built that way on the GPU box's host).  The reference itself
is pure Python, so there is no reference code to compile into
``oracle/_ref``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_xray.c"
LIB = HERE / "libxray_cpu.so"


def build(native: bool = False, lib: os.PathLike | str | None = None) -> Path:
    out = Path(lib) if lib is not None else LIB
    tmp = out.with_suffix(".so.tmp")
    march = "-march=native" if native else "-march=x86-64-v3"
    cmd = ["gcc", "-O3", march, "-fno-fast-math", "-ffp-contract=off", "-fopenmp",
           "-fPIC", "-shared", "-o", str(tmp), str(SRC), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def load(lib: os.PathLike | str | None = None) -> ctypes.CDLL:
    path = Path(lib) if lib is not None else LIB
    if not path.exists():
        build(lib=path)
    lib_ = ctypes.CDLL(str(path))
    P = ctypes.c_void_p
    d, i = ctypes.c_double, ctypes.c_int
    lib_.cpu_xr2d_weights.argtypes = [i, i, i, P, P, d, P, P, P]
    lib_.cpu_xr2d_forward.argtypes = [i, i, i, i, P, P, d, P, P, P]
    lib_.cpu_xr2d_adjoint.argtypes = [i, i, i, i, P, P, d, P, P, P]
    lib_.cpu_xr2d_forward_f64acc.argtypes = [i, i, i, i, P, P, d, P, P, P]
    lib_.cpu_xr2d_adjoint_f64acc.argtypes = [i, i, i, i, P, P, d, P, P, P]
    lib_.cpu_xr3d_sep_forward.argtypes = [i, i, i, i, i, i, P, i, i, i, P, P]
    lib_.cpu_xr3d_sep_adjoint.argtypes = [i, i, i, i, i, i, P, i, i, i, P, P]
    return lib_


if __name__ == "__main__":
    print(build())
