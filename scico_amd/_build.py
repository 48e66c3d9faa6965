"""Build ``libscico_amd.so`` (all HIP kernels + the C ABI) for gfx950.

Run ``python -m scico_amd._build``.  Objects go to ``_build/`` at the
repository root, the library next to this file (in-tree, so that it
travels with the source tree).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libscico_amd.so"

SOURCES = [
    "common.cpp",
    "xray2d.hip",
    "xray3d.hip",
    "batch2d.hip",
    "vec.hip",
    "solver.hip",
    "fft.hip",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the reference's weight arithmetic is restated operation
# by operation; FMA contraction would change detector bin indices.  Packed
# fp32 instructions are disabled for every kernel.
CFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I" + str(ROOT / "include"),
]


# Per-source additions.  The projector kernels are latency-bound at the
# occupancy their LDS rows allow, so they are scheduled for ILP.
SOURCE_FLAGS = {
    "xray2d.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def _compile(src: str) -> Path:
    s = CSRC / src
    obj = BUILD / (s.name + ".o")
    cmd = [HIPCC, *CFLAGS, *SOURCE_FLAGS.get(src, []), "-x", "hip", "-c", str(s), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        warn = "\n".join(
            ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()
        )
        if warn.strip():
            print(f"[{src}] {warn}", file=sys.stderr)
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} sources")
    return LIB


if __name__ == "__main__":
    build()
