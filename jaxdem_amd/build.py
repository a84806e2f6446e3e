"""Build libjdem_hip.so (the gfx950 HIP library) in-tree with hipcc.

Floating-point flags are part of the numerical contract of the library:
FMA contraction is disabled and float32 division / square root are IEEE
correctly rounded, so every per-pair and per-particle expression of the
kernels rounds exactly like the same expression evaluated op by op in numpy
float32.  Packed-FP32 instruction selection is disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libjdem_hip.so"
SOURCES = ["jdem_cells.hip", "jdem_step.hip", "jdem_resident.hip", "jdem_fire.hip", "jdem_abi.hip",
           "jdem_noalloc.hip"]
ARCH = os.environ.get("JDEM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-DJDEM_TARGET=\"{ARCH}\"",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build(verbose: bool = False, jobs: int = 3) -> Path:
    """Compile every HIP source for gfx950 and link the shared library."""
    cc = hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    incs = [f"-I{INCLUDE}", f"-I{CSRC}"]
    procs, objs = [], []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [cc, *HIP_FLAGS, *incs, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
            _drain(procs)
    _drain(procs, wait=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    return LIB


_done: set = set()


def _drain(procs, wait: bool = False) -> None:
    for src, p in procs:
        if src in _done:
            continue
        if wait:
            p.wait()
        if p.poll() is None:
            continue
        out = p.stdout.read().decode() if p.stdout else ""
        _done.add(src)
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{out}")


if __name__ == "__main__":
    _done.clear()
    path = build(verbose="-v" in sys.argv)
    print(path)
