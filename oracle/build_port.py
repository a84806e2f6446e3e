"""Build oracle/libjdem_port.so (TEST INFRASTRUCTURE: the CPU baseline port
and the float64 FIRE reference).

g++ with OpenMP; FMA contraction off and no fast-math so that every float32
operation rounds like the numpy oracle.
"""

from __future__ import annotations

import fcntl
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = [HERE / "jdem_port.cpp", HERE / "jdem_fire64.cpp"]
LIB = HERE / "libjdem_port.so"


def build(verbose: bool = False) -> Path:
    """Compile the library; concurrent callers (parallel test workers)
    build one at a time, each into its own temporary file, and the finished
    library replaces the old one in a single rename."""
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = LIB.with_suffix(f".so.tmp{os.getpid()}")
    cmd = [cxx, "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           *map(str, SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    with open(HERE / ".build_port.lock", "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"g++ failed:\n{r.stdout}")
            os.replace(tmp, LIB)
        finally:
            if tmp.exists():
                tmp.unlink()
            fcntl.flock(lock, fcntl.LOCK_UN)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
