"""Writes jamming_n1024.npz: the same-basin fixture of the precision-form
test.  1024 bidisperse spheres (0.5 / 0.7, 50 : 50) at volume fraction 0.66
from seed 3 (jaxdem_amd.workloads.bidisperse_jamming), relaxed by the float64
FIRE of oracle/jdem_fire64.cpp to max|F| <= 1e-4 (``start``), then by the
same to max|F| <= 1e-12 (``ref``); k = 1, dt0 = 1e-2.

    python tests/golden/make_jamming_fixture.py
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from jaxdem_amd.workloads import bidisperse_jamming  # noqa: E402
from tests.precision_vs_fp64 import fp64_run  # noqa: E402


def main():
    pos0, rad, L = bidisperse_jamming(1024, seed=3, phi=0.66)
    s1 = fp64_run(pos0, rad, L, force_tol=1e-4)
    s2 = fp64_run(s1["pos"], rad, L, force_tol=1e-12)
    np.savez(HERE / "jamming_n1024.npz", rad=rad, L=L, start=s1["pos"], ref=s2["pos"], ref_pe=s2["pe"],
             ref_pressure=s2["pressure"], ref_max_grad=s2["max_grad"], ref_steps=s2["steps"],
             start_steps=s1["steps"])
    print(f"start {s1['steps']} steps, ref {s2['steps']} steps, E/N {s2['pe']:.6e}, P {s2['pressure']:.6e}, "
          f"max|F| {s2['max_grad']:.2e}")


if __name__ == "__main__":
    main()
