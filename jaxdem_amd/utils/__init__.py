"""Helpers around the time step: initial states, energies, contact stress.

``grid_state`` follows jaxdem/utils/grid_state.py:21-127 with numpy's PCG64
generator in place of JAX's PRNG (whose bit streams cannot be reproduced
without JAX).  ``thermal`` and ``contacts`` follow utils/thermal.py:83-177 and
utils/contacts.py:89-164 with float64 reductions on the GPU.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

from ..state import State


def grid_state(
    *,
    n_per_axis: Sequence[int],
    spacing=1.0,
    radius: float = 1.0,
    mass: float = 1.0,
    jitter: float = 0.0,
    vel_range=None,
    radius_range=None,
    mass_range=None,
    seed: int = 0,
    device=None,
) -> State:
    """Spheres on a rectangular lattice, 'ij' order (grid_state.py:21-127)."""
    n_per_axis = tuple(int(n) for n in n_per_axis)
    dim = len(n_per_axis)
    sp = np.broadcast_to(np.asarray(spacing, np.float32), (dim,))
    axes = [np.arange(n, dtype=np.float32) * sp[a] for a, n in enumerate(n_per_axis)]
    coords = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, dim).astype(np.float32)
    N = coords.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    if jitter > 0.0:
        coords = coords + rng.uniform(-jitter, jitter, coords.shape).astype(np.float32)
    if radius_range is not None:
        lo, hi = np.asarray(radius_range, np.float64)
        rad = rng.uniform(lo, hi, N).astype(np.float32)
    else:
        rad = np.full(N, radius, np.float32)
    if mass_range is not None:
        lo, hi = np.asarray(mass_range, np.float64)
        m = rng.uniform(lo, hi, N).astype(np.float32)
    else:
        m = np.full(N, mass, np.float32)
    vr = (-1.0, 1.0) if vel_range is None else tuple(np.asarray(vel_range, np.float64).tolist())
    vel = rng.uniform(vr[0], vr[1], coords.shape).astype(np.float32)
    return State.create(pos=coords, vel=vel, rad=rad, mass=m, device=device)


from . import contacts, thermal  # noqa: E402
from .contacts import compute_contact_pressure, compute_contact_stress_tensor, compute_pressure  # noqa: E402
from .thermal import (  # noqa: E402
    compute_energy,
    compute_kinetic_energy,
    compute_potential_energy,
)

__all__ = [
    "grid_state",
    "thermal",
    "contacts",
    "compute_energy",
    "compute_kinetic_energy",
    "compute_potential_energy",
    "compute_contact_pressure",
    "compute_contact_stress_tensor",
    "compute_pressure",
]
