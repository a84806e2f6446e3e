"""Contact stress and pressure (jaxdem/utils/contacts.py:89-164).

sigma = (1/V) sum_{i<j} r_ij (x) F_ij, P = tr(sigma) / dim, with the pair sum
accumulated in float64 on the GPU.
"""

from __future__ import annotations

import numpy as np


def compute_contact_stress_tensor(state, system, cutoff=None, max_neighbors=None, volume=None) -> tuple:
    """``compute_contact_stress_tensor`` (contacts.py:89-147): the contact
    virial over the volume (the box volume unless ``volume`` is given), as
    (state, system, stress (dim, dim) float64).  The contacts come from the
    system's own collider, so ``cutoff`` and ``max_neighbors`` (the
    reference's neighbour-list sizing for this call) are not needed; a
    NeighborList that is unbuilt or stale is rebuilt first and the returned
    system carries the refreshed list, as in the reference (contacts.py:58-62)."""
    from ..system import _contact_virial

    W, system = _contact_virial(state, system)
    V = float(np.prod(np.asarray(system.domain.box_size, np.float64))) if volume is None else float(volume)
    return state, system, W / V


def compute_pressure(state, system) -> float:
    """trace(stress) / dim as a float."""
    return compute_contact_pressure(state, system)[2]


def compute_contact_pressure(state, system, cutoff=None, max_neighbors=None, volume=None) -> tuple:
    """``compute_contact_pressure`` (contacts.py:149-164): trace of the contact
    stress over dim, positive for repulsive overlaps; returns (state, system,
    pressure)."""
    state, system, stress = compute_contact_stress_tensor(state, system, cutoff, max_neighbors, volume)
    return state, system, float(np.trace(stress) / state.dim)


__all__ = ["compute_contact_pressure", "compute_contact_stress_tensor", "compute_pressure"]
