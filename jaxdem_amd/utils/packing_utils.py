"""Packing fraction helpers (utils/packing_utils.py:21-245) for spheres: one
sphere per body, so the affine rescale moves each sphere centre."""

from __future__ import annotations

from dataclasses import replace

import numpy as np

F32 = np.float32


def compute_particle_volume(state) -> F32:
    """Total particle volume (packing_utils.py:21-26; one sphere per body)."""
    if not state.is_sphere_state:
        raise NotImplementedError("jamming on the GPU path handles spheres only")
    return F32(state.volume.sum().item())


def compute_packing_fraction(state, system) -> F32:
    """Particle volume over the box volume, anchor at the origin
    (packing_utils.py:29-36)."""
    return F32(compute_particle_volume(state) / F32(np.prod(np.asarray(system.domain.box_size, F32), dtype=F32)))


def scale_to_packing_fraction(state, system, new_packing_fraction) -> tuple:
    """Rescale the box by one common factor, keeping its aspect ratio, and the
    sphere centres with it (packing_utils.py:67-113).  A NeighborList is
    marked for a rebuild.  The domain's inverse box size is refreshed with
    the box (the reference keeps the old one, which changes no contact
    displacement while the box is wider than twice the contact distance)."""
    from ..colliders import NeighborList

    box = np.asarray(system.domain.box_size, F32)
    vol = compute_particle_volume(state)
    scale = F32(F32(vol / (F32(new_packing_fraction) * F32(np.prod(box, dtype=F32)))) ** F32(1.0 / state.dim))
    new_box = (box * scale).astype(F32)
    # offset = com * scale - com, added to the centre (one sphere per body)
    s = replace(state, pos_c=state.pos_c + (state.pos_c * float(scale) - state.pos_c))
    dom = replace(system.domain, box_size=new_box, inv_box_size=(F32(1.0) / new_box).astype(F32))
    sys2 = replace(system, domain=dom)
    if isinstance(sys2.collider, NeighborList):
        sys2 = replace(sys2, collider=replace(sys2.collider, n_build_times=0))
    return s, sys2


def quasistatic_compress_to_packing_fraction(
    state,
    system,
    target_phi: float,
    *,
    step: float = 1e-3,
    phi_tolerance: float = 1e-10,
    pe_tol: float = 1e-16,
    pe_diff_tol: float = 1e-16,
    max_n_min_steps_per_outer: int = 1_000_000,
    max_n_outer_steps: int = 1_000_000,
    progress: bool = False,
) -> tuple:
    """Compress (or expand) toward ``target_phi`` in packing-fraction steps of
    at most ``step``, minimising after each rescale; the last step is cut so
    that the target is met within ``phi_tolerance`` (packing_utils.py:134-245).
    Returns (state, system, final_phi, final_pe), final_pe per particle after
    the last minimisation."""
    from ..minimizers import minimize

    def relax(s, y):
        return minimize(s, y, max_steps=max_n_min_steps_per_outer, pe_tol=pe_tol, pe_diff_tol=pe_diff_tol)

    state, system, _, pe = relax(state, system)
    current_phi = float(compute_packing_fraction(state, system))
    step_mag = abs(float(step))
    it = range(int(max_n_outer_steps))
    if progress:
        try:
            from tqdm import tqdm

            it = tqdm(it, total=int(max_n_outer_steps), desc="Compressing")
        except ImportError:
            pass
    for _ in it:
        remaining = float(target_phi) - current_phi
        if abs(remaining) <= phi_tolerance:
            break
        delta = (1.0 if remaining > 0.0 else -1.0) * min(step_mag, abs(remaining))
        new_phi = current_phi + delta
        state, system = scale_to_packing_fraction(state, system, new_phi)
        state, system, _, pe = relax(state, system)
        current_phi = new_phi
    return state, system, current_phi, float(pe)


__all__ = [
    "compute_packing_fraction",
    "compute_particle_volume",
    "quasistatic_compress_to_packing_fraction",
    "scale_to_packing_fraction",
]
