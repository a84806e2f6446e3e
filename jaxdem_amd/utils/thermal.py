"""Energies and temperature (jaxdem/utils/thermal.py:83-340), float64
reductions on the GPU."""

from __future__ import annotations

import numpy as np


def _bare_system(state):
    from ..system import System

    return System.create(state.shape, collider_type="")


def compute_kinetic_energy(state, system=None) -> float:
    """Translational + rotational kinetic energy (thermal.py:83-122)."""
    from ..system import _kinetic_energy

    return _kinetic_energy(state, system if system is not None else _bare_system(state))


def compute_translational_kinetic_energy(state) -> float:
    """sum 1/2 m v^2 (thermal.py:83-102), a float64 sum on the GPU."""
    from ..system import _kinetic_energy_parts

    return _kinetic_energy_parts(state, _bare_system(state))[0]


def compute_rotational_kinetic_energy(state) -> float:
    """sum 1/2 w_b . I w_b with the body-frame angular velocity
    (thermal.py:105-122), a float64 sum on the GPU."""
    from ..system import _kinetic_energy_parts

    return _kinetic_energy_parts(state, _bare_system(state))[1]


def compute_potential_energy(state, system) -> float:
    """Collider pair energy + ForceManager energy (thermal.py:127-150)."""
    return system.potential_energy(state)


def compute_energy(state, system) -> float:
    """Total energy (thermal.py:155-177)."""
    return compute_kinetic_energy(state, system) + compute_potential_energy(state, system)


def count_dynamic_dofs(state, can_rotate: bool, subtract_drift: bool) -> tuple:
    """Degrees of freedom of the free bodies (thermal.py:180-203): one body
    per clump id present with no fixed member; (n_dof, n_dof_v, n_dof_w)."""
    import torch

    clump = state.clump_id.reshape(-1).long()
    counts = torch.bincount(clump, minlength=state.N)
    fixed = torch.bincount(clump, weights=state.fixed.reshape(-1).to(torch.float64), minlength=state.N)
    free = int(((counts > 0) & (fixed == 0)).sum().item())
    n_v = (free - int(bool(subtract_drift))) * int(state.vel.shape[-1])
    n_w = free * int(state.ang_vel.shape[-1]) * int(bool(can_rotate))
    return n_v + n_w, n_v, n_w


def compute_temperature(state, can_rotate: bool, subtract_drift: bool, k_B: float = 1.0) -> float:
    """2 KE / (k_B n_dof) (thermal.py:250-271)."""
    from ..system import _kinetic_energy_parts

    n_dof, _, _ = count_dynamic_dofs(state, can_rotate, subtract_drift)
    ke_t, ke_r = _kinetic_energy_parts(state, _bare_system(state))
    ke = ke_t + (ke_r if can_rotate else 0.0)
    return 2.0 * ke / (k_B * n_dof)


def scale_to_temperature(state, target_temperature: float, can_rotate: bool, subtract_drift: bool,
                         k_B: float = 1.0):
    """Scale the velocities to ``target_temperature`` (thermal.py:306-360):
    with ``subtract_drift`` the mass-weighted drift of the free particles is
    removed first; the temperature is measured on those velocities and both
    velocities and (when rotations count) angular velocities are scaled by
    sqrt(T_target / T).  The state is returned with new arrays."""
    from dataclasses import replace

    import torch

    vel = state.vel
    if subtract_drift:
        free = (~state.fixed.bool()).to(vel.dtype)
        free_mass = state.mass * free
        total = free_mass.sum(dim=-1, keepdim=True)
        total = torch.where(total == 0, torch.ones_like(total), total)
        drift = (vel * free_mass[..., None]).sum(dim=-2, keepdim=True) / total[..., None]
        vel = vel - drift * free[..., None]
    t = compute_temperature(replace(state, vel=vel.contiguous()), can_rotate, subtract_drift, k_B)
    s = float(np.sqrt(target_temperature / t))
    ang = state.ang_vel * s if can_rotate else state.ang_vel.clone()
    return replace(state, vel=(vel * s).contiguous(), ang_vel=ang)


def set_temperature(state, target_temperature: float, can_rotate: bool, subtract_drift: bool, seed: int | None = 0,
                    k_B: float = 1.0):
    """Random velocities at ``target_temperature`` (thermal.py:206-303): one
    normal draw per free body (clump) for the velocity and for the body-frame
    angular velocity, the free bodies' mean velocity removed with
    ``subtract_drift``, the angular velocity rotated to the lab frame (3D),
    then ``scale_to_temperature``.  The draws come from numpy's PCG64 with
    ``seed`` (JAX's PRNG streams cannot be reproduced without JAX)."""
    from dataclasses import replace

    import torch

    from ..state import Quaternion

    rng = np.random.Generator(np.random.PCG64(0 if seed is None else int(seed)))
    n, dim, a = state.N, state.dim, int(state.ang_vel.shape[-1])
    clump = state.clump_id.reshape(-1).long().cpu().numpy()
    counts = np.bincount(clump, minlength=n)
    fixed = np.bincount(clump, weights=state.fixed.reshape(-1).cpu().numpy().astype(np.float64), minlength=n)
    free = (counts > 0) & (fixed == 0)
    v = rng.standard_normal((n, dim)) * free[:, None]
    if subtract_drift:
        v = v - (v.sum(axis=0) / max(int(free.sum()), 1)) * free[:, None]
    w = rng.standard_normal((n, a)) * free[:, None]
    dev = state.vel.device
    vel = torch.as_tensor(v[clump], dtype=torch.float32, device=dev).contiguous()
    wb = torch.as_tensor(w[clump], dtype=torch.float32, device=dev).contiguous()
    ang = wb if dim == 2 else Quaternion.rotate(state.q, wb).contiguous()
    return scale_to_temperature(replace(state, vel=vel, ang_vel=ang), target_temperature, can_rotate,
                                subtract_drift, k_B)


__all__ = [
    "set_temperature",
    "compute_energy",
    "compute_kinetic_energy",
    "compute_potential_energy",
    "compute_rotational_kinetic_energy",
    "compute_temperature",
    "compute_translational_kinetic_energy",
    "count_dynamic_dofs",
    "scale_to_temperature",
]
