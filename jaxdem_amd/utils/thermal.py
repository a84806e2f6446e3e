"""Energies (jaxdem/utils/thermal.py:83-177), float64 reductions on the GPU."""

from __future__ import annotations


def compute_kinetic_energy(state, system=None) -> float:
    """Translational + rotational kinetic energy (thermal.py:83-122)."""
    from ..system import _kinetic_energy

    if system is None:
        from ..system import System

        system = System.create(state.shape, collider_type="")
    return _kinetic_energy(state, system)


def compute_potential_energy(state, system) -> float:
    """Collider pair energy + ForceManager energy (thermal.py:127-150)."""
    return system.potential_energy(state)


def compute_energy(state, system) -> float:
    """Total energy (thermal.py:155-177)."""
    return compute_kinetic_energy(state, system) + compute_potential_energy(state, system)


__all__ = ["compute_kinetic_energy", "compute_potential_energy", "compute_energy"]
