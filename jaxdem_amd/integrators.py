"""Integrators (jaxdem/integrators/__init__.py:34-172, velocity_verlet.py,
direct_euler.py, velocity_verlet_spiral.py, spiral.py).

Each integrator's ``step_before_force`` / ``step_after_force`` runs as one
GPU stage (jdem_step_stage); inside ``System.step`` the same arithmetic is
fused into the per-step kernel.
"""

from __future__ import annotations

from dataclasses import dataclass

from . import _lib as L
from .factory import Factory


def free_mask(state):
    """integrators/__init__.py:21-31: (N, 1) True for free particles."""
    return (~state.fixed)[..., None]


@dataclass
class Integrator(Factory):
    kind = 0

    @staticmethod
    def step_before_force(state, system):
        return state, system

    @staticmethod
    def step_after_force(state, system):
        return state, system

    @staticmethod
    def initialize(state, system):
        return state, system


@dataclass
class LinearIntegrator(Integrator):
    """Translational integrators."""

    kind = L.LIN_NONE


@dataclass
class RotationIntegrator(Integrator):
    """Rotational integrators."""

    kind = L.ROT_NONE


LinearIntegrator.register("")(LinearIntegrator)
RotationIntegrator.register("")(RotationIntegrator)


def _stage(stage):
    def run(state, system):
        from .system import _run_stage

        return _run_stage(state, system, stage)

    return staticmethod(run)


@LinearIntegrator.register("verlet")
@dataclass
class VelocityVerlet(LinearIntegrator):
    """v += F dt/(2m); x += dt v | v += F dt/(2m) (velocity_verlet.py:57-94)."""

    kind = L.LIN_VERLET
    step_before_force = _stage(L.STAGE_LIN_BEFORE)
    step_after_force = _stage(L.STAGE_LIN_AFTER)


@LinearIntegrator.register("euler")
@dataclass
class DirectEuler(LinearIntegrator):
    """Symplectic Euler after the force (direct_euler.py:62-65)."""

    kind = L.LIN_EULER
    step_after_force = _stage(L.STAGE_LIN_AFTER)


@RotationIntegrator.register("verletspiral")
@dataclass
class VelocityVerletSpiral(RotationIntegrator):
    """SPIRAL leapfrog rotation (velocity_verlet_spiral.py:83-180)."""

    kind = L.ROT_VERLET_SPIRAL
    step_before_force = _stage(L.STAGE_ROT_BEFORE)
    step_after_force = _stage(L.STAGE_ROT_AFTER)


@RotationIntegrator.register("spiral")
@dataclass
class Spiral(RotationIntegrator):
    """Non-leapfrog SPIRAL rotation (spiral.py:43-144)."""

    kind = L.ROT_SPIRAL
    step_after_force = _stage(L.STAGE_ROT_AFTER)


__all__ = [
    "free_mask",
    "Integrator",
    "LinearIntegrator",
    "RotationIntegrator",
    "VelocityVerlet",
    "DirectEuler",
    "VelocityVerletSpiral",
    "Spiral",
]
