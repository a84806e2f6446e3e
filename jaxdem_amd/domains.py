"""Simulation domains (jaxdem/domains/__init__.py:28-215, free.py, periodic.py,
reflect.py).

A domain carries ``box_size``, ``inv_box_size`` and ``anchor`` as host
arrays of shape (dim,), float32 (float64 with ``jax_enable_x64``).  Its per-particle effects run on the GPU: the
periodic minimum image inside the pair kernels, the reflective wall impulse
and the free-domain bounding box inside the fused step.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np
import torch

from . import _lib as L
from .config import real_dtype
from .factory import Factory


def _arr(x, dim: int, default: float, name: str) -> np.ndarray:
    if x is None:
        return np.full(dim, default, real_dtype())
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    a = np.asarray(x, dtype=real_dtype())
    if a.shape != (dim,):
        raise ValueError(f"{name} must have shape ({dim},), got shape {a.shape}.")
    return a


@dataclass
class Domain(Factory):
    """Base domain: displacement r_i - r_j, no-op apply and shift."""

    box_size: np.ndarray
    inv_box_size: np.ndarray
    anchor: np.ndarray

    kind = L.DOMAIN_FREE

    @property
    def periodic(self) -> bool:
        return False

    @classmethod
    def Create(cls, dim: int, box_size=None, anchor=None, **kw: Any):
        """domains/__init__.py:62-120."""
        box = _arr(box_size, dim, 1.0, "box_size")
        anc = _arr(anchor, dim, 0.0, "anchor")
        ft = box.dtype.type
        return cls(box_size=box, inv_box_size=(ft(1.0) / box).astype(ft), anchor=anc, **kw)

    @staticmethod
    def displacement(ri: torch.Tensor, rj: torch.Tensor, system) -> torch.Tensor:
        """domains/__init__.py:128-150."""
        return ri - rj

    @staticmethod
    def _displacement(ri: torch.Tensor, rj: torch.Tensor, system) -> torch.Tensor:
        return system.domain.displacement(ri, rj, system)

    @staticmethod
    def apply(state, system):
        """domain.apply on the GPU (jdem_step_stage DOMAIN_APPLY)."""
        from .system import _run_stage

        return _run_stage(state, system, L.STAGE_DOMAIN_APPLY)

    @staticmethod
    def shift(state, system):
        return state, system

    def to_c(self, dim: int, f64: bool = False):
        """jdem_domain (jdem_domain_f64 with ``f64``)."""
        d = L.JdemDomainF64() if f64 else L.JdemDomain()
        d.kind = self.kind
        d.dim = dim
        for a in range(dim):
            d.box_size[a] = float(self.box_size[a])
            d.inv_box_size[a] = float(self.inv_box_size[a])
            d.anchor[a] = float(self.anchor[a])
        for a in range(dim, 3):
            d.box_size[a] = 1.0
            d.inv_box_size[a] = 1.0
            d.anchor[a] = 0.0
        d.restitution = float(getattr(self, "restitution_coefficient", 1.0))
        return d


Domain.register("")(Domain)


@Domain.register("free")
@dataclass
class FreeDomain(Domain):
    """Unbounded domain; apply() refits box and anchor to the particles
    (free.py:42-58)."""

    kind = L.DOMAIN_FREE


@Domain.register("periodic")
@dataclass
class PeriodicDomain(Domain):
    """Minimum-image displacement (periodic.py:39-79); no apply."""

    kind = L.DOMAIN_PERIODIC

    @property
    def periodic(self) -> bool:
        return True

    @staticmethod
    def displacement(ri: torch.Tensor, rj: torch.Tensor, system) -> torch.Tensor:
        box = torch.as_tensor(system.domain.box_size, device=ri.device)
        rij = ri - rj
        return rij - box * torch.round(rij / box)

    @staticmethod
    def _displacement(ri: torch.Tensor, rj: torch.Tensor, system) -> torch.Tensor:
        box = torch.as_tensor(system.domain.box_size, device=ri.device)
        inv = torch.as_tensor(system.domain.inv_box_size, device=ri.device)
        rij = ri - rj
        return rij - box * torch.round(rij * inv)

    @staticmethod
    def shift(state, system):
        """Wrap ``pos_c`` into the box (periodic.py:84-116)."""
        box = torch.as_tensor(system.domain.box_size, device=state.pos_c.device)
        anc = torch.as_tensor(system.domain.anchor, device=state.pos_c.device)
        new = replace(state, pos_c=state.pos_c - box * torch.floor((state.pos_c - anc) / box))
        return new, system


@Domain.register("reflect")
@dataclass
class ReflectDomain(Domain):
    """Reflective walls with Verlet time-of-collision correction
    (reflect.py:102-297, _toc.py:11-94)."""

    restitution_coefficient: float = 1.0
    kind = L.DOMAIN_REFLECT

    @classmethod
    def Create(cls, dim: int, box_size=None, anchor=None, restitution_coefficient: float = 1.0, **kw: Any):
        if not (0.0 < restitution_coefficient <= 1.0):
            raise ValueError(f"restitution_coefficient must be in (0, 1], got {restitution_coefficient}.")
        return super(ReflectDomain, cls).Create(
            dim, box_size=box_size, anchor=anchor, restitution_coefficient=float(real_dtype()(restitution_coefficient)),
            **kw
        )


__all__ = ["Domain", "FreeDomain", "PeriodicDomain", "ReflectDomain"]
