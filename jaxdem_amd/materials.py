"""Materials, matchmakers and the material table.

Reference: materials/__init__.py (Material), materials/elastic_mats.py
(``elastic``, ``elasticfrict``), material_matchmakers/harmonic.py:39 and
linear.py:32, materials/material_table.py:79-196.  Tables hold float32 numpy
arrays on the host (float64 with ``jax_enable_x64``); the force laws turn them
into small device tables.
"""

from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Any, Sequence

import numpy as np

from .config import real_dtype
from .factory import Factory


@dataclass
class Material(Factory):
    """Base material: every field becomes a column of the MaterialTable."""

    density: float


@Material.register("elastic")
@dataclass
class Elastic(Material):
    young: float
    poisson: float


@Material.register("elasticfrict")
@dataclass
class ElasticFriction(Material):
    young: float
    poisson: float
    mu: float
    e: float
    mu_r: float = 0.0


class MaterialMatchmaker(Factory):
    """Pair-property rule (material_matchmakers/__init__.py)."""

    @staticmethod
    def get_effective_property(p1: np.ndarray, p2: np.ndarray) -> np.ndarray:  # pragma: no cover
        raise NotImplementedError


@MaterialMatchmaker.register("harmonic")
class HarmonicMaterialMatchmaker(MaterialMatchmaker):
    """2 p1 p2 / (p1 + p2), 0 if either is 0 (harmonic.py:39-48)."""

    @staticmethod
    def get_effective_property(p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
        ft = np.result_type(p1, p2).type
        z = (p1 == 0.0) | (p2 == 0.0)
        s1 = np.where(z, ft(1.0), p1)
        s2 = np.where(z, ft(1.0), p2)
        return np.where(z, ft(0.0), 2.0 * s1 * s2 / (s1 + s2)).astype(ft)


@MaterialMatchmaker.register("linear")
class LinearMaterialMatchmaker(MaterialMatchmaker):
    """(p1 + p2) / 2 (linear.py:32-36)."""

    @staticmethod
    def get_effective_property(p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
        return ((p1 + p2) / 2).astype(np.result_type(p1, p2))


class MaterialTable:
    """Scalar properties (M,) and effective pair properties (M, M)."""

    def __init__(self, props: dict, pair: dict, matcher: MaterialMatchmaker):
        self.props = props
        self.pair = pair
        self.matcher = matcher
        self._device_cache: dict = {}

    @staticmethod
    def from_materials(
        mats: Sequence[Material], *, matcher: MaterialMatchmaker | None = None, fill: float = 0.0
    ) -> "MaterialTable":
        """material_table.py:79-128."""
        if not isinstance(mats, (list, tuple)) or not all(isinstance(m, Material) for m in mats):
            raise TypeError("mats must be a sequence of Material instances")
        keys: list[str] = []
        for m in mats:
            for f in fields(m):
                if f.name not in keys:
                    keys.append(f.name)
        scal = {k: np.asarray([getattr(m, k, fill) for m in mats], dtype=real_dtype()) for k in keys}
        if matcher is None:
            matcher = MaterialMatchmaker.create("harmonic")
        pair = {f"{k}_eff": matcher.get_effective_property(a[:, None], a[None, :]) for k, a in scal.items()}
        return MaterialTable(scal, pair, matcher)

    def __getattr__(self, item: str) -> Any:
        if item in ("props", "pair", "matcher", "_device_cache") or item.startswith("__"):
            raise AttributeError(item)
        if item in self.props:
            return self.props[item]
        if item in self.pair:
            return self.pair[item]
        raise AttributeError(item)

    def __len__(self) -> int:
        return int(next(iter(self.props.values())).shape[0])

    def has(self, name: str) -> bool:
        return name in self.props or name in self.pair


__all__ = [
    "Material",
    "Elastic",
    "ElasticFriction",
    "MaterialMatchmaker",
    "HarmonicMaterialMatchmaker",
    "LinearMaterialMatchmaker",
    "MaterialTable",
]
