"""Colliders (jaxdem/colliders/__init__.py, naive.py, cell_list.py,
neighbor_list.py) running on the GPU through libjdem_hip.so.

Grid parameters are computed here on the host in float32 with the operations
of ``_grid_params`` (colliders/_partition.py:55-99) and handed to the
kernels, so the cell grid is the reference's grid.  Cell coordinates are
clamped into the grid (a particle outside a non-periodic box lands in the
boundary cell instead of aliasing another cell's hash); for particles inside
the box this is the reference's hash.
"""

from __future__ import annotations

import copy
import ctypes
import math
import warnings
from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np
import torch

from . import _lib as L
from .factory import Factory, _normalize_key

F32 = np.float32


def grid_params(box_size: np.ndarray, cell_size, periodic: bool, ft=F32):
    """``_grid_params`` (_partition.py:55-99) in float32 (``ft`` float64
    for the float64 step).

    Returns (dims int64 (D,), strides int64 (D,), cell (D,), overflow).
    """
    box = np.asarray(box_size, ft)
    cs = ft(cell_size)
    if periodic:
        dims = np.maximum(np.floor(box / cs).astype(np.int64), 1)
        cell = (box / dims.astype(ft)).astype(ft)
    else:
        dims = np.maximum(np.ceil(box / cs).astype(np.int64), 1)
        cell = np.full(box.shape, cs, ft)
    strides = np.concatenate([[1], np.cumprod(dims[:-1])]).astype(np.int64)
    overflow = float(np.prod(dims.astype(np.float64))) > float(np.iinfo(np.int32).max)
    return dims, strides, cell, overflow


def grid_struct(box_size, cell_size, periodic: bool, search_range: int, f64: bool = False) -> tuple:
    """jdem_grid (jdem_grid_f64 with ``f64``) and the overflow flag."""
    dims, strides, cell, overflow = grid_params(box_size, cell_size, periodic, np.float64 if f64 else F32)
    g = L.JdemGridF64() if f64 else L.JdemGrid()
    D = len(dims)
    for a in range(3):
        g.dims[a] = int(dims[a]) if a < D else 1
        g.strides[a] = int(strides[a]) if a < D else 0
        g.cell_size[a] = float(cell[a]) if a < D else 1.0
    g.search_range = int(search_range)
    g.periodic = 1 if periodic else 0
    g.n_cells = int(np.prod(dims)) if not overflow else -1
    return g, overflow


def valid_interaction_mask(clump_i, clump_j, bond_id_i, index_j, interact_same_bond_id=False):
    """colliders/__init__.py:225-243 (torch)."""
    is_bonded = torch.any(bond_id_i == index_j[..., None], dim=-1)
    m1 = (clump_i != clump_j).to(torch.int32)
    m2 = ((~is_bonded) | bool(interact_same_bond_id)).to(torch.int32)
    return m1 * m2


@dataclass
class Collider(Factory):
    """Base collider; registered under "" as the no-op collider
    (colliders/__init__.py:22-222)."""

    overflow: bool = field(default=False, kw_only=True)
    kind = L.COLLIDER_NONE

    @staticmethod
    def compute_force(state, system):
        from .system import _collider_compute_force

        return _collider_compute_force(state, system)

    @staticmethod
    def compute_potential_energy(state, system):
        from .system import _collider_potential_energy

        return _collider_potential_energy(state, system)

    @staticmethod
    def create_neighbor_list(state, system, cutoff: float, max_neighbors: int):
        nl = torch.full((state.N, max_neighbors), -1, dtype=torch.int32, device=state.device)
        return state, system, nl, False


Collider.register("")(Collider)


@Collider.register("naive")
@dataclass
class NaiveSimulator(Collider):
    """O(N^2) all-pairs collider (naive.py:20-235)."""

    kind = L.COLLIDER_NAIVE

    @staticmethod
    def create_neighbor_list(state, system, cutoff: float, max_neighbors: int):
        """naive.py:122-185: rows in descending particle id."""
        from .system import _neighbor_list

        return _neighbor_list(state, system, L.COLLIDER_NAIVE, None, cutoff, max_neighbors)


@Collider.register("CellList")
@dataclass
class DynamicCellList(Collider):
    """Implicit cell list (cell_list.py:267-595): hash + stable sort by cell +
    cell ranges built on the GPU every call; 'ij'-ordered stencil."""

    neighbor_mask: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int64))
    cell_size: float = 1.0
    kind = L.COLLIDER_CELL_LIST

    @property
    def search_range(self) -> int:
        return max(int(np.max(np.abs(self.neighbor_mask))) if self.neighbor_mask.size else 1, 1)

    @classmethod
    def Create(cls, state, cell_size=None, search_range=None, box_size=None):
        """cell_list.py:375-432, in the dtype of the radii."""
        ft = np.float64 if state._rad.dtype == torch.float64 else F32
        rad = state._rad.detach().cpu().numpy().astype(ft)
        min_rad, max_rad = ft(rad.min()), ft(rad.max())
        alpha = max_rad / min_rad
        if cell_size is None:
            cell_size = ft(2.0) * max_rad if alpha < 2.5 else ft(0.5) * max_rad
        cell_size = ft(cell_size)
        if box_size is not None:
            box = np.asarray(box_size, ft)
            for _ in range(2):
                if search_range is None:
                    sr = max(1, int(np.ceil(ft(2.0) * max_rad / cell_size)))
                else:
                    sr = int(search_range)
                dims = np.maximum(np.floor(box / cell_size).astype(np.int64), 2 * sr + 1)
                cell_size = ft(np.min(box / dims.astype(ft)))
        if search_range is None:
            search_range = max(1, int(np.ceil(ft(2.0) * max_rad / cell_size)))
        sr = int(search_range)
        r = np.arange(-sr, sr + 1)
        mesh = np.meshgrid(*([r] * state.dim), indexing="ij")
        mask = np.stack([m.ravel() for m in mesh], axis=1).astype(np.int64)
        return cls(neighbor_mask=mask, cell_size=float(cell_size))

    def grid(self, system, f64: bool = False) -> tuple:
        return grid_struct(system.domain.box_size, self.cell_size, system.domain.periodic, self.search_range, f64)

    @staticmethod
    def create_neighbor_list(state, system, cutoff: float, max_neighbors: int):
        """cell_list.py:500-595: the stencil is inflated to the cutoff."""
        from .system import _neighbor_list

        col = system.collider
        sr = col.search_range
        cs = max(F32(col.cell_size), F32(F32(cutoff) / F32(sr)))
        g, ovf = grid_struct(system.domain.box_size, cs, system.domain.periodic, sr)
        st, sy, nl, o = _neighbor_list(state, system, L.COLLIDER_CELL_LIST, g, cutoff, max_neighbors)
        return st, sy, nl, bool(o or ovf)


class _OnDevice:
    """Marks a NeighborList buffer that lives in a resident working set on the
    device (the step kernel's layouts) until it is read."""

    __slots__ = ()

    def __repr__(self) -> str:
        return "<in the resident working set>"


ON_DEVICE = _OnDevice()
_RESIDENT_FIELDS = frozenset(("neighbor_list", "old_pos", "history"))


def _versions(*ts) -> tuple:
    out = []
    for t in ts:
        if isinstance(t, dict):
            out.extend((id(v), v._version) for v in t.values())
        elif t is not None:
            out.append((id(t), t._version))
    return tuple(out)


@Collider.register("NeighborList")
@dataclass
class NeighborList(Collider):
    """Verlet list with skin, rebuilt through a secondary collider when any
    particle moved more than skin/2 since the last build; the carrier of pair
    history (neighbor_list.py:137-669).

    After a step on the GPU the list, build positions and history stay in the
    step's resident working set (work order, slot-major rows, springs by
    rank); reading ``neighbor_list``, ``old_pos`` or ``history`` converts them
    to the reference's layouts then, and the next step starts from the working
    set as long as those arrays are unmodified."""

    secondary_collider: Any = None
    neighbor_list: torch.Tensor | None = None
    old_pos: torch.Tensor | None = None
    n_build_times: int = 0
    cutoff: float = 0.0
    skin: float = 0.0
    max_neighbors: int = 0
    history: Any = field(default=None, kw_only=True)
    kind = L.COLLIDER_NEIGHBOR_LIST

    @classmethod
    def Create(
        cls,
        state,
        cutoff: float,
        skin: float | None = None,
        skin_fraction: float | None = None,
        max_neighbors: int | None = None,
        number_density: float = 1.0,
        safety_factor: float = 1.2,
        secondary_collider_type: str = "CellList",
        secondary_collider_kw: dict | None = None,
    ):
        """neighbor_list.py:287-480."""
        if skin is not None and skin_fraction is not None:
            raise ValueError("Pass either `skin` (absolute distance) or `skin_fraction` (fraction of the cutoff), not both.")
        if skin is None:
            skin_fraction = 0.05 if skin_fraction is None else skin_fraction
            skin_val = float(skin_fraction) * float(cutoff)
        else:
            skin_val = float(skin)
        list_cutoff = float(cutoff) + skin_val
        rad = state._rad.detach().cpu().numpy().astype(np.float64)
        pos = state.pos.detach().cpu().numpy().astype(np.float64)
        dim = state.dim
        N = state.N
        if N > 0:
            max_rad = float(rad.max())
            box = np.maximum(pos.max(0) - pos.min(0) + 2.0 * max_rad, 1.0)
            density_est = N / float(np.prod(box))
            eff_density = max(number_density, density_est)
            r_eff = 0.9 * float(rad.min())
            packing_limit = 0.91 if dim == 2 else 0.74
            max_possible = int(math.ceil(packing_limit * ((list_cutoff + r_eff) / r_eff) ** dim))
            r_eff_mean = 0.9 * float(rad.mean())
            typical = int(math.ceil(((list_cutoff + r_eff_mean) / r_eff_mean) ** dim))
        else:
            eff_density, max_possible, typical = number_density, 0, 0
        user_max = max_neighbors is not None
        if max_neighbors is None:
            vol = math.pi * list_cutoff**dim * (1.0 if dim == 2 else 4.0 / 3.0)
            max_neighbors = max(int(math.ceil(safety_factor * vol * eff_density)), typical)
        requested = max_neighbors
        max_neighbors = max(min(max_neighbors, max_possible, N), 0)
        if user_max and max_neighbors < requested:
            warnings.warn(
                f"NeighborList max_neighbors={requested} clamped to {max_neighbors} (bounded by N={N} and the "
                f"physical packing limit of {max_possible} neighbors within the search radius).",
                stacklevel=2,
            )
        skw = dict(secondary_collider_kw or {})
        from inspect import signature

        sub = Collider._registry.get(_normalize_key(secondary_collider_type))
        fn = getattr(sub, "Create", None)
        if fn is not None and "state" in signature(fn).parameters:
            skw["state"] = state
        if _normalize_key(secondary_collider_type) == "celllist" and "cell_size" not in skw:
            skw["cell_size"] = list_cutoff
        sec = Collider.create(secondary_collider_type, **skw)
        ft = np.float64 if state.pos_c.dtype == torch.float64 else F32  # the state's precision
        return cls(
            secondary_collider=sec,
            neighbor_list=torch.full((N, max_neighbors), -1, dtype=torch.int32, device=state.device),
            old_pos=state.pos.clone(),
            n_build_times=0,
            cutoff=float(ft(cutoff)),
            skin=float(ft(skin_val)),
            overflow=False,
            max_neighbors=int(max_neighbors),
        )

    def __getattribute__(self, name):
        if name in _RESIDENT_FIELDS:
            d = object.__getattribute__(self, "__dict__")
            if d.get(name) is ON_DEVICE:
                object.__getattribute__(self, "_export_resident")()
        return object.__getattribute__(self, name)

    def _export_resident(self) -> None:
        d = self.__dict__
        nl, old, hist = d["_resident"].export()
        d["neighbor_list"], d["old_pos"] = nl, old
        if d.get("history") is ON_DEVICE:
            d["history"] = hist
        d["_resident_versions"] = _versions(nl, old, d["history"])

    def resident_for_step(self):
        """The working set a step may start from: the one the last step left,
        unless the exported arrays were replaced or modified since."""
        d = self.__dict__
        res = d.get("_resident")
        if res is None:
            return None
        if d.get("neighbor_list") is ON_DEVICE:
            return res
        if d.get("_resident_versions") == _versions(d["neighbor_list"], d["old_pos"], d["history"]):
            return res
        return None

    def with_resident(self, res, n_build_times: int, overflow: bool) -> "NeighborList":
        """This collider after a step whose list, build positions and history
        are the working set ``res``."""
        out = copy.copy(self)
        d = out.__dict__
        d["neighbor_list"] = ON_DEVICE
        d["old_pos"] = ON_DEVICE
        d["history"] = ON_DEVICE if res.hist else None
        d["_resident"] = res
        d["_resident_versions"] = None
        d["n_build_times"] = int(n_build_times)
        d["overflow"] = bool(overflow)
        return out

    def secondary_grid(self, system, f64: bool = False) -> tuple:
        """Grid of the secondary cell list inflated to cutoff + skin
        (cell_list.py:537-538), in float32 (float64 with ``f64``)."""
        ft = np.float64 if f64 else F32
        sec = self.secondary_collider
        lc = ft(ft(self.cutoff) + ft(self.skin))
        sr = sec.search_range
        cs = max(ft(sec.cell_size), ft(lc / ft(sr)))
        return grid_struct(system.domain.box_size, cs, system.domain.periodic, sr, f64)

    @staticmethod
    def create_neighbor_list(state, system, cutoff: float, max_neighbors: int):
        """neighbor_list.py:480-540: the cached list, refreshed if stale."""
        from .system import _nl_refresh

        state, system = _nl_refresh(state, system)
        col = system.collider
        return state, system, col.neighbor_list, col.overflow


def refresh_collider(state, collider: Collider) -> Collider:
    """colliders/__init__.py:246-321: rebuild size-dependent buffers."""
    from inspect import signature

    name = collider.type_name.lower()
    if name not in {"neighborlist", "celllist"}:
        return collider
    if name == "neighborlist":
        sec = collider.secondary_collider
        skw: dict[str, Any] = {"state": state}
        if hasattr(sec, "cell_size"):
            skw["cell_size"] = sec.cell_size
            skw["search_range"] = sec.search_range
        return type(collider).Create(
            state=state,
            cutoff=collider.cutoff,
            skin=collider.skin,
            max_neighbors=collider.max_neighbors,
            secondary_collider_type=sec.type_name,
            secondary_collider_kw=skw if sec.type_name.lower() == "celllist" else {},
        )
    kwargs: dict[str, Any] = {}
    for p in signature(type(collider).Create).parameters:
        if p == "state":
            kwargs[p] = state
        elif p == "search_range":
            kwargs[p] = collider.search_range
        elif hasattr(collider, p):
            kwargs[p] = getattr(collider, p)
    return type(collider).Create(**kwargs)


__all__ = [
    "Collider",
    "NaiveSimulator",
    "DynamicCellList",
    "NeighborList",
    "refresh_collider",
    "valid_interaction_mask",
    "grid_params",
]
