"""Per-particle state (jaxdem/state.py:103-867), held as torch tensors.

Field names, shapes and defaults follow the reference ``State``: ``pos_c``,
``pos_p``, ``vel``, ``force`` are (N, dim); ``q`` is a unit quaternion;
``ang_vel``, ``torque`` and ``inertia`` are (N, 3) in 3D and (N, 1) in 2D;
``rad``, ``_rad``, ``volume``, ``mass``, ``clump_id``, ``mat_id``,
``species_id`` and ``fixed`` are (N,); ``bond_id`` is (N, W) padded with -1.
Floats are float32 (the reference's default precision; float64 when
``jax_enable_x64`` is set in ``jaxdem_amd.config``), integers int32.

The GPU time step covers spheres (``pos_p == 0`` and every particle its own
clump; ``State.create`` records whether that holds in ``is_sphere_state``)
and clumps of spheres: rigid bodies whose members share a ``clump_id`` and
carry their offsets ``pos_p`` from the clump centre, each member holding the
clump's ``pos_c``, velocities, ``q``, total mass and principal inertia (every
collider and domain).  The reference's clump,
facet and mesh builders (``add_clump`` and friends) are out of scope.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, fields, replace
from typing import Any, Sequence

import numpy as np
import torch

from .config import real_dtype, torch_real


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


class Quaternion:
    """Orientation quaternions stored as one (N, 4) tensor [w, x, y, z].

    ``w`` and ``xyz`` are views with the shapes of the reference
    (utils/quaternion.py:20-60): (N, 1) and (N, 3).
    """

    __slots__ = ("data",)

    def __init__(self, data: torch.Tensor):
        if data.shape[-1] != 4:
            raise ValueError(f"quaternion data must have shape (..., 4), got {tuple(data.shape)}")
        self.data = data

    @staticmethod
    def create(w=None, xyz=None, *, device=None) -> "Quaternion":
        dev = device or default_device()
        ft = torch_real()
        w = torch.ones((1, 1)) if w is None else torch.as_tensor(np.asarray(w), dtype=ft)
        if w.ndim == 0:
            w = w[None]
        xyz = torch.zeros((1, 3)) if xyz is None else torch.as_tensor(np.asarray(xyz), dtype=ft)
        return Quaternion(torch.cat([w.to(ft), xyz.to(ft)], dim=-1).to(dev))

    @property
    def w(self) -> torch.Tensor:
        return self.data[..., 0:1]

    @property
    def xyz(self) -> torch.Tensor:
        return self.data[..., 1:4]

    def clone(self) -> "Quaternion":
        return Quaternion(self.data.clone())

    @staticmethod
    def rotate(q: "Quaternion", v: torch.Tensor) -> torch.Tensor:
        """Body -> lab rotation (quaternion.py:193-242)."""
        if v.shape[-1] == 2:
            qw, qz = q.w, q.data[..., 3:4]
            c = qw * qw - qz * qz
            s = 2.0 * qw * qz
            return torch.cat([c * v[..., 0:1] - s * v[..., 1:2], s * v[..., 0:1] + c * v[..., 1:2]], dim=-1)
        t = 2.0 * torch.linalg.cross(q.xyz, v, dim=-1)
        return v + q.w * t + torch.linalg.cross(q.xyz, t, dim=-1)

    @staticmethod
    def rotate_back(q: "Quaternion", v: torch.Tensor) -> torch.Tensor:
        """Lab -> body rotation: rotate by the conjugate (quaternion.py:247-268)."""
        return Quaternion.rotate(Quaternion(torch.cat([q.w, -q.xyz], dim=-1)), v)

    def __repr__(self) -> str:
        return f"Quaternion(shape={tuple(self.data.shape)})"


def _hypersphere_volume(rad: np.ndarray, dim: int) -> np.ndarray:
    """state.py: volume of a sphere (3D) or area of a disk (2D)."""
    if dim == 3:
        return (4.0 / 3.0) * math.pi * rad**3
    return math.pi * rad**2


def _np(x, dtype) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.asarray(x, dtype=dtype)


def _symmetric_bonds(bond_id, N: int) -> np.ndarray:
    """Bond rows made symmetric and -1 padded (state.py:705-765)."""
    try:
        conn = np.asarray(bond_id, dtype=np.int64)
        if conn.ndim == 1:
            conn = conn[:, None]
    except (ValueError, TypeError):
        rows = [[int(v) for v in np.asarray(r).ravel().tolist()] for r in list(bond_id)[:N]]
        width = max(max((len(r) for r in rows), default=0), 1)
        conn = np.full((N, width), -1, np.int64)
        for i, r in enumerate(rows):
            conn[i, : len(r)] = r
    w = max(conn.shape[-1], 1)
    rows_idx = np.repeat(np.arange(N), conn.shape[-1])
    cols = conn.reshape(-1)
    ok = (cols >= 0) & (cols < N)
    i_idx, j_idx = rows_idx[ok], cols[ok]
    if i_idx.size == 0:
        return np.full((N, w), -1, np.int64)
    pairs = np.unique(np.stack((np.concatenate((i_idx, j_idx)), np.concatenate((j_idx, i_idx))), axis=1), axis=0)
    counts = np.bincount(pairs[:, 0], minlength=N)
    max_deg = max(int(counts.max()), w)
    out = np.full((N, max_deg), -1, np.int64)
    starts = np.cumsum(counts) - counts
    slot = np.arange(pairs.shape[0]) - starts[pairs[:, 0]]
    out[pairs[:, 0], slot] = pairs[:, 1]
    return out


@dataclass
class State:
    """Sphere state; see the module docstring for shapes and dtypes."""

    pos_c: torch.Tensor
    pos_p: torch.Tensor
    vel: torch.Tensor
    force: torch.Tensor
    q: Quaternion
    ang_vel: torch.Tensor
    torque: torch.Tensor
    rad: torch.Tensor
    _rad: torch.Tensor
    volume: torch.Tensor
    mass: torch.Tensor
    inertia: torch.Tensor
    clump_id: torch.Tensor
    bond_id: torch.Tensor
    mat_id: torch.Tensor
    species_id: torch.Tensor
    fixed: torch.Tensor
    is_sphere_state: bool = field(default=True)
    has_bonds: bool = field(default=False)

    # -- shape --------------------------------------------------------------
    @property
    def N(self) -> int:
        return int(self.pos_c.shape[-2])

    @property
    def dim(self) -> int:
        return int(self.pos_c.shape[-1])

    @property
    def shape(self) -> tuple:
        return tuple(self.pos_c.shape)

    @property
    def batch_size(self) -> int:
        return 1 if self.pos_c.ndim < 3 else int(self.pos_c.shape[-3])

    @property
    def device(self) -> torch.device:
        return self.pos_c.device

    @property
    def _pos_p_rot(self) -> torch.Tensor:
        """R(q) pos_p (state.py:264-273)."""
        if self.is_sphere_state:
            return torch.zeros_like(self.pos_p)
        return Quaternion.rotate(self.q, self.pos_p)

    @property
    def pos(self) -> torch.Tensor:
        """Sphere centres ``pos_c + R(q) pos_p`` (state.py:295-304)."""
        if self.is_sphere_state:
            return self.pos_c
        return self.pos_c + self._pos_p_rot

    # -- construction -------------------------------------------------------
    @staticmethod
    def create(
        pos=None,
        *,
        dim: int | None = None,
        pos_p=None,
        vel=None,
        force=None,
        q=None,
        ang_vel=None,
        torque=None,
        rad=None,
        _rad=None,
        volume=None,
        mass=None,
        inertia=None,
        clump_id=None,
        bond_id=None,
        mat_id=None,
        species_id=None,
        fixed=None,
        mat_table=None,
        device=None,
    ) -> "State":
        """``State.create`` (state.py:371-867) for spheres."""
        ft = real_dtype()  # float32, or float64 with jax_enable_x64
        if pos is None:
            d = 3 if dim is None else int(dim)
            pos_c = np.zeros((0, d), ft)
        else:
            pos_c = _np(pos, ft)
        if pos_c.ndim != 2 or pos_c.shape[-1] not in (2, 3):
            raise ValueError(f"`pos` must have shape (N, 2|3); got {pos_c.shape}")
        N, D = pos_c.shape
        A = 1 if D == 2 else 3
        pos_p_np = np.zeros_like(pos_c) if pos_p is None else _np(pos_p, ft)
        vel_np = np.zeros_like(pos_c) if vel is None else _np(vel, ft)
        force_np = np.zeros_like(pos_c) if force is None else _np(force, ft)
        if q is None:
            q_np = np.zeros((N, 4), ft)
            q_np[:, 0] = 1.0
        elif isinstance(q, Quaternion):
            q_np = _np(q.data, ft)
        else:
            q_np = _np(q, ft)
            if q_np.ndim < 2 or q_np.shape[-1] != 4:
                raise ValueError(f"`q` array must have shape (..., N, 4). Got shape={q_np.shape}.")
        ang_np = np.zeros((N, A), ft) if ang_vel is None else _np(ang_vel, ft).reshape(N, A)
        torque_np = np.zeros((N, A), ft) if torque is None else _np(torque, ft).reshape(N, A)
        rad_np = np.ones(N, ft) if rad is None else _np(rad, ft).reshape(N)
        vol_np = _hypersphere_volume(rad_np.astype(ft), D).astype(ft) if volume is None else _np(volume, ft)
        clump_np = np.arange(N, dtype=np.int64) if clump_id is None else _np(clump_id, np.int64)
        bond_np = np.full((N, 1), -1, np.int64) if bond_id is None else _symmetric_bonds(bond_id, N)
        mat_np = np.zeros(N, np.int64) if mat_id is None else _np(mat_id, np.int64).reshape(N)
        spec_np = np.zeros(N, np.int64) if species_id is None else _np(species_id, np.int64).reshape(N)
        fixed_np = np.zeros(N, bool) if fixed is None else np.asarray(fixed, bool).reshape(N)
        if mat_table is not None:
            dens = np.asarray(mat_table.density, ft)
            mass_np = (dens[mat_np] * vol_np).astype(ft)
        else:
            mass_np = np.ones(N, ft) if mass is None else _np(mass, ft).reshape(N)
        coeff = ft(0.5 if D == 2 else 0.4)
        if inertia is None:
            inertia_np = ((coeff * mass_np * rad_np**2)[:, None] * np.ones((N, A), ft)).astype(ft)
        else:
            inertia_np = _np(inertia, ft).reshape(N, A)
        if N > 0:
            _, clump_np = np.unique(clump_np, return_inverse=True)
        _rad_np = rad_np.copy() if _rad is None else _np(_rad, ft)
        sphere = bool(np.all(pos_p_np == 0)) and (N == 0 or np.unique(clump_np).size == N)
        dev = torch.device(device) if device is not None else default_device()

        def T(a, dt=torch_real()):
            return torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=dt)

        return State(
            pos_c=T(pos_c),
            pos_p=T(pos_p_np),
            vel=T(vel_np),
            force=T(force_np),
            q=Quaternion(T(q_np)),
            ang_vel=T(ang_np),
            torque=T(torque_np),
            rad=T(rad_np),
            _rad=T(_rad_np),
            volume=T(vol_np),
            mass=T(mass_np),
            inertia=T(inertia_np),
            clump_id=T(clump_np, torch.int32),
            bond_id=T(bond_np, torch.int32),
            mat_id=T(mat_np, torch.int32),
            species_id=T(spec_np, torch.int32),
            fixed=T(fixed_np, torch.bool),
            is_sphere_state=sphere,
            has_bonds=bool(np.any(bond_np >= 0)),
        )

    def to(self, device) -> "State":
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, torch.Tensor):
                kw[f.name] = v.to(device)
            elif isinstance(v, Quaternion):
                kw[f.name] = Quaternion(v.data.to(device))
            else:
                kw[f.name] = v
        return State(**kw)

    def clone(self) -> "State":
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, torch.Tensor):
                kw[f.name] = v.clone()
            elif isinstance(v, Quaternion):
                kw[f.name] = v.clone()
            else:
                kw[f.name] = v
        return State(**kw)

    def with_kinematics(self) -> "State":
        """Copy of the fields one time step writes; the rest is shared."""
        return replace(
            self,
            pos_c=self.pos_c.clone(),
            vel=self.vel.clone(),
            force=self.force.clone(),
            q=self.q.clone(),
            ang_vel=self.ang_vel.clone(),
            torque=self.torque.clone(),
        )

    @staticmethod
    def stack(states: Sequence["State"]) -> "State":
        """Stack snapshots along a new leading axis (state.py:1132-1202)."""
        s0 = states[0]
        kw = {}
        for f in fields(s0):
            v = getattr(s0, f.name)
            if isinstance(v, torch.Tensor):
                kw[f.name] = torch.stack([getattr(s, f.name) for s in states])
            elif isinstance(v, Quaternion):
                kw[f.name] = Quaternion(torch.stack([getattr(s, f.name).data for s in states]))
            else:
                kw[f.name] = v
        return State(**kw)

    def numpy(self) -> dict:
        """Host copy of the fields (q as (N, 4))."""
        out: dict[str, Any] = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, torch.Tensor):
                out[f.name] = v.detach().cpu().numpy()
            elif isinstance(v, Quaternion):
                out[f.name] = v.data.detach().cpu().numpy()
        return out


__all__ = ["State", "Quaternion"]
