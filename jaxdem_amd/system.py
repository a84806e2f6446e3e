"""System: configuration and the drivers of the DEM time step
(jaxdem/system.py:60-748).

``System.step`` advances ``n`` steps in the fixed order of ``_step_once``
(system.py:60-82).  When every plugin is a built-in one, the whole call is a
single ``jdem_system_step``: one fused HIP kernel per time step (plus the
cell-list or neighbour-list rebuilds).  With user plugins, user force
functions or pre/post-step actions, it runs the same order stage by stage,
each built-in stage still a HIP kernel.

Semantics are functional as in the reference: ``step`` returns new
``(state, system)`` objects and leaves its arguments unchanged.
"""

from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass, field, replace
from typing import Any, Callable, Sequence

import numpy as np
import torch

from . import _lib as L
from .colliders import Collider, DynamicCellList, NaiveSimulator, NeighborList
from .domains import Domain, FreeDomain, PeriodicDomain, ReflectDomain
from .factory import _normalize_key
from .forces import (
    CundallStrackForce,
    ForceManager,
    ForceModel,
    ForceRouter,
    HertzFrictionForce,
    HertzianForce,
    LawCombiner,
    SpringForce,
    SpringFrictionForce,
    law_struct,
)
from .integrators import (
    DirectEuler,
    LinearIntegrator,
    RotationIntegrator,
    Spiral,
    VelocityVerlet,
    VelocityVerletSpiral,
)
from .config import real_dtype
from .materials import Material, MaterialTable
from .state import Quaternion, State

F32 = np.float32


def _save_state_system(state, system):
    return state, system


def save_state(state, system):
    """A ``save_fn`` for ``System.trajectory_rollout`` that keeps the state
    of every frame and not the system.  With the fused step (built-in
    integrators, domain other than FreeDomain, NeighborList collider, no
    user hooks) the frames are written on the device by one
    ``jdem_rollout`` call, with no host round trip between frames."""
    return state


@dataclass
class System:
    """The simulation configuration (system.py:123-226)."""

    linear_integrator: LinearIntegrator
    rotation_integrator: RotationIntegrator
    collider: Collider
    domain: Domain
    force_manager: ForceManager
    bonded_force_model: Any
    force_model: ForceModel
    mat_table: MaterialTable
    dt: Any
    time: Any
    dim: int
    step_count: int
    key: Any
    interact_same_bond_id: bool
    user_pre_step_actions: Callable = _save_state_system
    user_post_step_actions: Callable = _save_state_system
    minimizer: Any = None
    target_fn: Any = None

    # ------------------------------------------------------------------
    @staticmethod
    def create(
        state_shape: tuple | None = None,
        *,
        state: State | None = None,
        dt: float = 0.005,
        time: float = 0.0,
        linear_integrator_type: str | None = "verlet",
        rotation_integrator_type: str | None = "verletspiral",
        collider_type: str = "naive",
        domain_type: str = "free",
        bonded_force_model_type: str | None = None,
        bonded_force_model_kw: dict | None = None,
        bonded_force_manager_kw: dict | None = None,
        bonded_force_model: Any = None,
        force_model_type: str = "spring",
        force_manager_kw: dict | None = None,
        mat_table: MaterialTable | None = None,
        linear_integrator: LinearIntegrator | None = None,
        rotation_integrator: RotationIntegrator | None = None,
        collider: Collider | None = None,
        domain: Domain | None = None,
        force_model: ForceModel | None = None,
        force_manager: ForceManager | None = None,
        linear_integrator_kw: dict | None = None,
        rotation_integrator_kw: dict | None = None,
        collider_kw: dict | None = None,
        domain_kw: dict | None = None,
        force_model_kw: dict | None = None,
        seed: int = 0,
        key: Any = None,
        interact_same_bond_id: bool = False,
        user_pre_step_actions: Callable | None = None,
        user_post_step_actions: Callable | None = None,
        minimizer: Any = None,
        minimizer_kw: dict | None = None,
        target_fn: Callable | None = None,
    ) -> "System":
        """system.py:228-604."""
        if state is not None:
            if state_shape is not None and tuple(state_shape) != tuple(state.shape):
                raise ValueError(
                    f"state_shape {tuple(state_shape)} does not match the shape of the provided state {tuple(state.shape)}."
                )
            state_shape = tuple(state.shape)
        if state_shape is None:
            raise TypeError("System.create requires either `state_shape` or `state`.")
        dim = int(state_shape[-1])
        if linear_integrator_type is None:
            linear_integrator_type = ""
        if rotation_integrator_type is None:
            rotation_integrator_type = ""
        if bonded_force_model_type is not None or bonded_force_model is not None:
            raise NotImplementedError("bonded force models are outside the GPU DEM time step of this package")
        collider_kw = dict(collider_kw or {})
        domain_kw = dict(domain_kw or {})
        force_model_kw = dict(force_model_kw or {})
        fm_kw = {"gravity": None, "force_functions": ()} if force_manager_kw is None else dict(force_manager_kw)
        if mat_table is None:
            mat_table = MaterialTable.from_materials(
                [Material.create("elastic", density=0.27, young=1.0e4, poisson=0.3)]
            )
        if force_model is None:
            force_model = ForceModel.create(force_model_type, **force_model_kw)
        missing = [k for k in force_model.required_material_properties if not mat_table.has(k)]
        if missing:
            raise KeyError(f"MaterialTable lacks fields {missing}, required by the selected force model.")
        if force_manager is None:
            force_manager = ForceManager.create(tuple(state_shape), **fm_kw)
        if key is None:
            key = np.array([0, seed], dtype=np.uint32)
        if collider is None:
            if state is not None and "state" not in collider_kw:
                from inspect import signature

                sub = Collider._registry.get(_normalize_key(collider_type))
                fn = getattr(sub, "Create", None)
                if fn is not None and "state" in signature(fn).parameters:
                    collider_kw["state"] = state
            collider = Collider.create(collider_type, **collider_kw)
        if force_model.requires_history and collider.type_name.lower() != "neighborlist":
            raise ValueError(
                f"Force model '{force_model.type_name}' requires history tracking. "
                f"You must use the 'NeighborList' collider, but got '{collider.type_name}'."
            )
        if isinstance(collider, NeighborList) and collider.history is None and force_model.requires_history:
            collider = replace(collider, history=_alloc_history(collider, state_shape, force_model))
        return System(
            linear_integrator=(
                LinearIntegrator.create(linear_integrator_type, **dict(linear_integrator_kw or {}))
                if linear_integrator is None
                else linear_integrator
            ),
            rotation_integrator=(
                RotationIntegrator.create(rotation_integrator_type, **dict(rotation_integrator_kw or {}))
                if rotation_integrator is None
                else rotation_integrator
            ),
            collider=collider,
            domain=Domain.create(domain_type, dim=dim, **domain_kw) if domain is None else domain,
            force_manager=force_manager,
            bonded_force_model=None,
            force_model=force_model,
            mat_table=mat_table,
            dt=real_dtype()(dt),
            time=real_dtype()(time),
            dim=dim,
            step_count=0,
            key=key,
            interact_same_bond_id=bool(interact_same_bond_id),
            user_pre_step_actions=user_pre_step_actions or _save_state_system,
            user_post_step_actions=user_post_step_actions or _save_state_system,
            minimizer=(minimizer, dict(minimizer_kw or {})),
            target_fn=target_fn,
        )

    # ------------------------------------------------------------------
    @staticmethod
    def step(state: State, system: "System", *, n: int = 1) -> tuple:
        """Advance ``n`` steps (system.py:701-748)."""
        n = int(n)
        if state.batch_size > 1:
            raise NotImplementedError(
                "batched states step through jaxdem_amd.batching (segmented multi-system layout)"
            )
        if n <= 0:
            return state, system
        if _fusable(state, system):
            return _fused_step(state, system, n)
        for _ in range(n):
            state, system = _step_once(state, system)
        return state, system

    @staticmethod
    def trajectory_rollout(
        state: State,
        system: "System",
        *,
        n: int | None = None,
        stride: int = 1,
        strides=None,
        save_fn: Callable = _save_state_system,
        unroll: int = 2,
    ) -> tuple:
        """system.py:606-699: frames are saved after their steps; a leading 0
        stride records the initial state."""
        stride = int(stride)
        if strides is not None:
            ks = [int(k) for k in np.asarray(strides).reshape(-1).tolist()]
            if np.asarray(strides).ndim != 1:
                raise ValueError("`strides` must be 1D")
        else:
            if n is None:
                raise ValueError("`n` must be provided when `strides` is None.")
            ks = [stride] * int(n)
        if (
            save_fn is save_state
            and ks
            and min(ks) >= 0
            and state.batch_size == 1
            and state.N > 0
            and isinstance(system.collider, NeighborList)
            and not isinstance(system.domain, FreeDomain)
            and state.is_sphere_state
            and state.pos_c.dtype == torch.float32
            and _fusable(state, system)
        ):
            return _fused_rollout(state, system, ks)
        frames = []
        for k in ks:
            state, system = System.step(state, system, n=k)
            frames.append(save_fn(state, system))
        return state, system, _stack_frames(frames)

    def potential_energy(self, state: State) -> float:
        """Collider + force-manager potential energy in float64."""
        _, _, e = self.collider.compute_potential_energy(state, self)
        return e + ForceManager.compute_potential_energy(state, self)

    @staticmethod
    def minimize(state: State, system: "System", **kw):
        from .minimizers import minimize

        return minimize(state, system, **kw)


def _stack_frames(frames: list):
    if not frames:
        return []
    f0 = frames[0]
    if isinstance(f0, tuple) and len(f0) == 2 and isinstance(f0[0], State):
        return (State.stack([f[0] for f in frames]), [f[1] for f in frames])
    if isinstance(f0, State):
        return State.stack(frames)
    if isinstance(f0, torch.Tensor):
        return torch.stack(frames)
    if isinstance(f0, tuple):
        return tuple(_stack_frames([f[i] for f in frames]) for i in range(len(f0)))
    if isinstance(f0, dict):
        return {k: _stack_frames([f[k] for f in frames]) for k in f0}
    return frames


def _alloc_history(collider: NeighborList, state_shape, force_model) -> dict:
    N, D = int(state_shape[-2]), int(state_shape[-1])
    M = collider.max_neighbors
    dev = collider.neighbor_list.device
    return {
        "xi": torch.zeros((N, M, D), dtype=torch.float32, device=dev),
        "live": torch.zeros((N, max((M + 31) // 32, 1)), dtype=torch.int32, device=dev),
    }


def _clock_after(t0, dt, n: int):
    """time after n steps: n sequential additions of dt in the dtype of dt,
    float32 or float64 (system.py:73); ufunc.accumulate adds in order, each
    sum rounded."""
    if isinstance(dt, (np.float32, np.float64)):
        ft = type(dt)
    else:  # a Python number takes the default precision, as a weak type in jnp
        ft = real_dtype()
    if n <= 0:
        return ft(t0)
    seq = np.full(n + 1, ft(dt), ft)
    seq[0] = ft(t0)
    return ft(np.add.accumulate(seq, dtype=ft)[-1])


def _advance_clock(system: System, n: int) -> System:
    return replace(system, time=_clock_after(system.time, system.dt, n), step_count=system.step_count + n)


# ----------------------------------------------------------------------
# C structs
# ----------------------------------------------------------------------
_BUILTIN_LIN = (LinearIntegrator, VelocityVerlet, DirectEuler)
_BUILTIN_ROT = (RotationIntegrator, VelocityVerletSpiral, Spiral)
_BUILTIN_DOM = (Domain, FreeDomain, PeriodicDomain, ReflectDomain)
_BUILTIN_COL = (Collider, NaiveSimulator, DynamicCellList, NeighborList)
_BUILTIN_LAW = (SpringForce, HertzianForce, CundallStrackForce, SpringFrictionForce, HertzFrictionForce, LawCombiner,
                ForceRouter)


def _require_gpu(state: State) -> None:
    if not state.pos_c.is_cuda:
        raise RuntimeError(
            "jaxdem_amd runs the DEM time step as HIP kernels on a GPU; the state tensors are on "
            f"{state.pos_c.device} (create the State with device='cuda')"
        )


def _clump_arrays(state: State) -> tuple:
    """(pos_p, clump_id) pointers of a clump state, (None, None) for spheres.
    The kernels index per-clump tables by clump id, so the ids are checked
    to lie in [0, N) once per version of the array."""
    if state.is_sphere_state:
        return None, None
    cid = state.clump_id
    tag = (id(cid), cid._version)
    if state.__dict__.get("_clumps_checked") != tag:
        if state.N and (int(cid.min()) < 0 or int(cid.max()) >= state.N):
            raise ValueError("clump_id outside [0, N)")
        state.__dict__["_clumps_checked"] = tag
    return _ptr(state.pos_p), _ptr(cid)


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("state tensors must be contiguous")
    return t.data_ptr()


def _state_struct(state: State, cls=L.JdemState):
    """jdem_state (``cls`` L.JdemStateF64 for jdem_state_f64)."""
    s = cls()
    s.n = state.N
    s.dim = state.dim
    bw = int(state.bond_id.shape[-1]) if state.has_bonds else 0
    s.bond_width = bw
    s.pos = _ptr(state.pos_c)
    s.vel = _ptr(state.vel)
    s.ang_vel = _ptr(state.ang_vel)
    s.q = _ptr(state.q.data)
    s.force = _ptr(state.force)
    s.torque = _ptr(state.torque)
    s.rad = _ptr(state.rad)
    s.mass = _ptr(state.mass)
    s.inertia = _ptr(state.inertia)
    s.mat_id = _ptr(state.mat_id)
    s.bond_id = _ptr(state.bond_id) if bw else None
    s.fixed = _ptr(state.fixed)
    s.pos_p, s.clump_id = _clump_arrays(state)
    return s


def _collider_kind(col: Collider) -> int:
    return int(type(col).kind)


def _system_struct(state: State, system: System, keep: list, resident=None, for_collider: bool = True) -> L.JdemSystem:
    """The jdem_system of (state, system).  With ``resident`` a NeighborList's
    list, build positions and history come from that working set; without
    ``for_collider`` the collider is left out (kinetic quantities)."""
    _require_gpu(state)
    _require_f32(state, "this call")
    S = L.JdemSystem()
    S.state = _state_struct(state)
    S.domain = system.domain.to_c(state.dim)
    law, dev = law_struct(system.force_model, system.mat_table, state.device)
    keep.append(dev)
    _set_species(law, state)
    S.law = law
    S.lin_kind = int(type(system.linear_integrator).kind)
    S.rot_kind = int(type(system.rotation_integrator).kind)
    S.dt = float(F32(system.dt))
    S.interact_same_bond_id = 1 if system.interact_same_bond_id else 0
    g = np.zeros(3, F32)
    g[: state.dim] = np.asarray(system.force_manager.gravity, F32).reshape(-1)[: state.dim]
    for a in range(3):
        S.gravity[a] = float(g[a])
    fm = system.force_manager
    for name, cname in (("external_force", "ext_force"), ("external_force_com", "ext_force_com"), ("external_torque", "ext_torque")):
        t = getattr(fm, name)
        if t is not None:
            t = torch.as_tensor(t, dtype=torch.float32, device=state.device).contiguous()
            keep.append(t)
            setattr(S, cname, t.data_ptr())
    col = system.collider
    if not for_collider:
        S.collider_kind = L.COLLIDER_NONE
        return S
    S.collider_kind = _collider_kind(col)
    if isinstance(col, DynamicCellList):
        g, ovf = col.grid(system)
        if ovf:
            raise NotImplementedError("cell grid exceeds the int32 hash range (reference: hash_overflow)")
        S.grid = g
    if isinstance(col, NeighborList):
        nl = L.JdemNlist()
        nl.max_neighbors = col.max_neighbors
        nl.cutoff = float(col.cutoff)
        nl.skin = float(col.skin)
        if resident is None:
            nl.nl = _ptr(col.neighbor_list)
            nl.old_pos = _ptr(col.old_pos)
        else:
            nl.resident_in = resident.ptr
        nl.n_build_times = int(col.n_build_times)
        nl.overflow = 1 if col.overflow else 0
        sec = col.secondary_collider
        if isinstance(sec, DynamicCellList):
            nl.secondary = L.COLLIDER_CELL_LIST
            g2, ovf = col.secondary_grid(system)
            if ovf:
                raise NotImplementedError("secondary cell grid exceeds the int32 hash range")
            nl.grid = g2
        elif isinstance(sec, NaiveSimulator):
            nl.secondary = L.COLLIDER_NAIVE
        else:
            raise NotImplementedError(f"NeighborList secondary collider {type(sec).__name__} on the GPU")
        if resident is None and col.history is not None:
            nl.hist_xi = _ptr(col.history["xi"])
            nl.hist_live = _ptr(col.history["live"])
        S.nlist = nl
    return S


def _set_species(law, state: State) -> None:
    """Point a routed law with several species at State.species_id; the
    kernels index the router's table by species, so the array is checked
    once per version."""
    if law.kind == L.LAW_ROUTED and law.n_species > 1:
        sp = state.species_id
        tag = (id(sp), sp._version, int(law.n_species))
        if state.__dict__.get("_species_checked") != tag:
            if state.N and (int(sp.min()) < 0 or int(sp.max()) >= law.n_species):
                raise ValueError(f"species_id outside the router's {law.n_species} species")
            state.__dict__["_species_checked"] = tag
        law.species = sp.data_ptr()


_F64_FIELDS = ("pos_c", "vel", "ang_vel", "force", "torque", "rad", "mass", "inertia")


def _is_f64(state: State) -> bool:
    """Whether the state steps through the float64 kernels (a float64 state,
    as State.create makes with jax_enable_x64)."""
    if state.pos_c.dtype != torch.float64:
        return False
    bad = [f for f in _F64_FIELDS if getattr(state, f).dtype != torch.float64]
    if state.q.data.dtype != torch.float64:
        bad.append("q")
    if bad:
        raise TypeError(f"a float64 state needs every float field in float64; {bad} are not")
    return True


def _require_f32(state: State, what: str) -> None:
    if state.pos_c.dtype != torch.float32:
        raise NotImplementedError(
            f"{what} runs in float32; the float64 kernels cover System.step with built-in plugins and the laws "
            "without history (every collider), compute_force and the energies"
        )


def _system_struct64(state: State, system: System, keep: list, for_collider: bool = True) -> L.JdemSystemF64:
    """The jdem_system_f64 of a float64 (state, system): domain, law tables,
    dt and gravity in float64."""
    _require_gpu(state)
    S = L.JdemSystemF64()
    S.state = _state_struct(state, L.JdemStateF64)
    S.domain = system.domain.to_c(state.dim, f64=True)
    law, dev = law_struct(system.force_model, system.mat_table, state.device, f64=True)
    keep.append(dev)
    _set_species(law, state)
    S.law = law
    S.lin_kind = int(type(system.linear_integrator).kind)
    S.rot_kind = int(type(system.rotation_integrator).kind)
    S.dt = float(system.dt)
    S.interact_same_bond_id = 1 if system.interact_same_bond_id else 0
    g = np.zeros(3, np.float64)
    g[: state.dim] = np.asarray(system.force_manager.gravity, np.float64).reshape(-1)[: state.dim]
    for a in range(3):
        S.gravity[a] = float(g[a])
    fm = system.force_manager
    for name, cname in (("external_force", "ext_force"), ("external_force_com", "ext_force_com"), ("external_torque", "ext_torque")):
        t = getattr(fm, name)
        if t is not None:
            t = torch.as_tensor(t, dtype=torch.float64, device=state.device).contiguous()
            keep.append(t)
            setattr(S, cname, t.data_ptr())
    col = system.collider
    if not for_collider:
        S.collider_kind = L.COLLIDER_NONE
        return S
    S.collider_kind = _collider_kind(col)
    if isinstance(col, DynamicCellList):
        g, ovf = col.grid(system, f64=True)
        if ovf:
            raise NotImplementedError("cell grid exceeds the int32 hash range (reference: hash_overflow)")
        S.grid = g
    if isinstance(col, NeighborList):
        if col.history is not None:
            raise NotImplementedError("the history laws run in float32")
        if col.old_pos.dtype != torch.float64:
            raise TypeError("a float64 NeighborList needs float64 build positions (create it from the float64 state)")
        nl = L.JdemNlistF64()
        nl.max_neighbors = col.max_neighbors
        nl.cutoff = float(col.cutoff)
        nl.skin = float(col.skin)
        nl.nl = _ptr(col.neighbor_list)
        nl.old_pos = _ptr(col.old_pos)
        nl.n_build_times = int(col.n_build_times)
        nl.overflow = 1 if col.overflow else 0
        sec = col.secondary_collider
        if isinstance(sec, DynamicCellList):
            nl.secondary = L.COLLIDER_CELL_LIST
            g2, ovf = col.secondary_grid(system, f64=True)
            if ovf:
                raise NotImplementedError("secondary cell grid exceeds the int32 hash range")
            nl.grid = g2
        elif isinstance(sec, NaiveSimulator):
            nl.secondary = L.COLLIDER_NAIVE
        else:
            raise NotImplementedError(f"NeighborList secondary collider {type(sec).__name__} on the GPU")
        S.nlist = nl
    return S


def _domain_after(dom, S, state: State, f64: bool):
    """The domain after a fused call: a FreeDomain takes the box the call
    fitted, every domain the inverse box of its box, in the call's dtype."""
    ft = np.float64 if f64 else F32
    if isinstance(dom, FreeDomain) and state.N > 0:
        box = np.array([S.domain.box_size[a] for a in range(state.dim)], ft)
        anc = np.array([S.domain.anchor[a] for a in range(state.dim)], ft)
        return replace(dom, box_size=box, anchor=anc, inv_box_size=(ft(1.0) / box).astype(ft))
    return replace(dom, inv_box_size=(ft(1.0) / np.asarray(dom.box_size, ft)).astype(ft))


def _ctx(state: State) -> L.Context:
    return L.context_for(state.device)


def _fresh_collider(col: Collider) -> Collider:
    """Copy of the buffers a step writes (functional semantics)."""
    if isinstance(col, NeighborList):
        hist = None
        if col.history is not None:
            hist = {k: v.clone() for k, v in col.history.items()}
        return replace(col, neighbor_list=col.neighbor_list.clone(), old_pos=col.old_pos.clone(), history=hist)
    return col


def _collider_after(col: Collider, S: L.JdemSystem) -> Collider:
    if isinstance(col, NeighborList):
        return replace(col, n_build_times=int(S.nlist.n_build_times), overflow=bool(S.nlist.overflow))
    return col


def _fusable(state: State, system: System) -> bool:
    return (
        type(system.linear_integrator) in _BUILTIN_LIN
        and type(system.rotation_integrator) in _BUILTIN_ROT
        and type(system.domain) in _BUILTIN_DOM
        and type(system.collider) in _BUILTIN_COL
        and type(system.force_model) in _BUILTIN_LAW
        and not system.force_manager.has_functions
        and system.user_pre_step_actions is _save_state_system
        and system.user_post_step_actions is _save_state_system
    )


def _resident_call(state: State, system: System, keep: list):
    """(struct, output slot, collider) of a NeighborList call that keeps its
    working set: it starts from the collider's working set when that is
    current, else from the collider's arrays, which it only reads."""
    col = system.collider
    res_in = col.resident_for_step()
    S = _system_struct(state, system, keep, resident=res_in)
    slot = ctypes.c_void_p()
    keep.append(slot)
    S.nlist.resident_out = ctypes.addressof(slot)
    hist = res_in.hist if res_in is not None else col.history is not None
    return S, slot, col, hist


def _resident_after(ctx, col, S, slot, hist: bool, state: State):
    """The collider after a call made by _resident_call."""
    if not slot.value:  # nothing stepped (no particles, or n = 0)
        return _collider_after(_fresh_collider(col), S)
    res = L.Resident(ctx, slot.value, state.N, int(col.max_neighbors), state.dim, hist, state.device)
    return col.with_resident(res, S.nlist.n_build_times, bool(S.nlist.overflow))


def _fused_step(state: State, system: System, n: int) -> tuple:
    new_state = state.with_kinematics()
    stats = L.JdemStepStats()
    ctx = _ctx(new_state)
    keep: list = []
    if _is_f64(new_state):
        col = _fresh_collider(system.collider)
        sys2 = replace(system, collider=col, domain=replace(system.domain))
        S = _system_struct64(new_state, sys2, keep)
        ctx.check(ctx.lib.jdem_system_step_f64(ctx.ptr, ctypes.byref(S), int(n), ctypes.byref(stats)),
                  "jdem_system_step_f64")
        fm = replace(sys2.force_manager, external_force=None, external_force_com=None, external_torque=None)
        out_sys = replace(sys2, collider=_collider_after(col, S), domain=_domain_after(sys2.domain, S, new_state, True),
                          force_manager=fm)
        out_sys = _advance_clock(out_sys, n)
        out_sys.__dict__["_last_stats"] = stats
        return new_state, out_sys
    if isinstance(system.collider, NeighborList) and new_state.is_sphere_state:
        sys2 = replace(system, domain=replace(system.domain))
        S, slot, col, hist = _resident_call(new_state, sys2, keep)
        ctx.check(ctx.lib.jdem_system_step(ctx.ptr, ctypes.byref(S), int(n), ctypes.byref(stats)), "jdem_system_step")
        col_out = _resident_after(ctx, col, S, slot, hist, new_state)
    else:
        col = _fresh_collider(system.collider)
        sys2 = replace(system, collider=col, domain=replace(system.domain))
        S = _system_struct(new_state, sys2, keep)
        ctx.check(ctx.lib.jdem_system_step(ctx.ptr, ctypes.byref(S), int(n), ctypes.byref(stats)), "jdem_system_step")
        col_out = _collider_after(col, S)
    dom = _domain_after(sys2.domain, S, new_state, False)
    fm = replace(sys2.force_manager, external_force=None, external_force_com=None, external_torque=None)
    out_sys = replace(sys2, collider=col_out, domain=dom, force_manager=fm)
    out_sys = _advance_clock(out_sys, n)
    out_sys.__dict__["_last_stats"] = stats
    return new_state, out_sys


def _fused_rollout(state: State, system: System, ks: list) -> tuple:
    """trajectory_rollout with ``save_state`` frames in one jdem_rollout call:
    segment f runs System.step(ks[f]) on the resident working set and its
    end state is written to frame f."""
    new_state = state.with_kinematics()
    sys2 = replace(system, domain=replace(system.domain))
    keep: list = []
    S, slot, col, hist = _resident_call(new_state, sys2, keep)
    nf = len(ks)

    def frames_like(t: torch.Tensor) -> torch.Tensor:
        return torch.empty((nf,) + tuple(t.shape), dtype=t.dtype, device=t.device)

    out = {
        "pos_c": frames_like(new_state.pos_c),
        "vel": frames_like(new_state.vel),
        "ang_vel": frames_like(new_state.ang_vel),
        "q": frames_like(new_state.q.data),
        "force": frames_like(new_state.force),
        "torque": frames_like(new_state.torque),
    }
    steps = (ctypes.c_int64 * nf)(*ks)
    fr = L.JdemFrames()
    fr.n_frames = nf
    fr.steps = ctypes.cast(steps, ctypes.POINTER(ctypes.c_int64))
    fr.pos = out["pos_c"].data_ptr()
    fr.vel = out["vel"].data_ptr()
    fr.ang_vel = out["ang_vel"].data_ptr()
    fr.q = out["q"].data_ptr()
    fr.force = out["force"].data_ptr()
    fr.torque = out["torque"].data_ptr()
    stats = L.JdemStepStats()
    ctx = _ctx(new_state)
    ctx.check(ctx.lib.jdem_rollout(ctx.ptr, ctypes.byref(S), ctypes.byref(fr), ctypes.byref(stats)), "jdem_rollout")
    kw = {}
    for f in dataclasses.fields(new_state):
        v = getattr(new_state, f.name)
        if f.name in out:
            kw[f.name] = Quaternion(out["q"]) if f.name == "q" else out[f.name]
        elif isinstance(v, torch.Tensor):
            kw[f.name] = v.unsqueeze(0).expand((nf,) + tuple(v.shape)).contiguous()
        else:
            kw[f.name] = v
    traj = State(**kw)
    dom = replace(sys2.domain, inv_box_size=(F32(1.0) / np.asarray(sys2.domain.box_size, F32)).astype(F32))
    fm = replace(sys2.force_manager, external_force=None, external_force_com=None, external_torque=None)
    out_sys = replace(sys2, collider=_resident_after(ctx, col, S, slot, hist, new_state), domain=dom, force_manager=fm)
    out_sys = _advance_clock(out_sys, sum(ks))
    out_sys.__dict__["_last_stats"] = stats
    return new_state, out_sys, traj


def _step_once(state: State, system: System) -> tuple:
    """_step_once (system.py:60-82) stage by stage."""
    system = _advance_clock(system, 1)
    state, system = system.user_pre_step_actions(state, system)
    state, system = system.domain.apply(state, system)
    system = replace(
        system, domain=replace(system.domain, inv_box_size=(F32(1.0) / np.asarray(system.domain.box_size, F32)).astype(F32))
    )
    state, system = system.linear_integrator.step_before_force(state, system)
    state, system = system.rotation_integrator.step_before_force(state, system)
    state, system = system.collider.compute_force(state, system)
    state, system = system.force_manager.apply(state, system)
    state, system = system.linear_integrator.step_after_force(state, system)
    state, system = system.rotation_integrator.step_after_force(state, system)
    state, system = system.user_post_step_actions(state, system)
    return state, system


def _run_stage(state: State, system: System, stage: int) -> tuple:
    """One built-in stage as a GPU kernel on fresh copies of the state."""
    _require_f32(state, "a step with user plugins or actions (stage by stage)")
    new_state = state.with_kinematics()
    keep: list = []
    S = _system_struct(new_state, system, keep)
    ctx = _ctx(new_state)
    ctx.check(ctx.lib.jdem_step_stage(ctx.ptr, ctypes.byref(S), int(stage)), "jdem_step_stage")
    if stage == L.STAGE_DOMAIN_APPLY and isinstance(system.domain, FreeDomain) and new_state.N > 0:
        box = np.array([S.domain.box_size[a] for a in range(state.dim)], F32)
        anc = np.array([S.domain.anchor[a] for a in range(state.dim)], F32)
        system = replace(
            system, domain=replace(system.domain, box_size=box, anchor=anc, inv_box_size=(F32(1.0) / box).astype(F32))
        )
    return new_state, system


def _collider_compute_force(state: State, system: System) -> tuple:
    new_state = replace(state, force=state.force.clone(), torque=state.torque.clone())
    col = _fresh_collider(system.collider)
    sys2 = replace(system, collider=col)
    keep: list = []
    ctx = _ctx(new_state)
    if _is_f64(new_state):
        S = _system_struct64(new_state, sys2, keep)
        ctx.check(ctx.lib.jdem_compute_force_f64(ctx.ptr, ctypes.byref(S)), "jdem_compute_force_f64")
        return new_state, replace(sys2, collider=_collider_after(col, S))
    S = _system_struct(new_state, sys2, keep)
    ctx.check(ctx.lib.jdem_compute_force(ctx.ptr, ctypes.byref(S)), "jdem_compute_force")
    return new_state, replace(sys2, collider=_collider_after(col, S))


def _collider_potential_energy(state: State, system: System) -> tuple:
    col = _fresh_collider(system.collider)
    sys2 = replace(system, collider=col)
    keep: list = []
    ctx = _ctx(state)
    e = ctypes.c_double()
    if _is_f64(state):
        S = _system_struct64(state, sys2, keep)
        ctx.check(ctx.lib.jdem_potential_energy_f64(ctx.ptr, ctypes.byref(S), ctypes.byref(e)),
                  "jdem_potential_energy_f64")
        return state, replace(sys2, collider=_collider_after(col, S)), float(e.value)
    S = _system_struct(state, sys2, keep)
    ctx.check(ctx.lib.jdem_potential_energy(ctx.ptr, ctypes.byref(S), ctypes.byref(e)), "jdem_potential_energy")
    return state, replace(sys2, collider=_collider_after(col, S)), float(e.value)


def _contact_virial(state: State, system: System) -> tuple:
    """Float64 contact virial sum_{i<j} r_ij (x) F_ij and the system whose
    NeighborList was refreshed for it (a list that is unbuilt or stale is
    rebuilt first)."""
    _require_f32(state, "the contact virial")
    col = _fresh_collider(system.collider)
    sys2 = replace(system, collider=col)
    keep: list = []
    S = _system_struct(state, sys2, keep)
    ctx = _ctx(state)
    D = state.dim
    w = (ctypes.c_double * (D * D))()
    ctx.check(ctx.lib.jdem_contact_virial(ctx.ptr, ctypes.byref(S), w), "jdem_contact_virial")
    return np.array(list(w), dtype=np.float64).reshape(D, D), replace(sys2, collider=_collider_after(col, S))


def _kinetic_energy_parts(state: State, system: System) -> tuple:
    keep: list = []
    ctx = _ctx(state)
    t, r = ctypes.c_double(), ctypes.c_double()
    if _is_f64(state):
        S = _system_struct64(state, system, keep, for_collider=False)
        ctx.check(ctx.lib.jdem_kinetic_energy_parts_f64(ctx.ptr, ctypes.byref(S), ctypes.byref(t), ctypes.byref(r)),
                  "jdem_kinetic_energy_parts_f64")
        return float(t.value), float(r.value)
    S = _system_struct(state, system, keep, for_collider=False)
    ctx.check(ctx.lib.jdem_kinetic_energy_parts(ctx.ptr, ctypes.byref(S), ctypes.byref(t), ctypes.byref(r)),
              "jdem_kinetic_energy_parts")
    return float(t.value), float(r.value)


def _kinetic_energy(state: State, system: System) -> float:
    if _is_f64(state):
        t, r = _kinetic_energy_parts(state, system)
        return t + r
    keep: list = []
    S = _system_struct(state, system, keep, for_collider=False)
    ctx = _ctx(state)
    e = ctypes.c_double()
    ctx.check(ctx.lib.jdem_kinetic_energy(ctx.ptr, ctypes.byref(S), ctypes.byref(e)), "jdem_kinetic_energy")
    return float(e.value)


def _neighbor_list(state: State, system: System, kind: int, grid, cutoff: float, M: int) -> tuple:
    _require_gpu(state)
    _require_f32(state, "create_neighbor_list")
    M = int(M)
    nl = torch.full((state.N, M), -1, dtype=torch.int32, device=state.device)
    if M == 0 or state.N == 0:
        return state, system, nl, False
    keep: list = []
    S = _system_struct(state, system, keep)
    ctx = _ctx(state)
    ovf = ctypes.c_int32(0)
    g = grid if grid is not None else L.JdemGrid()
    ctx.check(
        ctx.lib.jdem_neighbor_list(
            ctx.ptr, ctypes.byref(S), int(kind), ctypes.byref(g), float(F32(cutoff)), M, nl.data_ptr(), ctypes.byref(ovf)
        ),
        "jdem_neighbor_list",
    )
    return state, system, nl, bool(ovf.value)


def _nl_refresh(state: State, system: System) -> tuple:
    """Refresh a NeighborList's cached list (neighbor_list.py:58-131)."""
    _, sys2 = _collider_compute_force(state, system)
    return state, sys2


__all__ = ["System", "save_state"]
