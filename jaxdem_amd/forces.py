"""Pair laws and the force manager.

ForceModel (forces/__init__.py:20-150) registers the laws by key:

* ``spring``        forces/spring.py:72-147
* ``hertz``         forces/hertz.py:72-162
* ``cundallstrack`` forces/cundall_strack.py:102-233
* ``springfriction`` / ``hertzfriction``: tangential spring with Coulomb cap
  and pair history.  JaxDEM has the history hook
  (``force_and_history``/``requires_history``/``init_history``,
  forces/__init__.py:110-140) but no law that uses it; these two are new and
  follow Cundall & Strack (1979) in the form of Luding (2008):
  ``F_n = max(0, k_n d - gamma_n v_n)`` with ``gamma_n = 2 zeta sqrt(k_n
  m_eff)``; the tangential spring ``xi`` is projected onto the contact plane,
  advanced by ``v_t dt``, and capped so ``|k_t xi| <= mu F_n``; it resets
  when the contact opens.  ``hertzfriction`` uses the Hertz-Mindlin
  stiffnesses ``k_n = 4/3 E* sqrt(R* d)``, ``k_t = 8 G* sqrt(R* d)``.

Each law converts a MaterialTable into the float32 device tables the kernels
read (``tables``), computed with the same float32 operations as the
reference's per-pair formulas so the kernels reproduce them bit for bit
(float64 tables, by the same operations, for the float64 step).
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Any, Callable, Sequence

import numpy as np
import torch

from . import _lib as L
from .config import real_dtype
from .factory import Factory

F32 = np.float32


def _hertz_estar(young: np.ndarray, poisson: np.ndarray, ft=F32) -> np.ndarray:
    """E* = 1/((1-nu_i^2)/E_i + (1-nu_j^2)/E_j) (hertz.py:107), in ``ft``."""
    one = ft(1.0)
    E = young.astype(ft)
    nu = poisson.astype(ft)
    Ei, Ej = E[:, None], E[None, :]
    ni, nj = nu[:, None], nu[None, :]
    return (one / ((one - ni * ni) / Ei + (one - nj * nj) / Ej)).astype(ft)


@dataclass
class ForceModel(Factory):
    """Base pair law (forces/__init__.py:20-150)."""

    laws: tuple = field(default=())
    kind = L.LAW_NONE

    @property
    def requires_history(self) -> bool:
        return False

    def init_history(self, shape: tuple) -> Any:
        return None

    @property
    def required_material_properties(self) -> tuple:
        return ()

    def tables(self, mat_table) -> dict:
        """Float32 host tables of include/jdem.h's jdem_law for this law."""
        raise NotImplementedError

    @staticmethod
    def force(i, j, pos, state, system):
        raise NotImplementedError(
            "pair laws are evaluated inside the collider kernels; call Collider.compute_force"
        )

    @staticmethod
    def energy(i, j, pos, state, system):
        raise NotImplementedError(
            "pair energies are evaluated inside the collider kernels; call Collider.compute_potential_energy"
        )


@ForceModel.register("spring")
@dataclass
class SpringForce(ForceModel):
    """k_eff d n, energy 1/2 k_eff d^2 (spring.py:72-147)."""

    kind = L.LAW_SPRING

    @property
    def required_material_properties(self) -> tuple:
        return ("young_eff",)

    def tables(self, mt, ft=F32) -> dict:
        return {"t0": np.asarray(mt.young_eff, ft)}


@ForceModel.register("hertz")
@dataclass
class HertzianForce(ForceModel):
    """4/3 E* sqrt(R*) d^1.5 n, energy 2/5 k d^2.5 (hertz.py:72-162)."""

    kind = L.LAW_HERTZ

    @property
    def required_material_properties(self) -> tuple:
        return ("young", "poisson")

    def tables(self, mt, ft=F32) -> dict:
        return {"t0": (ft(4.0 / 3.0) * _hertz_estar(np.asarray(mt.young), np.asarray(mt.poisson), ft)).astype(ft)}


@ForceModel.register("cundallstrack")
@dataclass
class CundallStrackForce(ForceModel):
    """Spring-dashpot normal, capped tangential dashpot, rolling friction
    (cundall_strack.py:102-233).  Stateless."""

    kind = L.LAW_CUNDALL_STRACK

    @property
    def required_material_properties(self) -> tuple:
        return ("young", "poisson", "e", "mu", "mu_r")

    def tables(self, mt, ft=F32) -> dict:
        one = ft(1.0)
        E = np.asarray(mt.young, ft)
        nu = np.asarray(mt.poisson, ft)
        G = (E / (2.0 * (one + nu))).astype(ft)
        e = np.asarray(mt.e, ft)
        e_eff = np.minimum(e[:, None], e[None, :])
        e_safe = np.where(e_eff > 0.0, e_eff, one)
        ln_e = np.log(e_safe)
        pi2 = ft(np.pi**2)
        beta = np.where(e_eff > 0.0, -ln_e / np.sqrt(pi2 + ln_e * ln_e), one).astype(ft)
        return {
            "t0": beta,
            "m0": E,
            "m1": G,
            "m2": np.asarray(mt.mu, ft),
            "m3": np.asarray(mt.mu_r, ft),
        }


@dataclass
class _HistoryLaw(ForceModel):
    kt_ratio: float = 2.0 / 7.0
    damping: float = 0.0

    @property
    def requires_history(self) -> bool:
        return True

    def init_history(self, shape: tuple) -> dict:
        """Tangential spring per (particle, neighbour slot): ``xi`` (N, M, dim)
        and ``live`` bits (N, ceil(M/32)); a slot without its bit set holds a
        zero spring."""
        return {"shape": tuple(shape)}


@ForceModel.register("springfriction")
@dataclass
class SpringFrictionForce(_HistoryLaw):
    """Linear spring normal + tangential spring with Coulomb cap + normal
    damping ratio ``damping`` (fraction of critical)."""

    kind = L.LAW_SPRING_FRICTION

    @property
    def required_material_properties(self) -> tuple:
        return ("young_eff", "mu_eff")

    def tables(self, mt, ft=F32) -> dict:
        kn = np.asarray(mt.young_eff, ft)
        return {
            "t0": kn,
            "t1": (ft(self.kt_ratio) * kn).astype(ft),
            "t2": np.asarray(mt.mu_eff, ft),
            "p0": float(ft(2.0 * self.damping)),
        }


@ForceModel.register("hertzfriction")
@dataclass
class HertzFrictionForce(_HistoryLaw):
    """Hertz normal + Mindlin tangential spring with Coulomb cap."""

    kind = L.LAW_HERTZ_FRICTION

    @property
    def required_material_properties(self) -> tuple:
        return ("young", "poisson", "mu_eff")

    def tables(self, mt, ft=F32) -> dict:
        one = ft(1.0)
        E = np.asarray(mt.young, ft)
        nu = np.asarray(mt.poisson, ft)
        G = (E / (2.0 * (one + nu))).astype(ft)
        G_star = (one / ((2.0 - nu[:, None]) / G[:, None] + (2.0 - nu[None, :]) / G[None, :])).astype(ft)
        return {
            "t0": (ft(4.0 / 3.0) * _hertz_estar(E, nu, ft)).astype(ft),
            "t1": (ft(8.0) * G_star).astype(ft),
            "t2": np.asarray(mt.mu_eff, ft),
            "p0": float(ft(2.0 * self.damping)),
        }


@ForceModel.register("lawcombiner")
@dataclass
class LawCombiner(ForceModel):
    """Sum of elementary laws (law_combiner.py:22-191): force, torque and
    energy of a pair are the zeros plus each law's, in order; an empty
    combiner gives none.  At most one of the laws may carry history."""

    laws: tuple = field(default=())
    kind = L.LAW_ROUTED

    @property
    def requires_history(self) -> bool:
        return any(law.requires_history for law in self.laws)

    def init_history(self, shape: tuple) -> Any:
        return {"shape": tuple(shape)} if self.requires_history else None

    @property
    def required_material_properties(self) -> tuple:
        return tuple(sorted({p for law in self.laws for p in law.required_material_properties}))


@ForceModel.register("forcerouter")
@dataclass
class ForceRouter(ForceModel):
    """Law by species pair (router.py:26-339): pair (i, j) takes the law at
    ``table[species_id[i]][species_id[j]]`` of an S x S table, of which the
    upper triangle is read (``from_dict`` builds it symmetric)."""

    table: tuple = field(default=())
    kind = L.LAW_ROUTED

    @property
    def requires_history(self) -> bool:
        return any(law.requires_history for row in self.table for law in row)

    def init_history(self, shape: tuple) -> Any:
        return {"shape": tuple(shape)} if self.requires_history else None

    @property
    def required_material_properties(self) -> tuple:
        return tuple(sorted({p for row in self.table for law in row for p in law.required_material_properties}))

    @staticmethod
    def from_dict(S: int, mapping: dict) -> "ForceRouter":
        """router.py:302-333: the mapping is symmetrised; pairs it does not
        name get an empty LawCombiner (no force)."""
        empty = LawCombiner()
        m = [[empty for _ in range(S)] for _ in range(S)]
        for (a, b), law in mapping.items():
            m[a][b] = m[b][a] = law
        return ForceRouter(table=tuple(tuple(r) for r in m))


def _flat_terms(model: ForceModel) -> list:
    """Elementary laws of one routed entry in evaluation order; None is the
    zeros a LawCombiner starts from (law_combiner.py:72-86).  Nested
    combiners flatten exactly: adding the zeros again changes no sum."""
    if isinstance(model, LawCombiner):
        out: list = [None]
        for law in model.laws:
            out += _flat_terms(law)
        return out
    if isinstance(model, ForceRouter):
        raise NotImplementedError("a ForceRouter inside a ForceRouter or LawCombiner is not supported")
    if model.kind == L.LAW_NONE:
        return [None]
    return [model]


def route_layout(model: ForceModel) -> tuple:
    """(n_species, route_off, terms) of a routed law: the terms of ordered
    species pair p = s_i * S + s_j are terms[route_off[p] : route_off[p + 1]]
    (one entry, n_species 0, for a LawCombiner on its own)."""
    if isinstance(model, ForceRouter):
        # the router evaluates the upper triangle and uses entry (a, b) for
        # both orders of the pair (router.py:113-131, 237-246)
        S = len(model.table)
        entries = [model.table[min(a, b)][max(a, b)] for a in range(S) for b in range(S)]
    else:
        S, entries = 0, [model]
    off, terms = [0], []
    for e in entries:
        t = _flat_terms(e)
        if sum(1 for x in t if x is not None and x.requires_history) > 1:
            raise NotImplementedError("a routed entry with more than one history law is not supported")
        terms += t
        off.append(len(terms))
    return S, np.asarray(off, np.int32), terms


def law_struct(model: ForceModel, mat_table, device, f64: bool = False) -> tuple:
    """Build the jdem_law struct (jdem_law_f64 with ``f64``); returns
    (struct, keepalive tensors)."""
    if model.kind == L.LAW_ROUTED:
        return _routed_struct(model, mat_table, device, f64)
    ft = np.float64 if f64 else F32
    law = L.JdemLawF64() if f64 else L.JdemLaw()
    law.kind = model.kind
    law.n_mat = len(mat_table)
    cache = mat_table._device_cache
    key = (type(model).__name__, getattr(model, "kt_ratio", None), getattr(model, "damping", None), str(device), f64)
    if key not in cache:
        tabs = model.tables(mat_table, ft)
        dev = {}
        for k, v in tabs.items():
            if k in ("p0", "p1"):
                dev[k] = v
            else:
                dev[k] = torch.as_tensor(np.ascontiguousarray(v, ft), device=device)
        cache[key] = dev
    dev = cache[key]
    for k in ("t0", "t1", "t2", "m0", "m1", "m2", "m3"):
        if k in dev:
            setattr(law, k, dev[k].data_ptr())
    law.p0 = float(dev.get("p0", 0.0))
    law.p1 = float(dev.get("p1", 0.0))
    return law, dev


def _routed_struct(model: ForceModel, mat_table, device, f64: bool = False) -> tuple:
    """jdem_law of a ForceRouter / LawCombiner: host arrays of the species
    offsets and the elementary terms (each with its own device tables); the
    caller sets ``species`` to State.species_id when n_species > 1."""
    S, off, terms = route_layout(model)
    cls = L.JdemLawF64 if f64 else L.JdemLaw
    arr = (cls * len(terms))()
    keep: list = [off, arr]
    for k, t in enumerate(terms):
        if t is None:
            arr[k].kind = L.LAW_NONE
            arr[k].n_mat = len(mat_table)
        else:
            sub, dev = law_struct(t, mat_table, device, f64)
            arr[k] = sub
            keep.append(dev)
    law = cls()
    law.kind = L.LAW_ROUTED
    law.n_mat = len(mat_table)
    law.n_species = S
    law.n_terms = len(terms)
    law.route_off = off.ctypes.data
    law.route_terms = arr
    return law, keep


# ---------------------------------------------------------------------------
# ForceManager (forces/force_manager.py:82-484)
# ---------------------------------------------------------------------------
def _default_energy(pos, state, system):
    return torch.zeros(pos.shape[0], device=pos.device)


@dataclass
class ForceManager:
    """External/gravity forces added to the contact forces each step.

    ``external_force``, ``external_force_com`` and ``external_torque`` are
    ``None`` while empty (zero); ``add_force`` and friends fill them and
    ``apply`` clears them, as force_manager.py:341-425.
    """

    gravity: np.ndarray
    external_force: torch.Tensor | None = None
    external_force_com: torch.Tensor | None = None
    external_torque: torch.Tensor | None = None
    force_functions: tuple = ()
    energy_functions: tuple = ()
    is_com_force: tuple = ()

    @staticmethod
    def create(state_shape: tuple, *, gravity=None, force_functions: Sequence = ()) -> "ForceManager":
        """force_manager.py:82-190."""
        dim = state_shape[-1]
        ft = real_dtype()
        g = np.zeros(dim, ft) if gravity is None else np.asarray(gravity, ft).reshape(dim)
        funcs, energies, com = [], [], []
        for entry in force_functions:
            args = tuple(entry) if isinstance(entry, (tuple, list)) else (entry,)
            e_func: Callable = _default_energy
            flag = False
            if len(args) == 1:
                f = args[0]
            elif len(args) == 2:
                f = args[0]
                if isinstance(args[1], bool):
                    flag = args[1]
                else:
                    e_func = args[1] if args[1] is not None else _default_energy
            elif len(args) == 3:
                f, e, flag = args
                e_func = e if e is not None else _default_energy
            else:
                raise ValueError(
                    f"Force function entry has invalid length: {len(args)}. "
                    "Expected (Force,), (Force, Bool), (Force, Energy), or (Force, Energy, Bool)."
                )
            if not callable(f):
                raise TypeError(f"First element must be a callable force function, got {type(f)}")
            funcs.append(f)
            energies.append(e_func)
            com.append(bool(flag))
        return ForceManager(
            gravity=g, force_functions=tuple(funcs), energy_functions=tuple(energies), is_com_force=tuple(com)
        )

    @property
    def has_functions(self) -> bool:
        return len(self.force_functions) > 0

    @staticmethod
    def add_force(state, system, force, *, is_com: bool = False):
        """force_manager.py:113-149 (spheres: clump count 1)."""
        f = torch.as_tensor(force, dtype=state.pos_c.dtype, device=state.device).expand(state.N, state.dim)
        fm = system.force_manager
        if is_com:
            fm.external_force_com = (fm.external_force_com if fm.external_force_com is not None else 0) + f
        else:
            fm.external_force = (fm.external_force if fm.external_force is not None else 0) + f
        return system

    @staticmethod
    def add_force_at(state, system, force, idx, *, is_com: bool = False):
        """force_manager.py:153-187."""
        idx_t = torch.as_tensor(idx, dtype=torch.long, device=state.device)
        add = torch.zeros((state.N, state.dim), dtype=state.pos_c.dtype, device=state.device)
        add.index_add_(0, idx_t.reshape(-1), torch.as_tensor(force, dtype=state.pos_c.dtype, device=state.device).reshape(-1, state.dim))
        return ForceManager.add_force(state, system, add, is_com=is_com)

    @staticmethod
    def add_torque(state, system, torque):
        """force_manager.py:191-221."""
        A = 3 if state.dim == 3 else 1
        t = torch.as_tensor(torque, dtype=state.pos_c.dtype, device=state.device).expand(state.N, A)
        fm = system.force_manager
        fm.external_torque = (fm.external_torque if fm.external_torque is not None else 0) + t
        return system

    @staticmethod
    def add_torque_at(state, system, torque, idx):
        A = 3 if state.dim == 3 else 1
        idx_t = torch.as_tensor(idx, dtype=torch.long, device=state.device)
        add = torch.zeros((state.N, A), dtype=state.pos_c.dtype, device=state.device)
        add.index_add_(0, idx_t.reshape(-1), torch.as_tensor(torque, dtype=state.pos_c.dtype, device=state.device).reshape(-1, A))
        return ForceManager.add_torque(state, system, add)

    @staticmethod
    def apply(state, system):
        """force_manager.py:341-425: user force functions (torch, on the
        state's device), then the fused add on the GPU, then clear."""
        from .system import _run_stage

        fm = system.force_manager
        if fm.force_functions:
            pos = state.pos
            for func, com in zip(fm.force_functions, fm.is_com_force):
                f, t = func(pos, state, system)
                f = torch.as_tensor(f, dtype=state.pos_c.dtype, device=state.device)
                t = torch.as_tensor(t, dtype=state.pos_c.dtype, device=state.device)
                if com:
                    fm.external_force_com = (fm.external_force_com if fm.external_force_com is not None else 0) + f
                else:
                    fm.external_force = (fm.external_force if fm.external_force is not None else 0) + f
                fm.external_torque = (fm.external_torque if fm.external_torque is not None else 0) + t
        state, system = _run_stage(state, system, L.STAGE_FORCE_MANAGER)
        system.force_manager = replace(
            system.force_manager, external_force=None, external_force_com=None, external_torque=None
        )
        return state, system

    @staticmethod
    def compute_potential_energy(state, system) -> float:
        """Gravity + custom energies (force_manager.py:430-479), float64."""
        g = torch.as_tensor(system.force_manager.gravity, dtype=torch.float64, device=state.device)
        pe = -float(((state.pos_c.double() @ g) * state.mass.double()).sum())
        for e in system.force_manager.energy_functions:
            pe += float(torch.as_tensor(e(state.pos, state, system)).double().sum())
        return pe


__all__ = [
    "ForceModel",
    "SpringForce",
    "HertzianForce",
    "CundallStrackForce",
    "SpringFrictionForce",
    "HertzFrictionForce",
    "ForceManager",
    "law_struct",
]
