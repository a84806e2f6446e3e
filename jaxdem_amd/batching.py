"""Independent systems of unequal size advanced together in one segmented
layout (``jdem_batch_step``).

The reference batches by ``vmap`` over systems padded to one particle count
(system.py:743-744, State.stack).  Here the systems' particles are
concatenated instead: every particle knows its system, cells never mix
systems, and one kernel per time step advances all of them.  Each system
keeps its own domain, cell grid, time step, cutoff and skin; the force
model, integrators and gravity are shared.  A system's trajectory in the
batch equals its trajectory alone, bit for bit.

    batch = Batch.pack([(state_0, system_0), (state_1, system_1), ...])
    batch = batch.step(n)
    pairs = batch.unpack()
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, fields, replace
from typing import Callable, Sequence

import numpy as np
import torch

from . import _lib as L
from .colliders import DynamicCellList, NeighborList
from .domains import PeriodicDomain, ReflectDomain
from .materials import MaterialTable
from .state import Quaternion, State

F32 = np.float32


def _table_key(t: MaterialTable) -> tuple:
    items = [(k, np.asarray(v, np.float32).tobytes()) for k, v in sorted(t.props.items())]
    items += [(k, np.asarray(v, np.float32).tobytes()) for k, v in sorted(t.pair.items())]
    return (type(t.matcher).__name__, tuple(items))


def _merge_tables(tables: Sequence[MaterialTable]) -> tuple[MaterialTable, np.ndarray]:
    """Block-diagonal union of the distinct material tables: system k's
    material m becomes mat_offsets[k] + m; systems with equal tables share a
    block, and pair entries across blocks are never used."""
    uniq: dict = {}
    blocks: list = []
    which = []
    for t in tables:
        key = _table_key(t)
        if key not in uniq:
            uniq[key] = len(blocks)
            blocks.append(t)
        which.append(uniq[key])
    sizes = [len(t) for t in blocks]
    boffs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    keys = sorted(set().union(*[set(t.props) for t in blocks]))
    props = {}
    for k in keys:
        props[k] = np.concatenate(
            [np.asarray(t.props[k], np.float32) if k in t.props else np.zeros(len(t), np.float32) for t in blocks]
        )
    pkeys = sorted(set().union(*[set(t.pair) for t in blocks]))
    M = int(boffs[-1])
    pair = {}
    for k in pkeys:
        a = np.zeros((M, M), np.float32)
        for t, o in zip(blocks, boffs[:-1]):
            if k in t.pair:
                n = len(t)
                a[o : o + n, o : o + n] = np.asarray(t.pair[k], np.float32)
        pair[k] = a
    offs = np.array([boffs[w] for w in which] + [M], np.int64)
    return MaterialTable(props, pair, tables[0].matcher), offs


def _pad_cols(t: torch.Tensor, width: int, fill) -> torch.Tensor:
    if t.shape[1] == width:
        return t
    pad = torch.full((t.shape[0], width - t.shape[1]) + tuple(t.shape[2:]), fill, dtype=t.dtype, device=t.device)
    return torch.cat([t, pad], dim=1)


def _shift(t: torch.Tensor, off: int) -> torch.Tensor:
    return torch.where(t >= 0, t + off, t)


@dataclass
class Batch:
    """A packed batch; ``systems`` keeps each system's own parameters."""

    state: State
    system: object  # the packed System (shared plugins, merged materials)
    offsets: np.ndarray
    mat_offsets: np.ndarray
    systems: list
    bond_widths: list
    n_build: list  # NeighborList builds of each system
    # how rebuild decisions reach the kernels (jdem_batch.launch_mode): 0 a
    # gated chain after every step with captured graphs for long calls, 1 the
    # chain step by step, 2 speculative chunks with one host read per chunk
    launch_mode: int = 0
    # particles per workgroup of the step kernel (256 or 128)
    tile_width: int = 256
    # launch mode 2 runs the rest of a call as mode 1 once two rebuilds in a
    # row come this many steps or fewer apart (0: 2)
    spec_min_gap: int = 0

    def __len__(self) -> int:
        return len(self.systems)

    @property
    def n_particles(self) -> int:
        return int(self.offsets[-1])

    # ------------------------------------------------------------------
    @staticmethod
    def pack(pairs: Sequence[tuple]) -> "Batch":
        from .system import _fusable

        if not pairs:
            raise ValueError("Batch.pack needs at least one (state, system) pair")
        states = [p[0] for p in pairs]
        systems = [p[1] for p in pairs]
        s0, y0 = states[0], systems[0]
        dev = s0.device
        for st, sy in pairs:
            if st.device != dev or st.dim != s0.dim:
                raise ValueError("batched systems must share the device and the dimension")
            if not st.is_sphere_state:
                raise NotImplementedError("batches hold sphere states only")
            if st.pos_c.dtype != torch.float32:
                raise NotImplementedError("batches run in float32 (the float64 kernels cover System.step)")
            if not _fusable(st, sy):
                raise NotImplementedError("batched systems must use built-in plugins without force functions")
            col = sy.collider
            if not isinstance(col, NeighborList) or not isinstance(col.secondary_collider, DynamicCellList):
                raise NotImplementedError("batched systems run a NeighborList over a CellList")
            if not isinstance(sy.domain, (PeriodicDomain, ReflectDomain)):
                raise NotImplementedError("batched systems need a periodic or reflect domain")
            same = (
                type(sy.force_model) is type(y0.force_model)
                and getattr(sy.force_model, "kt_ratio", None) == getattr(y0.force_model, "kt_ratio", None)
                and getattr(sy.force_model, "damping", None) == getattr(y0.force_model, "damping", None)
                and type(sy.linear_integrator) is type(y0.linear_integrator)
                and type(sy.rotation_integrator) is type(y0.rotation_integrator)
                and sy.interact_same_bond_id == y0.interact_same_bond_id
                and np.array_equal(np.asarray(sy.force_manager.gravity), np.asarray(y0.force_manager.gravity))
                and (col.history is None) == (y0.collider.history is None)
                and (sy.force_model.kind != L.LAW_ROUTED or sy.force_model == y0.force_model)
            )
            if not same:
                raise ValueError("batched systems must share the force model, integrators and gravity")
            if any(getattr(sy.force_manager, k) is not None for k in ("external_force", "external_force_com",
                                                                          "external_torque")):
                raise NotImplementedError("batched systems with pending external forces")
        sizes = [st.N for st in states]
        offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        mt, moff = _merge_tables([sy.mat_table for sy in systems])
        # concatenated state
        bw = [int(st.bond_id.shape[-1]) for st in states]
        width = max(bw)
        kw = {}
        for f in fields(s0):
            v = getattr(s0, f.name)
            if isinstance(v, Quaternion):
                kw[f.name] = Quaternion(torch.cat([getattr(s, f.name).data for s in states]).contiguous())
            elif isinstance(v, torch.Tensor):
                parts = []
                for k, s in enumerate(states):
                    t = getattr(s, f.name)
                    if f.name == "bond_id":
                        t = _shift(_pad_cols(t, width, -1), int(offsets[k]))
                    elif f.name == "clump_id":
                        t = t + int(offsets[k])
                    elif f.name == "mat_id":
                        t = t + int(moff[k])
                    parts.append(t)
                kw[f.name] = torch.cat(parts).contiguous()
            else:
                kw[f.name] = v
        kw["has_bonds"] = any(s.has_bonds for s in states)
        state = State(**kw)
        # packed NeighborList buffers (one row width M for every system)
        M = max(sy.collider.max_neighbors for sy in systems)
        W32 = max((M + 31) // 32, 1)
        nl = torch.cat(
            [_shift(_pad_cols(sy.collider.neighbor_list, M, -1), int(offsets[k])) for k, sy in enumerate(systems)]
        ).contiguous()
        old = torch.cat([sy.collider.old_pos for sy in systems]).contiguous()
        hist = None
        if y0.collider.history is not None:
            hist = {
                "xi": torch.cat([_pad_cols(sy.collider.history["xi"], M, 0.0) for sy in systems]).contiguous(),
                "live": torch.cat([_pad_cols(sy.collider.history["live"], W32, 0) for sy in systems]).contiguous(),
            }
        built = [int(sy.collider.n_build_times) for sy in systems]
        col = replace(
            y0.collider, neighbor_list=nl, old_pos=old, history=hist, max_neighbors=M,
            n_build_times=max(built), overflow=any(sy.collider.overflow for sy in systems),
        )
        packed = replace(y0, collider=col, mat_table=mt)
        return Batch(state, packed, offsets, moff, list(systems), bw, built)

    # ------------------------------------------------------------------
    def to(self, device) -> "Batch":
        """The batch with every tensor on ``device`` (pack on the host, then
        move once)."""
        st = self.state
        kw = {}
        for f in fields(st):
            v = getattr(st, f.name)
            if isinstance(v, Quaternion):
                kw[f.name] = Quaternion(v.data.to(device))
            elif isinstance(v, torch.Tensor):
                kw[f.name] = v.to(device)
            else:
                kw[f.name] = v
        col = self.system.collider
        hist = None if col.history is None else {k: v.to(device) for k, v in col.history.items()}
        col2 = replace(col, neighbor_list=col.neighbor_list.to(device), old_pos=col.old_pos.to(device), history=hist)
        return Batch(State(**kw), replace(self.system, collider=col2), self.offsets, self.mat_offsets, self.systems,
                     self.bond_widths, list(self.n_build), self.launch_mode, self.tile_width, self.spec_min_gap)

    def _c_arrays(self, keep: list) -> L.JdemBatch:
        """The jdem_batch of the systems (their domains, grids, dt, cutoffs,
        skins do not change as they step; built once per Batch)."""
        cached = self.__dict__.get("_c_cache")
        if cached is None:
            cached = self._c_build()
            self.__dict__["_c_cache"] = cached
        off, doms, grids, dt, cut, skin = cached
        nb = np.ascontiguousarray(self.n_build, np.int32)
        keep += [off, doms, grids, dt, cut, skin, nb]
        b = L.JdemBatch()
        b.n_systems = len(self.systems)
        b.launch_mode = int(self.launch_mode)
        b.tile_width = int(self.tile_width)
        b.spec_min_gap = int(self.spec_min_gap)
        b.n_build_times = nb.ctypes.data
        b.offsets = off.ctypes.data
        b.domains = ctypes.addressof(doms)
        b.grids = ctypes.addressof(grids)
        b.dt = dt.ctypes.data
        b.cutoff = cut.ctypes.data
        b.skin = skin.ctypes.data
        return b

    def _c_build(self) -> tuple:
        B = len(self.systems)
        off = np.ascontiguousarray(self.offsets, np.int64)
        doms = (L.JdemDomain * B)()
        grids = (L.JdemGrid * B)()
        dt = np.zeros(B, F32)
        cut = np.zeros(B, F32)
        skin = np.zeros(B, F32)
        for k, sy in enumerate(self.systems):
            doms[k] = sy.domain.to_c(self.state.dim)
            g, ovf = sy.collider.secondary_grid(sy)
            if ovf:
                raise NotImplementedError("secondary cell grid exceeds the int32 hash range")
            grids[k] = g
            dt[k] = F32(sy.dt)
            cut[k] = F32(sy.collider.cutoff)
            skin[k] = F32(sy.collider.skin)
        return off, doms, grids, dt, cut, skin

    def step(self, n: int = 1) -> "Batch":
        """Advance every system ``n`` steps; the batch itself is left
        untouched (functional semantics, like System.step)."""
        from .system import _clock_after, _ctx, _resident_after, _resident_call

        st = self.state.with_kinematics()
        keep: list = []
        # the batch's working set stays on the device between calls, as a
        # single system's does (unpack / pack export and import it)
        S, slot, col, hist = _resident_call(st, self.system, keep)
        Bc = self._c_arrays(keep)
        stats = L.JdemStepStats()
        ctx = _ctx(st)
        ctx.check(ctx.lib.jdem_batch_step(ctx.ptr, ctypes.byref(S), ctypes.byref(Bc), int(n), ctypes.byref(stats)),
                  "jdem_batch_step")
        n_build = [int(v) for v in keep[-1]]
        col2 = _resident_after(ctx, col, S, slot, hist, st)
        col2.__dict__["n_build_times"] = max(n_build)
        sy2 = replace(self.system, collider=col2)
        sy2.__dict__["_last_stats"] = stats
        # every system's clock: one float32 accumulation per distinct (time, dt)
        clocks: dict = {}
        systems = []
        for s in self.systems:
            key = (float(F32(s.time)), float(F32(s.dt)))
            if key not in clocks:
                clocks[key] = _clock_after(s.time, s.dt, int(n))
            systems.append(replace(s, time=clocks[key], step_count=s.step_count + int(n)))
        out = Batch(st, sy2, self.offsets, self.mat_offsets, systems, self.bond_widths, n_build, self.launch_mode,
                    self.tile_width, self.spec_min_gap)
        if "_c_cache" in self.__dict__:
            out.__dict__["_c_cache"] = self.__dict__["_c_cache"]
        return out

    def unpack(self) -> list:
        """The systems again, as (State, System) pairs.  A system's list keeps
        its own row width; entries past it set its overflow flag."""
        out = []
        st, col = self.state, self.system.collider
        for k, sy in enumerate(self.systems):
            a, b = int(self.offsets[k]), int(self.offsets[k + 1])
            kw = {}
            for f in fields(st):
                v = getattr(st, f.name)
                if isinstance(v, Quaternion):
                    kw[f.name] = Quaternion(v.data[a:b].clone())
                elif isinstance(v, torch.Tensor):
                    t = v[a:b].clone()
                    if f.name == "bond_id":
                        t = _shift(t, -a)[:, : self.bond_widths[k]].contiguous()
                    elif f.name == "clump_id":
                        t = t - a
                    elif f.name == "mat_id":
                        t = t - int(self.mat_offsets[k])
                    kw[f.name] = t
                else:
                    kw[f.name] = v
            kw["has_bonds"] = bool(self.bond_widths[k] > 0 and (kw["bond_id"] >= 0).any())
            sk = State(**kw)
            c0 = sy.collider
            Mk = c0.max_neighbors
            rows = _shift(col.neighbor_list[a:b], -a)
            overflow = bool(col.overflow) or bool((rows[:, Mk:] >= 0).any()) if rows.shape[1] > Mk else bool(col.overflow)
            hist = None
            if col.history is not None:
                w32 = c0.history["live"].shape[1]
                hist = {
                    "xi": col.history["xi"][a:b, :Mk].contiguous().clone(),
                    "live": col.history["live"][a:b, :w32].contiguous().clone(),
                }
                if Mk % 32:
                    hist["live"][:, -1] &= (1 << (Mk % 32)) - 1
            ck = replace(
                c0, neighbor_list=rows[:, :Mk].contiguous().clone(), old_pos=col.old_pos[a:b].clone(),
                history=hist, n_build_times=int(self.n_build[k]), overflow=overflow,
            )
            out.append((sk, replace(sy, collider=ck)))
        return out


def autotune_skin(make_pairs: Callable[[float], list], skins: Sequence[float], probe_steps: int = 200,
                  warm_steps: int = 20, device="cuda") -> tuple:
    """Rebuild-cadence autotuner of a batch.  For every candidate skin, packs
    the systems ``make_pairs(skin)`` returns, steps ``warm_steps`` untimed
    (first list builds) and ``probe_steps`` timed, and keeps the skin with
    the highest particle-steps/s.  The skin sets the list radius (cutoff +
    skin), hence the cell size of the list build and the rebuild cadence (a
    list is rebuilt once a sphere has moved skin/2): a small skin rebuilds
    often, a large one tests more listed pairs at every step.  It does not
    change which pairs touch, so every candidate gives the same forces up to
    the order of the sums.  Returns (best skin, [(skin, particle-steps/s,
    rebuilds), ...])."""
    table = []
    for skin in skins:
        pairs = make_pairs(float(skin))
        b = Batch.pack(pairs).to(device)
        del pairs
        b = b.step(warm_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b2 = b.step(probe_steps)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        st = b2.system.__dict__.get("_last_stats")
        table.append((float(skin), b.n_particles * probe_steps / secs, int(st.rebuilds) if st is not None else -1))
        del b, b2
        torch.cuda.empty_cache()
    best = max(table, key=lambda r: r[1])[0]
    return best, table


def with_cell_factor(pairs: list, factor: float) -> list:
    """The pairs with each NeighborList's secondary cell list built on cells
    of ``factor`` times its list radius r_l = cutoff + skin and a stencil of
    half-width ceil(1 / factor) cells, so every pair within r_l stays in the
    stencil (factor 1: 3^D cells of r_l; 0.5: 5^D cells of r_l / 2, a
    tighter candidate volume; 1.5: 3^D larger cells, fewer and fuller)."""
    if not factor > 0:
        raise ValueError(f"cell factor must be positive, got {factor}")
    sr = max(1, int(np.ceil(1.0 / factor - 1e-6)))
    r = np.arange(-sr, sr + 1)
    out = []
    for st, sy in pairs:
        col = sy.collider
        lc = F32(F32(col.cutoff) + F32(col.skin))
        mesh = np.meshgrid(*([r] * st.dim), indexing="ij")
        mask = np.stack([m.ravel() for m in mesh], axis=1).astype(np.int64)
        sec = DynamicCellList(neighbor_mask=mask, cell_size=float(F32(F32(factor) * lc)))
        out.append((st, replace(sy, collider=replace(col, secondary_collider=sec))))
    return out


def _probe(b: Batch, warm_steps: int, probe_steps: int) -> tuple:
    """(particle-steps/s, rebuilds) of ``probe_steps`` steps of ``b`` after
    ``warm_steps`` untimed."""
    b = b.step(warm_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b2 = b.step(probe_steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    st = b2.system.__dict__.get("_last_stats")
    return b.n_particles * probe_steps / secs, int(st.rebuilds) if st is not None else -1


def tune_launch(batch: Batch, tiles: Sequence[int] = (256, 128), modes: Sequence[int] = (0, 1, 2),
                probe_steps: int = 200, warm_steps: int = 20) -> tuple:
    """Tile width of the step kernel, then launch mode (how rebuild decisions
    reach the kernels), probed on ``batch`` itself: neither changes a bit of
    the result, so the batch is not repacked.  Sets the best values on
    ``batch`` and returns ({"tile_width", "launch_mode"}, [(knob, value,
    particle-steps/s, rebuilds), ...])."""
    table = []
    for knob, values in (("tile_width", tiles), ("launch_mode", modes)):
        best_v, best_r = getattr(batch, knob), -1.0
        for v in values:
            setattr(batch, knob, int(v))
            r, nb = _probe(batch, warm_steps, probe_steps)
            table.append((knob, int(v), r, nb))
            if r > best_r:
                best_v, best_r = int(v), r
        setattr(batch, knob, best_v)
    return {"tile_width": batch.tile_width, "launch_mode": batch.launch_mode}, table


def autotune(make_pairs: Callable[[float], list], skins: Sequence[float], cell_factors: Sequence[float] = (1.0, 0.5, 1.5),
             tiles: Sequence[int] = (256, 128), modes: Sequence[int] = (0, 1, 2), probe_steps: int = 200,
             warm_steps: int = 20, device="cuda") -> tuple:
    """Autotuner of a batch: rebuild cadence (the skin), cell size of the list
    build (``with_cell_factor``), then tile width and launch mode
    (``tune_launch``), each later knob at the best values of the earlier
    ones.  A probe packs ``make_pairs(skin)``, steps ``warm_steps`` untimed
    and ``probe_steps`` timed.  Returns (config dict, [(knob, value,
    particle-steps/s, rebuilds), ...]).  None of the knobs changes which
    pairs touch; skin and cell size change the order of the sums, tile width
    and launch mode nothing at all."""
    table = []
    cfg = {"skin": float(skins[0]), "cell_factor": float(cell_factors[0])}
    for knob, values in (("skin", skins), ("cell_factor", cell_factors)):
        best_v, best_r = cfg[knob], -1.0
        for v in values:
            c = dict(cfg)
            c[knob] = float(v)
            b = Batch.pack(with_cell_factor(make_pairs(c["skin"]), c["cell_factor"])).to(device)
            r, nb = _probe(b, warm_steps, probe_steps)
            del b
            torch.cuda.empty_cache()
            table.append((knob, float(v), r, nb))
            if r > best_r:
                best_v, best_r = float(v), r
        cfg[knob] = best_v
    if tiles or modes:
        b = Batch.pack(with_cell_factor(make_pairs(cfg["skin"]), cfg["cell_factor"])).to(device)
        lc, t2 = tune_launch(b, tiles or (b.tile_width,), modes or (b.launch_mode,), probe_steps, warm_steps)
        cfg.update(lc)
        table += t2
        del b
        torch.cuda.empty_cache()
    return cfg, table


__all__ = ["Batch", "autotune_skin", "autotune", "tune_launch", "with_cell_factor"]
