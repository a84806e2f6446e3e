"""CPU restatement of JaxDEM's DEM time step -- TEST INFRASTRUCTURE ONLY.

This module is the correctness oracle for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline -- never as the
thing that produces a product result.  The product path (``jaxdem_amd``) raises
when its HIP library is missing; it never falls back to this module.

It restates, in numpy, the reference algorithm of the DEM time step of
cdelv/JaxDEM (``jaxdem/system.py:60-82``) and the plugins it calls.  Every
function cites the reference ``file:line`` it follows.  Two dtypes are
supported: ``np.float32`` (the reference's default precision) and
``np.float64`` (the reference under ``jax_enable_x64``).

Operation order.  The arithmetic of each formula is written with an explicit
association (for example ``dot(a, b) = (a0*b0 + a1*b1) + a2*b2`` and
``rsqrt(x) = 1/sqrt(x)``) because the HIP kernels follow the same order with
FMA contraction disabled and IEEE division/square root.  Per-pair and
per-particle quantities of the f32 oracle and of the HIP path are therefore
bit-identical; per-particle sums are accumulated sequentially in the
reference's traversal order (stencil order, then ascending particle id within
a cell; ascending ``j`` for the naive collider; row order for neighbour lists).

Documented deviations from the reference (each also pinned by a test):

* cell coordinates are clamped into ``[0, dims-1]``.  For particles inside
  the box this is the reference's hash (``cell_list.py:55-62``); for a
  particle outside a non-periodic box, or for the periodic ``(x % 1) * dims``
  that rounds up to ``dims``, the reference's unclamped hash aliases another
  cell, while the clamped hash keeps the candidate set equal to the
  brute-force set (which the reference also equals for in-box particles).
* ``jnp.pow(delta, 1.5)`` / ``pow(delta, 2.5)`` of the Hertz law are
  evaluated as ``delta*sqrt(delta)`` / ``delta*delta*sqrt(delta)``
  (exact to rounding, and reproducible bit-for-bit on the GPU).
* ``springfriction`` / ``hertzfriction`` (tangential spring with Coulomb cap
  and history) have no reference implementation: JaxDEM's history hook
  (``forces/__init__.py:110-140``) is unused by every built-in law.  Their
  definition here is the Cundall & Strack (1979) incremental tangential
  spring in the form of Luding (2008); parity against the reference is
  unpinned and the laws are pinned by closed-form tests instead.

Clumps (rigid bodies of spheres, ``State.clump`` ids and body-frame offsets
``State.pos_p``): pair forces at the members' lab positions
``pos_c + R(q) pos_p`` with same-clump pairs masked, the torque lever
``R(q) pos_p x F`` added per member, and ``ForceManager.apply``'s segment sum
over the clump (force_manager.py:392-412), taken in ascending member index,
and the reflective domain's clump segment operations (reflect.py:160-297).
A state without ``clump`` is a sphere state: every particle its own clump and
``pos_p == 0``, for which those operations are identities.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

# --------------------------------------------------------------------------
# vector helpers (utils/linalg.py:15-183, utils/quaternion.py:75-367)
# --------------------------------------------------------------------------


def dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``jnp.vecdot`` (linalg.py:71) with left-to-right association."""
    s = a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]
    if a.shape[-1] == 3:
        s = s + a[..., 2] * b[..., 2]
    return s


def cross3(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """3D cross product (linalg.py:15-60)."""
    return np.stack(
        [
            a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
            a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
            a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0],
        ],
        axis=-1,
    )


def cross2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """2D cross product returning shape (..., 1) (linalg.py:54-55)."""
    return (a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0])[..., None]


def cross_w_r(w: np.ndarray, r: np.ndarray) -> np.ndarray:
    """``cross_3X3D_1X2D`` (linalg.py:186-231): w x r, w scalar in 2D."""
    if r.shape[-1] == 2:
        return np.stack([-w[..., 0] * r[..., 1], w[..., 0] * r[..., 0]], axis=-1)
    return cross3(w, r)


def rsqrt(x: np.ndarray) -> np.ndarray:
    """``lax.rsqrt`` restated as ``1/sqrt(x)`` (correctly rounded per op)."""
    one = np.asarray(1.0, dtype=x.dtype)
    return one / np.sqrt(x)


def q_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Hamilton product ``a @ b`` (quaternion.py:347-367); q = [w, x, y, z]."""
    aw, ax = a[..., 0], a[..., 1:]
    bw, bx = b[..., 0], b[..., 1:]
    w = aw * bw - dot(ax, bx)
    xyz = (aw[..., None] * bx + bw[..., None] * ax) + cross3(ax, bx)
    return np.concatenate([w[..., None], xyz], axis=-1)


def q_unit(q: np.ndarray) -> np.ndarray:
    """``Quaternion.unit`` (quaternion.py:75-94)."""
    n2 = q[..., 0] * q[..., 0] + dot(q[..., 1:], q[..., 1:])
    safe = np.where(n2 == 0.0, np.asarray(1.0, q.dtype), n2)
    inv = rsqrt(safe)
    return q * inv[..., None]


def q_conj(q: np.ndarray) -> np.ndarray:
    """``Quaternion.conj`` (quaternion.py:143-162): (w, -xyz), what
    ``rotate_back`` applies (quaternion.py:267-268); not ``inv``, which
    divides by |q|^2 (quaternion.py:164-187)."""
    return np.concatenate([q[..., 0:1], -q[..., 1:]], axis=-1)


def q_rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """``Quaternion.rotate`` body -> lab (quaternion.py:193-242)."""
    if v.shape[-1] == 2:
        qw = q[..., 0:1]
        qz = q[..., 3:4]
        c = qw * qw - qz * qz
        s = 2.0 * qw * qz
        vx, vy = v[..., 0:1], v[..., 1:2]
        return np.concatenate([c * vx - s * vy, s * vx + c * vy], axis=-1)
    t = 2.0 * cross3(q[..., 1:], v)
    return (v + q[..., 0:1] * t) + cross3(q[..., 1:], t)


def pos_p_rot(st) -> np.ndarray:
    """``State._pos_p_rot`` = R(q) pos_p (state.py:230-273, quaternion.py:193-242)."""
    return q_rotate(st.q, st.pos_p)


def lab_pos(st) -> np.ndarray:
    """``State.pos`` = pos_c + R(q) pos_p (state.py:295-304)."""
    return st.pos + pos_p_rot(st)


def norm(v: np.ndarray) -> np.ndarray:
    """``utils.norm`` (linalg.py:113-133): 0 at 0, else sqrt(max(|v|^2, 1e-16))."""
    n2 = dot(v, v)
    T_ = v.dtype
    return np.where(n2 == 0.0, np.asarray(0.0, T_), np.sqrt(np.maximum(n2, np.asarray(1e-16, T_))))


def q_rotate_back(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """``Quaternion.rotate_back`` (quaternion.py:247-268)."""
    return q_rotate(q_conj(q), v)


def q_from_small_rotvec(r: np.ndarray) -> np.ndarray:
    """``Quaternion.from_small_rotvec`` (quaternion.py:132-141)."""
    n2 = dot(r, r)
    ch = 1.0 - n2 / 8.0
    sc = 0.5 - n2 / 48.0
    return np.concatenate([ch[..., None], r * sc[..., None]], axis=-1)


# --------------------------------------------------------------------------
# parameters
# --------------------------------------------------------------------------


@dataclass
class Domain:
    """Domain parameters (domains/__init__.py:28-120)."""

    kind: str  # "free" | "periodic" | "reflect"
    box: np.ndarray
    anchor: np.ndarray
    restitution: float = 1.0

    @property
    def periodic(self) -> bool:
        return self.kind == "periodic"

    @property
    def inv_box(self) -> np.ndarray:
        return np.asarray(1.0, self.box.dtype) / self.box


@dataclass
class Law:
    """Pair-law parameters.

    ``kind`` is one of ``spring`` (spring.py:72-147), ``hertz``
    (hertz.py:72-162), ``cundallstrack`` (cundall_strack.py:102-233),
    ``springfriction`` and ``hertzfriction`` (new, history-carrying).
    Material-indexed arrays follow MaterialTable (material_table.py:79-128).
    """

    kind: str
    young_eff: np.ndarray | None = None  # (M, M)
    young: np.ndarray | None = None  # (M,)
    poisson: np.ndarray | None = None  # (M,)
    e: np.ndarray | None = None  # (M,)
    mu: np.ndarray | None = None  # (M,)
    mu_r: np.ndarray | None = None  # (M,)
    mu_eff: np.ndarray | None = None  # (M, M)
    kt_ratio: float = 2.0 / 7.0
    damping: float = 0.0  # normal damping ratio (fraction of critical)
    # kind "routed": ForceRouter (router.py:26-339) / LawCombiner
    # (law_combiner.py:22-191).  Ordered species pair p = s_i * S + s_j
    # (p = 0 for every pair when n_species <= 1) takes the elementary laws
    # terms[route_off[p] : route_off[p + 1]]; kind "none" is a combiner's
    # zeros.
    n_species: int = 0
    route_off: np.ndarray | None = None
    terms: tuple = ()


@dataclass
class Params:
    dim: int
    dt: float
    domain: Domain
    law: Law
    lin: str = "verlet"  # "verlet" | "euler" | ""
    rot: str = "verletspiral"  # "verletspiral" | "spiral" | ""
    collider: str = "naive"  # "naive" | "celllist" | "neighborlist"
    cell_size: float = 1.0
    search_range: int = 1
    gravity: np.ndarray | None = None
    interact_same_bond_id: bool = False
    # all-pairs rows found through a cell list (naive_near_candidates)
    naive_by_cells: bool = False
    # NeighborList collider (neighbor_list.py:287-480)
    nl_cutoff: float = 0.0
    nl_skin: float = 0.0
    nl_max_neighbors: int = 0
    nl_secondary: str = "celllist"
    nl_secondary_cell_size: float = 1.0
    nl_secondary_search_range: int = 1


@dataclass
class NLState:
    """NeighborList collider buffers (neighbor_list.py:137-286)."""

    nl: np.ndarray
    old_pos: np.ndarray
    n_build: int = 0
    overflow: bool = False
    xi: np.ndarray | None = None  # (N, M, dim) tangential spring history
    live: np.ndarray | None = None  # (N, M) bool: slot carries history
    n_rebuilds_this_call: int = 0


@dataclass
class State:
    """Sphere state (state.py:103-228); q is (N, 4) = [w, x, y, z]."""

    pos: np.ndarray
    vel: np.ndarray
    ang_vel: np.ndarray
    q: np.ndarray
    force: np.ndarray
    torque: np.ndarray
    rad: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    fixed: np.ndarray
    mat_id: np.ndarray
    bond_id: np.ndarray
    extra: dict = field(default_factory=dict)
    species: np.ndarray | None = None  # (N,) species_id, read by routed laws
    # clumps: (N,) ids in [0, K) and (N, D) body-frame offsets; None for spheres
    clump: np.ndarray | None = None
    pos_p: np.ndarray | None = None

    @property
    def N(self) -> int:
        return self.pos.shape[0]

    @property
    def dim(self) -> int:
        return self.pos.shape[1]

    def copy(self) -> "State":
        return State(
            **{
                k: (v.copy() if isinstance(v, np.ndarray) else (dict(v) if isinstance(v, dict) else v))
                for k, v in self.__dict__.items()
            }
        )


def make_state(
    pos,
    rad,
    vel=None,
    mass=None,
    inertia=None,
    ang_vel=None,
    q=None,
    fixed=None,
    mat_id=None,
    bond_id=None,
    dtype=np.float32,
) -> State:
    """Defaults of ``State.create`` (state.py:371-867) for spheres."""
    pos = np.asarray(pos, dtype=dtype)
    N, D = pos.shape
    A = 3 if D == 3 else 1
    rad = np.asarray(rad, dtype=dtype).reshape(N)
    vel = np.zeros((N, D), dtype) if vel is None else np.asarray(vel, dtype)
    mass = np.ones(N, dtype) if mass is None else np.asarray(mass, dtype).reshape(N)
    if inertia is None:
        coeff = 0.5 if D == 2 else 0.4
        inertia = (coeff * mass * rad**2)[:, None] * np.ones((N, A), dtype)
    inertia = np.asarray(inertia, dtype).reshape(N, A)
    ang_vel = np.zeros((N, A), dtype) if ang_vel is None else np.asarray(ang_vel, dtype)
    if q is None:
        q = np.zeros((N, 4), dtype)
        q[:, 0] = 1.0
    q = np.asarray(q, dtype)
    fixed = np.zeros(N, bool) if fixed is None else np.asarray(fixed, bool)
    mat_id = np.zeros(N, np.int64) if mat_id is None else np.asarray(mat_id, np.int64)
    if bond_id is None:
        bond_id = np.full((N, 1), -1, np.int64)
    bond_id = np.asarray(bond_id, np.int64).reshape(N, -1)
    return State(
        pos=pos.copy(),
        vel=vel.copy(),
        ang_vel=ang_vel.copy(),
        q=q.copy(),
        force=np.zeros((N, D), dtype),
        torque=np.zeros((N, A), dtype),
        rad=rad.copy(),
        mass=mass.copy(),
        inertia=inertia.copy(),
        fixed=fixed.copy(),
        mat_id=mat_id.copy(),
        bond_id=bond_id.copy(),
    )


# --------------------------------------------------------------------------
# domains (domains/__init__.py:128-189, periodic.py:39-116)
# --------------------------------------------------------------------------


def displacement_(a: np.ndarray, b: np.ndarray, dom: Domain) -> np.ndarray:
    """``Domain._displacement``: multiplies by ``inv_box`` (periodic.py:75-79)."""
    r = a - b
    if dom.periodic:
        r = r - dom.box * np.round(r * dom.inv_box)
    return r


def displacement(a: np.ndarray, b: np.ndarray, dom: Domain) -> np.ndarray:
    """``Domain.displacement``: divides by ``box`` (periodic.py:39-71)."""
    r = a - b
    if dom.periodic:
        r = r - dom.box * np.round(r / dom.box)
    return r


def periodic_shift(pos: np.ndarray, dom: Domain) -> np.ndarray:
    """``PeriodicDomain.shift`` (periodic.py:84-116)."""
    return pos - dom.box * np.floor((pos - dom.anchor) / dom.box)


# --------------------------------------------------------------------------
# interaction mask (colliders/__init__.py:225-243)
# --------------------------------------------------------------------------


def valid_mask(i: np.ndarray, j: np.ndarray, st: State, isb: bool) -> np.ndarray:
    """``valid_interaction_mask`` (colliders/__init__.py:225-243):
    ``clump_i != clump_j`` (``i != j`` for spheres, clump_id == arange); the
    bond test checks whether ``j`` is listed in ``bond_id[i]``.
    """
    m = (st.clump[i] != st.clump[j]) if st.clump is not None else (i != j)
    if not isb and st.bond_id.shape[1] > 0:
        bonded = np.any(st.bond_id[i] == j[..., None], axis=-1)
        m = m & ~bonded
    return m


# --------------------------------------------------------------------------
# cell list (colliders/_partition.py:55-99, cell_list.py:37-96, 188-261)
# --------------------------------------------------------------------------


def grid_params(box: np.ndarray, cell_size, periodic: bool):
    """``_grid_params`` (_partition.py:55-99) in the box dtype.

    Returns ``(dims int64 (D,), strides int64 (D,), cell (D,), overflow)``.
    """
    dt = box.dtype
    cs = np.asarray(cell_size, dt)
    if periodic:
        dims = np.floor(box / cs).astype(np.int64)
        dims = np.maximum(dims, 1)
        cell = box / dims.astype(dt)
    else:
        dims = np.ceil(box / cs).astype(np.int64)
        dims = np.maximum(dims, 1)
        cell = np.broadcast_to(cs, box.shape).astype(dt)
    strides = np.concatenate([[1], np.cumprod(dims[:-1])]).astype(np.int64)
    total = float(np.prod(dims.astype(np.float64)))
    overflow = total > float(np.iinfo(np.int32).max)
    return dims, strides, cell, overflow


def cell_coords(pos: np.ndarray, dom: Domain, dims: np.ndarray, cell: np.ndarray) -> np.ndarray:
    """Cell coordinates (cell_list.py:55-60), clamped into the grid."""
    if dom.periodic:
        t = (pos - dom.anchor) / dom.box
        t = np.remainder(t, np.asarray(1.0, pos.dtype))
        c = np.floor(t * dims.astype(pos.dtype)).astype(np.int64)
    else:
        c = np.floor((pos - dom.anchor) / cell).astype(np.int64)
    return np.clip(c, 0, dims - 1)


def stencil_axis_offsets(sr: int, d: int, periodic: bool) -> list[int]:
    """Offsets along one axis after ``_dedup_stencil_hashes`` (cell_list.py:92-96).

    For a periodic axis with fewer than ``2*sr+1`` cells several offsets wrap
    onto the same cell for every particle; only the first occurrence in
    stencil order is kept.
    """
    offs = list(range(-sr, sr + 1))
    if not periodic:
        return offs
    kept, seen = [], set()
    for o in offs:
        key = o % d
        if key in seen:
            continue
        seen.add(key)
        kept.append(o)
    return kept


def stencil_offsets(sr: int, dims: np.ndarray, periodic: bool) -> np.ndarray:
    """``neighbor_mask`` in 'ij' meshgrid order (cell_list.py:425-427), deduped."""
    axes = [stencil_axis_offsets(sr, int(d), periodic) for d in dims]
    mesh = np.meshgrid(*[np.asarray(a) for a in axes], indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1).astype(np.int64)


def cell_partition(pos: np.ndarray, dom: Domain, cell_size, sr: int):
    """Hash, stable sort and stencil (``_get_spatial_partition``, cell_list.py:37-87)."""
    dims, strides, cell, overflow = grid_params(dom.box, cell_size, dom.periodic)
    coords = cell_coords(pos, dom, dims, cell)
    h = coords @ strides
    perm = np.argsort(h, kind="stable")
    return dict(
        dims=dims,
        strides=strides,
        cell=cell,
        overflow=overflow,
        coords=coords,
        hash=h,
        perm=perm,
        sorted_hash=h[perm],
        offsets=stencil_offsets(sr, dims, dom.periodic),
    )


def cell_candidates(pos: np.ndarray, dom: Domain, cell_size, sr: int) -> tuple[np.ndarray, dict]:
    """Candidate ``j`` per particle in traversal order (cell_list.py:222-260).

    Returns an (N, K) int64 array padded with -1: stencil cells in order,
    then cell members in sorted (= ascending id) order.  ``i`` itself is a
    candidate (the reference masks it with the clump test).
    """
    part = cell_partition(pos, dom, cell_size, sr)
    dims, strides = part["dims"], part["strides"]
    coords, sh, perm = part["coords"], part["sorted_hash"], part["perm"]
    N = pos.shape[0]
    cols = []
    for o in part["offsets"]:
        nc = coords + o
        if dom.periodic:
            nc = nc - dims * np.floor_divide(nc, dims)
            valid = np.ones(N, bool)
        else:
            valid = np.all((nc >= 0) & (nc < dims), axis=1)
        nh = nc @ strides
        start = np.searchsorted(sh, nh, side="left")
        end = np.searchsorted(sh, nh, side="right")
        cnt = np.where(valid, end - start, 0)
        kmax = int(cnt.max()) if N else 0
        if kmax == 0:
            continue
        t = np.arange(kmax)
        idx = start[:, None] + t[None, :]
        ok = t[None, :] < cnt[:, None]
        cols.append(np.where(ok, perm[np.minimum(idx, N - 1)], -1))
    if not cols:
        return np.full((N, 0), -1, np.int64), part
    C = np.concatenate(cols, axis=1)
    # compact each row so valid entries come first, order preserved
    key = np.where(C >= 0, 0, 1)
    order = np.argsort(key, axis=1, kind="stable")
    C = np.take_along_axis(C, order, axis=1)
    K = int((C >= 0).sum(axis=1).max()) if N else 0
    return C[:, :K], part


def naive_candidates(N: int) -> np.ndarray:
    """All-pairs candidate order of ``NaiveSimulator.compute_force`` (naive.py:213-233)."""
    return np.broadcast_to(np.arange(N, dtype=np.int64), (N, N))


def neighbor_list_cells(st: State, dom: Domain, cell_size, sr: int, cutoff, M: int, isb=False):
    """``DynamicCellList.create_neighbor_list`` (cell_list.py:500-595).

    The stencil reach is inflated to the cutoff (cell_list.py:537-538); rows
    hold candidates with ``|displacement|^2 <= cutoff^2`` in traversal order,
    padded with -1, truncated to M; ``overflow`` as cell_list.py:593.
    """
    dt = st.pos.dtype
    cut = np.asarray(cutoff, dt)
    cutoff_sq = cut * cut
    sr_eff = max(int(sr), 1)
    cs = np.maximum(np.asarray(cell_size, dt), cut / np.asarray(sr_eff, dt))
    C, part = cell_candidates(st.pos, dom, cs, sr_eff)
    N = st.N
    i = np.arange(N)[:, None]
    jj = np.maximum(C, 0)
    d = displacement(st.pos[i], st.pos[jj], dom)
    ok = (C >= 0) & valid_mask(np.broadcast_to(i, C.shape), jj, st, isb) & (dot(d, d) <= cutoff_sq)
    return _pack_rows(C, ok, M, part["overflow"])


def _pack_rows(C, ok, M, extra_overflow=False):
    N = C.shape[0]
    nl = np.full((N, M), -1, np.int64)
    counts = ok.sum(axis=1)
    for r in range(N):
        v = C[r][ok[r]][:M]
        nl[r, : v.size] = v
    overflow = bool(np.any(counts > M)) or bool(extra_overflow)
    return nl, overflow


def neighbor_list_naive(st: State, dom: Domain, cutoff, M: int, isb=False):
    """``NaiveSimulator.create_neighbor_list`` (naive.py:122-185): descending ids."""
    dt = st.pos.dtype
    cut = np.asarray(cutoff, dt)
    N = st.N
    i = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    d = displacement_(st.pos[:, None, :], st.pos[None, :, :], dom)
    ok = valid_mask(np.broadcast_to(i, (N, N)), np.broadcast_to(j, (N, N)), st, isb) & (
        dot(d, d) <= cut * cut
    )
    C = np.broadcast_to(np.arange(N)[::-1], (N, N))
    return _pack_rows(C, ok[:, ::-1], M)


def brute_force_pairs(st: State, dom: Domain, cutoff, isb=False, use_inv=False) -> set:
    """Brute-force candidate set {(i, j)}: the pair-set oracle."""
    dt = st.pos.dtype
    cut = np.asarray(cutoff, dt)
    N = st.N
    fn = displacement_ if use_inv else displacement
    d = fn(st.pos[:, None, :], st.pos[None, :, :], dom)
    i = np.broadcast_to(np.arange(N)[:, None], (N, N))
    j = np.broadcast_to(np.arange(N)[None, :], (N, N))
    ok = valid_mask(i, j, st, isb) & (dot(d, d) <= cut * cut)
    ii, jj = np.nonzero(ok)
    return set(zip(ii.tolist(), jj.tolist()))


# --------------------------------------------------------------------------
# pair laws (vectorised over arrays of pairs)
# --------------------------------------------------------------------------


def _gather_pair(st: State, i: np.ndarray, j: np.ndarray, dom: Domain):
    rij = displacement_(st.pos[i], st.pos[j], dom)
    d_sq = dot(rij, rij)
    return rij, d_sq


def _route_index(law: Law, st: State, i, j) -> np.ndarray:
    i, j = np.broadcast_arrays(np.asarray(i), np.asarray(j))
    if law.n_species <= 1:
        return np.zeros(i.shape, np.int64)
    return st.species[i].astype(np.int64) * law.n_species + st.species[j].astype(np.int64)


def _routed_force(law: Law, st: State, i, j, dom: Domain, dt, xi_old):
    """ForceRouter.force_and_history (router.py:70-160): the pair takes the
    law of its species entry (select_n over the table, router.py:143-148);
    a LawCombiner entry (law_combiner.py:60-93) adds its laws' forces and
    torques to zeros, in order, and keeps the history of its history law."""
    T_ = st.pos.dtype
    D = st.dim
    A = 3 if D == 3 else 1
    p = _route_index(law, st, i, j)
    F = np.zeros(p.shape + (D,), T_)
    T = np.zeros(p.shape + (A,), T_)
    xi = np.zeros(p.shape + (D,), T_)
    contact = np.zeros(p.shape, bool)
    for e in np.unique(p):
        sel = p == e
        fe = te = None
        xe, ce = np.zeros_like(xi), np.zeros_like(contact)
        for sub in law.terms[law.route_off[e] : law.route_off[e + 1]]:
            if sub.kind == "none":
                f, t = np.zeros_like(F), np.zeros_like(T)
            else:
                f, t, xn, c = law_force(sub, st, i, j, dom, dt, xi_old)
                if xn is not None:
                    xe, ce = xn, np.asarray(c, bool)
            fe = f if fe is None else fe + f
            te = t if te is None else te + t
        if fe is None:
            continue
        F = np.where(sel[..., None], fe, F)
        T = np.where(sel[..., None], te, T)
        xi = np.where(sel[..., None], xe, xi)
        contact = np.where(sel, ce, contact)
    return F, T, xi, contact


def law_force(law: Law, st: State, i, j, dom: Domain, dt=None, xi_old=None):
    """Force/torque on ``i`` from ``j``; returns (F, T, xi_new, contact).

    spring: spring.py:97-108.  hertz: hertz.py:98-117 (pow restated as
    d*sqrt(d)).  cundallstrack: cundall_strack.py:123-193.  springfriction
    and hertzfriction: tangential spring with Coulomb cap (module docstring).
    """
    if law.kind == "routed":
        return _routed_force(law, st, i, j, dom, dt, xi_old)
    T_ = st.pos.dtype
    D = st.dim
    A = 3 if D == 3 else 1
    one = np.asarray(1.0, T_)
    zero = np.asarray(0.0, T_)
    ri, rj = st.rad[i], st.rad[j]
    mi, mj = st.mat_id[i], st.mat_id[j]
    rij, d_sq = _gather_pair(st, i, j, dom)
    ne = (i != j).astype(T_)
    shape_T = rij.shape[:-1] + (A,)
    xi_new = None
    if law.kind == "spring":
        R = ri + rj
        k = law.young_eff[mi, mj]
        inv = np.where(d_sq == 0.0, zero, rsqrt(np.maximum(d_sq, np.asarray(1e-16, T_))))
        r = d_sq * inv
        delta = np.maximum(zero, R - r) * ne
        mag = k * delta * inv
        F = mag[..., None] * rij
        return F, np.zeros(shape_T, T_), None, delta > 0
    # unit_and_norm (linalg.py:164-183)
    inv = np.where(d_sq == 0.0, zero, rsqrt(np.maximum(d_sq, np.asarray(1e-16, T_))))
    r = d_sq * inv
    n = rij * inv[..., None]
    if law.kind == "hertz":
        Ei, Ej = law.young[mi], law.young[mj]
        nui, nuj = law.poisson[mi], law.poisson[mj]
        R_star = (ri * rj) / (ri + rj)
        E_star = one / ((one - nui * nui) / Ei + (one - nuj * nuj) / Ej)
        k = (np.asarray(4.0 / 3.0, T_) * E_star) * np.sqrt(R_star)
        delta = np.maximum(zero, (ri + rj) - r) * ne
        mag = k * (delta * np.sqrt(delta))
        return mag[..., None] * n, np.zeros(shape_T, T_), None, delta > 0
    if law.kind == "cundallstrack":
        Ei, Ej = law.young[mi], law.young[mj]
        nui, nuj = law.poisson[mi], law.poisson[mj]
        Gi = Ei / (2.0 * (one + nui))
        Gj = Ej / (2.0 * (one + nuj))
        kn = (2.0 * Ei * ri * Ej * rj) / (Ei * ri + Ej * rj)
        kt = (2.0 * Gi * ri * Gj * rj) / (Gi * ri + Gj * rj)
        m_i, m_j = st.mass[i], st.mass[j]
        m_eff = (m_i * m_j) / (m_i + m_j)
        beta = cs_beta_table(law, T_)[mi, mj]
        mu_eff = np.minimum(law.mu[mi], law.mu[mj])
        gamma_n = 2.0 * beta * np.sqrt(kn * m_eff)
        gamma_t = 2.0 * beta * np.sqrt(kt * m_eff)
        delta = (ri + rj) - r
        is_contact = ((delta > 0) & (i != j)).astype(T_)
        delta = delta * is_contact
        r_ci = -ri[..., None] * n
        r_cj = rj[..., None] * n
        v_ci = st.vel[i] + cross_w_r(st.ang_vel[i], r_ci)
        v_cj = st.vel[j] + cross_w_r(st.ang_vel[j], r_cj)
        v_rel = v_ci - v_cj
        vn = dot(v_rel, n)
        vt_vec = v_rel - vn[..., None] * n
        vt2 = dot(vt_vec, vt_vec)
        tinv = np.where(vt2 == 0.0, zero, rsqrt(np.maximum(vt2, np.asarray(1e-16, T_))))
        t = vt_vec * tinv[..., None]
        vt = vt2 * tinv
        Fn = np.maximum(zero, kn * delta - gamma_n * vn) * is_contact
        Ft = gamma_t * vt
        Ft_max = mu_eff * Fn
        Ft = np.minimum(Ft, Ft_max) * is_contact
        F = Fn[..., None] * n - Ft[..., None] * t
        if D == 3:
            tq = cross3(r_ci, F)
        else:
            tq = cross2(r_ci, F)
        mu_r_eff = np.minimum(law.mu_r[mi], law.mu_r[mj])
        R_eff = (ri * rj) / (ri + rj)
        w_rel = st.ang_vel[i] - st.ang_vel[j]
        w2 = dot(w_rel, w_rel) if A == 3 else w_rel[..., 0] * w_rel[..., 0]
        wsafe = np.where(w2 == 0.0, one, w2)
        w_hat = w_rel * rsqrt(wsafe)[..., None]
        tq = tq - (mu_r_eff * R_eff * Fn)[..., None] * w_hat
        return F, tq, None, is_contact > 0
    if law.kind in ("springfriction", "hertzfriction"):
        return _friction_law(law, st, i, j, dom, dt, xi_old, rij, d_sq, inv, r, n, ne)
    raise ValueError(f"unknown law {law.kind}")


def cs_beta_table(law: Law, T_) -> np.ndarray:
    """Damping ratio beta(e_eff) of cundall_strack.py:141-149 per material pair."""
    e = np.asarray(law.e, T_)
    e_eff = np.minimum(e[:, None], e[None, :])
    e_safe = np.where(e_eff > 0.0, e_eff, np.asarray(1.0, T_))
    ln_e = np.log(e_safe)
    pi2 = np.asarray(np.pi**2, T_)
    return np.where(e_eff > 0.0, -ln_e / np.sqrt(pi2 + ln_e * ln_e), np.asarray(1.0, T_))


def friction_tables(law: Law, T_):
    """Per material-pair coefficients of the history laws.

    springfriction: kn = young_eff, kt = kt_ratio * kn.
    hertzfriction:  kn_h = 4/3 E* (times sqrt(R* delta) per contact),
                    kt_h = 8 G* (times sqrt(R* delta)), Mindlin (1949).
    """
    if law.kind == "springfriction":
        kn = np.asarray(law.young_eff, T_)
        kt = np.asarray(law.kt_ratio, T_) * kn
    else:
        E = np.asarray(law.young, T_)
        nu = np.asarray(law.poisson, T_)
        one = np.asarray(1.0, T_)
        E_star = one / ((one - nu[:, None] * nu[:, None]) / E[:, None] + (one - nu[None, :] * nu[None, :]) / E[None, :])
        G = E / (2.0 * (one + nu))
        G_star = one / ((2.0 - nu[:, None]) / G[:, None] + (2.0 - nu[None, :]) / G[None, :])
        kn = np.asarray(4.0 / 3.0, T_) * E_star
        kt = np.asarray(8.0, T_) * G_star
    mu = np.asarray(law.mu_eff, T_)
    return kn, kt, mu


def _friction_law(law, st, i, j, dom, dt, xi_old, rij, d_sq, inv, r, n, ne):
    T_ = st.pos.dtype
    D = st.dim
    zero = np.asarray(0.0, T_)
    one = np.asarray(1.0, T_)
    ri, rj = st.rad[i], st.rad[j]
    mi, mj = st.mat_id[i], st.mat_id[j]
    kn_t, kt_t, mu_t = friction_tables(law, T_)
    kn, kt, mu = kn_t[mi, mj], kt_t[mi, mj], mu_t[mi, mj]
    delta = np.maximum(zero, (ri + rj) - r) * ne
    contact = delta > 0
    if law.kind == "hertzfriction":
        R_star = (ri * rj) / (ri + rj)
        s = np.sqrt(R_star * delta)
        kn = kn * s  # normal stiffness 4/3 E* sqrt(R* d): F = kn * d
        kt = kt * s
    m_i, m_j = st.mass[i], st.mass[j]
    m_eff = (m_i * m_j) / (m_i + m_j)
    gamma_n = np.asarray(2.0 * law.damping, T_) * np.sqrt(kn * m_eff)
    r_ci = -ri[..., None] * n
    r_cj = rj[..., None] * n
    v_ci = st.vel[i] + cross_w_r(st.ang_vel[i], r_ci)
    v_cj = st.vel[j] + cross_w_r(st.ang_vel[j], r_cj)
    v_rel = v_ci - v_cj
    vn = dot(v_rel, n)
    vt = v_rel - vn[..., None] * n
    Fn = np.maximum(zero, kn * delta - gamma_n * vn)
    xi = xi_old - dot(xi_old, n)[..., None] * n
    xi = xi + vt * np.asarray(dt, T_)
    Ft = (-kt)[..., None] * xi
    ft2 = dot(Ft, Ft)
    fmax = mu * Fn
    slip = ft2 > fmax * fmax
    ft2s = np.where(slip, ft2, one)
    scale = fmax * rsqrt(ft2s)
    Ft = np.where(slip[..., None], Ft * scale[..., None], Ft)
    ktsafe = np.where(kt > 0, kt, one)
    xi = np.where((slip & (kt > 0))[..., None], Ft / (-ktsafe)[..., None], xi)
    F = Fn[..., None] * n + Ft
    if D == 3:
        tq = cross3(r_ci, F)
    else:
        tq = cross2(r_ci, F)
    c = contact[..., None]
    F = np.where(c, F, zero)
    tq = np.where(c, tq, zero)
    xi = np.where(c, xi, zero)
    return F, tq, xi, contact


def law_energy(law: Law, st: State, i, j, dom: Domain) -> np.ndarray:
    """Pair energy: spring.py:138-147, hertz.py:147-162, cundall_strack.py:220-233.

    The history laws store their normal part only (elastic normal energy).
    Routed laws: the entry's energies summed from zeros (law_combiner.py:
    180-188), selected by species (router.py:262-330).
    """
    if law.kind == "routed":
        p = _route_index(law, st, i, j)
        e_all = np.zeros(p.shape, st.pos.dtype)
        for e in np.unique(p):
            ee = None
            for sub in law.terms[law.route_off[e] : law.route_off[e + 1]]:
                ek = np.zeros(p.shape, st.pos.dtype) if sub.kind == "none" else law_energy(sub, st, i, j, dom)
                ee = ek if ee is None else ee + ek
            if ee is not None:
                e_all = np.where(p == e, ee, e_all)
        return e_all
    T_ = st.pos.dtype
    zero = np.asarray(0.0, T_)
    one = np.asarray(1.0, T_)
    ri, rj = st.rad[i], st.rad[j]
    mi, mj = st.mat_id[i], st.mat_id[j]
    rij, d_sq = _gather_pair(st, i, j, dom)
    ne = (i != j).astype(T_)
    if law.kind in ("spring", "springfriction"):
        k = law.young_eff[mi, mj]
        inv = np.where(d_sq == 0.0, zero, rsqrt(np.maximum(d_sq, np.asarray(1e-16, T_))))
        r = d_sq * inv
        s = np.maximum(zero, (ri + rj) - r) * ne
        return 0.5 * k * (s * s)
    # norm() (linalg.py:115-135)
    r = np.where(d_sq == 0.0, zero, np.sqrt(np.maximum(d_sq, np.asarray(1e-16, T_))))
    delta = (ri + rj) - r
    delta = delta * ((delta > 0) & (i != j)).astype(T_)
    if law.kind in ("hertz", "hertzfriction"):
        Ei, Ej = law.young[mi], law.young[mj]
        nui, nuj = law.poisson[mi], law.poisson[mj]
        R_star = (ri * rj) / (ri + rj)
        E_star = one / ((one - nui * nui) / Ei + (one - nuj * nuj) / Ej)
        k = (np.asarray(4.0 / 3.0, T_) * E_star) * np.sqrt(R_star)
        return 0.4 * k * (delta * delta * np.sqrt(delta))
    if law.kind == "cundallstrack":
        Ei, Ej = law.young[mi], law.young[mj]
        kn = (2.0 * Ei * ri * Ej * rj) / (Ei * ri + Ej * rj)
        return 0.5 * kn * (delta * delta)
    raise ValueError(law.kind)


# --------------------------------------------------------------------------
# collider force accumulation (cell_list.py:437-464, naive.py:190-235,
# neighbor_list.py:594-669)
# --------------------------------------------------------------------------


def accumulate(st: State, P: Params, C: np.ndarray, hist=None):
    """Sequential per-particle sum over candidate columns of ``C`` (N, K).

    ``hist = (xi, live)`` aligned with the columns for the history laws;
    returns (F, T, new_hist).
    """
    N, D = st.pos.shape
    A = 3 if D == 3 else 1
    T_ = st.pos.dtype
    F = np.zeros((N, D), T_)
    T = np.zeros((N, A), T_)
    i = np.arange(N)
    new_xi = new_live = None
    if hist is not None:
        xi, live = hist
        new_xi = np.zeros_like(xi)
        new_live = np.zeros_like(live)
    for k in range(C.shape[1]):
        j = C[:, k]
        ok = j >= 0
        js = np.where(ok, j, 0)
        ok = ok & valid_mask(i, js, st, P.interact_same_bond_id)
        if not np.any(ok):
            continue
        xo = None
        if hist is not None:
            xo = np.where(live[:, k][:, None], xi[:, k], np.asarray(0.0, T_))
        f, t, xn, contact = law_force(P.law, st, i, js, P.domain, P.dt, xo)
        # skipping a pair adds nothing; adding an exact zero leaves the sum
        # unchanged, so masking with where() matches the GPU's skip.
        F = np.where(ok[:, None], F + f, F)
        T = np.where(ok[:, None], T + t, T)
        if hist is not None:
            keep = ok & contact
            new_xi[:, k] = np.where(keep[:, None], xn, np.asarray(0.0, T_))
            new_live[:, k] = keep
    return F, T, (None if hist is None else (new_xi, new_live))


def remap_history(xi, live, old_nl, new_nl):
    """``_remap_history_array`` (neighbor_list.py:29-54): match partner ids."""
    N, M = new_nl.shape
    nxi = np.zeros_like(xi)
    nlive = np.zeros_like(live)
    for r in range(N):
        old = old_nl[r]
        for s in range(M):
            j = new_nl[r, s]
            if j < 0:
                continue
            hits = np.nonzero(old == j)[0]
            if hits.size:
                t = hits[0]
                nxi[r, s] = xi[r, t]
                nlive[r, s] = live[r, t]
    return nxi, nlive


def drop_far_candidates(st: State, P: Params, C: np.ndarray) -> np.ndarray:
    """Remove candidates with ``|r_ij|^2 > 1.0001 (r_i + r_j)^2``, keeping order.

    Every law gives an exactly zero force, torque and energy for such a pair
    (the overlap ``R - |r_ij|`` is negative by far more than the rounding of
    ``d_sq * rsqrt(d_sq)``), and adding an exact zero leaves a sum unchanged,
    so the sums over the compacted candidates equal the full sums.  This only
    makes the all-pairs oracle fast.
    """
    N = st.N
    T_ = st.pos.dtype
    out = np.full_like(C, -1)
    width = 0
    for a in range(0, N, 2048):
        b = min(N, a + 2048)
        Cb = C[a:b]
        js = np.maximum(Cb, 0)
        ii = np.arange(a, b)[:, None]
        d = displacement_(st.pos[ii], st.pos[js], P.domain)
        d2 = dot(d, d)
        R = st.rad[ii] + st.rad[js]
        keep = (Cb >= 0) & ~(d2 > (R * R) * np.asarray(1.0001, T_))
        cnt = keep.sum(axis=1)
        order = np.argsort(~keep, axis=1, kind="stable")
        sel = np.take_along_axis(Cb, order, axis=1)
        w = int(cnt.max()) if cnt.size else 0
        width = max(width, w)
        t = np.arange(sel.shape[1])[None, :]
        out[a:b] = np.where(t < cnt[:, None], sel, -1)
    return out[:, : max(width, 1)]


def naive_near_candidates(st: State, P: Params) -> np.ndarray:
    """The rows of ``drop_far_candidates(naive_candidates(N))`` without the
    N x N distance matrix.

    A kept pair has ``|r_ij| <= sqrt(1.0001) (r_i + r_j) < 2.0004 r_max``, so a
    cell list with cells of 2.0004 r_max (search range 1) holds every kept
    pair; its candidates pruned by the same test and sorted by ``j`` are the
    all-pairs rows, which list the kept ``j`` in ascending order.
    """
    N = st.N
    if N == 0:
        return np.full((0, 1), -1, np.int64)
    T_ = st.pos.dtype
    cs = np.asarray(2.0004, T_) * np.asarray(st.rad.max(), T_)
    C, _ = cell_candidates(st.pos, P.domain, cs, 1)
    C = drop_far_candidates(st, P, C)
    key = np.where(C < 0, np.iinfo(np.int64).max, C)
    return np.take_along_axis(C, np.argsort(key, axis=1, kind="stable"), axis=1)


def compute_force(st: State, P: Params, nls: NLState | None = None):
    """``collider.compute_force`` for naive / CellList / NeighborList."""
    if st.clump is not None:
        _compute_force_clumps(st, P, nls)
        return
    if P.collider == "naive":
        if st.N > 1024 or P.naive_by_cells:
            C = naive_near_candidates(st, P)
        else:
            C = drop_far_candidates(st, P, naive_candidates(st.N))
        F, T, _ = accumulate(st, P, C)
    elif P.collider == "celllist":
        C, _ = cell_candidates(st.pos, P.domain, P.cell_size, P.search_range)
        C = drop_far_candidates(st, P, C)
        F, T, _ = accumulate(st, P, C)
    elif P.collider == "none":
        N, D = st.pos.shape
        F = np.zeros((N, D), st.pos.dtype)
        T = np.zeros((N, 3 if D == 3 else 1), st.pos.dtype)
    elif P.collider == "neighborlist":
        nl_check_and_rebuild(st, P, nls)
        hist = (nls.xi, nls.live) if nls.xi is not None else None
        F, T, h = accumulate(st, P, nls.nl, hist)
        if h is not None:
            nls.xi, nls.live = h
    else:
        raise ValueError(P.collider)
    st.force, st.torque = F, T


def _compute_force_clumps(st: State, P: Params, nls: NLState | None):
    """Clump members: forces at the lab positions (state.pos), same-clump pairs
    masked, then ``torque += pos_p_rot x force`` per member (naive.py:210-233,
    cell_list.py:461-462, neighbor_list.py:624-648)."""
    r = pos_p_rot(st)
    lab = st.copy()
    lab.pos = st.pos + r
    if P.collider == "neighborlist":
        # the list is checked and built on the lab positions, with the clump
        # mask (neighbor_list.py:58-131, 594-669)
        nl_check_and_rebuild(lab, P, nls)
        hist = (nls.xi, nls.live) if nls.xi is not None else None
        F, T, h = accumulate(lab, P, nls.nl, hist)
        if h is not None:
            nls.xi, nls.live = h
    else:
        if P.collider == "naive":
            if st.N > 1024 or P.naive_by_cells:
                C = naive_near_candidates(lab, P)
            else:
                C = drop_far_candidates(lab, P, naive_candidates(st.N))
        elif P.collider == "celllist":
            C, _ = cell_candidates(lab.pos, P.domain, P.cell_size, P.search_range)
            C = drop_far_candidates(lab, P, C)
        else:
            raise ValueError(P.collider)
        F, T, _ = accumulate(lab, P, C)
    lever = cross3(r, F) if st.dim == 3 else cross2(r, F)
    st.force, st.torque = F, T + lever


def nl_init(st: State, P: Params, history: bool) -> NLState:
    M = P.nl_max_neighbors
    nls = NLState(nl=np.full((st.N, M), -1, np.int64), old_pos=st.pos.copy())
    if history:
        nls.xi = np.zeros((st.N, M, st.dim), st.pos.dtype)
        nls.live = np.zeros((st.N, M), bool)
    return nls


def nl_check_and_rebuild(st: State, P: Params, nls: NLState) -> bool:
    """``_check_and_rebuild`` (neighbor_list.py:58-131)."""
    T_ = st.pos.dtype
    d = st.pos - nls.old_pos
    max_d2 = np.max(dot(d, d)) if st.N else np.asarray(0.0, T_)
    skin = np.asarray(P.nl_skin, T_)
    trigger = (skin * skin) / np.asarray(4.0, T_)
    if not (max_d2 > trigger or nls.n_build == 0):
        return False
    cutoff = np.asarray(P.nl_cutoff, T_) + skin
    if P.nl_secondary == "celllist":
        new_nl, ovf = neighbor_list_cells(
            st, P.domain, P.nl_secondary_cell_size, P.nl_secondary_search_range,
            cutoff, P.nl_max_neighbors, P.interact_same_bond_id,
        )
    else:
        new_nl, ovf = neighbor_list_naive(st, P.domain, cutoff, P.nl_max_neighbors, P.interact_same_bond_id)
    if nls.xi is not None:
        if nls.n_build == 0:
            nls.xi = np.zeros_like(nls.xi)
            nls.live = np.zeros_like(nls.live)
        else:
            nls.xi, nls.live = remap_history(nls.xi, nls.live, nls.nl, new_nl)
    nls.nl = new_nl
    nls.old_pos = st.pos.copy()
    nls.n_build += 1
    nls.overflow = ovf
    nls.n_rebuilds_this_call += 1
    return True


# --------------------------------------------------------------------------
# force manager (force_manager.py:341-425), spheres only
# --------------------------------------------------------------------------


def force_manager_apply(st: State, P: Params, ext: dict | None = None):
    F_part = ext.get("force") if ext else None
    F_com = ext.get("force_com") if ext else None
    T_part = ext.get("torque") if ext else None
    T_ = st.pos.dtype
    N, D = st.pos.shape
    A = 3 if D == 3 else 1
    F_part = np.zeros((N, D), T_) if F_part is None else F_part
    F_com = np.zeros((N, D), T_) if F_com is None else F_com
    T_part = np.zeros((N, A), T_) if T_part is None else T_part
    g = np.zeros(D, T_) if P.gravity is None else np.asarray(P.gravity, T_)
    if st.clump is None:
        F_com = F_com + g * (st.mass / np.asarray(1.0, T_))[:, None]
        F_total = (st.force + F_part) + F_com
        T_total = st.torque + T_part
        st.force, st.torque = F_total, T_total
        return
    # clumps (force_manager.py:392-412): every member holds the clump's total
    # mass, so gravity takes mass / member count; particle forces add their
    # lever torque; the totals are summed over the clump and broadcast
    count = np.bincount(st.clump, minlength=N)[st.clump].astype(T_)
    r = pos_p_rot(st)
    F_com = F_com + g * (st.mass / count)[:, None]
    T_part = T_part + (cross3(r, F_part) if D == 3 else cross2(r, F_part))
    F_total = (st.force + F_part) + F_com
    T_total = st.torque + T_part
    F_c = np.zeros((N, D), T_)
    T_c = np.zeros((N, A), T_)
    np.add.at(F_c, st.clump, F_total)  # sequential, ascending member index
    np.add.at(T_c, st.clump, T_total)
    st.force, st.torque = F_c[st.clump], T_c[st.clump]


# --------------------------------------------------------------------------
# integrators (velocity_verlet.py:57-94, direct_euler.py:62-65,
# velocity_verlet_spiral.py:83-180, spiral.py:24-144)
# --------------------------------------------------------------------------


def _free(st: State) -> np.ndarray:
    return (~st.fixed)[:, None].astype(st.pos.dtype)


def lin_before(st: State, P: Params):
    T_ = st.pos.dtype
    dt = np.asarray(P.dt, T_)
    if P.lin == "verlet":
        s = (dt * 0.5) / st.mass
        st.vel = st.vel + (st.force * s[:, None]) * _free(st)
        st.pos = st.pos + dt * st.vel


def lin_after(st: State, P: Params):
    T_ = st.pos.dtype
    dt = np.asarray(P.dt, T_)
    if P.lin == "verlet":
        s = (dt * 0.5) / st.mass
        st.vel = st.vel + (st.force * s[:, None]) * _free(st)
    elif P.lin == "euler":
        s = dt / st.mass
        st.vel = st.vel + (st.force * s[:, None]) * _free(st)
        st.pos = st.pos + dt * st.vel


def omega_dot(w, tq, I, invI):
    """``spiral.omega_dot`` (spiral.py:24-40)."""
    if w.shape[-1] == 1:
        return tq * invI
    wx, wy, wz = w[..., 0], w[..., 1], w[..., 2]
    ix, iy, iz = I[..., 0], I[..., 1], I[..., 2]
    cx = wy * (iz * wz) - wz * (iy * wy)
    cy = wz * (ix * wx) - wx * (iz * wz)
    cz = wx * (iy * wy) - wy * (ix * wx)
    c = np.stack([cx, cy, cz], axis=-1)
    return (tq - c) * invI


def _to_rotvec3(v):
    if v.shape[-1] == 1:
        z = np.zeros_like(v[..., 0])
        return np.stack([z * v[..., 0], z * v[..., 0], v[..., 0]], axis=-1)
    return v


def rot_before(st: State, P: Params):
    if P.rot != "verletspiral":
        return
    T_ = st.pos.dtype
    dt = np.asarray(P.dt, T_)
    dt_2 = dt / np.asarray(2.0, T_)
    invI = np.asarray(1.0, T_) / st.inertia
    mask = _free(st)
    if st.dim == 3:
        w = q_rotate_back(st.q, st.ang_vel)
        tq = q_rotate_back(st.q, st.torque)
    else:
        w, tq = st.ang_vel, st.torque
    k1 = dt_2 * omega_dot(w, tq, st.inertia, invI)
    k2 = dt_2 * omega_dot(w + k1, tq, st.inertia, invI)
    k3 = dt_2 * omega_dot(w + 0.25 * (k1 + k2), tq, st.inertia, invI)
    w = w + (mask * ((k1 + k2) + 4.0 * k3)) / np.asarray(6.0, T_)
    rotvec = _to_rotvec3((2.0 * dt_2) * w)
    dq = q_from_small_rotvec(rotvec)
    st.q = q_unit(q_mul(st.q, dq))
    st.ang_vel = q_rotate(st.q, w) if st.dim == 3 else w


def rot_after(st: State, P: Params):
    T_ = st.pos.dtype
    dt = np.asarray(P.dt, T_)
    if P.rot == "verletspiral":
        dt_2 = dt / np.asarray(2.0, T_)
        invI = np.asarray(1.0, T_) / st.inertia
        mask = _free(st)
        if st.dim == 3:
            w = q_rotate_back(st.q, st.ang_vel)
            tq = q_rotate_back(st.q, st.torque)
        else:
            w, tq = st.ang_vel, st.torque
        k1 = dt_2 * omega_dot(w, tq, st.inertia, invI)
        k2 = dt_2 * omega_dot(w + k1, tq, st.inertia, invI)
        k3 = dt_2 * omega_dot(w + 0.25 * (k1 + k2), tq, st.inertia, invI)
        w = w + (mask * ((k1 + k2) + 4.0 * k3)) / np.asarray(6.0, T_)
        st.ang_vel = q_rotate(st.q, w) if st.dim == 3 else w
    elif P.rot == "spiral":
        dt_2 = dt / np.asarray(2.0, T_)
        invI = np.asarray(1.0, T_) / st.inertia
        if st.dim == 3:
            w = q_rotate_back(st.q, st.ang_vel)
            tq = q_rotate_back(st.q, st.torque)
        else:
            w, tq = st.ang_vel, st.torque
        w_dot = omega_dot(w, tq, st.inertia, invI)
        rotvec1 = _to_rotvec3((2.0 * dt_2) * w)
        rotvec2 = _to_rotvec3(((2.0 * dt_2) * dt_2) * w_dot)
        dq = q_mul(q_from_small_rotvec(rotvec1), q_from_small_rotvec(rotvec2))
        st.q = q_unit(q_mul(st.q, dq))
        k1 = dt * w_dot
        k2 = dt * omega_dot(w + k1, tq, st.inertia, invI)
        k3 = dt * omega_dot(w + 0.25 * (k1 + k2), tq, st.inertia, invI)
        w = w + (_free(st) * ((k1 + k2) + 4.0 * k3)) / np.asarray(6.0, T_)
        st.ang_vel = q_rotate(st.q, w) if st.dim == 3 else w


# --------------------------------------------------------------------------
# domain.apply (reflect.py:102-297, _toc.py:11-94, free.py:42-58)
# --------------------------------------------------------------------------


def verlet_collision_fraction(v, acc, delta, ws, dt):
    """``verlet_collision_fraction`` (_toc.py:11-94)."""
    T_ = v.dtype
    zero = np.asarray(0.0, T_)
    one = np.asarray(1.0, T_)
    v0 = v - dt * acc
    vmid = v - (0.5 * dt) * acc
    v0n = v0 * ws
    vmidn = vmid * ws
    accn = acc * ws
    A = ((0.5 * accn) * dt) * dt
    B = v0n * dt
    dw = delta + vmidn * dt
    disc = B * B + (4.0 * A) * dw
    disc = np.maximum(zero, disc)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        sq = np.where(disc > 0.0, np.sqrt(np.where(disc > 0.0, disc, one)), zero)
        den1 = B - sq
        a1 = (2.0 * dw) / np.where(den1 < np.asarray(-1e-10, T_), den1, -one)
        twoA = 2.0 * A
        a2 = ((-B) - sq) / np.where(np.abs(twoA) > np.asarray(1e-10, T_), twoA, one)
        alpha = np.where(B < 0, a1, a2)
    return np.clip(alpha, zero, one)


def domain_apply(st: State, P: Params):
    dom = P.domain
    T_ = st.pos.dtype
    if dom.kind == "free":
        # free.py:42-58: bounding radius |pos_p| + rad (rad for spheres)
        br = st.rad[:, None] if st.pos_p is None else (norm(st.pos_p) + st.rad)[:, None]
        pmin = np.min(st.pos - br, axis=0)
        pmax = np.max(st.pos + br, axis=0)
        dom.box = pmax - pmin
        dom.anchor = pmin
        return
    if dom.kind != "reflect":
        return
    if st.clump is not None:
        _reflect_clumps(st, P)
        return
    dt = np.asarray(P.dt, T_)
    zero = np.asarray(0.0, T_)
    one = np.asarray(1.0, T_)
    e = np.asarray(dom.restitution, T_)
    pos = st.pos
    rad = st.rad[:, None]
    lo = dom.anchor + rad
    hi = (dom.anchor + dom.box) - rad
    over_lo = np.maximum(zero, lo - pos)
    over_hi = np.maximum(zero, pos - hi)
    inv_mass = one / st.mass
    denom = inv_mass[:, None] + zero  # r_p x n == 0 for spheres
    denom = np.where(denom == 0.0, one, denom)
    delta = np.maximum(over_lo, over_hi)
    wall_sign = (over_lo > 0).astype(T_) - (over_hi > 0).astype(T_)
    active = np.abs(wall_sign)
    acc = st.force * inv_mass[:, None]
    alpha = verlet_collision_fraction(st.vel, acc, delta, wall_sign, dt)
    a_min = np.min(np.where(active > 0, alpha, one), axis=-1, keepdims=True)
    dt_factor = (a_min - 1.0) * dt
    v_contact = st.vel + dt_factor * acc
    j = ((-(one + e)) * v_contact) / denom
    closing = ((v_contact * wall_sign) < 0.0).astype(T_)
    weight = active / np.maximum(active, one)
    j = j * (weight * closing)
    dv = j * inv_mass[:, None]
    dv = np.where(st.fixed[:, None], zero, dv)
    st.vel = st.vel + dv
    st.q = q_unit(st.q)
    dt_rem = (1.0 - a_min) * dt
    st.pos = st.pos + dv * dt_rem


def _seg(op, x: np.ndarray, ids: np.ndarray, init) -> np.ndarray:
    """``jax.ops.segment_{max,min,sum}`` over clump ids, broadcast back to the
    members; sums in ascending member index."""
    out = np.full((ids.max() + 1,) + x.shape[1:], init, x.dtype)
    op.at(out, ids, x)
    return out[ids]


def _reflect_clumps(st: State, P: Params):
    """``ReflectDomain.apply`` for clumps (reflect.py:160-297): each member's
    overlap past a wall, the deepest member per clump and axis takes the
    impulse, the clump's time of collision is its members' minimum, and the
    impulses are summed over the clump into linear and angular velocity
    changes with the position and orientation corrected over the remaining
    time.  The matrix products (R^T v, R v, the einsum of the effective
    mass) sum in index order."""
    dom = P.domain
    T_ = st.pos.dtype
    N, D = st.pos.shape
    A = 3 if D == 3 else 1
    dt = np.asarray(P.dt, T_)
    zero = np.asarray(0.0, T_)
    one = np.asarray(1.0, T_)
    e = np.asarray(dom.restitution, T_)
    ids = st.clump
    r = pos_p_rot(st)
    pos = st.pos + r
    rad = st.rad[:, None]
    lo = dom.anchor + rad
    hi = (dom.anchor + dom.box) - rad
    over_lo = np.maximum(zero, lo - pos)
    over_hi = np.maximum(zero, pos - hi)
    max_lo = _seg(np.maximum, over_lo, ids, -np.inf)
    max_hi = _seg(np.maximum, over_hi, ids, -np.inf)
    inv_mass = one / st.mass
    inv_I = one / st.inertia
    # n_prime[:, k] = R^T e_k, which is row k of R (body -> lab)
    eye = np.eye(D, dtype=T_)
    Rm = np.stack([q_rotate_back(st.q, np.broadcast_to(eye[k], (N, D))) for k in range(D)], axis=1)
    cr = cross3 if D == 3 else cross2
    rpn = np.stack([cr(st.pos_p, Rm[:, k]) for k in range(D)], axis=1)  # (N, D, A)
    sq = rpn * rpn
    denom_rot = sq[..., 0] * inv_I[:, None, 0]
    for i in range(1, A):
        denom_rot = denom_rot + sq[..., i] * inv_I[:, None, i]
    denom = inv_mass[:, None] + denom_rot
    denom = np.where(denom == 0.0, one, denom)
    delta = np.maximum(max_lo, max_hi)
    deep_lo = (over_lo > 0) & (over_lo == max_lo)
    deep_hi = (over_hi > 0) & (over_hi == max_hi)
    ws = deep_lo.astype(T_) - deep_hi.astype(T_)
    active = np.abs(ws)
    v_step = st.vel + cross_w_r(st.ang_vel, r)
    acc_clump = st.force * inv_mass[:, None]

    def mat(M, v, transpose):  # R v or R^T v, summing over j in order
        out = []
        for i in range(D):
            acc = None
            for j in range(D):
                m = M[:, j, i] if transpose else M[:, i, j]
                t = m * v[:, j]
                acc = t if acc is None else acc + t
            out.append(acc)
        return np.stack(out, axis=1)

    if D == 3:
        tb = mat(Rm, st.torque, True)
        wb = mat(Rm, st.ang_vel, True)
        ab = (tb - cross3(wb, st.inertia * wb)) * inv_I
        alpha_rot = mat(Rm, ab, False)
    else:
        alpha_rot = st.torque * inv_I
    acc_contact = acc_clump + cross_w_r(alpha_rot, r)
    alpha = verlet_collision_fraction(v_step, acc_contact, delta, ws, dt)
    a_min = np.min(np.where(active > 0, alpha, one), axis=-1, keepdims=True)
    a_clump = _seg(np.minimum, a_min, ids, np.inf)
    dt_factor = (a_clump - 1.0) * dt
    v_contact = v_step + dt_factor * acc_contact
    j = ((-(one + e)) * v_contact) / denom
    closing = ((v_contact * ws) < 0.0).astype(T_)
    count = _seg(np.add, active, ids, 0.0)
    weight = active / np.maximum(count, one)
    j = j * (weight * closing)
    dv = _seg(np.add, j * inv_mass[:, None], ids, 0.0)
    dv = np.where(st.fixed[:, None], zero, dv)
    st.vel = st.vel + dv
    jb = mat(Rm, j, True)
    moment = cr(st.pos_p, jb)
    if D == 2:
        dw = moment[..., -1:] * inv_I
    else:
        dw = mat(Rm, moment * inv_I, False)
    dw = _seg(np.add, dw, ids, 0.0)
    dw = np.where(st.fixed[:, None], zero, dw)
    st.ang_vel = st.ang_vel + dw
    dt_rem = (1.0 - a_clump) * dt
    dth = dw * dt_rem
    if D == 2:
        dth = np.concatenate([np.zeros_like(dth), np.zeros_like(dth), dth], axis=-1)
    st.q = q_unit(q_mul(q_from_small_rotvec(dth), st.q))
    st.pos = st.pos + dv * dt_rem


# --------------------------------------------------------------------------
# the time step (system.py:60-82) and drivers (system.py:85-120, 606-748)
# --------------------------------------------------------------------------


def step_once(st: State, P: Params, nls: NLState | None = None, ext: dict | None = None):
    domain_apply(st, P)
    lin_before(st, P)
    rot_before(st, P)
    compute_force(st, P, nls)
    force_manager_apply(st, P, ext)
    lin_after(st, P)
    rot_after(st, P)


def step(st: State, P: Params, n: int, nls: NLState | None = None) -> State:
    st = st.copy()
    for _ in range(int(n)):
        step_once(st, P, nls)
    return st


def potential_energy(st: State, P: Params, C: np.ndarray | None = None) -> float:
    """``compute_potential_energy`` (cell_list.py:469-496, naive.py:76-117):
    1/2 sum of pair energies; pair terms in the state dtype, sum in f64.
    Clump members interact at their lab positions (state.pos)."""
    if st.clump is not None:
        lab = st.copy()
        lab.pos = lab_pos(st)
        lab.clump = st.clump
        st = lab
    if C is None:
        if P.collider == "celllist":
            C, _ = cell_candidates(st.pos, P.domain, P.cell_size, P.search_range)
        else:
            C = naive_candidates(st.N)
    N = st.N
    i = np.arange(N)
    tot = np.zeros(N, np.float64)
    for k in range(C.shape[1]):
        j = C[:, k]
        ok = j >= 0
        js = np.where(ok, j, 0)
        ok = ok & valid_mask(i, js, st, P.interact_same_bond_id)
        e = law_energy(P.law, st, i, js, P.domain)
        tot += np.where(ok, 0.5 * e.astype(np.float64), 0.0)
    return float(tot.sum())


def kinetic_energy(st: State) -> float:
    """Translational + rotational KE (utils/thermal.py:20-122) in f64; a
    clump member counts 1/(member count) of its clump's energy."""
    m = st.mass.astype(np.float64)
    cnt = np.ones(st.N) if st.clump is None else np.bincount(st.clump, minlength=st.N)[st.clump].astype(np.float64)
    v = st.vel.astype(np.float64)
    w = st.ang_vel.astype(np.float64)
    if st.dim == 3:
        wb = q_rotate_back(st.q.astype(np.float64), w)
    else:
        wb = w
    I = st.inertia.astype(np.float64)
    return float(0.5 * np.sum((m / cnt) * np.sum(v * v, axis=1)) + 0.5 * np.sum(np.sum(I * wb * wb, axis=1) / cnt))


def pair_virial(st: State, P: Params, C: np.ndarray) -> np.ndarray:
    """Contact virial sum_{i<j} r_ij (x) F_ij (contacts.py:89-164), f64 sum."""
    N, D = st.pos.shape
    i = np.arange(N)
    W = np.zeros((D, D), np.float64)
    for k in range(C.shape[1]):
        j = C[:, k]
        ok = (j >= 0)
        js = np.where(ok, j, 0)
        ok = ok & valid_mask(i, js, st, P.interact_same_bond_id)
        f, _, _, _ = law_force(P.law, st, i, js, P.domain, P.dt,
                               np.zeros((N, D), st.pos.dtype))
        rij = displacement_(st.pos[i], st.pos[js], P.domain)
        contrib = rij[:, :, None].astype(np.float64) * f[:, None, :].astype(np.float64)
        W += 0.5 * np.sum(np.where(ok[:, None, None], contrib, 0.0), axis=0)
    return W


__all__ = [name for name in dir() if not name.startswith("_")]


# --------------------------------------------------------------------------
# FIRE energy minimisation (minimizers/optimizers.py:150-339,
# minimizers/routines.py:24-383), spheres
# --------------------------------------------------------------------------


@dataclass
class FireParams:
    """Arguments of ``fire(dt, ...)`` (optimizers.py:150-160)."""

    dt: float
    alpha_init: float = 0.1
    f_inc: float = 1.1
    f_dec: float = 0.5
    f_alpha: float = 0.99
    N_min: int = 5
    N_bad_max: int = 10
    dt_max_scale: float = 10.0
    dt_min_scale: float = 1e-3


@dataclass
class FireState:
    """``FIREState`` (optimizers.py:222-230): velocities of the two parameter
    leaves (pos_c, rotvec) and the scalars, in the parameter dtype."""

    vel_pos: np.ndarray
    vel_rot: np.ndarray
    dt: Any
    alpha: Any
    N_good: int = 0
    N_bad: int = 0


def fire_init(st: State, fp: FireParams) -> FireState:
    T_ = st.pos.dtype
    return FireState(
        vel_pos=np.zeros_like(st.pos),
        vel_rot=np.zeros_like(st.torque),
        dt=T_.type(fp.dt),
        alpha=T_.type(fp.alpha_init),
    )


def safe_norm_rows(x: np.ndarray, min_norm: float) -> np.ndarray:
    """``optax.safe_norm(x, min_norm, axis=-1, keepdims=True)``: the row norm,
    or ``min_norm`` where the norm is not above it."""
    T_ = x.dtype
    if x.shape[-1] == 1:
        n = np.sqrt(x[..., 0] * x[..., 0])
    elif x.shape[-1] == 2:
        n = np.sqrt(x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1])
    else:
        n = np.sqrt(dot(x, x))
    return np.where(n <= T_.type(min_norm), T_.type(min_norm), n)[..., None]


def fire_power(Fp: np.ndarray, Fr: np.ndarray, fs: FireState) -> float:
    """Power ``sum(F . v_old)`` over both leaves (optimizers.py:245-248).
    Each float32 product is exact in float64; the sum is taken in float64."""
    dt = fs.dt
    vp = fs.vel_pos + Fp * dt / Fp.dtype.type(2.0)
    vr = fs.vel_rot + Fr * dt / Fr.dtype.type(2.0)
    return float(np.sum(Fp.astype(np.float64) * vp.astype(np.float64)) +
                 np.sum(Fr.astype(np.float64) * vr.astype(np.float64)))


def fire_update(Fp: np.ndarray, Fr: np.ndarray, fs: FireState, fp: FireParams, power: float | None = None):
    """``fire().update`` (optimizers.py:241-318) with F = -grad for the two
    leaves.  Returns (update_pos, update_rot, new FireState)."""
    T_ = Fp.dtype.type
    two = T_(2.0)
    dt = fs.dt
    v_old_p = fs.vel_pos + Fp * dt / two
    v_old_r = fs.vel_rot + Fr * dt / two
    if power is None:
        power = fire_power(Fp, Fr, fs)
    dt_cand_inc = np.minimum(dt * T_(fp.f_inc), T_(fp.dt * fp.dt_max_scale))
    dt_cand_dec = np.maximum(dt * T_(fp.f_dec), T_(fp.dt * fp.dt_min_scale))
    if power > 0.0:  # downhill (optimizers.py:278-292)
        N_good = fs.N_good + 1
        new_dt = dt_cand_inc if N_good > fp.N_min else dt
        new_alpha = fs.alpha * T_(fp.f_alpha) if N_good > fp.N_min else fs.alpha
        N_bad = 0
        dt_rev = T_(0.0)
        vs = T_(1.0)
    else:  # uphill (optimizers.py:258-276)
        N_bad = fs.N_bad + 1
        exceeded = N_bad > fp.N_bad_max
        new_dt = T_(fp.dt) if exceeded else dt_cand_dec
        N_bad = 0 if exceeded else N_bad
        new_alpha = T_(fp.alpha_init)
        N_good = 0
        dt_rev = -new_dt
        vs = T_(0.0)
    out = []
    for F, v_old in ((Fp, v_old_p), (Fr, v_old_r)):
        v_half = v_old * vs + F * new_dt / two
        vn = safe_norm_rows(v_half, 1e-16)
        fn = safe_norm_rows(F, 1e-16)
        mix = np.where(fn > T_(1e-16), vn / fn * new_alpha, T_(0.0))
        v_half = (v_half * (T_(1.0) - new_alpha) + F * mix) * vs
        upd = v_old * dt_rev / two + v_half * new_dt / two
        out.append((upd, v_half))
    ns = FireState(vel_pos=out[0][1], vel_rot=out[1][1], dt=new_dt, alpha=new_alpha, N_good=N_good, N_bad=N_bad)
    return out[0][0], out[1][0], ns


def q_from_rotvec(r: np.ndarray) -> np.ndarray:
    """``Quaternion.from_rotvec`` (quaternion.py:99-127)."""
    T_ = r.dtype.type
    n2 = dot(r, r)
    safe = np.maximum(n2, T_(1e-16))
    theta = np.sqrt(safe)
    half = T_(0.5) * theta
    ch = np.cos(half)
    sc = np.sin(half) / theta
    return np.concatenate([ch[..., None], r * sc[..., None]], axis=-1)


def apply_fire_params(st: State, upd_pos: np.ndarray, upd_rot: np.ndarray) -> None:
    """``optax.apply_updates`` + ``jnp.where(mask, new, old)`` (routines.py:
    351-360) and ``_delta_params_to_state`` (routines.py:37-65): positions
    move, the orientation turns by the delta rotation vector about itself."""
    m = (~st.fixed)[:, None]
    st.pos = np.where(m, st.pos + upd_pos * m.astype(st.pos.dtype), st.pos)
    zero = np.zeros_like(upd_rot)
    rv = np.where(m, zero + upd_rot * m.astype(st.pos.dtype), zero)
    if st.dim == 2:
        rv = np.concatenate([np.zeros_like(st.pos), rv], axis=-1)
    st.q = q_unit(q_mul(q_from_rotvec(rv), st.q))


def gravity_energy(st: State, P: Params) -> float:
    """``ForceManager.compute_potential_energy`` gravity term
    (force_manager.py:450-459) for spheres: -sum (g . x) m."""
    if P.gravity is None:
        return 0.0
    g = np.asarray(P.gravity, st.pos.dtype)
    e = -(dot(np.broadcast_to(g, st.pos.shape), st.pos) * st.mass)
    return float(e.astype(np.float64).sum())


def fire_eval(st: State, P: Params, nls: NLState | None = None) -> tuple[float, float]:
    """``eval_step`` (routines.py:228-245): forces, torques and the potential
    energy at the current parameters; returns (pe, max |grad| component)."""
    compute_force(st, P, nls)
    force_manager_apply(st, P)
    if P.collider == "neighborlist":
        pe = potential_energy(st, P, nls.nl)
    else:
        pe = potential_energy(st, P)
    pe += gravity_energy(st, P)
    g = max(float(np.max(np.abs(st.force), initial=0.0)), float(np.max(np.abs(st.torque), initial=0.0)))
    return pe, g


def minimize(st: State, P: Params, fp: FireParams, max_steps: int = 10000, pe_tol: float = 1e-16,
             pe_diff_tol: float = 1e-16, force_tol: float = 0.0, nls: NLState | None = None):
    """``minimize`` (routines.py:150-383) with FIRE; the potential energy and
    the power are float64 sums.  Returns (steps, pe per particle)."""
    N = st.N
    fs = fire_init(st, fp)
    pe, gmax = fire_eval(st, P, nls)
    prev = math.inf
    steps = 0
    tiny = np.finfo(np.float64).tiny
    while steps < max_steps:
        conv_pe = abs(pe / N) <= pe_tol
        denom = max(abs(pe), abs(prev), tiny)
        conv_rel = abs(pe - prev) / denom < pe_diff_tol
        conv_f = gmax <= force_tol
        if conv_pe or conv_rel or conv_f:
            break
        m = (~st.fixed)[:, None].astype(st.pos.dtype)
        up, ur, fs = fire_update(st.force * m, st.torque * m, fs, fp)
        apply_fire_params(st, up, ur)
        prev = pe
        pe, gmax = fire_eval(st, P, nls)
        steps += 1
    return steps, pe / N


def _selfcheck() -> None:  # pragma: no cover - manual use
    st = make_state([[0.0, 0.0, 0.0], [0.9, 0.0, 0.0]], [0.5, 0.5])
    P = Params(dim=3, dt=1e-3, domain=Domain("free", np.ones(3, np.float32), np.zeros(3, np.float32)),
               law=Law("spring", young_eff=np.array([[1e4]], np.float32)))
    compute_force(st, P)
    assert math.isclose(float(st.force[0, 0]), -1e4 * 0.1, rel_tol=1e-5)
