"""Seeded synthetic inputs of the DEM workloads, as host numpy arrays.

Both the GPU path and any checker consume these same arrays.  numpy's PCG64
stands in for JAX's PRNG.

* ``spheres_1k``: 10x10x10 lattice, spacing 0.5, radius 0.1, unit density,
  velocities N(0, 0.1^2) (seed 0), reflective 20^3 box at the origin,
  linear spring k = 1e4, VV + VV-SPIRAL, dt = 1e-3, float32 (the README
  example of JaxDEM, README.md:34-63).
* ``spheres_1m_friction``: N = 2^20 bidisperse spheres (r = 0.5 / 0.7,
  50:50 by seeded permutation) on the first N sites of a 102^3 lattice with
  +/-5 % jitter, periodic cube at volume fraction 0.50, velocities N(0, 1)
  with zero net momentum, unit density, spring k_n = 1e4 + tangential spring
  (k_t = 2/7 k_n) with Coulomb mu = 0.5 and normal damping ratio 0.1,
  NeighborList over a CellList with cell size 1.4 (the largest diameter),
  VV + VV-SPIRAL, dt = t_c / 50 with t_c = pi sqrt(m_eff / k_n) of the
  lightest pair, float32.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    pos: np.ndarray
    vel: np.ndarray
    rad: np.ndarray
    mass: np.ndarray
    box: np.ndarray
    anchor: np.ndarray
    domain: str
    dt: float
    law: str
    young: float
    mu: float = 0.0
    damping: float = 0.0
    kt_ratio: float = 2.0 / 7.0
    collider: str = "naive"
    cell_size: float | None = None
    cutoff: float | None = None
    skin: float | None = None
    max_neighbors: int | None = None
    extra: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.pos.shape[0])

    @property
    def dim(self) -> int:
        return int(self.pos.shape[1])


def sphere_volume(r: np.ndarray) -> np.ndarray:
    return ((4.0 / 3.0) * math.pi * r.astype(np.float64) ** 3).astype(np.float32)


def spheres_1k(seed: int = 0) -> Workload:
    n = 10
    sp = np.float32(0.5)
    ax = np.arange(n, dtype=np.float32) * sp
    pos = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    N = pos.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = (0.1 * rng.standard_normal((N, 3))).astype(np.float32)
    rad = np.full(N, 0.1, np.float32)
    return Workload(
        name="spheres_1k",
        pos=pos,
        vel=vel,
        rad=rad,
        mass=sphere_volume(rad),
        box=np.full(3, 20.0, np.float32),
        anchor=np.zeros(3, np.float32),
        domain="reflect",
        dt=1e-3,
        law="spring",
        young=1e4,
        collider="naive",
    )


def spheres_friction(N: int = 1 << 20, seed: int = 0, phi: float = 0.50, skin: float = 0.1) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = np.where(rng.permutation(N) < N // 2, 0.5, 0.7).astype(np.float32)
    vbar = float(np.mean((4.0 / 3.0) * math.pi * rad.astype(np.float64) ** 3))
    L = (N * vbar / phi) ** (1.0 / 3.0)
    side = int(math.ceil(N ** (1.0 / 3.0)))
    spacing = L / side
    idx = np.arange(N)
    ijk = np.stack([idx // (side * side), (idx // side) % side, idx % side], axis=1).astype(np.float64)
    pos = (ijk + 0.5) * spacing
    pos = pos + rng.uniform(-0.05 * spacing, 0.05 * spacing, pos.shape)
    pos = pos.astype(np.float32)
    vel = rng.standard_normal((N, 3))
    vel = (vel - vel.mean(axis=0)).astype(np.float32)
    mass = sphere_volume(rad)
    k = 1e4
    m_small = float(sphere_volume(np.array([0.5], np.float32))[0])
    m_eff = 0.5 * m_small
    t_c = math.pi * math.sqrt(m_eff / k)
    return Workload(
        name=f"spheres_{N}_friction",
        pos=pos,
        vel=vel,
        rad=rad,
        mass=mass,
        box=np.full(3, L, np.float32),
        anchor=np.zeros(3, np.float32),
        domain="periodic",
        dt=t_c / 50.0,
        law="springfriction",
        young=k,
        mu=0.5,
        damping=0.1,
        collider="neighborlist",
        cell_size=1.4,
        cutoff=1.4,
        skin=skin,
        max_neighbors=None,
        extra={"spacing": spacing, "L": L, "t_c": t_c},
    )


def disks_hertz_friction(N: int, seed: int, phi: float, skin: float = 0.1, E: float = 1e4, mu: float = 0.3,
                         dt: float = 2e-4) -> Workload:
    """One 2D system of the batching workload: N disks, radii U[0.4, 0.6],
    periodic square box at area fraction phi, on the first N sites of a
    jittered square lattice (+/-5 % of the spacing), velocities N(0, 1) with
    zero net momentum, unit density, Hertz normal contact (E, nu = 0.3) with
    the tangential spring and Coulomb friction mu, NeighborList over a
    CellList, VV + VV-SPIRAL, float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = rng.uniform(0.4, 0.6, N).astype(np.float32)
    area = float(np.sum(math.pi * rad.astype(np.float64) ** 2))
    L = math.sqrt(area / phi)
    side = int(math.ceil(math.sqrt(N)))
    spacing = L / side
    idx = np.arange(N)
    ij = np.stack([idx // side, idx % side], axis=1).astype(np.float64)
    pos = (ij + 0.5) * spacing + rng.uniform(-0.05 * spacing, 0.05 * spacing, (N, 2))
    vel = rng.standard_normal((N, 2))
    vel = (vel - vel.mean(axis=0)).astype(np.float32)
    mass = (math.pi * rad.astype(np.float64) ** 2).astype(np.float32)
    return Workload(
        name=f"disks_{N}_hertzfriction",
        pos=pos.astype(np.float32),
        vel=vel,
        rad=rad,
        mass=mass,
        box=np.full(2, L, np.float32),
        anchor=np.zeros(2, np.float32),
        domain="periodic",
        dt=dt,
        law="hertzfriction",
        young=E,
        mu=mu,
        damping=0.0,
        collider="neighborlist",
        cell_size=1.2,
        cutoff=1.2,
        skin=skin,
        extra={"phi": phi, "L": L},
    )


def batch_disks(n_systems: int = 4096, seed: int = 0, n_min: int = 256, n_max: int = 16384, skin: float = 0.1,
                every: int = 1) -> list:
    """The batching workload: n_systems independent disk systems, N
    log-uniform in [n_min, n_max] and area fraction uniform in [0.3, 0.8]
    (both seeded); system k uses seed (seed, k) for its own arrays.  With
    ``every`` > 1 only systems 0, every, 2 every, ... are made (a sample of
    the same systems)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ns = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_systems))).astype(int)
    phis = rng.uniform(0.3, 0.8, n_systems)
    return [disks_hertz_friction(int(N), seed=seed * 1000003 + k, phi=float(p), skin=skin)
            for k, (N, p) in enumerate(zip(Ns, phis)) if k % every == 0]


def polydisperse_spheres(N: int, seed: int, phi: float, domain: str, mu: float, skin: float = 0.1,
                         k: float = 1e4, damping: float = 0.1) -> Workload:
    """One request of the serving workload: N spheres with log-normal radii
    (mean 0.5, log-space sigma 0.2, clipped to [0.25, 1.0]) on the first N
    sites of a jittered cubic lattice in a cube at volume fraction phi
    (periodic or reflect), velocities N(0, 1) with zero net momentum, unit
    density, spring k_n = k with the tangential spring, Coulomb mu and normal
    damping ratio, NeighborList over a CellList, VV + VV-SPIRAL, dt = t_c/50
    of the lightest pair, float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s2 = 0.2
    rad = np.exp(rng.normal(math.log(0.5) - 0.5 * s2 * s2, s2, N)).clip(0.25, 1.0).astype(np.float32)
    vol = float(np.sum((4.0 / 3.0) * math.pi * rad.astype(np.float64) ** 3))
    L = (vol / phi) ** (1.0 / 3.0)
    side = int(math.ceil(N ** (1.0 / 3.0)))
    spacing = L / side
    idx = np.arange(N)
    ijk = np.stack([idx // (side * side), (idx // side) % side, idx % side], axis=1).astype(np.float64)
    pos = (ijk + 0.5) * spacing + rng.uniform(-0.05 * spacing, 0.05 * spacing, (N, 3))
    vel = rng.standard_normal((N, 3))
    vel = (vel - vel.mean(axis=0)).astype(np.float32)
    mass = sphere_volume(rad)
    m_min = float(mass.min())
    t_c = math.pi * math.sqrt(0.5 * m_min / k)
    rmax = float(rad.max())
    return Workload(
        name=f"poly_{N}_{domain}",
        pos=pos.astype(np.float32),
        vel=vel,
        rad=rad,
        mass=mass,
        box=np.full(3, L, np.float32),
        anchor=np.zeros(3, np.float32),
        domain=domain,
        dt=t_c / 50.0,
        law="springfriction",
        young=k,
        mu=mu,
        damping=damping,
        collider="neighborlist",
        cell_size=2.0 * rmax,
        cutoff=2.0 * rmax,
        skin=skin,
        extra={"phi": phi, "L": L},
    )


def serving_requests(n_requests: int, rate: float, seed: int = 0, n_min: int = 1000, n_max: int = 200000,
                     steps_min: int = 100, steps_max: int = 10000) -> list:
    """The serving workload: Poisson arrivals at ``rate`` per second; per
    request N log-uniform in [n_min, n_max], steps uniform in [steps_min,
    steps_max], periodic or reflect (50:50), volume fraction U[0.3, 0.55],
    friction mu U[0, 0.5].  Returns (arrival time, steps, Workload) tuples."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for k in range(n_requests):
        N = int(np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        steps = int(rng.integers(steps_min, steps_max + 1))
        dom = "periodic" if rng.random() < 0.5 else "reflect"
        phi = float(rng.uniform(0.3, 0.55))
        mu = float(rng.uniform(0.0, 0.5))
        w = polydisperse_spheres(N, seed=seed * 1000003 + 7 * k + 1, phi=phi, domain=dom, mu=mu)
        if dom == "reflect":
            w.box = w.box * np.float32(1.02)  # lattice sites stay inside the walls
        out.append((float(t[k]), steps, w))
    return out


def bidisperse_jamming(N: int = 1 << 18, seed: int = 0, phi: float = 0.66) -> tuple:
    """Energy-minimisation start: N spheres, radii 0.5 and 0.7 (1 : 1.4)
    alternating 50 : 50, uniform random positions (float64) in a periodic
    cube at volume fraction phi (N = 2^18 at 0.66 gives L = 73.02).
    Returns (pos, rad, L)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = np.where(np.arange(N) % 2 == 0, 0.5, 0.7)
    L = float((np.sum(4.0 / 3.0 * math.pi * rad**3) / phi) ** (1.0 / 3.0))
    return rng.uniform(0.0, L, (N, 3)), rad, L


def make_gpu_system(w: Workload, device="cuda", max_neighbors: int | None = None, collider: str | None = None):
    """(State, System) of the jaxdem_amd API for a workload."""
    from . import Material, MaterialMatchmaker, MaterialTable, State, System

    st = State.create(pos=w.pos, vel=w.vel, rad=w.rad, mass=w.mass, device=device)
    col = collider or w.collider
    if w.law in ("springfriction", "hertzfriction"):
        mats = [Material.create("elasticfrict", density=1.0, young=w.young, poisson=0.3, mu=w.mu, e=1.0)]
        fm_kw = {"kt_ratio": w.kt_ratio, "damping": w.damping}
    else:
        mats = [Material.create("elastic", density=1.0, young=w.young, poisson=0.3)]
        fm_kw = {}
    mt = MaterialTable.from_materials(mats, matcher=MaterialMatchmaker.create("harmonic"))
    ckw: dict = {}
    if col == "neighborlist":
        ckw = dict(
            state=st,
            cutoff=w.cutoff,
            skin=w.skin,
            secondary_collider_type="CellList",
            secondary_collider_kw=dict(state=st, cell_size=w.cell_size),
        )
        if max_neighbors is not None or w.max_neighbors is not None:
            ckw["max_neighbors"] = max_neighbors or w.max_neighbors
    elif col == "celllist":
        ckw = dict(state=st) if w.cell_size is None else dict(state=st, cell_size=w.cell_size)
    sysm = System.create(
        st.shape,
        dt=w.dt,
        domain_type=w.domain,
        domain_kw=dict(box_size=w.box, anchor=w.anchor),
        collider_type={"naive": "naive", "celllist": "CellList", "neighborlist": "NeighborList"}[col],
        collider_kw=ckw,
        force_model_type=w.law,
        force_model_kw=fm_kw,
        mat_table=mt,
    )
    return st, sysm


__all__ = [
    "Workload",
    "spheres_1k",
    "spheres_friction",
    "disks_hertz_friction",
    "batch_disks",
    "polydisperse_spheres",
    "serving_requests",
    "make_gpu_system",
    "sphere_volume",
]
