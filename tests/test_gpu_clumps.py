"""Clumps (rigid bodies of spheres) on the GPU against the oracle on the same
seeded inputs: dimers, trimers and single spheres with random orientations
and spins; forces and torques bit for bit (naive and CellList, 2D and 3D),
whole steps of every integrator pair with gravity (periodic, free and
reflective domains), NeighborList runs with rebuilds on the lab positions, potential and
kinetic energy to 1e-12, the float64 form within 1e-12 of the float64
oracle; the combinations off the GPU path refuse."""

from __future__ import annotations

import contextlib
import dataclasses
import math

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from oracle import jdem_oracle as O
from tests import oracle_bridge as B
from tests.test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


def clump_scene(n_side, dim, seed, spacing=1.7, r=0.5, d=0.45, device="cuda", dtype=np.float32):
    """A jittered lattice of bodies: dimers, trimers and single spheres with
    random orientations, velocities and spins.  Returns (state, box)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    A = 3 if dim == 3 else 1
    pos, pp, cid, mass, inertia, q, vel, w = [], [], [], [], [], [], [], []
    k = 0
    for idx in np.ndindex(*([n_side] * dim)):
        c = (np.array(idx) + 0.5) * spacing + rng.uniform(-0.25, 0.25, dim)
        kind = k % 3  # 0 sphere, 1 dimer, 2 trimer
        offs = {0: [[0.0, 0.0, 0.0]], 1: [[-d, 0.0, 0.0], [d, 0.0, 0.0]],
                2: [[-d, 0.0, 0.0], [d, 0.0, 0.0], [0.0, d, 0.0]]}[kind]
        offs = np.array(offs)[:, :dim]
        offs = offs - offs.mean(0)
        m1 = (4.0 / 3.0 * math.pi * r**3) if dim == 3 else math.pi * r * r
        M = m1 * len(offs)
        if dim == 3:
            qq = rng.standard_normal(4)
            qq /= np.linalg.norm(qq)
            # principal inertia of the body about its centre (member spheres + parallel axes), body frame
            I = np.zeros(3)
            for o in offs:
                I += 0.4 * m1 * r * r + m1 * (np.dot(o, o) - o * o)
        else:
            th = rng.uniform(0, 2 * np.pi)
            qq = np.array([np.cos(th / 2), 0.0, 0.0, np.sin(th / 2)])
            I = np.array([sum(0.5 * m1 * r * r + m1 * np.dot(o, o) for o in offs)])
        v = rng.standard_normal(dim)
        om = rng.standard_normal(A)
        for o in offs:
            pos.append(c)
            pp.append(o)
            cid.append(k)
            mass.append(M)
            inertia.append(I)
            q.append(qq)
            vel.append(v)
            w.append(om)
        k += 1
    box = np.full(dim, n_side * spacing)
    st = jd.State.create(pos=np.array(pos, dtype), pos_p=np.array(pp, dtype), clump_id=np.array(cid),
                         mass=np.array(mass, dtype), inertia=np.array(inertia, dtype), q=np.array(q, dtype),
                         vel=np.array(vel, dtype), ang_vel=np.array(w, dtype), rad=np.full(len(pos), r, dtype),
                         device=device)
    assert not st.is_sphere_state
    return st, box


def clump_system(st, box, law, collider, domain="periodic", dt=1e-4, lin="verlet", rot="verletspiral", g=None):
    mats = [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]
    ckw = dict(state=st) if collider == "CellList" else {}
    fm = {"gravity": g} if g is not None else None
    return jd.System.create(st.shape, dt=dt, domain_type=domain,
                            domain_kw=dict(box_size=box, anchor=np.zeros(st.dim)), collider_type=collider,
                            collider_kw=ckw, force_model_type=law, mat_table=jd.MaterialTable.from_materials(mats),
                            linear_integrator_type=lin, rotation_integrator_type=rot, force_manager_kw=fm)


@pytest.mark.parametrize("collider", ["naive", "CellList"])
@pytest.mark.parametrize("law", ["spring", "cundallstrack"])
@pytest.mark.parametrize("dim", [3, 2])
def test_clump_compute_force_bitwise(device, dim, law, collider):
    print(f"[clumps] force {law} {collider} dim={dim}", flush=True)
    st, box = clump_scene(7 if dim == 3 else 16, dim, seed=4)
    sy = clump_system(st, box, law, collider)
    st2, _ = sy.collider.compute_force(st, sy)
    ost = B.state_of(st)
    O.compute_force(ost, B.params_of(st, sy))
    assert np.abs(ost.force).max() > 0
    assert_bitwise("force", st2.force.cpu().numpy(), ost.force)
    assert_bitwise("torque", st2.torque.cpu().numpy(), ost.torque)


CASES = [("naive", "periodic", "verlet", "verletspiral"), ("naive", "free", "verlet", "verletspiral"),
         ("CellList", "periodic", "euler", "spiral"), ("CellList", "periodic", "verlet", "verletspiral"),
         ("naive", "reflect", "verlet", "verletspiral"), ("CellList", "reflect", "euler", "spiral")]


@pytest.mark.parametrize("collider,domain,lin,rot", CASES)
@pytest.mark.parametrize("dim", [3, 2])
def test_clump_steps_bitwise(device, dim, collider, domain, lin, rot):
    """30 steps with gravity: every member's state bit for bit, members of a
    clump identical, the free box fitted with |pos_p| + rad."""
    print(f"[clumps] step {collider} {domain} {lin}/{rot} dim={dim}", flush=True)
    st, box = clump_scene(6 if dim == 3 else 14, dim, seed=7)
    g = np.array([0.0, -2.0, 0.0][:dim])
    sy = clump_system(st, box, "cundallstrack", collider, domain, lin=lin, rot=rot, g=g)
    n = 30
    st2, sy2 = jd.System.step(st, sy, n=n)
    P = B.params_of(st, sy)
    ost = O.step(B.state_of(st), P, n)
    for f, of in (("pos_c", "pos"), ("vel", "vel"), ("ang_vel", "ang_vel"), ("force", "force"),
                  ("torque", "torque")):
        assert_bitwise(f, getattr(st2, f).cpu().numpy(), getattr(ost, of))
    assert_bitwise("q", st2.q.data.cpu().numpy(), ost.q)
    cid = st.clump_id.cpu().numpy()
    F = st2.force.cpu().numpy()
    for c in np.unique(cid)[:20]:
        assert np.all(F[cid == c] == F[cid == c][0])
    if domain == "free":
        np.testing.assert_array_equal(sy2.domain.box_size, P.domain.box)
        np.testing.assert_array_equal(sy2.domain.anchor, P.domain.anchor)


def test_clump_energies(device):
    print("[clumps] energies", flush=True)
    st, box = clump_scene(7, 3, seed=9)
    sy = clump_system(st, box, "spring", "CellList")
    e = sy.potential_energy(st)
    ke = jd.utils.compute_kinetic_energy(st, sy)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    lab = ost.copy()
    lab.pos = O.lab_pos(ost)
    C, _ = O.cell_candidates(lab.pos, P.domain, P.cell_size, P.search_range)
    oe = O.potential_energy(ost, P, C)
    assert e > 0 and e == pytest.approx(oe, rel=1e-12)
    assert ke == pytest.approx(O.kinetic_energy(ost), rel=1e-12)


def test_clump_f64(device):
    """The float64 form: 20 steps within 1e-12 of the float64 oracle."""
    print("[clumps] float64", flush=True)
    old = jd.config.jax_enable_x64
    jd.config.update("jax_enable_x64", True)
    try:
        st, box = clump_scene(6, 3, seed=12, dtype=np.float64)
        assert st.pos_c.dtype == torch.float64
        sy = clump_system(st, box, "hertz", "naive", g=np.array([0.0, 0.0, -1.0]))
        st2, _ = jd.System.step(st, sy, n=20)
    finally:
        jd.config.update("jax_enable_x64", old)
    ost = O.step(B.state_of(st, np.float64), B.params_of(st, sy, np.float64), 20)
    for f, of in (("pos_c", "pos"), ("vel", "vel"), ("ang_vel", "ang_vel"), ("force", "force")):
        a = getattr(st2, f).cpu().numpy()
        b = getattr(ost, of)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), f


def _nl_clump_system(st, box, secondary, law="cundallstrack", skin=0.05, g=None):
    mats = [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]
    skw = dict(state=st) if secondary == "CellList" else {}
    return jd.System.create(
        st.shape, dt=1e-4, domain_type="periodic", domain_kw=dict(box_size=box, anchor=np.zeros(st.dim)),
        collider_type="NeighborList", force_model_type=law, mat_table=jd.MaterialTable.from_materials(mats),
        collider_kw=dict(state=st, cutoff=float(2 * st.rad.max().item()), skin=skin,
                         secondary_collider_type=secondary, secondary_collider_kw=skw),
        force_manager_kw={"gravity": g} if g is not None else None)


@pytest.mark.parametrize("secondary", ["CellList", "naive"])
@pytest.mark.parametrize("dim", [3, 2])
def test_clump_neighborlist_bitwise(device, dim, secondary):
    """60 steps on a NeighborList built and checked on the members' lab
    positions (pairs of one clump unlisted), with rebuilds: state, list rows,
    build positions and build count bit for bit."""
    print(f"[clumps] neighborlist {secondary} dim={dim}", flush=True)
    st, box = clump_scene(6 if dim == 3 else 14, dim, seed=21)
    sy = _nl_clump_system(st, box, secondary, g=np.array([0.0, -1.0, 0.0][:dim]))
    n = 60
    st2, sy2 = jd.System.step(st, sy, n=n)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    nls = B.nl_state_of(st, sy)
    ost = O.step(ost, P, n, nls)
    for f, of in (("pos_c", "pos"), ("vel", "vel"), ("ang_vel", "ang_vel"), ("force", "force"),
                  ("torque", "torque")):
        assert_bitwise(f, getattr(st2, f).cpu().numpy(), getattr(ost, of))
    assert sy2.collider.n_build_times == nls.n_build and nls.n_build >= 2
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    assert_bitwise("old_pos", sy2.collider.old_pos.cpu().numpy(), nls.old_pos)
    # no row lists a member of the particle's own clump
    nl = sy2.collider.neighbor_list.cpu().numpy()
    cid = st.clump_id.cpu().numpy()
    i, k = np.nonzero(nl >= 0)
    assert not np.any(cid[i] == cid[nl[i, k]])


def test_clump_neighborlist_f64(device):
    print("[clumps] neighborlist float64", flush=True)
    old = jd.config.jax_enable_x64
    jd.config.update("jax_enable_x64", True)
    try:
        st, box = clump_scene(6, 3, seed=22, dtype=np.float64)
        sy = _nl_clump_system(st, box, "CellList", law="hertz")
        st2, sy2 = jd.System.step(st, sy, n=40)
        _, _, frames = jd.System.trajectory_rollout(st, sy, n=2, stride=20)
    finally:
        jd.config.update("jax_enable_x64", old)
    nls = B.nl_state_of(st, sy, np.float64)
    ost = O.step(B.state_of(st, np.float64), B.params_of(st, sy, np.float64), 40, nls)
    for f, of in (("pos_c", "pos"), ("vel", "vel"), ("ang_vel", "ang_vel"), ("force", "force")):
        a = getattr(st2, f).cpu().numpy()
        b = getattr(ost, of)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), f
    assert sy2.collider.n_build_times == nls.n_build
    # a rollout with save_state frames (two segments of 20) ends where 40 steps do
    states, _ = frames
    assert torch.equal(states.pos_c[-1], st2.pos_c)


def test_clump_unsupported_combinations_raise(device):
    st, box = clump_scene(4, 3, seed=1)
    mats = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=1.0, young=1e4, poisson=0.3)])
    sy = jd.System.create(st.shape, dt=1e-4, domain_type="periodic", domain_kw=dict(box_size=box),
                          collider_type="naive", mat_table=mats)
    with pytest.raises(NotImplementedError, match="spheres only"):
        jd.minimize(st, sy, max_steps=2)


# ---------------------------------------------------------------------------
# The reference's clump tests (test_colliders_invariance.py:113-233, 405-434;
# test_energy_conservation.py:108-124), restated: clumps of nv asperities of
# radius 0.1 on a ring (2D, nv = 16) or a Fibonacci sphere (3D, nv = 42) of
# radius 0.3, total mass 1 per clump, on a lattice of spacing 1.0 with the
# lattice velocities of grid_state(seed 42).  The reference takes volume,
# centre, principal inertia and frame of the asperity union from a Monte
# Carlo estimate (utils/clumps.py), a builder outside this package; here the
# centre is the asperities' mean, the frame the lab frame and the inertia that
# of a uniform ball (disk) of radius 0.4 -- a consistent rigid body, which is
# all these tests check.
# ---------------------------------------------------------------------------
def set_up_clump(dim, N=350, collider="naive", domain="periodic", seed=42, dt=1e-3, device="cuda"):
    if dim == 3:
        N = int(np.ceil(N * 1.5))
    nv = 16 if dim == 2 else 42
    per = int((N // nv) ** (1.0 / dim))
    n = per**dim
    R_p, r_a = 0.4, 0.1
    if dim == 2:
        t = 2 * np.pi * np.arange(nv) / nv
        shell = np.stack([np.cos(t), np.sin(t)], 1)
    else:
        k = np.arange(nv) + 0.5
        phi = np.arccos(1 - 2 * k / nv)
        th = np.pi * (1 + 5**0.5) * k
        shell = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    shell = (R_p - r_a) * shell
    shell -= shell.mean(0)
    spacing = 2.0 * (R_p + r_a)
    grid = jd.utils.grid_state(n_per_axis=(per,) * dim, spacing=spacing, radius_range=(R_p, R_p),
                               vel_range=(-1.0, 1.0), seed=seed, device="cpu")
    centres = grid.pos_c.numpy()[:n].astype(np.float64)
    vel_c = grid.vel.numpy()[:n]
    I = (0.4 if dim == 3 else 0.5) * 1.0 * R_p * R_p
    A = 3 if dim == 3 else 1
    st = jd.State.create(pos=np.repeat(centres, nv, 0), pos_p=np.tile(shell, (n, 1)),
                         vel=np.repeat(vel_c, nv, 0), rad=np.full(n * nv, r_a), mass=np.ones(n * nv),
                         inertia=np.full((n * nv, A), I), clump_id=np.repeat(np.arange(n), nv), device=device)
    mt = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=0.27, young=1.0e3, poisson=0.3)])
    ckw = dict(state=st) if collider.endswith("CellList") else {}
    ctype = collider
    if collider.startswith("nl_"):
        ckw = dict(state=st, cutoff=float(2 * r_a), skin=0.5, secondary_collider_type=collider[3:],
                   secondary_collider_kw=ckw)
        ctype = "NeighborList"
    sy = jd.System.create(st.shape, domain_type=domain,
                          domain_kw=dict(box_size=np.full(dim, spacing * per), anchor=np.full(dim, -R_p - r_a)),
                          collider_type=ctype, collider_kw=ckw, dt=dt, mat_table=mt)
    return st, sy


@contextlib.contextmanager
def precision(name):
    old = jd.config.jax_enable_x64
    jd.config.update("jax_enable_x64", name == "float64")
    try:
        yield
    finally:
        jd.config.update("jax_enable_x64", old)


@pytest.mark.parametrize("prec", ["float32", "float64"])
@pytest.mark.parametrize("collider,domain", [("CellList", "periodic"), ("nl_naive", "periodic"),
                                             ("nl_CellList", "periodic"), ("nl_naive", "free"),
                                             ("CellList", "reflect"), ("nl_CellList", "reflect")])
@pytest.mark.parametrize("dim", [2, 3])
def test_clumps_invariance(device, dim, collider, domain, prec):
    """naive vs CellList and NeighborList over 200 steps, atol = rtol = 5e-2
    (test_colliders_invariance.py:405-434; the domains each collider takes
    with clumps)."""
    print(f"[clumps] invariance dim={dim} {collider} {domain} {prec}", flush=True)
    with precision(prec):
        s1, y1 = set_up_clump(dim, N=128, collider="naive", domain=domain)
        s2, y2 = set_up_clump(dim, N=128, collider=collider, domain=domain)
        s1, y1 = jd.System.step(s1, y1, n=200)
        s2, y2 = jd.System.step(s2, y2, n=200)
        pe1 = jd.utils.compute_potential_energy(s1, y1)
        pe2 = jd.utils.compute_potential_energy(s2, y2)
    for f in ("pos", "vel", "force", "ang_vel", "torque"):
        np.testing.assert_allclose(getattr(s1, f).cpu().numpy(), getattr(s2, f).cpu().numpy(), atol=5e-2, rtol=5e-2,
                                   err_msg=f)
    np.testing.assert_allclose(s1.q.data.cpu().numpy(), s2.q.data.cpu().numpy(), atol=5e-2, rtol=5e-2)
    np.testing.assert_allclose(pe1, pe2, atol=5e-2, rtol=5e-2)


@pytest.mark.parametrize("domain", ["periodic", "reflect", "free"])
@pytest.mark.parametrize("integrator", ["euler", "verlet"])
@pytest.mark.parametrize("dim", [2, 3])
def test_clumps_energy_conservation(device, dim, integrator, domain):
    """std(E) over 100 frames at five dt in [3e-5, 2e-3], total time 5, in
    float64 as the reference runs it: slope > 1.8 (velocity Verlet +
    VV-SPIRAL), > 0.7 (Euler + SPIRAL)."""
    print(f"[clumps] energy dim={dim} {integrator} {domain}", flush=True)
    from jaxdem_amd.utils import compute_energy

    lin, rot = ("euler", "spiral") if integrator == "euler" else ("verlet", "verletspiral")
    dts = np.logspace(np.log10(2e-3), np.log10(3e-5), 5)
    stds = []
    with precision("float64"):
        st, sy = set_up_clump(dim, domain=domain)
        for dt in dts:
            steps = int(5.0 / dt)
            stride = max(steps // 100, 1)
            s = dataclasses.replace(sy, dt=float(dt), linear_integrator=jd.LinearIntegrator.create(lin),
                                    rotation_integrator=jd.RotationIntegrator.create(rot))
            _, _, e = jd.System.trajectory_rollout(st, s, n=100, stride=stride, save_fn=compute_energy)
            stds.append(float(np.std(np.asarray(e, np.float64))))
    slope, _ = np.polyfit(np.log(dts), np.log(stds), 1)
    print(f"[clumps] energy slope {slope:.3f}", flush=True)
    assert slope > (0.7 if integrator == "euler" else 1.8), (slope, stds)
