"""The plain step in float64 (``jax_enable_x64``, the reference's own test
precision, tests/conftest.py:4) against the float64 oracle on the same seeded
inputs: forces and torques within 1e-12 relative (max |GPU - oracle| over the
largest component), whole steps of every integrator pair and domain with the
naive, CellList and NeighborList colliders (list rebuilds inside the run),
energies to 1e-12, routed laws, gravity.
Each test also checks that the result is not a float32 computation: the
float32 oracle differs from it by far more than 1e-12.
"""

from __future__ import annotations

import contextlib

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from jaxdem_amd import _lib as L
from oracle import jdem_oracle as O
from tests import oracle_bridge as B
from tests.test_gpu_parity import make_system, packed

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative, float64 (the reference's contract for its x64 runs)


@contextlib.contextmanager
def x64():
    old = jd.config.jax_enable_x64
    jd.config.update("jax_enable_x64", True)
    try:
        yield
    finally:
        jd.config.update("jax_enable_x64", old)


def rel(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = float(np.abs(b).max()) if b.size else 0.0
    return float(np.abs(a - b).max()) / scale if scale > 0 else float(np.abs(a - b).max())


def check(tag, gpu, ref, ref32=None):
    e = rel(gpu, ref)
    assert e <= TOL, f"{tag}: relative error {e:.3e} > {TOL}"
    if ref32 is not None and np.abs(ref).max() > 0:
        # a float32 computation would sit near 1e-7, not 1e-12
        assert rel(ref32, ref) > 1e3 * TOL, f"{tag}: the float32 oracle agrees to {rel(ref32, ref):.1e}"


def compare_states(st, ost, tag, ost32=None):
    assert st.pos_c.dtype == torch.float64
    for f, of in (("pos_c", "pos"), ("vel", "vel"), ("ang_vel", "ang_vel"), ("force", "force"), ("torque", "torque")):
        check(f"{tag} {f}", getattr(st, f).cpu().numpy(), getattr(ost, of),
              None if ost32 is None or f in ("ang_vel", "torque") else getattr(ost32, of))
    check(f"{tag} q", st.q.data.cpu().numpy(), ost.q)


@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("collider", ["naive", "CellList"])
@pytest.mark.parametrize("law", ["spring", "hertz", "cundallstrack"])
@pytest.mark.parametrize("dim", [3, 2])
def test_f64_compute_force(device, dim, law, collider, domain):
    print(f"[f64] force {law} {collider} {domain} dim={dim}", flush=True)
    with x64():
        st, box = packed(2500 if collider == "CellList" else 700, dim, seed=3, domain=domain)
        sy = make_system(st, box, domain, law, collider)
        st2, _ = sy.collider.compute_force(st, sy)
    assert st2.force.dtype == torch.float64
    ost = B.state_of(st, np.float64)
    O.compute_force(ost, B.params_of(st, sy, np.float64))
    ost32 = B.state_of(st, np.float32)
    O.compute_force(ost32, B.params_of(st, sy, np.float32))
    assert np.abs(ost.force).max() > 0
    check("force", st2.force.cpu().numpy(), ost.force, ost32.force)
    check("torque", st2.torque.cpu().numpy(), ost.torque)


CASES = [  # (collider, domain, integrators)
    ("naive", "periodic", ("verlet", "verletspiral")),
    ("naive", "reflect", ("verlet", "verletspiral")),
    ("naive", "free", ("verlet", "verletspiral")),
    ("naive", "reflect", ("euler", "spiral")),
    ("CellList", "periodic", ("verlet", "verletspiral")),
    ("CellList", "reflect", ("euler", "spiral")),
]


@pytest.mark.parametrize("collider,domain,integ", CASES)
@pytest.mark.parametrize("dim", [3, 2])
def test_f64_step(device, dim, collider, domain, integ):
    """30 fused steps of Cundall-Strack (forces and torques) in float64."""
    print(f"[f64] step {collider} {domain} {integ} dim={dim}", flush=True)
    n = 30
    with x64():
        st, box = packed(1500 if collider == "CellList" else 600, dim, seed=11, domain=domain)
        sy = make_system(st, box, domain, "cundallstrack", collider, dt=1e-4,
                         linear_integrator_type=integ[0], rotation_integrator_type=integ[1])
        st2, sy2 = jd.System.step(st, sy, n=n)
    ost = O.step(B.state_of(st, np.float64), B.params_of(st, sy, np.float64), n)
    ost32 = O.step(B.state_of(st, np.float32), B.params_of(st, sy, np.float32), n)
    compare_states(st2, ost, f"{collider}/{domain}/{integ}", ost32)
    # the clock: n float64 additions of dt
    t = np.float64(0.0)
    for _ in range(n):
        t = t + np.float64(1e-4)
    assert isinstance(sy2.time, np.float64) and sy2.time == t
    if domain == "free":
        # FreeDomain.apply (free.py:42-58) fitted the box in float64
        assert sy2.domain.box_size.dtype == np.float64


def test_f64_energies(device):
    print("[f64] energies", flush=True)
    with x64():
        st, box = packed(3000, 3, seed=5, domain="periodic")
        sy = make_system(st, box, "periodic", "hertz", "CellList")
        e = sy.potential_energy(st)
        ke = jd.utils.compute_kinetic_energy(st, sy)
    ost = B.state_of(st, np.float64)
    P = B.params_of(st, sy, np.float64)
    C, _ = O.cell_candidates(ost.pos, P.domain, P.cell_size, P.search_range)
    oe = O.potential_energy(ost, P, C)
    assert e > 0 and abs(e - oe) <= TOL * abs(oe)
    m = ost.mass.astype(np.float64)
    v = ost.vel.astype(np.float64)
    w = ost.ang_vel.astype(np.float64)
    ke_ref = 0.5 * float((m * (v * v).sum(1)).sum()) + 0.5 * float((ost.inertia * w * w).sum())
    assert abs(ke - ke_ref) <= 1e-12 * ke_ref


def test_f64_routed_and_gravity(device):
    """A ForceRouter over three species and gravity in a reflective box."""
    print("[f64] routed + gravity", flush=True)
    fm = jd.ForceModel.create
    mapping = {(0, 0): fm("spring"), (0, 1): fm("hertz"), (1, 1): fm("cundallstrack"),
               (0, 2): jd.LawCombiner(laws=(fm("spring"), fm("hertz"))), (2, 2): fm("spring")}
    router = jd.ForceRouter.from_dict(3, mapping)
    with x64():
        st, box = packed(800, 3, seed=8, domain="reflect")
        sp = np.random.Generator(np.random.PCG64(2)).integers(0, 3, st.N).astype(np.int32)
        st = jd.State.create(pos=st.pos_c.cpu().numpy(), vel=st.vel.cpu().numpy(), rad=st.rad.cpu().numpy(),
                             mass=st.mass.cpu().numpy(), ang_vel=st.ang_vel.cpu().numpy(), species_id=sp,
                             device=device)
        mats = [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]
        sy = jd.System.create(st.shape, dt=1e-4, domain_type="reflect",
                              domain_kw=dict(box_size=box, anchor=np.zeros(3)), collider_type="naive",
                              mat_table=jd.MaterialTable.from_materials(mats), force_model_type="forcerouter",
                              force_model_kw={"table": router.table},
                              force_manager_kw={"gravity": np.array([0.0, -9.81, 0.0])})
        st2, _ = jd.System.step(st, sy, n=20)
    ost = O.step(B.state_of(st, np.float64), B.params_of(st, sy, np.float64), 20)
    compare_states(st2, ost, "routed")


@pytest.mark.parametrize("secondary", ["CellList", "naive"])
@pytest.mark.parametrize("dim", [3, 2])
def test_f64_neighborlist_step(device, dim, secondary):
    """60 steps on a float64 NeighborList with a small skin (rebuilds inside
    the run): state, list, build positions and build count."""
    print(f"[f64] neighborlist {secondary} dim={dim}", flush=True)
    n = 60
    with x64():
        st, box = packed(2000 if secondary == "CellList" else 600, dim, seed=13, domain="periodic", phi=0.6)
        ckw = dict(state=st) if secondary == "CellList" else {}
        sy = jd.System.create(
            st.shape, dt=1e-4, domain_type="periodic", domain_kw=dict(box_size=box, anchor=np.zeros(dim)),
            collider_type="NeighborList", force_model_type="cundallstrack",
            collider_kw=dict(state=st, cutoff=float(2 * st.rad.max().item()), skin=0.02,
                             secondary_collider_type=secondary, secondary_collider_kw=ckw),
            mat_table=jd.MaterialTable.from_materials(
                [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]))
        assert sy.collider.old_pos.dtype == torch.float64
        st2, sy2 = jd.System.step(st, sy, n=n)
        f_only, _ = sy2.collider.compute_force(st2, sy2)
    ost = B.state_of(st, np.float64)
    P = B.params_of(st, sy, np.float64)
    nls = B.nl_state_of(st, sy, np.float64)
    ost = O.step(ost, P, n, nls)
    ost32 = O.step(B.state_of(st, np.float32), B.params_of(st, sy, np.float32), n, B.nl_state_of(st, sy))
    compare_states(st2, ost, f"nl/{secondary}", ost32)
    assert sy2.collider.n_build_times == nls.n_build and nls.n_build >= 2
    assert np.array_equal(sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    check("old_pos", sy2.collider.old_pos.cpu().numpy(), nls.old_pos)
    # compute_force on the stepped state (a list check first; velocities after the last kick)
    oc = ost.copy()
    O.compute_force(oc, P, nls)
    check("compute_force", f_only.force.cpu().numpy(), oc.force)
    check("compute_force torque", f_only.torque.cpu().numpy(), oc.torque)


def test_f64_unsupported_paths_raise(device):
    with x64():
        st, box = packed(200, 3, seed=1)
        hist = make_system(st, box, "periodic", "springfriction", "NeighborList")
        with pytest.raises(NotImplementedError, match="float32"):
            jd.System.step(st, hist, n=1)
        sy = make_system(st, box, "periodic", "spring", "naive")
        stf, boxf = packed(200, 3, seed=1, domain="free")
        fc = make_system(stf, boxf, "free", "spring", "CellList")
        with pytest.raises(L.JdemError, match="FreeDomain"):
            jd.System.step(stf, fc, n=2)
    import dataclasses

    mixed = dataclasses.replace(st, vel=st.vel.float())
    with pytest.raises(TypeError, match="float64"):
        jd.System.step(mixed, sy, n=1)
    with pytest.raises(NotImplementedError, match="float32"):
        jd.minimize(st, sy, max_steps=2)
