"""Routed laws on the GPU -- ForceRouter (forces/router.py:26-339) and
LawCombiner (forces/law_combiner.py:22-191) -- against the oracle on the
same seeded inputs: forces and torques bit for bit (naive and CellList
colliders, 2D and 3D), potential energies to 1e-12 (float64 sums), whole
NeighborList rollouts with a history law on one species pair (state, list
and springs bit for bit, rebuilds inside the run), FIRE with a router, and a
batch of routed systems equal to each system alone.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from oracle import jdem_oracle as O
from tests import oracle_bridge as B
from tests.test_gpu_parity import _compare_states, assert_bitwise, packed

pytestmark = pytest.mark.gpu


def _species(st, S, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    sp = torch.as_tensor(rng.integers(0, S, st.N).astype(np.int32), device=st.pos_c.device)
    return dataclasses.replace(st, species_id=sp)


def _router(with_history: bool):
    fm = jd.ForceModel.create
    mapping = {
        (0, 0): fm("springfriction", damping=0.1) if with_history else fm("spring"),
        (0, 1): fm("hertz"),
        (1, 1): fm("cundallstrack"),
        (0, 2): jd.LawCombiner(laws=(fm("spring"), fm("hertz"))),
        (2, 2): fm("spring"),
    }  # (1, 2): no interaction
    return jd.ForceRouter.from_dict(3, mapping)


def _system(st, box, domain, collider, model, dt=1e-4, skin=0.1):
    mats = [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]
    mt = jd.MaterialTable.from_materials(mats)
    ckw = {}
    if collider == "CellList":
        ckw = dict(state=st)
    if collider == "NeighborList":
        ckw = dict(state=st, cutoff=float(2 * st.rad.max().item()), skin=skin, secondary_collider_type="CellList",
                   secondary_collider_kw=dict(state=st))
    sy = jd.System.create(st.shape, dt=dt, domain_type=domain,
                          domain_kw=dict(box_size=box, anchor=np.zeros(st.dim, np.float32)),
                          collider_type=collider, collider_kw=ckw, mat_table=mt, force_model_type="forcerouter",
                          force_model_kw={"table": model.table})
    return sy


@pytest.mark.parametrize("collider", ["naive", "CellList"])
@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("dim", [3, 2])
def test_routed_force_and_energy(device, collider, domain, dim):
    print(f"[routed] force {collider} {domain} dim={dim}", flush=True)
    st, box = packed(2000 if collider == "CellList" else 700, dim, seed=21, domain=domain)
    st = _species(st, 3, seed=5)
    sy = _system(st, box, domain, collider, _router(False))
    st2, _ = sy.collider.compute_force(st, sy)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    O.compute_force(ost, P)
    assert np.abs(ost.force).max() > 0
    assert_bitwise("force", st2.force.cpu().numpy(), ost.force)
    assert_bitwise("torque", st2.torque.cpu().numpy(), ost.torque)
    _, _, e = sy.collider.compute_potential_energy(st, sy)
    C = O.naive_candidates(st.N) if collider == "naive" else O.cell_candidates(ost.pos, P.domain, P.cell_size,
                                                                                P.search_range)[0]
    assert e == pytest.approx(O.potential_energy(ost, P, C), rel=1e-12)


@pytest.mark.parametrize("dim", [3, 2])
def test_routed_neighborlist_with_history_bitwise(device, dim):
    """(0, 0) spring friction with damping: 60 steps with rebuilds; springs
    live only on (0, 0) contacts."""
    print(f"[routed] neighborlist history dim={dim}", flush=True)
    st, box = packed(3000, dim, seed=22, domain="periodic", phi=0.6)
    st = _species(st, 3, seed=6)
    sy = _system(st, box, "periodic", "NeighborList", _router(True), skin=0.05)
    n = 60
    st2, sy2 = jd.System.step(st, sy, n=n)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    nls = B.nl_state_of(st, sy)
    ost = O.step(ost, P, n, nls)
    _compare_states(st2, ost, "routed-nl")
    assert sy2.collider.n_build_times == nls.n_build and nls.n_build >= 2
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    xi, live = B.history_of(sy2.collider)
    assert_bitwise("live", live, nls.live)
    assert_bitwise("xi", xi, nls.xi)
    assert live.sum() > 0
    sp = st.species_id.cpu().numpy()
    i, k = np.nonzero(live)
    assert np.all(sp[i] == 0) and np.all(sp[nls.nl[i, k]] == 0)


def test_routed_minimize_bitwise(device):
    print("[routed] fire", flush=True)
    st, box = packed(700, 3, seed=23, phi=0.70, ang=False)
    st = _species(st, 3, seed=7)
    sy = _system(st, box, "periodic", "CellList", _router(False), dt=1e-3)
    n = 30
    st2, sy2, steps, pe = jd.minimize(st, sy, max_steps=n, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=-1.0)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    osteps, ope = O.minimize(ost, P, O.FireParams(dt=float(np.float32(1e-3))), max_steps=n, pe_tol=0.0,
                             pe_diff_tol=-1.0, force_tol=-1.0)
    assert steps == osteps == n
    assert_bitwise("fire pos", st2.pos_c.cpu().numpy(), ost.pos)
    assert_bitwise("fire force", st2.force.cpu().numpy(), ost.force)
    assert pe == pytest.approx(ope, rel=1e-12)


def test_routed_batch_equals_alone(device):
    print("[routed] batch", flush=True)
    pairs = []
    for k, (N, phi) in enumerate([(900, 0.6), (400, 0.4), (1500, 0.55)]):
        st, box = packed(N, 3, seed=30 + k, domain="periodic", phi=phi)
        st = _species(st, 3, seed=40 + k)
        pairs.append((st, _system(st, box, "periodic", "NeighborList", _router(True), skin=0.05)))
    out = jd.Batch.pack(pairs).step(50).unpack()
    for k, ((st, sy), (sb, yb)) in enumerate(zip(pairs, out)):
        sa, ya = jd.System.step(st, sy, n=50)
        for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
            assert torch.equal(getattr(sb, f), getattr(sa, f)), (k, f)
        assert torch.equal(yb.collider.neighbor_list, ya.collider.neighbor_list)
