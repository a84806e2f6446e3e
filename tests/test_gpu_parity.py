"""GPU parity: the HIP path against the CPU oracle on identical seeded inputs.

Per-pair and per-particle arithmetic of the kernels and of the float32 oracle
round identically (same operation order, no FMA contraction, IEEE division
and square root), and both sum in the same order, so forces, torques and
whole time steps are compared bit for bit.  Candidate pair sets are compared
as sets against brute force and row by row against the oracle's traversal
order.  Energies are float64 sums: relative tolerance 1e-12.
"""

from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from jaxdem_amd import _lib as L
from jaxdem_amd import workloads as W
from oracle import jdem_oracle as O
from tests import oracle_bridge as B

pytestmark = pytest.mark.gpu


def assert_bitwise(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    if np.any(bad):
        idx = np.argwhere(bad)[:5]
        diff = np.max(np.abs(got.astype(np.float64) - want.astype(np.float64)))
        raise AssertionError(
            f"{name}: {int(bad.sum())} of {bad.size} entries differ (max |diff| {diff:.3e}); first at "
            f"{idx.tolist()}: got {got[tuple(idx[0])]!r} want {want[tuple(idx[0])]!r}"
        )


def packed(N, dim, seed, domain="periodic", phi=0.55, rr=(0.4, 0.6), device="cuda", ang=True, bonds=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = rng.uniform(rr[0], rr[1], N).astype(np.float32)
    vol = (4.0 / 3.0 * math.pi * rad.astype(np.float64) ** 3) if dim == 3 else (math.pi * rad.astype(np.float64) ** 2)
    Lb = (vol.sum() / phi) ** (1.0 / dim)
    pos = rng.uniform(0.0, Lb, (N, dim)).astype(np.float32)
    if domain == "periodic":
        # exercise unwrapped coordinates (the reference never wraps)
        pos[: N // 8] += np.float32(Lb)
        pos[N // 8 : N // 4] -= np.float32(Lb)
    vel = rng.standard_normal((N, dim)).astype(np.float32)
    A = 3 if dim == 3 else 1
    av = rng.standard_normal((N, A)).astype(np.float32) if ang else None
    mass = (vol * 1.0).astype(np.float32)
    st = jd.State.create(pos=pos, vel=vel, rad=rad, mass=mass, ang_vel=av, device=device, bond_id=bonds)
    return st, np.full(dim, Lb, np.float32)


def make_system(st, box, domain, law, collider, dt=1e-4, cell_size=None, **kw):
    if law in ("cundallstrack", "springfriction", "hertzfriction"):
        mats = [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05)]
    else:
        mats = [jd.Material.create("elastic", density=1.0, young=1e4, poisson=0.3)]
    mt = jd.MaterialTable.from_materials(mats)
    ckw = {}
    if collider == "CellList":
        ckw = dict(state=st) if cell_size is None else dict(state=st, cell_size=cell_size)
    if collider == "NeighborList":
        ckw = dict(state=st, cutoff=float(2 * st.rad.max().item()), skin=kw.pop("skin", 0.1),
                   secondary_collider_type="CellList", secondary_collider_kw=dict(state=st))
    anchor = np.zeros(st.dim, np.float32)
    fm_kw = kw.pop("force_model_kw", {})
    return jd.System.create(
        st.shape, dt=dt, domain_type=domain, domain_kw=dict(box_size=box, anchor=anchor),
        collider_type=collider, collider_kw=ckw, force_model_type=law, force_model_kw=fm_kw,
        mat_table=mt, **kw,
    )


# ---------------------------------------------------------------------------
# (a) cell list: hash + stable sort by cell + cell ranges
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 70001])
@pytest.mark.parametrize("periodic", [True, False])
def test_cell_build_matches_stable_sort(device, n, periodic):
    print(f"[cell_build] n={n} periodic={periodic}", flush=True)
    st, box = packed(n, 3, seed=n, domain="periodic" if periodic else "reflect")
    sy = make_system(st, box, "periodic" if periodic else "reflect", "spring", "CellList")
    col = sy.collider
    g, _ = col.grid(sy)
    ncell = int(g.n_cells)
    perm = torch.empty(n, dtype=torch.int32, device=device)
    sh = torch.empty(n, dtype=torch.int32, device=device)
    cs = torch.empty(ncell, dtype=torch.int32, device=device)
    ce = torch.empty(ncell, dtype=torch.int32, device=device)
    keep: list = []
    S = jd.system._system_struct(st, sy, keep)
    dom = sy.domain.to_c(3)
    ctx = L.context_for(st.device)
    ctx.check(ctx.lib.jdem_cell_build(ctx.ptr, ctypes.byref(S.state), ctypes.byref(dom), ctypes.byref(g),
                                      perm.data_ptr(), sh.data_ptr(), cs.data_ptr(), ce.data_ptr()), "cell_build")
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    part = O.cell_partition(ost.pos, P.domain, P.cell_size, P.search_range)
    assert_bitwise("perm", perm.cpu().numpy(), part["perm"].astype(np.int32))
    assert_bitwise("sorted_hash", sh.cpu().numpy(), part["sorted_hash"].astype(np.int32))
    hs = part["sorted_hash"]
    cells = np.arange(ncell)
    start = np.searchsorted(hs, cells, "left")
    end = np.searchsorted(hs, cells, "right")
    occ = end > start
    assert_bitwise("cell_start", cs.cpu().numpy()[occ], start[occ].astype(np.int32))
    assert_bitwise("cell_end", ce.cpu().numpy()[occ], end[occ].astype(np.int32))
    assert np.all(cs.cpu().numpy()[~occ] == ce.cpu().numpy()[~occ])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("collider", ["CellList", "naive"])
def test_neighbor_list_pair_set(device, dim, domain, collider):
    print(f"[nlist] dim={dim} domain={domain} collider={collider}", flush=True)
    N = 3000 if dim == 3 else 2000
    st, box = packed(N, dim, seed=11 + dim, domain=domain, phi=0.4)
    sy = make_system(st, box, domain, "spring", collider)
    cutoff = float(np.float32(1.15))
    M = 96
    _, _, nl, ovf = sy.collider.create_neighbor_list(st, sy, cutoff, M)
    assert not ovf
    got = nl.cpu().numpy().astype(np.int64)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    brute = O.brute_force_pairs(ost, P.domain, cutoff, use_inv=(collider == "naive"))
    pairs = {(i, int(j)) for i in range(N) for j in got[i] if j >= 0}
    assert pairs == brute, f"{len(pairs ^ brute)} pairs differ"
    if collider == "CellList":
        want, _ = O.neighbor_list_cells(ost, P.domain, P.cell_size, P.search_range, cutoff, M)
    else:
        want, _ = O.neighbor_list_naive(ost, P.domain, cutoff, M)
    assert_bitwise("row order", got, want)


def test_neighbor_list_overflow_flag(device):
    st, box = packed(2000, 3, seed=5, phi=0.5)
    sy = make_system(st, box, "periodic", "spring", "CellList")
    _, _, nl, ovf = sy.collider.create_neighbor_list(st, sy, 1.2, 2)
    assert ovf
    assert nl.shape == (2000, 2)


# ---------------------------------------------------------------------------
# (b) forces
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("law", ["spring", "hertz", "cundallstrack"])
@pytest.mark.parametrize("collider", ["naive", "CellList"])
@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("dim", [3, 2])
def test_compute_force_bitwise(device, law, collider, domain, dim):
    print(f"[force] law={law} collider={collider} domain={domain} dim={dim}", flush=True)
    N = 2500 if collider == "CellList" else 700
    st, box = packed(N, dim, seed=3, domain=domain)
    sy = make_system(st, box, domain, law, collider)
    st2, _ = sy.collider.compute_force(st, sy)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    O.compute_force(ost, P)
    assert_bitwise("force", st2.force.cpu().numpy(), ost.force)
    assert_bitwise("torque", st2.torque.cpu().numpy(), ost.torque)
    # float32 vs float64 restatement: norm-wise relative error
    ost64 = B.state_of(st, np.float64)
    P64 = B.params_of(st, sy, np.float64)
    O.compute_force(ost64, P64)
    err = np.linalg.norm(st2.force.cpu().numpy() - ost64.force) / np.linalg.norm(ost64.force)
    assert err < 1e-5, err


@pytest.mark.parametrize("dim,n", [(3, 5), (3, 14), (2, 3), (2, 5)])
def test_compute_force_small_periodic_grid(device, dim, n):
    """Periodic boxes of two cells per axis, smaller than the 3-cell
    stencil: each neighbour cell is visited once (cell_list.py:420-427),
    forces bit for bit against the oracle, and the pair set equals brute
    force."""
    print(f"[force] small periodic grid dim={dim} n={n}", flush=True)
    st, box = packed(n, dim, seed=50 + n, domain="periodic", phi=0.35)
    sy = make_system(st, box, "periodic", "spring", "CellList")
    g, _ = sy.collider.grid(sy)
    assert max(int(g.dims[a]) for a in range(dim)) <= 2, [int(g.dims[a]) for a in range(dim)]
    st2, _ = sy.collider.compute_force(st, sy)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    O.compute_force(ost, P)
    assert_bitwise("force", st2.force.cpu().numpy(), ost.force)
    cutoff = float(np.float32(1.1))
    _, _, nl, ovf = sy.collider.create_neighbor_list(st, sy, cutoff, 16)
    assert not ovf
    got = nl.cpu().numpy().astype(np.int64)
    pairs = {(i, int(j)) for i in range(n) for j in got[i] if j >= 0}
    assert pairs == O.brute_force_pairs(ost, P.domain, cutoff, use_inv=False)


@pytest.mark.parametrize("dim,n", [(3, 14), (2, 5)])
def test_neighborlist_small_periodic_grid(device, dim, n):
    """The resident NeighborList engine on periodic grids of at most two
    cells per axis (its stencil runs wrap onto themselves): 40 steps bit for
    bit against the oracle, rows included."""
    print(f"[step] neighborlist small periodic grid dim={dim} n={n}", flush=True)
    st, box = packed(n, dim, seed=60 + n, domain="periodic", phi=0.35)
    sy = make_system(st, box, "periodic", "springfriction", "NeighborList", skin=0.1)
    g2, _ = sy.collider.secondary_grid(sy)
    assert max(int(g2.dims[a]) for a in range(dim)) <= 2
    st2, sy2 = jd.System.step(st, sy, n=40)
    nls = B.nl_state_of(st, sy)
    ost = O.step(B.state_of(st), B.params_of(st, sy), 40, nls)
    _compare_states(st2, ost, "nl-small")
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)


def test_excluded_pairs_known_answer(device):
    """tests/test_excluded_pairs.py:11-61 of the reference."""
    st = jd.State.create(pos=np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0]], np.float32),
                         rad=np.array([1.1, 1.1, 1.1], np.float32), bond_id=[[1], [0, 2], [1]], device=device)
    mats = [jd.Material.create("elastic", young=1000.0, poisson=0.3, density=1.0)]
    mt = jd.MaterialTable.from_materials(mats, matcher=jd.MaterialMatchmaker.create("linear"))
    for name in ["naive", "CellList", "NeighborList"]:
        ckw = {}
        if name == "CellList":
            ckw = {"state": st}
        elif name == "NeighborList":
            ckw = {"state": st, "cutoff": 3.0, "skin": 0.5, "secondary_collider_type": "CellList",
                   "secondary_collider_kw": {"state": st}}
        sy = jd.System.create(st.shape, dt=0.001, collider_type=name, collider_kw=ckw,
                              force_model_type="spring", mat_table=mt)
        f = sy.collider.compute_force(st, sy)[0].force.cpu().numpy()
        assert np.allclose(f[1], 0.0, atol=1e-5), (name, f)
        assert abs(f[0, 0]) > 0.1 and abs(f[2, 0]) > 0.1, (name, f)
        assert np.allclose(f[0], -f[2], atol=1e-5), (name, f)
        # closed form: F_0 = -k (2.2 - 2.0) x_hat = -200 x_hat
        assert np.isclose(f[0, 0], -1000.0 * (2.2 - 2.0), rtol=1e-5), (name, f)


# ---------------------------------------------------------------------------
# (c) whole steps
# ---------------------------------------------------------------------------
def _compare_states(gst, ost, tag):
    assert_bitwise(f"{tag} pos", gst.pos_c.cpu().numpy(), ost.pos)
    assert_bitwise(f"{tag} vel", gst.vel.cpu().numpy(), ost.vel)
    assert_bitwise(f"{tag} q", gst.q.data.cpu().numpy(), ost.q)
    assert_bitwise(f"{tag} ang_vel", gst.ang_vel.cpu().numpy(), ost.ang_vel)
    assert_bitwise(f"{tag} force", gst.force.cpu().numpy(), ost.force)
    assert_bitwise(f"{tag} torque", gst.torque.cpu().numpy(), ost.torque)


def test_step_spheres_1k_bitwise(device):
    """README workload: naive collider, reflective box, 40 fused steps."""
    print("[step] spheres_1k", flush=True)
    w = W.spheres_1k()
    st, sy = W.make_gpu_system(w, device=device)
    st2, sy2 = jd.System.step(st, sy, n=40)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    ost = O.step(ost, P, 40)
    _compare_states(st2, ost, "spheres_1k")
    assert sy2.step_count == 40
    # the clock is 40 sequential float32 additions of dt (system.py:73)
    t = np.float32(0.0)
    for _ in range(40):
        t = np.float32(t + np.float32(w.dt))
    assert np.float32(sy2.time) == t
    assert t != np.float32(0.04)  # 0.039999988: a product 40 * dt would differ


@pytest.mark.parametrize("collider", ["naive", "celllist"])
def test_spheres_1k_rollout_every_frame(device, collider):
    """The CPU-runnable plain workload in full: 1000 spheres on a 10^3
    lattice (spacing 0.5, r 0.1, unit density), N(0, 0.1^2) velocities from
    seed 0, a reflecting box of 20, spring k = 1e4, dt = 1e-3,
    trajectory_rollout(n=100, stride=10) = 1000 steps.  Each of the 100
    frames equals the oracle's state after 10 (f + 1) steps, bit for bit."""
    print(f"[c1] rollout 100 x 10 collider={collider}", flush=True)
    w = W.spheres_1k()
    st, sy = W.make_gpu_system(w, device=device, collider=collider)
    fs, fy, traj = jd.System.trajectory_rollout(st, sy, n=100, stride=10)
    fr = traj[0]
    got = {k: getattr(fr, a).cpu().numpy() for k, a in (("pos", "pos_c"), ("vel", "vel"), ("ang_vel", "ang_vel"),
                                                       ("force", "force"), ("torque", "torque"))}
    got["q"] = fr.q.data.cpu().numpy()
    assert got["pos"].shape == (100, 1000, 3)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    P.naive_by_cells = True
    for f in range(100):
        ost = O.step(ost, P, 10)
        for k in ("pos", "vel", "q", "ang_vel", "force", "torque"):
            assert_bitwise(f"frame {f} {k}", got[k][f], getattr(ost, k))
    _compare_states(fs, ost, "end")
    assert fy.step_count == 1000


@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("law", ["hertz", "cundallstrack"])
def test_step_celllist_bitwise(device, domain, law):
    print(f"[step] celllist domain={domain} law={law}", flush=True)
    st, box = packed(3000, 3, seed=21, domain=domain, phi=0.45)
    sy = make_system(st, box, domain, law, "CellList", dt=2e-4)
    st2, _ = jd.System.step(st, sy, n=12)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    ost = O.step(ost, P, 12)
    _compare_states(st2, ost, f"celllist/{domain}/{law}")


def test_step_2d_bitwise(device):
    print("[step] 2d periodic hertz celllist", flush=True)
    st, box = packed(2000, 2, seed=9, domain="periodic", phi=0.6)
    sy = make_system(st, box, "periodic", "hertz", "CellList", dt=2e-4)
    st2, _ = jd.System.step(st, sy, n=10)
    ost = O.step(B.state_of(st), B.params_of(st, sy), 10)
    _compare_states(st2, ost, "2d")


@pytest.mark.parametrize("rot", ["spiral", "verletspiral"])
@pytest.mark.parametrize("lin", ["euler", "verlet"])
def test_integrator_variants_bitwise(device, lin, rot):
    print(f"[step] integrators lin={lin} rot={rot}", flush=True)
    st, box = packed(600, 3, seed=2, domain="reflect", phi=0.3)
    sy = make_system(st, box, "reflect", "cundallstrack", "naive", dt=1e-4,
                     linear_integrator_type=lin, rotation_integrator_type=rot,
                     force_manager_kw={"gravity": [0.0, 0.0, -9.81]})
    st2, _ = jd.System.step(st, sy, n=8)
    ost = O.step(B.state_of(st), B.params_of(st, sy), 8)
    _compare_states(st2, ost, f"{lin}/{rot}")


def test_stagewise_equals_fused(device):
    """The plugin-by-plugin path (user pre-step hook) equals the fused step."""
    st, box = packed(800, 3, seed=4, domain="reflect", phi=0.35)
    sy = make_system(st, box, "reflect", "hertz", "CellList", dt=1e-4)
    sa, _ = jd.System.step(st, sy, n=5)
    sy_hook = jd.System.create(st.shape, dt=1e-4, domain=sy.domain, collider=sy.collider,
                               force_model=sy.force_model, mat_table=sy.mat_table,
                               user_pre_step_actions=lambda s, y: (s, y))
    sb, _ = jd.System.step(st, sy_hook, n=5)
    _compare_states(sa, B.state_of(sb), "stagewise")


def test_neighborlist_friction_bitwise(device):
    """NeighborList + tangential-spring history, including rebuilds."""
    print("[step] neighborlist friction", flush=True)
    w = W.spheres_friction(N=4096, seed=1, skin=0.05)
    st, sy = W.make_gpu_system(w, device=device)
    n = 60
    st2, sy2 = jd.System.step(st, sy, n=n)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    nls = B.nl_state_of(st, sy)
    ost = O.step(ost, P, n, nls)
    _compare_states(st2, ost, "nl-friction")
    assert sy2.collider.n_build_times == nls.n_build, (sy2.collider.n_build_times, nls.n_build)
    assert nls.n_build >= 2
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    xi, live = B.history_of(sy2.collider)
    assert_bitwise("live", live, nls.live)
    assert_bitwise("xi", xi, nls.xi)
    assert live.sum() > 0


def test_trajectory_rollout_frames(device):
    w = W.spheres_1k()
    st, sy = W.make_gpu_system(w, device=device)
    fs, fy, traj = jd.System.trajectory_rollout(st, sy, n=3, stride=2)
    assert traj[0].pos_c.shape == (3, 1000, 3)
    s6, _ = jd.System.step(st, sy, n=6)
    assert torch.equal(traj[0].pos_c[-1], s6.pos_c)
    _, _, t0 = jd.System.trajectory_rollout(st, sy, strides=np.array([0, 1, 1, 1]))
    assert t0[0].pos_c.shape[0] == 4
    assert torch.equal(t0[0].pos_c[0], st.pos_c)


def test_rollout_abi_rejects(device):
    """jdem_rollout refuses what it does not cover, with an error and no
    launch: another collider, a negative stride, missing frame arrays."""
    print("[rollout] abi errors", flush=True)
    st, box = packed(500, 3, seed=4)
    sy = make_system(st, box, "periodic", "spring", "CellList")
    keep: list = []
    S = jd.system._system_struct(st, sy, keep)
    w = W.spheres_friction(N=1024, seed=5, skin=0.05)
    st2, sy2 = W.make_gpu_system(w, device=device)
    S2 = jd.system._system_struct(st2, sy2, keep)
    # frames sized for the larger system, so no call below could write past them
    frames = {k: torch.empty((2,) + tuple(v.shape), device=device)
              for k, v in (("pos", st2.pos_c), ("vel", st2.vel), ("ang_vel", st2.ang_vel), ("q", st2.q.data),
                           ("force", st2.force), ("torque", st2.torque))}
    steps = (ctypes.c_int64 * 2)(3, 4)
    fr = L.JdemFrames()
    fr.n_frames = 2
    fr.steps = ctypes.cast(steps, ctypes.POINTER(ctypes.c_int64))
    for k, t in frames.items():
        setattr(fr, k, t.data_ptr())
    ctx = L.context_for(st.device)
    with pytest.raises(L.JdemError, match="NeighborList"):
        ctx.check(ctx.lib.jdem_rollout(ctx.ptr, ctypes.byref(S), ctypes.byref(fr), None), "jdem_rollout")
    steps[1] = -1
    with pytest.raises(L.JdemError, match="negative"):
        ctx.check(ctx.lib.jdem_rollout(ctx.ptr, ctypes.byref(S2), ctypes.byref(fr), None), "jdem_rollout")
    steps[1] = 4
    fr.q = None
    with pytest.raises(L.JdemError, match="null frame"):
        ctx.check(ctx.lib.jdem_rollout(ctx.ptr, ctypes.byref(S2), ctypes.byref(fr), None), "jdem_rollout")


@pytest.mark.parametrize("strides", [[4, 0, 7, 0, 0, 12], [25]])
def test_device_rollout_frames_reflect(device, strides):
    """Device frames in a reflecting box with friction, repeated frames
    (zero strides) and a single frame: each equals System.step of its
    cumulative steps, bit for bit."""
    print(f"[step] device rollout reflect {strides}", flush=True)
    w = W.polydisperse_spheres(2500, seed=41, phi=0.45, domain="reflect", mu=0.3)
    w.box = w.box * np.float32(1.02)
    st, sy = W.make_gpu_system(w, device=device)
    _, fy, traj = jd.System.trajectory_rollout(st, sy, strides=np.array(strides), save_fn=jd.save_state)
    assert traj.pos_c.shape[0] == len(strides)
    total = 0
    for f, k in enumerate(strides):
        total += k
        ref = jd.System.step(st, sy, n=total)[0]
        for name in ("pos_c", "vel", "ang_vel", "force", "torque"):
            assert_bitwise(f"frame {f} {name}", getattr(traj, name)[f].cpu().numpy(), getattr(ref, name).cpu().numpy())
        assert_bitwise(f"frame {f} q", traj.q.data[f].cpu().numpy(), ref.q.data.cpu().numpy())
    assert fy.step_count == sy.step_count + total


@pytest.mark.parametrize("case", ["spheres_friction", "disks_friction"])
def test_device_rollout_frames_bitwise(device, case):
    """save_state frames written on the device by jdem_rollout equal, bit for
    bit, System.step(cumulative steps) from the initial state, across
    NeighborList rebuilds; the end state and list equal System.step(total)."""
    print(f"[step] device rollout {case}", flush=True)
    if case == "spheres_friction":
        w = W.spheres_friction(N=4096, seed=1, skin=0.05)
    else:
        w = W.disks_hertz_friction(N=3000, seed=2, phi=0.8, skin=0.05)
    st, sy = W.make_gpu_system(w, device=device)
    strides = [0, 9, 20, 1, 30]
    fs, fy, traj = jd.System.trajectory_rollout(st, sy, strides=np.array(strides), save_fn=jd.save_state)
    assert fy._last_stats.rebuilds >= 2, fy._last_stats.rebuilds
    assert traj.pos_c.shape == (len(strides),) + tuple(st.pos_c.shape)
    assert torch.equal(traj.rad[2], st.rad)
    total = 0
    for f, k in enumerate(strides):
        total += k
        ref, ref_sy = jd.System.step(st, sy, n=total) if total else (st, sy)
        for name in ("pos_c", "vel", "ang_vel", "force", "torque"):
            assert_bitwise(f"frame {f} {name}", getattr(traj, name)[f].cpu().numpy(), getattr(ref, name).cpu().numpy())
        assert_bitwise(f"frame {f} q", traj.q.data[f].cpu().numpy(), ref.q.data.cpu().numpy())
    _compare_states(fs, B.state_of(ref), "rollout end")
    assert fy.collider.n_build_times == ref_sy.collider.n_build_times
    assert fy.step_count == ref_sy.step_count and fy.time == ref_sy.time
    assert_bitwise("rollout nl", fy.collider.neighbor_list.cpu().numpy(), ref_sy.collider.neighbor_list.cpu().numpy())
    xa, la = B.history_of(fy.collider)
    xb, lb = B.history_of(ref_sy.collider)
    assert_bitwise("rollout live", la, lb)
    assert_bitwise("rollout xi", xa, xb)
    # the generic path (one System.step call per frame) gives the same frames
    _, _, gen = jd.System.trajectory_rollout(st, sy, strides=np.array(strides), save_fn=lambda s, y: (s, y))
    assert_bitwise("generic frames pos", gen[0].pos_c.cpu().numpy(), traj.pos_c.cpu().numpy())
    assert_bitwise("generic frames vel", gen[0].vel.cpu().numpy(), traj.vel.cpu().numpy())


# ---------------------------------------------------------------------------
# energies
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("law", ["spring", "hertz", "cundallstrack"])
@pytest.mark.parametrize("collider", ["naive", "CellList"])
def test_potential_energy(device, law, collider):
    st, box = packed(1500, 3, seed=8, domain="periodic")
    sy = make_system(st, box, "periodic", law, collider)
    _, _, e = sy.collider.compute_potential_energy(st, sy)
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    if collider == "naive":
        C = O.naive_candidates(st.N)
    else:
        C, _ = O.cell_candidates(ost.pos, P.domain, P.cell_size, P.search_range)
    want = O.potential_energy(ost, P, C)
    assert e == pytest.approx(want, rel=1e-12)
    ke = jd.utils.compute_kinetic_energy(st, sy)
    assert ke == pytest.approx(O.kinetic_energy(ost), rel=1e-12)


@pytest.mark.parametrize("collider", ["naive", "CellList", "NeighborList", "NeighborList-rescaled"])
@pytest.mark.parametrize("dim", [3, 2])
def test_contact_stress_and_pressure(device, collider, dim):
    """Contact stress (contacts.py:89-164) against the oracle's float64
    virial sum of the same float32 pair forces: 1e-12 of the largest
    component (different summation order).  A NeighborList that was never
    built, or was marked stale by ``scale_to_packing_fraction``, is rebuilt
    first, as ``get_pair_forces_and_ids`` does (contacts.py:58-62)."""
    from jaxdem_amd.utils.contacts import compute_contact_pressure, compute_contact_stress_tensor
    from jaxdem_amd.utils.packing_utils import compute_packing_fraction, scale_to_packing_fraction

    print(f"[stress] collider={collider} dim={dim}", flush=True)
    st, box = packed(1500 if dim == 3 else 1200, dim, seed=31 + dim, domain="periodic")
    sy = make_system(st, box, "periodic", "hertz", collider.split("-")[0])
    if collider == "NeighborList-rescaled":
        # build the list on the loose packing, then compress by 3 %
        sy = compute_contact_stress_tensor(st, sy)[1]
        assert sy.collider.n_build_times == 1
        st, sy = scale_to_packing_fraction(st, sy, float(compute_packing_fraction(st, sy)) * 1.03)
        box = np.asarray(sy.domain.box_size, np.float32)
        assert sy.collider.n_build_times == 0
    _, sy_out, stress = compute_contact_stress_tensor(st, sy)
    if collider.startswith("NeighborList"):
        assert sy_out.collider.n_build_times == 1  # refreshed for this call
        collider = "naive"  # all pairs: the list must hold every contact
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    if collider == "naive":
        C = O.naive_candidates(st.N)
    else:
        C, _ = O.cell_candidates(ost.pos, P.domain, P.cell_size, P.search_range)
    want = O.pair_virial(ost, P, C) / float(np.prod(box.astype(np.float64)))
    assert stress.shape == (dim, dim)
    assert np.max(np.abs(stress - want)) <= 1e-12 * np.max(np.abs(want))
    assert np.trace(want) > 0  # repulsive overlaps: positive pressure
    _, _, p = compute_contact_pressure(st, sy)
    assert p == pytest.approx(np.trace(want) / dim, rel=1e-12)
    _, _, s2 = compute_contact_stress_tensor(st, sy, volume=2.0)
    assert np.allclose(s2 * 2.0, stress * float(np.prod(box.astype(np.float64))), rtol=1e-12)


# ---------------------------------------------------------------------------
# resident NeighborList engine
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["spring-periodic", "hertzfriction-2d", "springfriction-reflect", "naive-secondary"])
def test_neighborlist_variants_bitwise(device, case):
    print(f"[step] neighborlist {case}", flush=True)
    dim = 2 if case.endswith("2d") else 3
    domain = "reflect" if "reflect" in case else "periodic"
    law = {"spring-periodic": "spring", "hertzfriction-2d": "hertzfriction",
           "springfriction-reflect": "springfriction", "naive-secondary": "springfriction"}[case]
    N = 600 if case == "naive-secondary" else 3000
    st, box = packed(N, dim, seed=17, domain=domain, phi=0.5)
    fkw = {"force_model_kw": {"kt_ratio": 0.3, "damping": 0.05}} if "friction" in law else {}
    sy = make_system(st, box, domain, law, "NeighborList", dt=2e-4, skin=0.05, **fkw)
    if case == "naive-secondary":
        col = jd.NeighborList.Create(st, cutoff=float(2 * st.rad.max().item()), skin=0.05,
                                     secondary_collider_type="naive")
        from dataclasses import replace as _r
        sy = _r(sy, collider=jd.system._fresh_collider(_r(col, history=jd.system._alloc_history(col, st.shape, sy.force_model))))
    n = 25
    st2, sy2 = jd.System.step(st, sy, n=n)
    nls = B.nl_state_of(st, sy)
    ost = O.step(B.state_of(st), B.params_of(st, sy), n, nls)
    _compare_states(st2, ost, f"nl/{case}")
    assert sy2.collider.n_build_times == nls.n_build
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    if sy2.collider.history is not None:
        xi, live = B.history_of(sy2.collider)
        assert_bitwise("live", live, nls.live)
        assert_bitwise("xi", xi, nls.xi)


def test_neighborlist_split_calls_equal_one_call(device):
    """step(n=30) == step(n=12) then step(n=18): the list and the history
    survive the round trip through the caller's layouts."""
    w = W.spheres_friction(N=4096, seed=3, skin=0.05)
    st, sy = W.make_gpu_system(w, device=device)
    a, ya = jd.System.step(st, sy, n=30)
    b1, yb = jd.System.step(st, sy, n=12)
    b, yb = jd.System.step(b1, yb, n=18)
    _compare_states(a, B.state_of(b), "split")
    assert ya.collider.n_build_times == yb.collider.n_build_times
    assert torch.equal(ya.collider.neighbor_list, yb.collider.neighbor_list)
    xa, la = B.history_of(ya.collider)
    xb, lb = B.history_of(yb.collider)
    assert_bitwise("live", la, lb)
    assert_bitwise("xi", xa, xb)
    # the inputs are left untouched (functional semantics)
    assert int(sy.collider.n_build_times) == 0


def test_resident_working_set_semantics(device):
    """After a step the list, build positions and springs stay in the step's
    working set on the device.  The next step starts from it; stepping twice
    from the same (state, system) gives the same result (the working set is
    immutable); reading the arrays exports them in the reference's layouts;
    an exported array modified in place sends the next step back to those
    arrays, with the same result.  All bit for bit, and equal to the oracle."""
    print("[resident] semantics", flush=True)
    w = W.spheres_friction(N=4096, seed=5, skin=0.05)
    st0, sy0 = W.make_gpu_system(w, device=device)
    st, sy = jd.System.step(st0, sy0, n=20)
    assert sy.collider.resident_for_step() is not None
    a1, ya1 = jd.System.step(st, sy, n=15)
    a2, ya2 = jd.System.step(st, sy, n=15)
    c1, yc1 = jd.System.step(st, sy, n=7)
    c, yc = jd.System.step(c1, yc1, n=8)
    ref = B.state_of(a1)
    for other, tag in ((a2, "again"), (c, "7+8")):
        _compare_states(other, ref, tag)
    ost = B.state_of(st0)
    nls = B.nl_state_of(st0, sy0)
    ost = O.step(ost, B.params_of(st0, sy0), 35, nls)
    _compare_states(a1, ost, "oracle")
    assert nls.n_build >= 3
    for y in (ya1, ya2, yc):
        assert_bitwise("nl", y.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
        xi, live = B.history_of(y.collider)
        assert_bitwise("live", live, nls.live)
        assert_bitwise("xi", xi, nls.xi)
    # exported and unmodified: still the working set
    _ = sy.collider.neighbor_list
    assert sy.collider.resident_for_step() is not None
    e1, _ = jd.System.step(st, sy, n=15)
    _compare_states(e1, ref, "exported")
    # modified in place (same values): the next step reads the arrays
    sy.collider.neighbor_list.add_(0)
    assert sy.collider.resident_for_step() is None
    f1, yf1 = jd.System.step(st, sy, n=15)
    _compare_states(f1, ref, "from arrays")
    assert torch.equal(yf1.collider.neighbor_list, ya1.collider.neighbor_list)


# ---------------------------------------------------------------------------
# (d) FIRE energy minimisation (minimizers/routines.py:150-383)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("collider", ["naive", "CellList", "NeighborList"])
def test_minimize_fire_bitwise(device, collider):
    """A fixed number of FIRE iterations on a jammed periodic packing: every
    position, orientation, force and torque equals the float32 oracle's; the
    float64 energy agrees to rounding of the sum order."""
    print(f"[fire] collider={collider}", flush=True)
    st, box = packed(700, 3, seed=9, phi=0.70, ang=False)
    sy = make_system(st, box, "periodic", "spring", collider, dt=1e-3)
    n = 40
    pos0 = st.pos_c.clone()
    st2, sy2, steps, pe = jd.minimize(st, sy, max_steps=n, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=-1.0)
    assert steps == n
    ost = B.state_of(st)
    P = B.params_of(st, sy)
    nls = B.nl_state_of(st, sy) if collider == "NeighborList" else None
    osteps, ope = O.minimize(ost, P, O.FireParams(dt=float(np.float32(1e-3))), max_steps=n, pe_tol=0.0,
                             pe_diff_tol=-1.0, force_tol=-1.0, nls=nls)
    assert osteps == n
    assert_bitwise("fire pos", st2.pos_c.cpu().numpy(), ost.pos)
    assert_bitwise("fire q", st2.q.data.cpu().numpy(), ost.q)
    assert_bitwise("fire force", st2.force.cpu().numpy(), ost.force)
    assert_bitwise("fire torque", st2.torque.cpu().numpy(), ost.torque)
    assert pe == pytest.approx(ope, rel=1e-12)
    assert torch.equal(st.pos_c, pos0)  # the inputs are untouched


def test_minimize_fire_converges_and_stops(device):
    """Stopping rules of routines.py:284-314 on a 20k-sphere jammed packing:
    force_tol stops the loop once max |F| <= tol; pe_tol stops at once when
    |pe|/N is already below it; max_steps caps the count."""
    print("[fire] convergence", flush=True)
    st, box = packed(20000, 3, seed=4, phi=0.68, ang=False)
    sy = make_system(st, box, "periodic", "spring", "CellList", dt=1e-3)
    tol = 1.0  # contact forces start near 1e3 (k = 1e4)
    s1, y1, steps, pe = jd.minimize(st, sy, max_steps=20000, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=tol)
    res = y1.__dict__["_last_minimize"]
    print(f"[fire] steps={steps} max_grad={res.max_grad:.3e} pe/N={pe:.6e}", flush=True)
    assert 0 < steps < 20000 and res.max_grad <= tol
    assert float(s1.force.abs().max()) == pytest.approx(res.max_grad)
    e0 = sy.potential_energy(st) / st.N
    assert 0.0 < pe < e0
    _, _, s0, _ = jd.minimize(st, sy, max_steps=100, pe_tol=1e30)
    assert s0 == 0
    _, _, s3, _ = jd.minimize(st, sy, max_steps=3, pe_tol=0.0, pe_diff_tol=-1.0)
    assert s3 == 3


# ---------------------------------------------------------------------------
# (e) segmented batches of independent systems (jdem_batch_step)
# ---------------------------------------------------------------------------
def _kinematics_equal(a, b, tag):
    for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
        assert_bitwise(f"{tag} {f}", getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy())
    assert_bitwise(f"{tag} q", a.q.data.cpu().numpy(), b.q.data.cpu().numpy())


def test_batch_2d_hertz_friction_equals_alone(device):
    """Four disk systems of unequal size and density in one batch: each
    trajectory equals the same system stepped alone, bit for bit."""
    print("[batch] 2d hertzfriction", flush=True)
    ws = [W.disks_hertz_friction(N, seed=k, phi=p) for k, (N, p) in
          enumerate([(300, 0.5), (1500, 0.78), (256, 0.3), (4100, 0.65)])]
    pairs = [W.make_gpu_system(w, device=device) for w in ws]
    n = 60
    alone = [jd.System.step(st, sy, n=n) for st, sy in pairs]
    out = jd.Batch.pack(pairs).step(n).unpack()
    for k, ((sa, ya), (sb, yb)) in enumerate(zip(alone, out)):
        _kinematics_equal(sa, sb, f"sys{k}")
        assert yb.step_count == ya.step_count and yb.time == ya.time


def test_batch_3d_mixed_domains_equals_alone(device):
    """3D spring-friction systems, periodic and reflect, different skins and
    time steps; one of them was stepped before and joins with its list and
    history, the others are fresh."""
    print("[batch] 3d mixed", flush=True)
    w0 = W.spheres_friction(N=2000, seed=5, skin=0.05)
    w1 = W.spheres_friction(N=700, seed=6, phi=0.4, skin=0.15)
    w2 = W.spheres_friction(N=3000, seed=7, phi=0.55, skin=0.1)
    w2.domain = "reflect"
    w2.box = w2.box * np.float32(1.05)
    w1.dt = w1.dt * 0.5
    pairs = [W.make_gpu_system(w, device=device) for w in (w0, w1, w2)]
    s0_10, y0_10 = jd.System.step(*pairs[0], n=10)
    assert y0_10.collider.n_build_times > 0
    n = 50
    ref = [jd.System.step(*pairs[0], n=n + 10), jd.System.step(*pairs[1], n=n), jd.System.step(*pairs[2], n=n)]
    out = jd.Batch.pack([(s0_10, y0_10), pairs[1], pairs[2]]).step(n).unpack()
    for k, ((sa, _), (sb, _)) in enumerate(zip(ref, out)):
        _kinematics_equal(sa, sb, f"sys{k}")


def test_batch_ragged_sizes_equal_alone(device):
    """Systems of 1, 63, 64, 65 and 200 disks in one batch (boundaries inside
    and on wavefront edges, a system with no pair): each equals the same
    system stepped alone, bit for bit, after rebuilds."""
    print("[batch] ragged sizes", flush=True)
    ws = [W.disks_hertz_friction(N, seed=70 + k, phi=0.6, skin=0.05) for k, N in enumerate([1, 63, 64, 65, 200])]
    for w in ws:
        w.vel = w.vel * np.float32(20.0)  # rebuilds within the run
    pairs = [W.make_gpu_system(w, device=device) for w in ws]
    out = jd.Batch.pack(pairs).step(40).unpack()
    for k, ((s0, y0), (sb, yb)) in enumerate(zip(pairs, out)):
        sa, ya = jd.System.step(s0, y0, n=40)
        _kinematics_equal(sa, sb, f"sys{k}")
        assert ya.collider.n_build_times == yb.collider.n_build_times


def test_batch_graph_replays_equal_eager(device):
    """A batch call of 260 steps replays a captured 50-step graph four times
    (steps 1-200; the rest launch one by one).  Its result equals the same
    batch advanced in eager calls of 40 steps (no graph), bit for bit, and
    a second graph call from the same start, after eager work and host
    synchronisations, gives it again; list rebuilds fall inside the graphs
    (skin 0.02)."""
    print("[batch] graph replays", flush=True)
    ws = [W.disks_hertz_friction(N, seed=40 + k, phi=p, skin=0.02) for k, (N, p) in
          enumerate([(700, 0.6), (2000, 0.75), (300, 0.4)])]
    b = jd.Batch.pack([W.make_gpu_system(w, device=device) for w in ws])
    g1 = b.step(260)
    assert g1.system.__dict__["_last_stats"].rebuilds >= 5
    e = b
    for n in (40, 40, 40, 40, 40, 40, 20):
        e = e.step(n)
        torch.cuda.synchronize()
    g2 = b.step(260)
    for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
        assert_bitwise(f"graph vs eager {f}", getattr(g1.state, f).cpu().numpy(), getattr(e.state, f).cpu().numpy())
        assert_bitwise(f"graph again {f}", getattr(g2.state, f).cpu().numpy(), getattr(g1.state, f).cpu().numpy())
    assert torch.equal(g1.system.collider.neighbor_list, e.system.collider.neighbor_list)
    assert g1.n_build == e.n_build == g2.n_build
    for (sa, _), (sb, _) in zip(g1.unpack(), [jd.System.step(*W.make_gpu_system(w, device=device), n=260) for w in ws]):
        _kinematics_equal(sa, sb, "graph vs alone")


def test_batch_launch_modes_equal(device):
    """The three ways rebuild decisions reach the kernels -- a gated chain
    after every step with captured graphs, the same chain launched step by
    step, speculative chunks with a host read per chunk (handing over to the
    gated chain once rebuilds come close together, as they do at skin 0.02)
    -- give the same batch bit for bit, with rebuilds inside the run and a
    system joining without a list; so do speculative chunks that never hand
    over (skin 0.3, rare rebuilds)."""
    print("[batch] launch modes", flush=True)
    ws = [W.disks_hertz_friction(N, seed=60 + k, phi=p, skin=0.02) for k, (N, p) in
          enumerate([(900, 0.7), (256, 0.35), (3000, 0.6)])]
    pairs = [W.make_gpu_system(w, device=device) for w in ws]
    pairs[0] = jd.System.step(*pairs[0], n=7)
    out = []
    for mode in (0, 1, 2):
        b = jd.Batch.pack(pairs)
        b.launch_mode = mode
        out.append(b.step(230))
    for mode, o in zip((1, 2), out[1:]):
        for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
            assert_bitwise(f"mode {mode} {f}", getattr(o.state, f).cpu().numpy(), getattr(out[0].state, f).cpu().numpy())
        assert torch.equal(o.system.collider.neighbor_list, out[0].system.collider.neighbor_list)
        assert o.n_build == out[0].n_build
    assert out[2].system.__dict__["_last_stats"].host_syncs > 1
    assert max(out[0].n_build) >= 5
    ws = [W.disks_hertz_friction(N, seed=64 + k, phi=p, skin=0.3) for k, (N, p) in enumerate([(900, 0.7), (2000, 0.5)])]
    pairs = [W.make_gpu_system(w, device=device) for w in ws]
    out = []
    for mode in (1, 2):
        b = jd.Batch.pack(pairs)
        b.launch_mode = mode
        out.append(b.step(150))
    for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
        assert_bitwise(f"rare rebuilds {f}", getattr(out[1].state, f).cpu().numpy(), getattr(out[0].state, f).cpu().numpy())
    assert torch.equal(out[1].system.collider.neighbor_list, out[0].system.collider.neighbor_list)
    assert out[1].n_build == out[0].n_build


def test_batch_tile_width_and_cell_factor(device):
    """The autotuner's knobs: the step kernel's tile width (128 or 256
    particles per workgroup) changes nothing, bit for bit; cells of half the
    list radius (stencil half-width 2) and of 1.5 list radii give lists and
    states equal to the oracle's with the same cell lists."""
    from jaxdem_amd.batching import with_cell_factor

    print("[batch] tile width / cell factor", flush=True)
    ws = [W.disks_hertz_friction(N, seed=70 + k, phi=p, skin=0.02) for k, (N, p) in
          enumerate([(900, 0.7), (256, 0.35), (3000, 0.6)])]
    pairs = [W.make_gpu_system(w, device=device) for w in ws]
    out = []
    for tile in (256, 128):
        b = jd.Batch.pack(pairs)
        b.tile_width = tile
        out.append(b.step(120))
    for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
        assert_bitwise(f"tile 128 {f}", getattr(out[1].state, f).cpu().numpy(), getattr(out[0].state, f).cpu().numpy())
    assert torch.equal(out[1].system.collider.neighbor_list, out[0].system.collider.neighbor_list)
    with pytest.raises(Exception, match="tile_width must be 128 or 256"):
        b = jd.Batch.pack(pairs)
        b.tile_width = 64
        b.step(1)
    for factor in (0.5, 1.5):
        fp = with_cell_factor(pairs, factor)
        assert fp[0][1].collider.secondary_collider.search_range == (2 if factor < 1 else 1)
        res = jd.Batch.pack(fp).step(80).unpack()
        for k, ((st, sy), (sb, yb)) in enumerate(zip(fp, res)):
            print(f"[batch] cell factor {factor} system {k}", flush=True)
            ost, nls = _oracle_system(st, sy, 80)
            _compare_states(sb, ost, f"factor {factor} system {k}")
            assert yb.collider.n_build_times == nls.n_build
            assert_bitwise("nl", yb.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)


def test_batch_split_steps(device):
    """batch.step(30) == batch.step(12).step(18)."""
    ws = [W.disks_hertz_friction(N, seed=10 + k, phi=0.7) for k, N in enumerate([500, 900])]
    b = jd.Batch.pack([W.make_gpu_system(w, device=device) for w in ws])
    one = b.step(30).unpack()
    two = b.step(12).step(18).unpack()
    for k, ((sa, _), (sb, _)) in enumerate(zip(one, two)):
        _kinematics_equal(sa, sb, f"sys{k}")


def _oracle_system(st, sy, n):
    ost = B.state_of(st)
    nls = B.nl_state_of(st, sy)
    return O.step(ost, B.params_of(st, sy), n, nls), nls


def test_batching_workload_sample_against_oracle(device):
    """Systems of the batching workload (4096 2D Hertz-friction disk
    systems, N log-uniform in [256, 16384], area fraction U[0.3, 0.8]) in one
    batch: the smallest, the largest and the 10/25/50/75/90 % quantiles by N,
    80 steps with list rebuilds (skin 0.02).  Each system's state, list and
    springs equal the oracle's bit for bit."""
    print("[c3] batching workload sample vs oracle", flush=True)
    ws = W.batch_disks(4096, seed=0, skin=0.02)
    Ns = np.array([w.N for w in ws])
    order = np.argsort(Ns, kind="stable")
    picks = sorted({int(order[int(round(q * (len(order) - 1)))]) for q in (0.0, 0.1, 0.25, 0.5, 0.75, 0.9, 1.0)})
    sel = [ws[k] for k in picks]
    assert max(w.N for w in sel) == Ns.max() and min(w.N for w in sel) == Ns.min()
    pairs = [W.make_gpu_system(w, device=device) for w in sel]
    n = 80
    out = jd.Batch.pack(pairs).step(n).unpack()
    builds = []
    for k, ((st, sy), (sb, yb)) in enumerate(zip(pairs, out)):
        print(f"[c3] system {picks[k]} N={sel[k].N}", flush=True)
        ost, nls = _oracle_system(st, sy, n)
        _compare_states(sb, ost, f"system {picks[k]} (N={sel[k].N})")
        assert yb.collider.n_build_times == nls.n_build
        assert_bitwise("nl", yb.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
        xi, live = B.history_of(yb.collider)
        assert_bitwise("live", live, nls.live)
        assert_bitwise("xi", xi, nls.xi)
        builds.append(nls.n_build)
    assert max(builds) >= 3  # rebuilds inside the run


def test_batching_workload_full_size(device):
    """The whole batching workload (4096 systems, ~15.6 M disks) in one
    batch: step(12) then step(18) equals step(30) bit for bit for every
    disk; every system keeps its momentum (periodic boxes, equal and
    opposite pair forces) to float32 rounding; sampled systems equal the
    same system stepped alone."""
    print("[c3] full batching workload", flush=True)
    ws = W.batch_disks(4096, seed=0, skin=0.1)
    pairs = [W.make_gpu_system(w, device="cpu") for w in ws]
    b = jd.Batch.pack(pairs).to(device)
    one = b.step(30)
    two = b.step(12).step(18)
    for f in ("pos_c", "vel", "ang_vel", "force", "torque"):
        assert torch.equal(getattr(one.state, f), getattr(two.state, f)), f
    assert torch.equal(one.state.q.data, two.state.q.data)
    seg = torch.repeat_interleave(torch.arange(len(ws), device=device),
                                  torch.as_tensor(np.diff(b.offsets), device=device))

    def momenta(st):
        p = torch.zeros((len(ws), 2), dtype=torch.float64, device=device)
        p.index_add_(0, seg, (st.mass[:, None] * st.vel).double())
        a = torch.zeros(len(ws), dtype=torch.float64, device=device)
        a.index_add_(0, seg, (st.mass[:, None] * st.vel).double().abs().sum(1))
        return p, a
    p0, a0 = momenta(b.state)
    p1, _ = momenta(one.state)
    assert float(((p1 - p0).abs().max(1).values / a0).max()) < 1e-5
    for k in (0, 1234, 4095):
        sa, _ = jd.System.step(*W.make_gpu_system(ws[k], device=device), n=30)
        a, z = int(b.offsets[k]), int(b.offsets[k + 1])
        for f in ("pos_c", "vel", "ang_vel"):
            assert_bitwise(f"system {k} {f}", getattr(sa, f).cpu().numpy(), getattr(one.state, f)[a:z].cpu().numpy())


# ---------------------------------------------------------------------------
# (f) serving: continuous batching of rollout requests
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("budget,policy", [(None, "fcfs"), (1, "fcfs"), (None, "sjf")])
def test_serving_results_equal_alone(device, budget, policy):
    """Requests of different sizes, lengths, domains and friction served
    together (and, with a one-byte budget, one at a time by admission; with
    least-work-first admission, at most two at a time in order of work):
    every result equals the request run alone with System.step, bit for
    bit."""
    from jaxdem_amd import serving as SV

    print(f"[serving] budget={budget} policy={policy}", flush=True)
    spec = [(1200, 37, "periodic", 0.4, 0.1), (2500, 80, "reflect", 0.5, 0.3), (900, 15, "reflect", 0.35, 0.0),
            (3000, 64, "periodic", 0.45, 0.45), (1500, 50, "periodic", 0.3, 0.2)]
    reqs = []
    for k, (N, steps, dom, phi, mu) in enumerate(spec):
        w = W.polydisperse_spheres(N, seed=100 + k, phi=phi, domain=dom, mu=mu)
        if dom == "reflect":
            w.box = w.box * np.float32(1.02)
        st, sy = W.make_gpu_system(w, device=device)
        reqs.append(SV.Request(k, st, sy, steps, arrival=0.0))
    server = SV.Server(device, hbm_budget=budget if budget is not None else 1 << 36, chunk=16, policy=policy,
                       max_active=2 if policy == "sjf" else 4096)
    done = server.run(reqs, realtime=False)
    assert sorted(r.rid for r in done) == list(range(len(spec)))
    if budget == 1:
        assert max(n for _, n, _, _ in server.log) == 1  # one request at a time
    elif policy == "sjf":
        assert max(n for _, n, _, _ in server.log) == 2
        # the two least-work requests were admitted first
        first = sorted(range(len(spec)), key=lambda k: spec[k][0] * spec[k][1])[:2]
        assert sorted(r.rid for r in sorted(done, key=lambda r: r.admitted)[:2]) == sorted(first)
    else:
        assert max(n for _, n, _, _ in server.log) == len(spec)
    for r in done:
        sa, ya = jd.System.step(reqs[r.rid].state, reqs[r.rid].system, n=r.steps)
        sb, yb = r.result
        _kinematics_equal(sa, sb, f"request {r.rid}")
        assert yb.step_count == r.steps
    st = SV.latency_stats(done)
    assert st["p50_latency_s"] <= st["p99_latency_s"] and st["particle_steps"] == sum(N * s for N, s, *_ in spec)


def test_serving_workload_requests_against_oracle(device):
    """Requests of the serving workload (Poisson arrivals; 3D polydisperse
    spheres, log-normal radii, periodic or reflect, volume fraction
    U[0.3, 0.55], spring + friction mu U[0, 0.5]) capped at N <= 4000 and
    <= 80 steps, served together by continuous batching: each result equals
    the oracle's run of that request bit for bit (state, list, springs)."""
    from jaxdem_amd import serving as SV

    print("[c5] serving workload vs oracle", flush=True)
    spec = W.serving_requests(4, rate=50.0, seed=3, n_min=1000, n_max=4000, steps_min=20, steps_max=80)
    assert {w.domain for _, _, w in spec} == {"periodic", "reflect"}
    reqs = []
    for k, (t, steps, w) in enumerate(spec):
        st, sy = W.make_gpu_system(w, device=device)
        reqs.append(SV.Request(k, st, sy, steps, arrival=t))
    server = SV.Server(device, hbm_budget=1 << 36, chunk=16)
    done = server.run(reqs, realtime=False)
    assert max(n for _, n, _, _ in server.log) >= 2  # served together
    for r in sorted(done, key=lambda r: r.rid):
        q = reqs[r.rid]
        print(f"[c5] request {r.rid} N={q.state.N} steps={q.steps}", flush=True)
        ost, nls = _oracle_system(q.state, q.system, q.steps)
        sb, yb = r.result
        _compare_states(sb, ost, f"request {r.rid}")
        assert_bitwise("nl", yb.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
        xi, live = B.history_of(yb.collider)
        assert_bitwise("live", live, nls.live)
        assert_bitwise("xi", xi, nls.xi)


def test_serving_full_size_request_in_a_mix(device):
    """A full-size request of the serving workload (2e5 spheres) served
    beside two small ones arriving while it runs: its result equals the same
    request run alone with System.step, bit for bit."""
    from jaxdem_amd import serving as SV

    print("[c5] full-size request in a mix", flush=True)
    big = W.serving_requests(1, rate=1.0, seed=11, n_min=200000, n_max=200000, steps_min=150, steps_max=150)[0][2]
    small = W.serving_requests(2, rate=1.0, seed=12, n_min=1000, n_max=3000, steps_min=60, steps_max=100)
    reqs = [SV.Request(0, *W.make_gpu_system(big, device=device), 150, arrival=0.0)]
    for k, (_, steps, w) in enumerate(small):
        reqs.append(SV.Request(k + 1, *W.make_gpu_system(w, device=device), steps, arrival=0.0))
    server = SV.Server(device, hbm_budget=1 << 38, chunk=16)
    done = {r.rid: r for r in server.run(reqs, realtime=False)}
    assert max(n for _, n, _, _ in server.log) == 3
    for k, q in enumerate(reqs):
        sa, _ = jd.System.step(q.state, q.system, n=q.steps)
        _kinematics_equal(sa, done[k].result[0], f"request {k} (N={q.state.N})")


# ---------------------------------------------------------------------------
# (g) precision form: integer cell + float32 offset, float64 reductions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_precise_fire_against_fp64(device, skin):
    """FIRE on the integer-cell + float32-offset positions started at the
    float64 minimum of tests/golden/jamming_n1024.npz (1024 bidisperse
    spheres at volume fraction 0.66, max|F| < 1e-12): the run must stay at
    that minimum to within what float32 pair arithmetic resolves.
    Tolerances: float64 re-evaluation of the final positions E/N 1e-8 and
    P 2e-5 relative; the run's own float64 reductions of float32 pair terms
    E/N and P 2e-5 (the float32 radii alone shift E/N by 4.4e-6 and P by
    3.3e-6); non-rattler positions 1e-5 after the mean shift; residual floor
    (no new low of max|F| or of the energy for 5000 steps) below 1e-5.
    skin 0: the 27-cell stencil every evaluation; skin 0.1: the Verlet list."""
    from tests import precision_vs_fp64 as PV

    fx = np.load(Path(__file__).parent / "golden" / "jamming_n1024.npz")
    rad, L = fx["rad"], float(fx["L"])
    print(f"[precision] N=1024 from the float64 minimum, skin {skin}", flush=True)
    got = PV.gpu_precise(fx["ref"], rad, L, max_steps=20000, stall_steps=5000, device=device, skin=skin)
    res = PV.compare({"pos": fx["ref"]}, got, rad, L)
    print(f"[precision] steps {got['steps']} max|F| {got['max_grad']:.2e} {res}", flush=True)
    assert got["max_grad"] < 1e-5
    assert res["pe_rel_err"] < 1e-8 and res["pressure_rel_err"] < 2e-5
    assert res["reported_pe_rel_err"] < 2e-5 and res["reported_pressure_rel_err"] < 2e-5
    assert res["pos_max_err"] < 1e-5
    assert abs(res["ref_pe"] - float(fx["ref_pe"])) <= 1e-12 * float(fx["ref_pe"])


def test_precise_verlet_list_matches_stencil(device):
    """The precision form's Verlet list against its cell stencil on FIRE from
    the random bidisperse start (4096 spheres, many list rebuilds in 200
    steps): the same pairs in another summation order, so positions agree to
    float32 rounding of the forces (1e-5 of the cell edge), and the energy
    to 1e-5 relative; a row bound too small for the packing is an error."""
    from jaxdem_amd import _lib as Lb
    from jaxdem_amd import precision as PR

    pos, rad, L = W.bidisperse_jamming(4096, seed=3, phi=0.66)
    kw = dict(k=1.0, dt=1e-2, max_steps=200, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=0.0, device=device)
    print("[precision] list vs stencil, 4096 spheres, 200 steps", flush=True)
    a = PR.minimize_precise(pos, rad, L, **kw)
    b = PR.minimize_precise(pos, rad, L, skin=0.05, **kw)
    assert a["steps"] == b["steps"] == 200
    h = L / a["dims"][0]
    dx = PR.min_image_diff(a["pos"], b["pos"], L)
    assert np.abs(dx).max() < 1e-5 * h, np.abs(dx).max()
    assert abs(a["pe"] - b["pe"]) <= 1e-5 * abs(a["pe"])
    with pytest.raises(Lb.JdemError, match="max_neighbors"):
        PR.minimize_precise(pos, rad, L, skin=0.05, max_neighbors=4, **kw)


def test_skin_autotuner_system(device):
    """The plain-system skin autotuner probes every candidate and keeps the
    fastest; every candidate's list fits its rows."""
    from jaxdem_amd.tuning import autotune_skin_system

    print("[tuning] skin autotuner", flush=True)

    def make_pair(sk):
        return W.make_gpu_system(W.spheres_friction(N=4096, seed=6, skin=sk), device=device)

    best, table = autotune_skin_system(make_pair, [0.05, 0.1], probe_steps=30, warm_steps=5)
    assert [r[0] for r in table] == [0.05, 0.1]
    assert best in (0.05, 0.1) and best == max(table, key=lambda r: r[1])[0]
    assert all(r[1] > 0 and r[2] >= 0 and r[3] > 0 for r in table)


def test_thermal_parts_and_temperature(device):
    """Translational and rotational kinetic energy against float64 sums of
    the same state (rotational: the body-frame form, which for the isotropic
    inertia of a sphere equals 1/2 I |w|^2), temperature = 2 KE / n_dof, and
    scale_to_temperature hits its target."""
    from jaxdem_amd.utils import thermal as T

    print("[thermal] kinetic parts", flush=True)
    st, _ = packed(3000, 3, seed=21, domain="reflect")
    m = st.mass.double().cpu().numpy()
    v = st.vel.double().cpu().numpy()
    w = st.ang_vel.double().cpu().numpy()
    inert = st.inertia.double().cpu().numpy()
    assert np.allclose(inert[:, 0], inert[:, 1]) and np.allclose(inert[:, 0], inert[:, 2])
    kt = 0.5 * float(np.sum(m * np.sum(v * v, axis=1)))
    kr = 0.5 * float(np.sum(inert[:, 0] * np.sum(w * w, axis=1)))
    assert abs(T.compute_translational_kinetic_energy(st) - kt) < 1e-12 * kt
    assert abs(T.compute_rotational_kinetic_energy(st) - kr) < 1e-6 * kr  # float32 quaternion rotation
    n_dof, n_v, n_w = T.count_dynamic_dofs(st, can_rotate=True, subtract_drift=True)
    assert (n_v, n_w) == ((3000 - 1) * 3, 3000 * 3)
    temp = T.compute_temperature(st, can_rotate=True, subtract_drift=True)
    assert abs(temp - 2.0 * (T.compute_translational_kinetic_energy(st) + T.compute_rotational_kinetic_energy(st))
               / n_dof) < 1e-12 * temp
    s2 = T.scale_to_temperature(st, 0.25, can_rotate=True, subtract_drift=True)
    assert abs(T.compute_temperature(s2, can_rotate=True, subtract_drift=True) - 0.25) < 1e-5
    drift = (s2.vel.double() * s2.mass.double()[:, None]).sum(dim=0).cpu().numpy()
    assert np.all(np.abs(drift) < 1e-3)
    s3 = T.set_temperature(st, 1.5, can_rotate=True, subtract_drift=True, seed=3)
    assert abs(T.compute_temperature(s3, can_rotate=True, subtract_drift=True) - 1.5) < 1e-5 * 1.5
    assert not torch.equal(s3.vel, st.vel)
    s4 = T.set_temperature(st, 1.5, can_rotate=False, subtract_drift=False, seed=3)
    assert abs(T.compute_temperature(s4, can_rotate=False, subtract_drift=False) - 1.5) < 1e-5 * 1.5


@pytest.mark.parametrize("collider", ["naive", "CellList", "NeighborList"])
def test_empty_and_single_sphere(device, collider):
    """N = 0: stepping, energies and rollout frames are no-ops of the right
    shapes (naive collider; a cell list cannot be created for no spheres,
    as in the reference, whose CellList.Create takes the min of the radii,
    cell_list.py:400-402).  N = 1 in a periodic box: no pair, so velocity
    Verlet moves the sphere ballistically, x += v dt per step in float32,
    bit for bit."""
    print(f"[edge] empty and single sphere, {collider}", flush=True)
    box = np.full(3, 6.0, np.float32)
    for n in (0, 1):
        pos = np.full((n, 3), 1.25, np.float32)
        vel = np.tile(np.array([[0.3, -0.7, 1.1]], np.float32), (n, 1))
        st = jd.State.create(pos=pos, vel=vel, rad=np.full(n, 0.5, np.float32), mass=np.ones(n, np.float32),
                             device=device)
        mt = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=1.0, young=1e4, poisson=0.3)])
        ckw = {}
        if collider == "CellList":
            ckw = dict(state=st, cell_size=1.0)
        if collider == "NeighborList":
            ckw = dict(state=st, cutoff=1.0, skin=0.1, max_neighbors=8, secondary_collider_type="CellList",
                       secondary_collider_kw=dict(state=st, cell_size=1.0))
        kw = dict(dt=1e-2, domain_type="periodic", domain_kw=dict(box_size=box, anchor=np.zeros(3, np.float32)),
                  collider_type=collider, collider_kw=ckw, force_model_type="spring", mat_table=mt)
        if n == 0 and collider != "naive":
            with pytest.raises(ValueError):
                jd.System.create(st.shape, **kw)
            continue
        sy = jd.System.create(st.shape, **kw)
        s2, y2 = jd.System.step(st, sy, n=25)
        assert s2.pos_c.shape == (n, 3) and y2.step_count == sy.step_count + 25
        assert sy.collider.compute_potential_energy(s2, y2)[2] == 0.0
        want = pos.copy()
        for _ in range(25):
            want = want + vel * np.float32(1e-2)
        assert_bitwise(f"n={n} pos", s2.pos_c.cpu().numpy(), want)
        assert_bitwise(f"n={n} vel", s2.vel.cpu().numpy(), vel)
        if collider == "NeighborList":
            _, _, traj = jd.System.trajectory_rollout(st, sy, strides=np.array([5, 20]), save_fn=jd.save_state)
            assert traj.pos_c.shape == (2, n, 3)
            if n:
                assert_bitwise("frame 1", traj.pos_c[1].cpu().numpy(), want)


def test_neighborlist_wide_rows_friction_bitwise(device):
    """Rows of more than 64 slots (a large skin): the step kernel's 4-word
    row form and several contact-history words per sphere, across
    rebuilds, bit for bit against the oracle."""
    print("[step] neighborlist wide rows", flush=True)
    w = W.spheres_friction(N=4096, seed=7, skin=0.3)
    w.vel = w.vel * np.float32(40.0)  # fast enough to rebuild every few steps
    st, sy = W.make_gpu_system(w, device=device)
    assert sy.collider.max_neighbors > 64, sy.collider.max_neighbors
    n = 60
    st2, sy2 = jd.System.step(st, sy, n=n)
    nls = B.nl_state_of(st, sy)
    ost = O.step(B.state_of(st), B.params_of(st, sy), n, nls)
    _compare_states(st2, ost, "nl-wide")
    assert sy2.collider.n_build_times == nls.n_build
    assert nls.n_build >= 3, nls.n_build
    assert_bitwise("nl", sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl)
    xi, live = B.history_of(sy2.collider)
    assert_bitwise("live", live, nls.live)
    assert_bitwise("xi", xi, nls.xi)


def test_full_size_properties(device):
    """The benchmark workload at full size (2^20 frictional spheres), where
    the oracle is too slow: step(12) then step(18) equals step(30) bit for
    bit (rebuilds included), rollout frames equal the split steps, and the
    total momentum (periodic box, pair forces equal and opposite) is kept to
    float32 rounding of the per-sphere updates."""
    print("[full size] spheres_1048576_friction", flush=True)
    w = W.spheres_friction(N=1 << 20, seed=0, skin=0.2)
    w.vel = w.vel * np.float32(10.0)  # rebuilds within the run
    st, sy = W.make_gpu_system(w, device=device)
    a, ya = jd.System.step(st, sy, n=30)
    b1, yb1 = jd.System.step(st, sy, n=12)
    b, yb = jd.System.step(b1, yb1, n=18)
    _kinematics_equal(a, b, "split")
    assert ya.__dict__["_last_stats"].rebuilds >= 1
    _, _, traj = jd.System.trajectory_rollout(st, sy, strides=np.array([12, 18]), save_fn=jd.save_state)
    assert_bitwise("frame 0", traj.pos_c[0].cpu().numpy(), b1.pos_c.cpu().numpy())
    assert_bitwise("frame 1", traj.pos_c[1].cpu().numpy(), a.pos_c.cpu().numpy())
    m = st.mass.double()[:, None]
    p0 = (m * st.vel.double()).sum(dim=0)
    p1 = (m * a.vel.double()).sum(dim=0)
    scale = (m * st.vel.double().abs()).sum()
    assert float((p1 - p0).abs().max()) < 1e-5 * float(scale), (p0, p1)
