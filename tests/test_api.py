"""Host-side API of jaxdem_amd on the CPU: the plugin registry behaves like
jaxdem/factory.py (tests/test_public_api.py:24-79 of the reference), System
and State construction validate like the reference, and the GPU entry points
refuse CPU tensors instead of falling back to anything."""

from __future__ import annotations

import warnings

import numpy as np
import pytest

import jaxdem_amd as jd


def two_spheres(device="cpu"):
    st = jd.State.create(pos=np.array([[0.0, 0.0], [1.5, 0.0]]), rad=np.ones(2), device=device)
    return st, jd.System.create(st.shape)


def test_registry_keys_are_normalized():
    st, _ = two_spheres()
    classes = {type(jd.Collider.create(k, state=st)) for k in ["CellList", "cell_list", "celllist", "Cell-List"]}
    assert classes == {jd.DynamicCellList}
    with pytest.raises(KeyError, match="Unknown Collider"):
        jd.Collider.create("not_a_collider")


def test_registered_keys_of_the_hot_path():
    assert set(jd.Collider._registry) >= {"", "naive", "celllist", "neighborlist"}
    assert set(jd.ForceModel._registry) >= {"spring", "hertz", "cundallstrack", "springfriction", "hertzfriction"}
    assert set(jd.LinearIntegrator._registry) >= {"", "verlet", "euler"}
    assert set(jd.RotationIntegrator._registry) >= {"", "verletspiral", "spiral"}
    assert set(jd.Domain._registry) >= {"free", "periodic", "reflect"}
    assert set(jd.Material._registry) >= {"elastic", "elasticfrict"}
    assert set(jd.MaterialMatchmaker._registry) >= {"harmonic", "linear"}


def test_register_conflict_and_idempotence():
    class Dummy(jd.Collider):
        pass

    jd.Collider.register("dummy_test_key")(Dummy)
    jd.Collider.register("dummy_test_key")(Dummy)  # same class again: fine

    class Other(jd.Collider):
        pass

    with pytest.raises(ValueError, match="already registered"):
        jd.Collider.register("dummy_test_key")(Other)
    del jd.Collider._registry["dummytestkey"]


def test_create_may_not_be_overridden():
    with pytest.raises(TypeError, match="not allowed to override"):

        class Bad(jd.Collider):  # noqa: F841
            @classmethod
            def create(cls, key, **kw):
                return None


def test_unknown_keyword_warns_and_bad_keyword_raises():
    st, _ = two_spheres()
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        c = jd.Collider.create("CellList", state=st, not_a_kw=3)  # Create has no **kw: dropped
    assert isinstance(c, jd.DynamicCellList)
    assert any("ignoring unknown keyword" in str(x.message) for x in w)
    with pytest.raises(TypeError):  # Domain.Create forwards **kw to the constructor
        jd.Domain.create("periodic", dim=2, box_size=[1.0, 1.0], not_a_kw=3)
    with pytest.raises(ValueError):
        jd.Domain.create("reflect", dim=2, restitution_coefficient=1.5)
    with pytest.raises(ValueError, match="box_size must have shape"):
        jd.Domain.create("periodic", dim=3, box_size=[1.0, 1.0])


def test_type_name_and_defaults():
    st, sy = two_spheres()
    assert sy.collider.type_name == "naive"
    assert sy.linear_integrator.type_name == "verlet"
    assert sy.rotation_integrator.type_name == "verletspiral"
    assert sy.domain.type_name == "free"
    assert sy.force_model.type_name == "spring"
    assert float(sy.mat_table.young_eff[0, 0]) == pytest.approx(1e4)
    assert sy.dt == np.float32(0.005)


def test_system_create_validation():
    st, _ = two_spheres()
    with pytest.raises(ValueError, match="does not match"):
        jd.System.create((3, 2), state=st)
    with pytest.raises(TypeError):
        jd.System.create()
    mt = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=1.0, young=1.0, poisson=0.3)])
    with pytest.raises(KeyError, match="lacks fields"):
        jd.System.create(st.shape, force_model_type="springfriction", mat_table=mt)
    mt2 = jd.MaterialTable.from_materials(
        [jd.Material.create("elasticfrict", density=1.0, young=1.0, poisson=0.3, mu=0.5, e=1.0)]
    )
    with pytest.raises(ValueError, match="requires history"):
        jd.System.create(st.shape, force_model_type="springfriction", mat_table=mt2, collider_type="naive")


def test_neighborlist_create_and_history_allocation():
    st = jd.State.create(pos=np.random.default_rng(0).uniform(0, 5, (50, 3)), rad=np.full(50, 0.5), device="cpu")
    mt = jd.MaterialTable.from_materials(
        [jd.Material.create("elasticfrict", density=1.0, young=1e4, poisson=0.3, mu=0.5, e=1.0)]
    )
    sy = jd.System.create(
        st.shape,
        collider_type="NeighborList",
        collider_kw=dict(state=st, cutoff=1.0, skin=0.1),
        force_model_type="springfriction",
        mat_table=mt,
    )
    col = sy.collider
    assert col.max_neighbors > 0 and col.neighbor_list.shape == (50, col.max_neighbors)
    assert col.history["xi"].shape == (50, col.max_neighbors, 3)
    assert col.secondary_collider.type_name == "celllist"
    assert col.secondary_collider.cell_size == pytest.approx(1.1)
    with pytest.raises(ValueError):
        jd.NeighborList.Create(st, cutoff=1.0, skin=0.1, skin_fraction=0.1)


def test_cell_list_create_defaults():
    """cell_list.py:400-427: cell 2 r_max (alpha < 2.5) or r_max/2, 'ij' stencil."""
    st = jd.State.create(pos=np.zeros((3, 3)), rad=np.array([0.5, 0.6, 0.7]), device="cpu")
    c = jd.Collider.create("CellList", state=st)
    assert c.cell_size == pytest.approx(1.4)
    assert c.neighbor_mask.shape == (27, 3)
    assert list(c.neighbor_mask[0]) == [-1, -1, -1] and list(c.neighbor_mask[1]) == [-1, -1, 0]
    st2 = jd.State.create(pos=np.zeros((2, 2)), rad=np.array([0.1, 1.0]), device="cpu")
    c2 = jd.Collider.create("CellList", state=st2)
    assert c2.cell_size == pytest.approx(0.5) and c2.search_range == 4


def test_grid_params_match_reference_rules():
    dims, strides, cell, ovf = jd.colliders.grid_params(np.array([10.0, 7.0, 3.0], np.float32), 1.4, True)
    assert list(dims) == [7, 5, 2] and list(strides) == [1, 7, 35]
    np.testing.assert_allclose(cell, np.array([10 / 7, 7 / 5, 3 / 2], np.float32))
    dims, _, cell, _ = jd.colliders.grid_params(np.array([10.0, 7.0], np.float32), 1.4, False)
    assert list(dims) == [8, 5] and np.all(cell == np.float32(1.4))


def test_state_create_defaults_and_bonds():
    st = jd.State.create(pos=np.zeros((3, 3)), rad=np.array([1.0, 2.0, 3.0]), bond_id=[[1], [2], []], device="cpu")
    assert st.mass.tolist() == [1.0, 1.0, 1.0]
    np.testing.assert_allclose(st.inertia[:, 0].numpy(), 0.4 * np.array([1.0, 4.0, 9.0]))
    assert st.q.w.shape == (3, 1) and st.q.xyz.shape == (3, 3)
    b = st.bond_id.numpy()
    assert {1} <= set(b[0]) and {0, 2} <= set(b[1]) and {1} <= set(b[2])  # symmetrised
    assert st.is_sphere_state and st.has_bonds
    st2 = jd.State.create(pos=np.zeros((2, 2)), device="cpu")
    assert st2.ang_vel.shape == (2, 1) and st2.inertia[0, 0] == pytest.approx(0.5)


def test_material_table_matchers():
    mats = [
        jd.Material.create("elastic", density=1.0, young=1.0, poisson=0.3),
        jd.Material.create("elastic", density=2.0, young=3.0, poisson=0.2),
    ]
    h = jd.MaterialTable.from_materials(mats)
    assert float(h.young_eff[0, 1]) == pytest.approx(2 * 1 * 3 / 4)
    lin = jd.MaterialTable.from_materials(mats, matcher=jd.MaterialMatchmaker.create("linear"))
    assert float(lin.young_eff[0, 1]) == pytest.approx(2.0)
    assert len(h) == 2


def test_gpu_entry_points_refuse_cpu_tensors():
    st, sy = two_spheres()
    with pytest.raises(RuntimeError, match="GPU"):
        jd.System.step(st, sy, n=1)
    with pytest.raises(RuntimeError, match="GPU"):
        sy.collider.compute_force(st, sy)


def test_batched_states_are_rejected_by_plain_step():
    st, sy = two_spheres()
    bst = jd.State.stack([st, st])
    assert bst.batch_size == 2
    with pytest.raises(NotImplementedError):
        jd.System.step(bst, sy, n=1)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_batch_pack_unpack_round_trip():
    """Batch.pack concatenates states, merges material tables block-diagonally
    and pads the neighbour lists; unpack gives every system back unchanged."""
    import torch

    from jaxdem_amd import workloads as W

    ws = [W.disks_hertz_friction(N, seed=k, phi=p) for k, (N, p) in enumerate([(300, 0.5), (700, 0.7), (256, 0.3)])]
    pairs = [W.make_gpu_system(w, device="cpu") for w in ws]
    b = jd.Batch.pack(pairs)
    assert b.n_particles == 1256 and list(b.offsets) == [0, 300, 1000, 1256]
    assert len(b.system.mat_table) == 1  # the three systems share one material: one block
    assert b.system.collider.max_neighbors == max(p[1].collider.max_neighbors for p in pairs)
    assert int(b.state.mat_id.max()) == 0
    w3 = W.disks_hertz_friction(200, seed=9, phi=0.5, mu=0.1)  # another material
    b2 = jd.Batch.pack(pairs + [W.make_gpu_system(w3, device="cpu")])
    assert len(b2.system.mat_table) == 2 and list(b2.mat_offsets) == [0, 0, 0, 1, 2]
    assert float(b2.system.mat_table.mu_eff[1, 1]) == pytest.approx(0.1)
    for (s0, y0), (s1, y1) in zip(pairs, b.unpack()):
        for f in ("pos_c", "vel", "rad", "mass", "mat_id", "clump_id", "bond_id", "fixed"):
            assert torch.equal(getattr(s0, f), getattr(s1, f)), f
        assert torch.equal(s0.q.data, s1.q.data)
        assert torch.equal(y0.collider.neighbor_list, y1.collider.neighbor_list)
        assert torch.equal(y0.collider.history["xi"], y1.collider.history["xi"])
    with pytest.raises(RuntimeError, match="GPU"):
        b.step(1)


def test_serving_footprint_and_latency_stats():
    from jaxdem_amd import serving as SV
    from jaxdem_amd import workloads as W

    small = W.make_gpu_system(W.polydisperse_spheres(500, seed=1, phi=0.4, domain="periodic", mu=0.2), device="cpu")
    big = W.make_gpu_system(W.polydisperse_spheres(4000, seed=1, phi=0.4, domain="periodic", mu=0.2), device="cpu")
    assert 0 < SV.footprint_bytes(*small) < SV.footprint_bytes(*big)
    rs = []
    for k, (a, f) in enumerate([(0.0, 1.0), (0.5, 1.5), (1.0, 4.0), (1.0, 2.0)]):
        r = SV.Request(k, small[0], small[1], 10, arrival=a)
        r.finished = f
        rs.append(r)
    s = SV.latency_stats(rs)
    assert s["requests"] == 4 and s["makespan_s"] == 4.0
    assert s["p50_latency_s"] == pytest.approx(1.0) and s["p99_latency_s"] == pytest.approx(2.94)
    assert s["particle_steps"] == 4 * 500 * 10


def test_precision_cell_offset_round_trip():
    """x = c h + o with 0 <= o < h; the float64 reconstruction is within one
    float32 rounding of the offset."""
    from jaxdem_amd import precision as PR

    rng = np.random.default_rng(3)
    box = np.array([73.0, 73.0, 73.0])
    pos = rng.uniform(-80.0, 150.0, (5000, 3))
    dims, h = PR.grid_for(box, np.array([0.5, 0.7]))
    assert list(dims) == [52, 52, 52]
    c, o = PR.to_cell_offset(pos, box, dims)
    assert c.dtype == np.int32 and o.dtype == np.float32
    assert (c >= 0).all() and (c < 52).all() and (o >= 0).all() and (o < h.astype(np.float32)).all()
    back = PR.from_cell_offset(c, o, box, dims)
    err = np.abs(PR.min_image_diff(back, pos, box))
    assert err.max() < 2e-7  # float32 resolution of an offset below 1.41, not of a coordinate near 73


def test_server_rejects_bad_policy_and_cap():
    from jaxdem_amd import serving as SV

    with pytest.raises(ValueError, match="policy"):
        SV.Server("cpu", hbm_budget=1, policy="lifo")
    with pytest.raises(ValueError, match="max_active"):
        SV.Server("cpu", hbm_budget=1, max_active=0)
    assert SV.Server("cpu", hbm_budget=1, policy="sjf", max_active=3).policy == "sjf"


def test_server_finishes_zero_step_requests_at_admission():
    """A request of 0 steps returns its own state at once and never enters
    the batch (no step of 0 and no division by it)."""
    from jaxdem_amd import serving as SV
    from jaxdem_amd import workloads as W

    st, sy = W.make_gpu_system(W.polydisperse_spheres(300, seed=2, phi=0.4, domain="periodic", mu=0.2), device="cpu")
    reqs = [SV.Request(k, st, sy, 0, arrival=0.0) for k in range(3)]
    done = SV.Server("cpu", hbm_budget=1).run(reqs, realtime=False)
    assert sorted(r.rid for r in done) == [0, 1, 2]
    for r in done:
        assert r.result[0] is st and r.result[1] is sy and r.latency >= 0.0


def test_trajectory_rollout_save_state_signature():
    # save_state keeps the state of a frame and drops the system
    assert jd.save_state("s", "y") == "s"
    assert jd.system.save_state is jd.save_state


def test_clock_is_sequential_float32_sum():
    """System.time after n steps is n sequential float32 additions of dt
    (system.py:73), for any start time and step count."""
    from jaxdem_amd.system import _clock_after

    rng = np.random.default_rng(7)
    for _ in range(20):
        t0 = np.float32(rng.uniform(0, 5))
        dt = np.float32(10.0 ** rng.uniform(-6, -2))
        n = int(rng.integers(0, 3000))
        t = t0
        for _ in range(n):
            t = np.float32(t + dt)
        assert _clock_after(t0, dt, n) == t
    # a product n * dt would differ
    assert _clock_after(np.float32(0), np.float32(1e-3), 40) != np.float32(0.04)
