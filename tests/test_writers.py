"""VTK output (writers/vtk_writer.py, vtk_spheres_writer.py,
vtk_domain_writer.py, async_base.py): files and collections laid out as the
reference writes them, every array read back equal to what was saved, in
binary and ASCII form (CPU)."""

from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import replace

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from jaxdem_amd.writers import read_vtp, write_polydata


def _system(st, domain="periodic", box=6.0):
    return jd.System.create(st.shape, domain_type=domain,
                            domain_kw=dict(box_size=np.full(st.dim, box, np.float32),
                                           anchor=np.zeros(st.dim, np.float32)))


def _state(dim=3, n=7, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(0.5, 5.5, (n, dim)).astype(np.float32)
    fixed = np.zeros(n, bool)
    fixed[2] = True
    return jd.State.create(pos=pos, vel=rng.standard_normal((n, dim)).astype(np.float32),
                           rad=rng.uniform(0.2, 0.4, n).astype(np.float32), fixed=fixed, device="cpu")


@pytest.mark.parametrize("binary", [True, False])
def test_polydata_round_trip(tmp_path, binary):
    pts = np.arange(12, dtype=np.float32).reshape(4, 3) / 7.0
    data = {"a": np.array([1.5, -2.25, 3.0, 1e-7], np.float32), "b": np.arange(8, dtype=np.int32).reshape(4, 2),
            "f": np.array([True, False, True, False])}
    fn = tmp_path / "x.vtp"
    write_polydata(fn, pts, data, polys=[[0, 1, 2], [1, 2, 3]], binary=binary)
    back = read_vtp(fn)
    assert np.array_equal(back["Points"], pts)
    assert np.array_equal(back["a"], data["a"]) and back["a"].dtype == np.float32
    assert np.array_equal(back["b"], data["b"])
    assert back["f"].dtype == np.int8 and np.array_equal(back["f"], data["f"].astype(np.int8))
    assert back["connectivity"].tolist() == [0, 1, 2, 1, 2, 3] and back["offsets"].tolist() == [3, 6]
    root = ET.fromstring(fn.read_bytes().split(b"<AppendedData")[0] + (b"</VTKFile>" if binary else b""))
    piece = next(root.iter("Piece"))
    assert piece.get("NumberOfPoints") == "4" and piece.get("NumberOfPolys") == "2"


@pytest.mark.parametrize("dim", [2, 3])
def test_spheres_and_domain_files(tmp_path, dim):
    st = _state(dim)
    sy = _system(st)
    st = replace(st, pos_c=st.pos_c + torch.tensor([6.0] + [0.0] * (dim - 1)))
    w = jd.VTKWriter(directory=tmp_path / "frames", binary=True, max_workers=2)
    w.save(st, sy)
    w.close()
    d = tmp_path / "frames" / "batch_00000000"
    sph = read_vtp(d / "spheres_00000000.vtp")
    # periodic positions are wrapped into the box before writing (domain.shift)
    want = np.mod(st.pos_c.numpy(), 6.0)
    pts = sph["Points"]
    assert pts.shape == (st.N, 3)
    assert np.allclose(pts[:, :dim], want, atol=1e-6) and not np.any(pts[:, dim:])
    assert np.array_equal(sph["rad"], st.rad.numpy())
    vel = sph["vel"]
    assert vel.shape == (st.N, 3) and np.array_equal(vel[:, :dim], st.vel.numpy())
    assert np.array_equal(sph["fixed"], st.fixed.numpy().astype(np.int8))
    assert np.array_equal(sph["q.w"], st.q.w.numpy()[:, 0]) and sph["q.xyz"].shape == (st.N, 3)
    assert "pos_c" not in sph and "_rad" not in sph
    dom = read_vtp(d / "domain_00000000.vtp")
    corners = dom["Points"]
    assert corners.shape == (8, 3)
    assert np.allclose(corners[:, :dim].min(0), 0.0) and np.allclose(corners[:, :dim].max(0), 6.0)
    assert dom["offsets"].tolist() == [4, 8, 12, 16, 20, 24]
    pvd = ET.parse(tmp_path / "frames" / "batch_00000000_spheres.pvd").getroot()
    ds = list(pvd.iter("DataSet"))
    assert [x.get("file") for x in ds] == ["batch_00000000/spheres_00000000.vtp"]


def test_trajectory_batches_and_skipping(tmp_path):
    st = _state(3, n=5)
    sy = _system(st, domain="reflect")
    frames, systems = [], []
    for k in range(3):
        frames.append(replace(st, pos_c=st.pos_c + 0.1 * k))
        systems.append(replace(sy, step_count=10 * (k + 1), time=0.5 * (k + 1)))
    stacked = jd.State.stack(frames)
    w = jd.VTKWriter(directory=tmp_path / "traj", writers=["spheres"], binary=False, max_workers=1)
    w.save(stacked, systems, trajectory=True)
    w.save(stacked, sy)  # three batches of one frame
    w.close()
    files = sorted(p.name for p in (tmp_path / "traj" / "batch_00000000").iterdir())
    assert files == ["spheres_00000000.vtp", "spheres_00000010.vtp", "spheres_00000020.vtp", "spheres_00000030.vtp"]
    assert (tmp_path / "traj" / "batch_00000002" / "spheres_00000000.vtp").exists()
    pvd = ET.parse(tmp_path / "traj" / "batch_00000000_spheres.pvd").getroot()
    steps = [(x.get("timestep"), x.get("file")) for x in pvd.iter("DataSet")]
    assert steps[1:] == [("0.5", "batch_00000000/spheres_00000010.vtp"), ("1", "batch_00000000/spheres_00000020.vtp"),
                         ("1.5", "batch_00000000/spheres_00000030.vtp")]
    back = read_vtp(tmp_path / "traj" / "batch_00000000" / "spheres_00000020.vtp")
    assert np.allclose(back["Points"], frames[1].pos_c.numpy(), atol=1e-6)
    # save_every: the first call and every k-th after it
    w2 = jd.VTKWriter(directory=tmp_path / "skip", writers=["domain"], save_every=2, max_workers=1)
    for k in range(5):
        w2.save(st, replace(sy, step_count=k))
    w2.close()
    assert sorted(p.name for p in (tmp_path / "skip" / "batch_00000000").iterdir()) == [
        "domain_00000000.vtp", "domain_00000002.vtp", "domain_00000004.vtp"]


def test_writer_registry_and_safety(tmp_path, monkeypatch):
    with pytest.raises(KeyError):
        jd.VTKWriter(directory=tmp_path / "x", writers=["nope"])
    assert set(jd.VTKBaseWriter._registry) >= {"spheres", "domain"}
    # the working directory is never cleaned
    monkeypatch.chdir(tmp_path)
    (tmp_path / "keep.txt").write_text("x")
    jd.VTKWriter(directory=tmp_path, writers=["domain"], max_workers=1).close()
    assert (tmp_path / "keep.txt").exists()


@pytest.mark.gpu
def test_device_rollout_frames_to_files(device, tmp_path):
    """Frames of a rollout written on the device (save_state) go to one
    file per frame; each file holds that frame's positions (wrapped into
    the periodic box) and velocities."""
    from jaxdem_amd import workloads as W

    st, sy = W.make_gpu_system(W.spheres_friction(N=2048, seed=3, skin=0.1), device=device)
    st2, sy2, frames = jd.System.trajectory_rollout(st, sy, n=4, stride=5, save_fn=jd.save_state)
    w = jd.VTKWriter(directory=tmp_path / "run", max_workers=2)
    systems = [replace(sy, step_count=5 * (k + 1)) for k in range(4)]
    w.save(frames, systems, trajectory=True)
    w.close()
    box = np.asarray(sy.domain.box_size, np.float32)
    for k in range(4):
        back = read_vtp(tmp_path / "run" / "batch_00000000" / f"spheres_{5 * (k + 1):08d}.vtp")
        pos = frames.pos_c[k].cpu().numpy()
        wrapped = pos - box * np.floor(pos / box)
        assert np.allclose(back["Points"], wrapped, atol=1e-5)
        assert np.array_equal(back["vel"], frames.vel[k].cpu().numpy())
