"""VTK output of simulation frames (writers/vtk_writer.py:26-298,
writers/vtk_spheres_writer.py:23-142, writers/vtk_domain_writer.py:23-66,
writers/async_base.py:17-173).

``VTKWriter(directory=..., writers=[...], binary=True).save(state, system)``
writes one ``.vtp`` file (VTK XML PolyData) per registered writer and frame
under ``directory/batch_XXXXXXXX/`` and keeps a ``.pvd`` collection per
batch and writer, as the reference does.  Leading axes of the state are
batches, or time with ``trajectory=True`` (a stacked trajectory from
``System.trajectory_rollout``); a trajectory's systems may be one System or
the list of per-frame Systems the rollout returns.  Device arrays are copied
to the host in ``save``; the files are written by background threads.

The reference writes through the ``vtk`` package (not installed here); this
module writes the same XML format itself: ``binary=True`` stores the arrays
raw in an appended block (UInt64 size headers, little endian), ``False`` as
ASCII.  ``read_vtp`` reads both back.
"""

from __future__ import annotations

import atexit
import logging
import queue
import shutil
import threading
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Any

import numpy as np
import torch

from .factory import Factory, _normalize_key

_log = logging.getLogger(__name__)

_VTK_TYPES = {
    np.dtype(np.float32): "Float32",
    np.dtype(np.float64): "Float64",
    np.dtype(np.int8): "Int8",
    np.dtype(np.uint8): "UInt8",
    np.dtype(np.int32): "Int32",
    np.dtype(np.int64): "Int64",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def _as_vtk_array(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype == np.bool_:
        a = a.astype(np.int8)
    if a.dtype not in _VTK_TYPES:
        a = a.astype(np.float64 if np.issubdtype(a.dtype, np.floating) else np.int64)
    return np.ascontiguousarray(a)


def write_polydata(filename, points: np.ndarray, point_data: dict | None = None, polys: list | None = None,
                   binary: bool = True) -> None:
    """Write a VTK XML PolyData file: ``points`` (P, 3), per-point arrays
    ``point_data`` (name -> (P,) or (P, C)), optional polygons (lists of
    point indices)."""
    points = _as_vtk_array(points)
    n = points.shape[0]
    arrays = [("Points", None, points)]
    for name, a in (point_data or {}).items():
        a = _as_vtk_array(a)
        if a.shape[0] != n:
            raise ValueError(f"point array {name!r} has {a.shape[0]} entries for {n} points")
        arrays.append(("PointData", name, a))
    if polys:
        conn = np.asarray([i for p in polys for i in p], np.int64)
        offs = np.cumsum([len(p) for p in polys]).astype(np.int64)
        arrays += [("Polys", "connectivity", conn), ("Polys", "offsets", offs)]
    appended = bytearray()

    def data_array(name, a) -> str:
        ncomp = 1 if a.ndim == 1 else int(np.prod(a.shape[1:]))
        attrs = f'type="{_VTK_TYPES[a.dtype]}"' + (f' Name="{name}"' if name else "") + \
            f' NumberOfComponents="{ncomp}"'
        if binary:
            off = len(appended)
            raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
            appended.extend(np.uint64(len(raw)).tobytes())
            appended.extend(raw)
            return f'<DataArray {attrs} format="appended" offset="{off}"/>'
        body = " ".join(repr(float(x)) if np.issubdtype(a.dtype, np.floating) else str(int(x)) for x in a.ravel())
        return f'<DataArray {attrs} format="ascii">{body}</DataArray>'

    npoly = len(polys) if polys else 0
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
             "<PolyData>",
             f'<Piece NumberOfPoints="{n}" NumberOfVerts="0" NumberOfLines="0" NumberOfStrips="0" '
             f'NumberOfPolys="{npoly}">']
    pd = [data_array(nm, a) for sec, nm, a in arrays if sec == "PointData"]
    parts += ["<PointData>", *pd, "</PointData>"]
    parts += ["<Points>", data_array(None, points), "</Points>"]
    if polys:
        parts += ["<Polys>", *[data_array(nm, a) for sec, nm, a in arrays if sec == "Polys"], "</Polys>"]
    parts += ["</Piece>", "</PolyData>"]
    head = "\n".join(parts).encode()
    with open(filename, "wb") as f:
        f.write(head)
        if binary:
            f.write(b'\n<AppendedData encoding="raw">\n_')
            f.write(bytes(appended))
            f.write(b"\n</AppendedData>")
        f.write(b"\n</VTKFile>\n")


def read_vtp(filename) -> dict:
    """Arrays of a file written by ``write_polydata``: "Points", each point
    array by name, and "connectivity" / "offsets" of the polygons."""
    blob = Path(filename).read_bytes()
    raw = b""
    cut = blob.find(b'<AppendedData encoding="raw">')
    if cut >= 0:
        start = blob.index(b"_", cut) + 1
        end = blob.rindex(b"</AppendedData>")
        raw = blob[start:end]
        xml = blob[:cut] + b"</VTKFile>"
    else:
        xml = blob
    root = ET.fromstring(xml)
    out = {}
    for sec in ("PointData", "Points", "Polys"):
        for node in root.iter(sec):
            for da in node.iter("DataArray"):
                dt = _NP_TYPES[da.get("type")]
                ncomp = int(da.get("NumberOfComponents", "1"))
                if da.get("format") == "appended":
                    off = int(da.get("offset"))
                    size = int(np.frombuffer(raw[off : off + 8], "<u8")[0])
                    a = np.frombuffer(raw[off + 8 : off + 8 + size], dt.newbyteorder("<")).astype(dt)
                else:
                    a = np.array((da.text or "").split(), dtype=dt)
                if ncomp > 1:
                    a = a.reshape(-1, ncomp)
                out[da.get("Name") or sec] = a
    return out


def _host(x: Any) -> Any:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return x


class VTKBaseWriter(Factory):
    """Base of the per-frame VTK writers (writers/__init__.py:25-98):
    ``is_active(state, system)`` and ``write(state, system, filename,
    binary)``, registered by key."""

    @classmethod
    def is_active(cls, state, system) -> bool:
        return True

    @classmethod
    def write(cls, state, system, filename: Path, binary: bool) -> None:
        raise NotImplementedError


def _pad3(a: np.ndarray) -> np.ndarray:
    if a.ndim == 2 and a.shape[1] == 2 and np.issubdtype(a.dtype, np.floating):
        return np.pad(a, ((0, 0), (0, 1)), "constant")
    return a


@VTKBaseWriter.register("spheres")
class VTKSpheresWriter(VTKBaseWriter):
    """Sphere centres as points with every per-particle State field as point
    data (vtk_spheres_writer.py:23-99): 2-component float vectors padded to
    3D, booleans as Int8, the orientation as "q.xyz" and "q.w"."""

    @classmethod
    def write(cls, state, system, filename: Path, binary: bool) -> None:
        pos = _pad3(_host(state.pos))
        n = pos.shape[0]
        pdata = {}
        for f in fields(state):
            if f.name in ("pos_c", "q") or f.name.startswith("_"):
                continue
            a = getattr(state, f.name)
            if isinstance(a, torch.Tensor) and a.ndim >= 1 and a.shape[0] == n:
                pdata[f.name] = _pad3(_host(a))
        q = _host(state.q.data)
        pdata["q.xyz"] = q[:, 1:4]
        pdata["q.w"] = q[:, 0:1]
        write_polydata(filename, pos, pdata, binary=binary)


@VTKBaseWriter.register("domain")
class VTKDomainWriter(VTKBaseWriter):
    """The domain box as an axis-aligned cuboid (vtk_domain_writer.py:23-63):
    its 8 corners and 6 faces; 2D boxes get zero depth."""

    @classmethod
    def write(cls, state, system, filename: Path, binary: bool) -> None:
        box = np.zeros(3)
        anc = np.zeros(3)
        b = np.asarray(system.domain.box_size, np.float64).reshape(-1)
        a = np.asarray(system.domain.anchor, np.float64).reshape(-1)
        box[: b.size] = b
        anc[: a.size] = a
        corners = np.array([[(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], np.float64)
        pts = (anc + corners * box).astype(np.float32)
        faces = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2], [1, 3, 7, 5]]
        write_polydata(filename, pts, None, polys=faces, binary=binary)


@dataclass
class _HostDomain:
    box_size: np.ndarray
    anchor: np.ndarray


@dataclass
class _HostSystem:
    """The parts of a System the writers read, on the host."""

    step_count: int
    time: float
    domain: _HostDomain


def _frame_host(state, system) -> tuple:
    """(state with host tensors, host system) of one frame."""
    sy = _HostSystem(int(system.step_count), float(system.time),
                     _HostDomain(np.asarray(system.domain.box_size, np.float64).copy(),
                                 np.asarray(system.domain.anchor, np.float64).copy()))
    return state.to("cpu"), sy


def _slice(state, k: int):
    """Frame k of a state stacked along a leading axis."""
    from .state import Quaternion, State

    kw = {}
    for f in fields(state):
        v = getattr(state, f.name)
        if isinstance(v, torch.Tensor):
            kw[f.name] = v[k]
        elif isinstance(v, Quaternion):
            kw[f.name] = Quaternion(v.data[k])
        else:
            kw[f.name] = v
    return State(**kw)


def _frames(state, system, trajectory: bool) -> list:
    """(batch offset, state, system) frames of a save call."""
    lead = state.pos_c.ndim - 2
    if lead == 0:
        return [(0, state, system)]
    if lead > 1:
        raise NotImplementedError("states with more than one leading axis")
    n = state.pos_c.shape[0]
    systems = system if isinstance(system, (list, tuple)) else [system] * n
    out = []
    for k in range(n):
        out.append((0 if trajectory else k, _slice(state, k), systems[k]))
    return out


@dataclass
class VTKWriter:
    """Front end (vtk_writer.py:26-298, async_base.py:17-173): writers by
    registry key (all registered ones when empty), ``save_every`` skipping,
    background threads for the file writes, ``.pvd`` collections."""

    directory: Path = Path("./frames")
    save_every: int = 1
    clean: bool = True
    max_workers: int = 8
    max_queue_size: int = 512
    writers: list = field(default_factory=list)
    binary: bool = True

    def __post_init__(self) -> None:
        self.directory = Path(self.directory)
        self.save_every = int(self.save_every)
        if not self.writers:
            self.writers = list(VTKBaseWriter._registry.keys())
        unknown = [w for w in self.writers if _normalize_key(w) not in VTKBaseWriter._registry]
        if unknown:
            raise KeyError(f"Unknown VTKBaseWriter name(s) {unknown}. Available: {list(VTKBaseWriter._registry)}")
        if self.clean and self.directory.exists() and self._is_safe_to_clean(self.directory):
            shutil.rmtree(self.directory)
        self.directory.mkdir(parents=True, exist_ok=True)
        self._lock = threading.Lock()
        self._manifest: dict = {}
        self._save_calls = 0
        self._queue: queue.Queue = queue.Queue(maxsize=max(0, int(self.max_queue_size)))
        self._threads = []
        for _ in range(max(1, int(self.max_workers))):
            t = threading.Thread(target=self._worker, daemon=True)
            t.start()
            self._threads.append(t)
        self._closed = False
        atexit.register(self.close)

    @staticmethod
    def _is_safe_to_clean(path: Path) -> bool:
        """async_base.py:83-100: never the working directory, one of its
        parents or the root."""
        p = path.resolve()
        cwd = Path.cwd().resolve()
        if p == cwd or p in cwd.parents:
            return False
        return p != Path(p.anchor)

    def _worker(self) -> None:
        while True:
            item = self._queue.get()
            if item is None:
                self._queue.task_done()
                break
            func, args = item
            try:
                func(*args)
            except Exception:
                _log.exception("VTK writer task failed")
            finally:
                self._queue.task_done()

    def _should_save(self) -> bool:
        count = self._save_calls
        self._save_calls = count + 1
        return self.save_every <= 1 or count % self.save_every == 0

    def save(self, state, system, *, trajectory: bool = False, trajectory_axis: int = 0, batch0: int = 0) -> None:
        """Copy the frame(s) to the host and queue their files
        (vtk_writer.py:123-169).  Periodic domains wrap positions into the
        box first (``domain.shift``)."""
        if not self._should_save():
            return
        if trajectory and trajectory_axis != 0:
            raise NotImplementedError("trajectories stacked along axis 0")
        frames = []
        for k, st, sy in _frames(state, system, trajectory):
            st, sy = sy.domain.shift(st, sy)
            frames.append((batch0 + k, *_frame_host(st, sy)))
        self._queue.put((self._write_frames, (frames,)))

    def _write_frames(self, frames: list) -> None:
        dirty = set()
        for b, st, sy in frames:
            batch = f"batch_{int(b):08d}"
            d = self.directory / batch
            d.mkdir(parents=True, exist_ok=True)
            for name in self.writers:
                cls = VTKBaseWriter._registry[_normalize_key(name)]
                if not cls.is_active(st, sy):
                    continue
                fn = d / f"{name}_{sy.step_count:08d}.vtp"
                try:
                    cls.write(st, sy, fn, self.binary)
                except Exception:
                    _log.exception("VTK writer %r failed for %s", name, fn)
                    continue
                with self._lock:
                    self._manifest.setdefault(batch, {}).setdefault(name, {})[sy.step_count] = sy.time
                dirty.add((batch, name))
        for batch, name in dirty:
            self._update_pvd(batch, name)

    def _update_pvd(self, batch: str, name: str) -> None:
        """vtk_writer.py:275-298."""
        with self._lock:
            data = self._manifest[batch][name]
            root = ET.Element("VTKFile", type="Collection", version="0.1")
            coll = ET.SubElement(root, "Collection")
            for step in sorted(data):
                ET.SubElement(coll, "DataSet", timestep=format(data[step], ".12g"),
                              file=(Path(batch) / f"{name}_{step:08d}.vtp").as_posix())
            ET.ElementTree(root).write(self.directory / f"{batch}_{name}.pvd", encoding="utf-8",
                                       xml_declaration=True)

    def flush(self) -> None:
        """Block until every queued frame is written."""
        self._queue.join()

    def close(self) -> None:
        """Write what is queued and stop the threads (async_base.py:147-173)."""
        if getattr(self, "_closed", True):
            return
        self._closed = True
        self._queue.join()
        for _ in self._threads:
            self._queue.put(None)
        for t in self._threads:
            t.join()


__all__ = ["VTKBaseWriter", "VTKSpheresWriter", "VTKDomainWriter", "VTKWriter", "write_polydata", "read_vtp"]
