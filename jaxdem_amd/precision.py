"""Precision form: energy minimisation of harmonic spheres with positions
held as an integer cell index plus a float32 offset inside the cell.

A float32 coordinate in a box of edge L resolves ~6e-8 L; near jamming FIRE
moves spheres by far less than that, so its updates would round away.  With
x = c h + o (c an integer cell, 0 <= o < h) the resolution is ~6e-8 h
everywhere, h being the cell edge (about a diameter).  Pair arithmetic stays
float32; the energy, the virial, the FIRE power and the residual are float64
sums on the GPU (``jdem_precise_minimize``).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .minimizers import FireConfig, fire

F32 = np.float32


def grid_for(box, rad) -> tuple:
    """Cells per axis (at least 3) whose edge is at least the largest
    diameter, and the edges."""
    box = np.asarray(box, np.float64)
    dmax = 2.0 * float(np.max(rad))
    dims = np.maximum(np.floor(box / dmax).astype(np.int64), 3)
    return dims, (box / dims)


def cell_edges(box, dims) -> tuple:
    """The cell edge box / dims as a float32 pair (hi, lo) and their sum in
    float64, the lattice constant both the GPU and the conversions use."""
    h = np.asarray(box, np.float64) / np.asarray(dims, np.int64)
    hi = h.astype(F32)
    lo = (h - hi.astype(np.float64)).astype(F32)
    return hi, lo, hi.astype(np.float64) + lo.astype(np.float64)


def to_cell_offset(pos, box, dims) -> tuple:
    """float64 positions -> (cell index int32, offset float32), the offset
    rounded once and kept in [0, h)."""
    pos = np.asarray(pos, np.float64)
    box = np.asarray(box, np.float64)
    dims = np.asarray(dims, np.int64)
    hi, _, h = cell_edges(box, dims)
    L = dims * h
    x = pos - L * np.floor(pos / L)
    c = np.clip(np.floor(x / h).astype(np.int64), 0, dims - 1)
    o = (x - c * h).astype(F32)
    over = o >= hi
    o = np.where(over, F32(0.0), np.maximum(o, F32(0.0))).astype(F32)
    c = np.where(over, (c + 1) % dims, c)
    return c.astype(np.int32), o


def from_cell_offset(cidx, off, box, dims) -> np.ndarray:
    """Positions in [0, L) as float64 from the cell index and offset."""
    _, _, h = cell_edges(box, dims)
    return np.asarray(cidx, np.float64) * h + np.asarray(off, np.float64)


def minimize_precise(pos, rad, box, k: float = 1.0, fire_cfg: FireConfig | None = None, dt: float = 1e-2,
                     max_steps: int = 10000, pe_tol: float = 1e-16, pe_diff_tol: float = 1e-16,
                     force_tol: float = 0.0, stall_steps: int = 0, skin: float = 0.0,
                     max_neighbors: int | None = None, device="cuda") -> dict:
    """FIRE (minimizers/routines.py:150-383) of harmonic spheres in a periodic
    box on the precision-form representation.  Returns float64 positions, the
    energy per particle, the largest force component, the contact virial and
    the pressure tr(virial)/(D V), the step count and the forces.

    The loop stops on the reference's three tests (pe_tol, pe_diff_tol,
    force_tol) or at ``max_steps``; with ``stall_steps`` > 0 it also stops
    once neither max |F| nor the energy (by more than 1e-13 of it) has
    reached a new low for that many steps.  With ``skin`` > 0 the pairs come
    from a Verlet list of reach r_i + r_j + skin, rebuilt (with the cell
    re-sort) once a sphere has moved skin / 2; ``max_neighbors`` defaults to
    a packing bound, ((3 r_max + skin) / r_min)^D spheres of radius >= r_min
    around a centre.  The float32
    pair arithmetic and the float32 offsets leave a residual floor near
    1e-7 k h, below which force_tol cannot be met."""
    pos = np.asarray(pos, np.float64)
    N, D = pos.shape
    rad = np.asarray(rad, F32)
    box = np.broadcast_to(np.asarray(box, np.float64), (D,)).copy()
    dims, _ = grid_for(box, rad)
    c, o = to_cell_offset(pos, box, dims)
    dev = torch.device(device)
    tc = torch.as_tensor(c, device=dev).contiguous()
    to = torch.as_tensor(o, device=dev).contiguous()
    tr = torch.as_tensor(rad, device=dev).contiguous()
    tf = torch.zeros((N, D), dtype=torch.float32, device=dev)
    p = L.JdemPrecise()
    p.n = N
    p.dim = D
    hi, lo, _ = cell_edges(box, dims)
    for a in range(3):
        p.dims[a] = int(dims[a]) if a < D else 1
        p.cell[a] = float(hi[a]) if a < D else 1.0
        p.cell_lo[a] = float(lo[a]) if a < D else 0.0
    p.cidx = tc.data_ptr()
    p.off = to.data_ptr()
    p.rad = tr.data_ptr()
    p.force = tf.data_ptr()
    p.k = float(F32(k))
    p.stall_steps = int(stall_steps)
    p.skin = float(F32(max(skin, 0.0)))
    if skin > 0 and max_neighbors is None and N > 0:
        rmin, rmax = float(rad.min()), float(rad.max())
        max_neighbors = int(np.ceil(((3 * rmax + skin) / rmin) ** D))
    p.max_neighbors = int(max_neighbors or 0) if skin > 0 else 0
    cfg = fire_cfg or fire(dt)
    fs = cfg.struct()
    stats = L.JdemMinimizeStats()
    vir = (ctypes.c_double * 9)()
    ctx = L.context_for(dev)
    ctx.check(
        ctx.lib.jdem_precise_minimize(ctx.ptr, ctypes.byref(p), ctypes.byref(fs), int(max_steps), float(pe_tol),
                                      float(pe_diff_tol), float(force_tol), ctypes.byref(stats), vir),
        "jdem_precise_minimize",
    )
    W = np.array(vir[: D * D], np.float64).reshape(D, D)
    V = float(np.prod(box))
    return {
        "pos": from_cell_offset(tc.cpu().numpy(), to.cpu().numpy(), box, dims),
        "cidx": tc,
        "off": to,
        "force": tf,
        "steps": int(stats.steps),
        "pe": float(stats.pe),
        "max_grad": float(stats.max_grad),
        "virial": W,
        "pressure": float(np.trace(W) / (D * V)),
        "dims": dims,
    }


def min_image_diff(a, b, box) -> np.ndarray:
    """Per-coordinate minimum-image difference of two position arrays."""
    box = np.asarray(box, np.float64)
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return d - box * np.rint(d / box)


__all__ = ["grid_for", "cell_edges", "to_cell_offset", "from_cell_offset", "minimize_precise", "min_image_diff"]
