"""ctypes driver of oracle/libjdem_port.so -- TEST INFRASTRUCTURE ONLY.

Runs the C++/OpenMP port of the oracle on host numpy arrays; used by the CPU
tests (checked against the numpy oracle) and as bench.py's CPU baseline.
"""

from __future__ import annotations

import ctypes
import time
from ctypes import c_float, c_int, c_int64, c_void_p
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().with_name("libjdem_port.so")


class PortParams(ctypes.Structure):
    _fields_ = [
        ("dt", c_float),
        ("box", c_float * 3),
        ("inv_box", c_float * 3),
        ("anchor", c_float * 3),
        ("periodic", c_int),
        ("reflect", c_int),
        ("restitution", c_float),
        ("law", c_int),
        ("kn", c_float),
        ("kt", c_float),
        ("mu", c_float),
        ("damp2", c_float),
        ("cutoff", c_float),
        ("skin", c_float),
        ("M", c_int),
        ("cell_size", c_float),
        ("sr", c_int),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build_port

            build_port.build()
        lib = ctypes.CDLL(str(LIB))
        lib.jdem_port_steps.restype = c_int
        lib.jdem_port_steps.argtypes = [ctypes.POINTER(PortParams), c_int64] + [c_void_p] * 13 + [ctypes.POINTER(ctypes.c_int32), c_int64, c_int]
        _lib = lib
    return _lib


class PortRun:
    """Host arrays of one system advanced by the port."""

    def __init__(self, pos, vel, rad, mass, box, anchor, *, domain, dt, law, kn, kt=0.0, mu=0.0, damping=0.0,
                 cutoff, skin, M, cell_size, sr=1, restitution=1.0, inertia=None, ang_vel=None, q=None):
        f = np.float32
        N = pos.shape[0]
        self.N = N
        self.pos = np.ascontiguousarray(pos, f).copy()
        self.vel = np.ascontiguousarray(vel, f).copy()
        self.rad = np.ascontiguousarray(rad, f)
        self.mass = np.ascontiguousarray(mass, f)
        if inertia is None:
            inertia = ((f(0.4) * self.mass * self.rad**2)[:, None] * np.ones((N, 3), f)).astype(f)
        self.inertia = np.ascontiguousarray(inertia, f)
        self.angv = np.zeros((N, 3), f) if ang_vel is None else np.ascontiguousarray(ang_vel, f).copy()
        if q is None:
            q = np.zeros((N, 4), f)
            q[:, 0] = 1
        self.q = np.ascontiguousarray(q, f).copy()
        self.force = np.zeros((N, 3), f)
        self.torque = np.zeros((N, 3), f)
        self.nl = np.full((N, M), -1, np.int32)
        self.old_pos = self.pos.copy()
        self.xi = np.zeros((N, M, 3), f)
        self.live = np.zeros((N, M), np.uint8)
        self.n_build = ctypes.c_int32(0)
        p = PortParams()
        p.dt = float(f(dt))
        box = np.asarray(box, f)
        for a in range(3):
            p.box[a] = float(box[a])
            p.inv_box[a] = float(f(1.0) / box[a])
            p.anchor[a] = float(np.asarray(anchor, f)[a])
        p.periodic = 1 if domain == "periodic" else 0
        p.reflect = 1 if domain == "reflect" else 0
        p.restitution = float(f(restitution))
        p.law = {"spring": 1, "springfriction": 4}[law]
        p.kn = float(f(kn))
        p.kt = float(f(kt))
        p.mu = float(f(mu))
        p.damp2 = float(f(2.0 * damping))
        p.cutoff = float(f(cutoff))
        p.skin = float(f(skin))
        p.M = int(M)
        p.cell_size = float(f(cell_size))
        p.sr = int(sr)
        self.p = p

    def steps(self, n: int, threads: int = 0) -> float:
        lib = load()
        ptr = lambda a: a.ctypes.data_as(c_void_p)  # noqa: E731
        t0 = time.perf_counter()
        rc = lib.jdem_port_steps(
            ctypes.byref(self.p), self.N, ptr(self.pos), ptr(self.vel), ptr(self.angv), ptr(self.q), ptr(self.force),
            ptr(self.torque), ptr(self.rad), ptr(self.mass), ptr(self.inertia), ptr(self.nl), ptr(self.old_pos),
            ptr(self.xi), ptr(self.live), ctypes.byref(self.n_build), int(n), int(threads),
        )
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"jdem_port_steps returned {rc}")
        return dt


class Fire64Params(ctypes.Structure):
    _fields_ = [
        ("box", ctypes.c_double),
        ("k", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("alpha_init", ctypes.c_double),
        ("f_inc", ctypes.c_double),
        ("f_dec", ctypes.c_double),
        ("f_alpha", ctypes.c_double),
        ("dt_max_scale", ctypes.c_double),
        ("dt_min_scale", ctypes.c_double),
        ("n_min", c_int),
        ("n_bad_max", c_int),
        ("max_steps", c_int64),
        ("pe_tol", ctypes.c_double),
        ("pe_diff_tol", ctypes.c_double),
        ("force_tol", ctypes.c_double),
    ]


class Fire64Result(ctypes.Structure):
    _fields_ = [
        ("steps", c_int64),
        ("pe", ctypes.c_double),
        ("max_grad", ctypes.c_double),
        ("virial", ctypes.c_double * 9),
    ]


def fire64(pos, rad, box: float, k: float = 1.0, dt: float = 1e-2, max_steps: int = 10000, pe_tol: float = 1e-16,
           pe_diff_tol: float = 1e-16, force_tol: float = 0.0, threads: int = 0, **fire_kw) -> dict:
    """float64 FIRE of harmonic spheres in a periodic cube (jdem_fire64.cpp)."""
    lib = ctypes.CDLL(str(LIB))
    lib.fire64_minimize.restype = c_int
    lib.fire64_minimize.argtypes = [ctypes.POINTER(Fire64Params), c_int64, c_void_p, c_void_p,
                                    ctypes.POINTER(Fire64Result), c_int]
    p = Fire64Params()
    p.box, p.k, p.dt = float(box), float(k), float(dt)
    p.alpha_init = fire_kw.get("alpha_init", 0.1)
    p.f_inc = fire_kw.get("f_inc", 1.1)
    p.f_dec = fire_kw.get("f_dec", 0.5)
    p.f_alpha = fire_kw.get("f_alpha", 0.99)
    p.dt_max_scale = fire_kw.get("dt_max_scale", 10.0)
    p.dt_min_scale = fire_kw.get("dt_min_scale", 1e-3)
    p.n_min = fire_kw.get("N_min", 5)
    p.n_bad_max = fire_kw.get("N_bad_max", 10)
    p.max_steps, p.pe_tol, p.pe_diff_tol, p.force_tol = int(max_steps), pe_tol, pe_diff_tol, force_tol
    x = np.ascontiguousarray(pos, np.float64).copy()
    r = np.ascontiguousarray(rad, np.float64)
    out = Fire64Result()
    t = time.perf_counter()
    lib.fire64_minimize(ctypes.byref(p), x.shape[0], x.ctypes.data, r.ctypes.data, ctypes.byref(out), int(threads))
    return {"pos": x, "steps": int(out.steps), "pe": float(out.pe), "max_grad": float(out.max_grad),
            "virial": np.array(out.virial[:]).reshape(3, 3), "seconds": time.perf_counter() - t}
