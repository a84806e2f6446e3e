"""Energy minimisation with FIRE on the GPU (jaxdem/minimizers/).

``fire(dt, ...)`` has the signature of the reference optimiser factory
(optimizers.py:150-160) and returns its configuration; ``minimize`` runs the
loop of routines.py:150-383 in ``libjdem_hip.so`` (``jdem_minimize``): one
force/energy evaluation per iteration, the FIRE update on the device, and the
stopping rules max_steps / pe_tol / pe_diff_tol / force_tol.  The potential
energy and the FIRE power are float64 sums on the GPU.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace

import numpy as np

from . import _lib as L
from .state import State

F32 = np.float32


@dataclass(frozen=True)
class FireConfig:
    """Arguments of ``fire`` (optimizers.py:150-160)."""

    dt: float
    alpha_init: float = 0.1
    f_inc: float = 1.1
    f_dec: float = 0.5
    f_alpha: float = 0.99
    N_min: int = 5
    N_bad_max: int = 10
    dt_max_scale: float = 10.0
    dt_min_scale: float = 1e-3
    type_name: str = "fire"

    def struct(self) -> L.JdemFire:
        f = L.JdemFire()
        f.dt = float(self.dt)
        f.alpha_init = float(self.alpha_init)
        f.f_inc = float(self.f_inc)
        f.f_dec = float(self.f_dec)
        f.f_alpha = float(self.f_alpha)
        f.n_min = int(self.N_min)
        f.n_bad_max = int(self.N_bad_max)
        f.dt_max_scale = float(self.dt_max_scale)
        f.dt_min_scale = float(self.dt_min_scale)
        return f


def fire(
    dt: float,
    alpha_init: float = 0.1,
    f_inc: float = 1.1,
    f_dec: float = 0.5,
    f_alpha: float = 0.99,
    N_min: int = 5,
    N_bad_max: int = 10,
    dt_max_scale: float = 10.0,
    dt_min_scale: float = 1e-3,
) -> FireConfig:
    """Fast Inertial Relaxation Engine (Bitzek et al., PRL 97, 170201 (2006))
    with the reference's uphill-counter reset (optimizers.py:150-339)."""
    return FireConfig(dt, alpha_init, f_inc, f_dec, f_alpha, N_min, N_bad_max, dt_max_scale, dt_min_scale)


def _resolve(system) -> FireConfig:
    """``System.create`` stores ``(minimizer, minimizer_kw)``; a missing
    minimizer is ``fire`` and its ``dt`` defaults to the system's
    (system.py:517-531)."""
    fn, kw = system.minimizer if system.minimizer is not None else (None, {})
    kw = dict(kw or {})
    if isinstance(fn, FireConfig):
        return fn
    if fn is None or fn is fire:
        kw.setdefault("dt", float(system.dt))
        return fire(**kw)
    raise NotImplementedError(
        f"minimizer {getattr(fn, '__name__', fn)!r}: only FIRE runs on the GPU path of this package"
    )


def minimize(
    state: State,
    system,
    max_steps: int = 10000,
    pe_tol: float = 1e-16,
    pe_diff_tol: float = 1e-16,
    force_tol: float = 0.0,
) -> tuple:
    """``minimize`` (routines.py:150-383).  Returns (state, system, steps,
    potential energy per particle); the inputs are left untouched."""
    from .system import _collider_after, _ctx, _fresh_collider, _require_gpu, _system_struct

    _require_gpu(state)
    if not state.is_sphere_state:
        raise NotImplementedError("minimize on the GPU path handles spheres only")
    fm = system.force_manager
    if fm.has_functions:
        raise NotImplementedError("minimize with Python force functions is not on the GPU path")
    if any(getattr(fm, k, None) is not None for k in ("external_force", "external_force_com", "external_torque")):
        raise NotImplementedError("minimize with pending external forces is not on the GPU path")
    cfg = _resolve(system)
    new_state = state.with_kinematics()
    col = _fresh_collider(system.collider)
    sys2 = replace(system, collider=col)
    keep: list = []
    S = _system_struct(new_state, sys2, keep)
    F = cfg.struct()
    out = L.JdemMinimizeStats()
    ctx = _ctx(new_state)
    ctx.check(
        ctx.lib.jdem_minimize(
            ctx.ptr, ctypes.byref(S), ctypes.byref(F), int(max_steps), float(pe_tol), float(pe_diff_tol),
            float(force_tol), ctypes.byref(out),
        ),
        "jdem_minimize",
    )
    out_sys = replace(sys2, collider=_collider_after(col, S))
    out_sys.__dict__["_last_minimize"] = out
    return new_state, out_sys, int(out.steps), float(out.pe)


__all__ = ["FireConfig", "fire", "minimize"]
