"""ctypes binding of libjdem_hip.so (declared in include/jdem.h).

The structures below mirror the C structs field by field; a CPU test checks
their sizes and offsets against the header compiled with gcc.  Loading the
library never falls back to anything: if it is missing, every GPU entry
point raises.
"""

from __future__ import annotations

import ctypes
import sys
import threading
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libjdem_hip.so")

JDEM_OK = 0
ERRORS = {-1: "JDEM_EINVAL", -2: "JDEM_EHIP", -3: "JDEM_ENOMEM", -4: "JDEM_EUNSUPPORTED"}

DOMAIN_FREE, DOMAIN_PERIODIC, DOMAIN_REFLECT = 0, 1, 2
LAW_NONE, LAW_SPRING, LAW_HERTZ, LAW_CUNDALL_STRACK, LAW_SPRING_FRICTION, LAW_HERTZ_FRICTION, LAW_ROUTED = range(7)
LIN_NONE, LIN_VERLET, LIN_EULER = 0, 1, 2
ROT_NONE, ROT_VERLET_SPIRAL, ROT_SPIRAL = 0, 1, 2
COLLIDER_NONE, COLLIDER_NAIVE, COLLIDER_CELL_LIST, COLLIDER_NEIGHBOR_LIST = 0, 1, 2, 3
(
    STAGE_DOMAIN_APPLY,
    STAGE_LIN_BEFORE,
    STAGE_ROT_BEFORE,
    STAGE_FORCE_MANAGER,
    STAGE_LIN_AFTER,
    STAGE_ROT_AFTER,
) = range(1, 7)


class JdemState(ctypes.Structure):
    _fields_ = [
        ("n", c_int64),
        ("dim", c_int32),
        ("bond_width", c_int32),
        ("pos", c_void_p),
        ("vel", c_void_p),
        ("ang_vel", c_void_p),
        ("q", c_void_p),
        ("force", c_void_p),
        ("torque", c_void_p),
        ("rad", c_void_p),
        ("mass", c_void_p),
        ("inertia", c_void_p),
        ("mat_id", c_void_p),
        ("bond_id", c_void_p),
        ("fixed", c_void_p),
        ("pos_p", c_void_p),
        ("clump_id", c_void_p),
    ]


class JdemDomain(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("dim", c_int32),
        ("box_size", c_float * 3),
        ("inv_box_size", c_float * 3),
        ("anchor", c_float * 3),
        ("restitution", c_float),
    ]


class JdemLaw(ctypes.Structure):
    pass


JdemLaw._fields_ = [
    ("kind", c_int32),
    ("n_mat", c_int32),
    ("t0", c_void_p),
    ("t1", c_void_p),
    ("t2", c_void_p),
    ("m0", c_void_p),
    ("m1", c_void_p),
    ("m2", c_void_p),
    ("m3", c_void_p),
    ("p0", c_float),
    ("p1", c_float),
    ("n_species", c_int32),
    ("n_terms", c_int32),
    ("route_off", c_void_p),
    ("route_terms", ctypes.POINTER(JdemLaw)),
    ("species", c_void_p),
]


class JdemGrid(ctypes.Structure):
    _fields_ = [
        ("dims", c_int32 * 3),
        ("strides", c_int32 * 3),
        ("cell_size", c_float * 3),
        ("search_range", c_int32),
        ("periodic", c_int32),
        ("n_cells", c_int64),
    ]


class JdemNlist(ctypes.Structure):
    _fields_ = [
        ("max_neighbors", c_int32),
        ("cutoff", c_float),
        ("skin", c_float),
        ("nl", c_void_p),
        ("old_pos", c_void_p),
        ("n_build_times", c_int32),
        ("overflow", c_int32),
        ("secondary", c_int32),
        ("grid", JdemGrid),
        ("hist_xi", c_void_p),
        ("hist_live", c_void_p),
        ("resident_in", c_void_p),
        ("resident_out", c_void_p),
    ]


class JdemSystem(ctypes.Structure):
    _fields_ = [
        ("state", JdemState),
        ("domain", JdemDomain),
        ("law", JdemLaw),
        ("lin_kind", c_int32),
        ("rot_kind", c_int32),
        ("dt", c_float),
        ("interact_same_bond_id", c_int32),
        ("gravity", c_float * 3),
        ("ext_force", c_void_p),
        ("ext_force_com", c_void_p),
        ("ext_torque", c_void_p),
        ("collider_kind", c_int32),
        ("grid", JdemGrid),
        ("nlist", JdemNlist),
    ]


class JdemStepStats(ctypes.Structure):
    _fields_ = [
        ("steps", c_int64),
        ("rebuilds", c_int64),
        ("host_syncs", c_int64),
        ("overflow", c_int32),
        ("reserved", c_int32),
        ("free_box", c_double * 3),
        ("free_anchor", c_double * 3),
    ]


class JdemFrames(ctypes.Structure):
    _fields_ = [
        ("n_frames", c_int64),
        ("steps", POINTER(c_int64)),
        ("pos", c_void_p),
        ("vel", c_void_p),
        ("ang_vel", c_void_p),
        ("q", c_void_p),
        ("force", c_void_p),
        ("torque", c_void_p),
    ]


class JdemTiming(ctypes.Structure):
    _fields_ = [
        ("step_kernel_ms", c_double),
        ("step_kernels", c_int64),
        ("idle_kernel_ms", c_double),
        ("idle_kernels", c_int64),
        ("rebuild_ms", c_double),
        ("rebuilds", c_int64),
    ]


class JdemBatch(ctypes.Structure):
    _fields_ = [
        ("n_systems", c_int32),
        ("launch_mode", c_int32),
        ("n_build_times", c_void_p),
        ("offsets", c_void_p),
        ("domains", c_void_p),
        ("grids", c_void_p),
        ("dt", c_void_p),
        ("cutoff", c_void_p),
        ("skin", c_void_p),
        ("tile_width", c_int32),
        ("spec_min_gap", c_int32),
    ]


class JdemFire(ctypes.Structure):
    _fields_ = [
        ("dt", c_double),
        ("alpha_init", c_double),
        ("f_inc", c_double),
        ("f_dec", c_double),
        ("f_alpha", c_double),
        ("n_min", c_int32),
        ("n_bad_max", c_int32),
        ("dt_max_scale", c_double),
        ("dt_min_scale", c_double),
    ]


class JdemMinimizeStats(ctypes.Structure):
    _fields_ = [
        ("steps", c_int64),
        ("pe", c_double),
        ("max_grad", c_double),
        ("power", c_double),
        ("overflow", c_int32),
        ("reserved", c_int32),
    ]


class JdemPrecise(ctypes.Structure):
    _fields_ = [
        ("n", c_int64),
        ("dim", c_int32),
        ("dims", c_int32 * 3),
        ("cell", c_float * 3),
        ("cell_lo", c_float * 3),
        ("cidx", c_void_p),
        ("off", c_void_p),
        ("rad", c_void_p),
        ("force", c_void_p),
        ("k", c_float),
        ("stall_steps", c_int32),
        ("skin", c_float),
        ("max_neighbors", c_int32),
    ]


# ---- the plain step in float64 (jdem_*_f64) ---------------------------------
class JdemStateF64(ctypes.Structure):
    _fields_ = JdemState._fields_


class JdemDomainF64(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("dim", c_int32),
        ("box_size", c_double * 3),
        ("inv_box_size", c_double * 3),
        ("anchor", c_double * 3),
        ("restitution", c_double),
    ]


class JdemLawF64(ctypes.Structure):
    pass


JdemLawF64._fields_ = [
    ("kind", c_int32),
    ("n_mat", c_int32),
    ("t0", c_void_p),
    ("t1", c_void_p),
    ("t2", c_void_p),
    ("m0", c_void_p),
    ("m1", c_void_p),
    ("m2", c_void_p),
    ("m3", c_void_p),
    ("p0", c_double),
    ("p1", c_double),
    ("n_species", c_int32),
    ("n_terms", c_int32),
    ("route_off", c_void_p),
    ("route_terms", ctypes.POINTER(JdemLawF64)),
    ("species", c_void_p),
]


class JdemGridF64(ctypes.Structure):
    _fields_ = [
        ("dims", c_int32 * 3),
        ("strides", c_int32 * 3),
        ("cell_size", c_double * 3),
        ("search_range", c_int32),
        ("periodic", c_int32),
        ("n_cells", c_int64),
    ]


class JdemNlistF64(ctypes.Structure):
    _fields_ = [
        ("max_neighbors", c_int32),
        ("cutoff", c_double),
        ("skin", c_double),
        ("nl", c_void_p),
        ("old_pos", c_void_p),
        ("n_build_times", c_int32),
        ("overflow", c_int32),
        ("secondary", c_int32),
        ("grid", JdemGridF64),
    ]


class JdemSystemF64(ctypes.Structure):
    _fields_ = [
        ("state", JdemStateF64),
        ("domain", JdemDomainF64),
        ("law", JdemLawF64),
        ("lin_kind", c_int32),
        ("rot_kind", c_int32),
        ("dt", c_double),
        ("interact_same_bond_id", c_int32),
        ("gravity", c_double * 3),
        ("ext_force", c_void_p),
        ("ext_force_com", c_void_p),
        ("ext_torque", c_void_p),
        ("collider_kind", c_int32),
        ("grid", JdemGridF64),
        ("nlist", JdemNlistF64),
    ]


STRUCTS = {
    "jdem_state_f64": JdemStateF64,
    "jdem_domain_f64": JdemDomainF64,
    "jdem_law_f64": JdemLawF64,
    "jdem_grid_f64": JdemGridF64,
    "jdem_nlist_f64": JdemNlistF64,
    "jdem_system_f64": JdemSystemF64,
    "jdem_timing": JdemTiming,
    "jdem_state": JdemState,
    "jdem_domain": JdemDomain,
    "jdem_law": JdemLaw,
    "jdem_grid": JdemGrid,
    "jdem_nlist": JdemNlist,
    "jdem_system": JdemSystem,
    "jdem_step_stats": JdemStepStats,
    "jdem_batch": JdemBatch,
    "jdem_fire": JdemFire,
    "jdem_minimize_stats": JdemMinimizeStats,
    "jdem_precise": JdemPrecise,
    "jdem_frames": JdemFrames,
}

# name -> (restype, argtypes) for every entry point of include/jdem.h
SIGNATURES = {
    "jdem_abi_version": (c_int, []),
    "jdem_build_target": (ctypes.c_char_p, []),
    "jdem_ctx_create": (c_int, [c_int, c_void_p, POINTER(c_void_p)]),
    "jdem_ctx_destroy": (c_int, [c_void_p]),
    "jdem_ctx_set_stream": (c_int, [c_void_p, c_void_p]),
    "jdem_last_error": (ctypes.c_char_p, [c_void_p]),
    "jdem_ctx_scratch_bytes": (c_int, [c_void_p, POINTER(c_size_t)]),
    "jdem_ctx_set_timing": (c_int, [c_void_p, c_int]),
    "jdem_ctx_get_timing": (c_int, [c_void_p, POINTER(JdemTiming)]),
    "jdem_cell_build": (
        c_int,
        [c_void_p, POINTER(JdemState), POINTER(JdemDomain), POINTER(JdemGrid), c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "jdem_cell_build_noalloc": (
        c_int,
        [c_void_p, POINTER(JdemState), POINTER(JdemDomain), POINTER(JdemGrid), c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "jdem_noalloc_scratch_bytes": (c_size_t, []),
    "jdem_neighbor_list": (
        c_int,
        [c_void_p, POINTER(JdemSystem), c_int32, POINTER(JdemGrid), c_float, c_int32, c_void_p, POINTER(c_int32)],
    ),
    "jdem_compute_force": (c_int, [c_void_p, POINTER(JdemSystem)]),
    "jdem_potential_energy": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(c_double)]),
    "jdem_contact_virial": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(c_double)]),
    "jdem_kinetic_energy": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(c_double)]),
    "jdem_kinetic_energy_parts": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(c_double), POINTER(c_double)]),
    "jdem_step_stage": (c_int, [c_void_p, POINTER(JdemSystem), c_int32]),
    "jdem_system_step": (c_int, [c_void_p, POINTER(JdemSystem), c_int64, POINTER(JdemStepStats)]),
    "jdem_rollout": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(JdemFrames), POINTER(JdemStepStats)]),
    "jdem_batch_step": (c_int, [c_void_p, POINTER(JdemSystem), POINTER(JdemBatch), c_int64, POINTER(JdemStepStats)]),
    "jdem_precise_minimize": (
        c_int,
        [c_void_p, POINTER(JdemPrecise), POINTER(JdemFire), c_int64, c_double, c_double, c_double,
         POINTER(JdemMinimizeStats), c_void_p],
    ),
    "jdem_resident_release": (c_int, [c_void_p, c_void_p]),
    "jdem_resident_export": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jdem_ctx_trim": (c_int, [c_void_p]),
    "jdem_minimize": (
        c_int,
        [c_void_p, POINTER(JdemSystem), POINTER(JdemFire), c_int64, c_double, c_double, c_double,
         POINTER(JdemMinimizeStats)],
    ),
    "jdem_system_step_f64": (c_int, [c_void_p, POINTER(JdemSystemF64), c_int64, POINTER(JdemStepStats)]),
    "jdem_compute_force_f64": (c_int, [c_void_p, POINTER(JdemSystemF64)]),
    "jdem_potential_energy_f64": (c_int, [c_void_p, POINTER(JdemSystemF64), POINTER(c_double)]),
    "jdem_kinetic_energy_parts_f64": (
        c_int, [c_void_p, POINTER(JdemSystemF64), POINTER(c_double), POINTER(c_double)]),
}

_lib = None
_lock = threading.Lock()


class JdemError(RuntimeError):
    """A libjdem_hip.so call returned an error status."""


def load() -> ctypes.CDLL:
    """Load libjdem_hip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise JdemError(
                f"{LIB_PATH.name} is not built (expected at {LIB_PATH}); build it with "
                "`python -m jaxdem_amd.build` -- the DEM step has no CPU fallback"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


class Context:
    """One jdem_ctx per (device, stream)."""

    def __init__(self, device: int, stream: int):
        self.lib = load()
        self.ptr = c_void_p()
        rc = self.lib.jdem_ctx_create(int(device), c_void_p(stream), ctypes.byref(self.ptr))
        if rc != JDEM_OK:
            raise JdemError(f"jdem_ctx_create(device={device}) failed: {ERRORS.get(rc, rc)}")
        self.device = device
        self.stream = stream

    def check(self, rc: int, what: str) -> None:
        if rc != JDEM_OK:
            msg = self.lib.jdem_last_error(self.ptr)
            raise JdemError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_stream(self, stream: int) -> None:
        if stream != self.stream:
            self.check(self.lib.jdem_ctx_set_stream(self.ptr, c_void_p(stream)), "jdem_ctx_set_stream")
            self.stream = stream

    def set_timing(self, on: bool) -> None:
        self.check(self.lib.jdem_ctx_set_timing(self.ptr, 1 if on else 0), "jdem_ctx_set_timing")

    def get_timing(self) -> dict:
        t = JdemTiming()
        self.check(self.lib.jdem_ctx_get_timing(self.ptr, ctypes.byref(t)), "jdem_ctx_get_timing")
        return {name: getattr(t, name) for name, _ in JdemTiming._fields_}

    def scratch_bytes(self) -> int:
        n = c_size_t()
        self.check(self.lib.jdem_ctx_scratch_bytes(self.ptr, ctypes.byref(n)), "jdem_ctx_scratch_bytes")
        return int(n.value)

    def trim(self) -> None:
        """Free the device buffers pooled by the context (released working
        sets, grown scratch)."""
        self.check(self.lib.jdem_ctx_trim(self.ptr), "jdem_ctx_trim")

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            if self.ptr:
                self.lib.jdem_ctx_destroy(self.ptr)
        except Exception:
            pass


class Resident:
    """A NeighborList working set kept on the device between steps
    (jdem_resident): immutable once made; released to the context's pool when
    the last reference goes."""

    def __init__(self, ctx: Context, ptr: int, n: int, M: int, dim: int, hist: bool, device):
        self.ctx = ctx  # keeps the context alive while the working set is
        self.ptr = c_void_p(ptr)
        self.n, self.M, self.dim, self.hist = int(n), int(M), int(dim), bool(hist)
        self.device = device

    def export(self) -> tuple:
        """(neighbor_list (N, M) int32, old_pos (N, D), history dict or None)
        in the reference's layouts (neighbor_list.py:137-286)."""
        import torch

        N, M, D = self.n, self.M, self.dim
        nl = torch.empty((N, M), dtype=torch.int32, device=self.device)
        old = torch.empty((N, D), dtype=torch.float32, device=self.device)
        hist = None
        if self.hist:
            hist = {
                "xi": torch.zeros((N, M, D), dtype=torch.float32, device=self.device),
                "live": torch.zeros((N, max((M + 31) // 32, 1)), dtype=torch.int32, device=self.device),
            }
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        self.ctx.check(
            self.ctx.lib.jdem_resident_export(
                self.ctx.ptr, self.ptr, nl.data_ptr(), old.data_ptr(),
                hist["xi"].data_ptr() if hist else None, hist["live"].data_ptr() if hist else None,
            ),
            "jdem_resident_export",
        )
        return nl, old, hist

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            if self.ptr and not sys.is_finalizing():
                self.ctx.lib.jdem_resident_release(self.ctx.ptr, self.ptr)
                self.ptr = c_void_p()
        except Exception:
            pass


_contexts: dict = {}


def context_for(device) -> Context:
    """Context for a torch cuda device on torch's current stream."""
    import torch

    idx = device.index if device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(idx).cuda_stream
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = Context(idx, stream)
        _contexts[idx] = ctx
    else:
        ctx.set_stream(stream)
    return ctx
