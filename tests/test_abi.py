"""The C ABI (include/jdem.h) of libjdem_hip.so, checked on the CPU:

* the library loads with every symbol resolved (RTLD_NOW),
* it exports every function the header declares,
* the ctypes mirrors of the header's structs have the C sizes and offsets
  (the header compiled with gcc prints them),
* context creation without a GPU fails with an error code, not a crash.
"""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from jaxdem_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "jdem.h"


def header_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(jdem_\w+)\s*\(", text, flags=re.M)))


def test_library_built_and_loads_now():
    assert L.LIB_PATH.exists(), "run `python -m jaxdem_amd.build` first"
    ctypes.CDLL(str(L.LIB_PATH), mode=ctypes.RTLD_LOCAL | 2)  # RTLD_NOW: every symbol resolved


def test_exports_every_declared_function():
    lib = ctypes.CDLL(str(L.LIB_PATH))
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_version_and_target():
    lib = L.load()
    assert lib.jdem_abi_version() == 1
    assert lib.jdem_build_target() == b"gfx950"


STRUCT_FIELDS = {
    "jdem_state": ["n", "dim", "bond_width", "pos", "vel", "ang_vel", "q", "force", "torque", "rad", "mass",
                   "inertia", "mat_id", "bond_id", "fixed", "pos_p", "clump_id"],
    "jdem_domain": ["kind", "dim", "box_size", "inv_box_size", "anchor", "restitution"],
    "jdem_law": ["kind", "n_mat", "t0", "t1", "t2", "m0", "m1", "m2", "m3", "p0", "p1", "n_species", "n_terms",
                 "route_off", "route_terms", "species"],
    "jdem_grid": ["dims", "strides", "cell_size", "search_range", "periodic", "n_cells"],
    "jdem_nlist": ["max_neighbors", "cutoff", "skin", "nl", "old_pos", "n_build_times", "overflow", "secondary",
                   "grid", "hist_xi", "hist_live", "resident_in", "resident_out"],
    "jdem_system": ["state", "domain", "law", "lin_kind", "rot_kind", "dt", "interact_same_bond_id", "gravity",
                    "ext_force", "ext_force_com", "ext_torque", "collider_kind", "grid", "nlist"],
    "jdem_step_stats": ["steps", "rebuilds", "host_syncs", "overflow", "reserved", "free_box", "free_anchor"],
    "jdem_timing": ["step_kernel_ms", "step_kernels", "idle_kernel_ms", "idle_kernels", "rebuild_ms", "rebuilds"],
    "jdem_batch": ["n_systems", "launch_mode", "n_build_times", "offsets", "domains", "grids", "dt", "cutoff", "skin",
                   "tile_width", "spec_min_gap"],
    "jdem_fire": ["dt", "alpha_init", "f_inc", "f_dec", "f_alpha", "n_min", "n_bad_max", "dt_max_scale",
                  "dt_min_scale"],
    "jdem_minimize_stats": ["steps", "pe", "max_grad", "power", "overflow", "reserved"],
    "jdem_precise": ["n", "dim", "dims", "cell", "cell_lo", "cidx", "off", "rad", "force", "k", "stall_steps",
                     "skin", "max_neighbors"],
    "jdem_frames": ["n_frames", "steps", "pos", "vel", "ang_vel", "q", "force", "torque"],
    "jdem_state_f64": ["n", "dim", "bond_width", "pos", "vel", "ang_vel", "q", "force", "torque", "rad", "mass",
                       "inertia", "mat_id", "bond_id", "fixed", "pos_p", "clump_id"],
    "jdem_domain_f64": ["kind", "dim", "box_size", "inv_box_size", "anchor", "restitution"],
    "jdem_law_f64": ["kind", "n_mat", "t0", "t1", "t2", "m0", "m1", "m2", "m3", "p0", "p1", "n_species", "n_terms",
                     "route_off", "route_terms", "species"],
    "jdem_grid_f64": ["dims", "strides", "cell_size", "search_range", "periodic", "n_cells"],
    "jdem_nlist_f64": ["max_neighbors", "cutoff", "skin", "nl", "old_pos", "n_build_times", "overflow", "secondary",
                       "grid"],
    "jdem_system_f64": ["state", "domain", "law", "lin_kind", "rot_kind", "dt", "interact_same_bond_id", "gravity",
                        "ext_force", "ext_force_com", "ext_torque", "collider_kind", "grid", "nlist"],
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for s, fields in STRUCT_FIELDS.items():
        src.append(f'  printf("{s} size %zu\\n", sizeof({s}));')
        for f in fields:
            src.append(f'  printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    src.append("  return 0;\n}")
    d = tmp_path_factory.mktemp("abi")
    c = d / "layout.c"
    c.write_text("\n".join(src))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c99", str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        k, v = line.rsplit(" ", 1)
        res[k] = int(v)
    return res


@pytest.mark.parametrize("name", sorted(STRUCT_FIELDS))
def test_struct_layout_matches_header(c_layout, name):
    cls = L.STRUCTS[name]
    assert ctypes.sizeof(cls) == c_layout[f"{name} size"], name
    for f in STRUCT_FIELDS[name]:
        assert getattr(cls, f).offset == c_layout[f"{name}.{f}"], (name, f)
    assert [f for f, _ in cls._fields_] == STRUCT_FIELDS[name]


def test_ctx_create_without_gpu_is_an_error_not_a_crash():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = L.load()
    p = ctypes.c_void_p()
    rc = lib.jdem_ctx_create(0, None, ctypes.byref(p))
    assert rc != 0 and not p.value


def test_null_arguments_are_rejected():
    lib = L.load()
    assert lib.jdem_ctx_create(0, None, None) == -1
    assert lib.jdem_ctx_destroy(None) == 0
    assert lib.jdem_system_step(None, None, 1, None) == -1
    assert lib.jdem_compute_force(None, None) == -1
    assert lib.jdem_ctx_set_timing(None, 1) == -1
