"""jdem_cell_build_noalloc (the cell build with its scratch in one fixed
device array) against jdem_cell_build, through the C ABI.

Two calls in a row, the second on a smaller input, on a counting-sort grid
and on a grid much larger than N (the radix sort); every output must equal
jdem_cell_build's bit for bit.  A size whose scratch does not fit the fixed
array is refused with an error before any GPU work.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from jaxdem_amd import _lib as L
from jaxdem_amd.colliders import grid_struct

pytestmark = pytest.mark.gpu


def _state(n, box, seed, device):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    st = jd.State.create(pos=torch.as_tensor(pos, device=device), rad=torch.full((n,), 0.5, device=device))
    return st


def _build(entry, st, box, cell, device):
    g, ovf = grid_struct(np.full(3, box, np.float32), np.float32(cell), True, 1)
    assert not ovf
    n, ncell = st.N, int(g.n_cells)
    outs = [torch.full((max(n, 1),), -7, dtype=torch.int32, device=device) for _ in range(2)]
    outs += [torch.full((ncell,), -7, dtype=torch.int32, device=device) for _ in range(2)]
    sy = jd.System.create(st.shape, domain_type="periodic",
                          domain_kw=dict(box_size=np.full(3, box, np.float32), anchor=np.zeros(3, np.float32)))
    keep: list = []
    S = jd.system._system_struct(st, sy, keep)
    dom = sy.domain.to_c(3)
    ctx = L.context_for(device)
    rc = getattr(ctx.lib, entry)(ctx.ptr, ctypes.byref(S.state), ctypes.byref(dom), ctypes.byref(g),
                                 *[o.data_ptr() for o in outs])
    ctx.check(rc, entry)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs], ncell


@pytest.mark.parametrize("cell,label", [(1.0, "counting sort"), (0.2, "radix sort")])
def test_noalloc_cell_build_equals_cell_build(device, cell, label):
    box = 24.0
    for n, seed in ((60000, 1), (9000, 2)):  # the second call is smaller
        print(f"[noalloc] {label}: n={n} seed={seed}", flush=True)
        st = _state(n, box, seed, device)
        want, ncell = _build("jdem_cell_build", st, box, cell, device)
        got, _ = _build("jdem_cell_build_noalloc", st, box, cell, device)
        if label == "counting sort":
            assert ncell <= 4 * n + 65536
        else:
            assert ncell > 4 * n + 65536
        for name, g, w in zip(("perm", "sorted_hash", "cell_start", "cell_end"), got, want):
            assert np.array_equal(g, w), f"{label} n={n}: {name} differs from jdem_cell_build"
        assert np.array_equal(np.sort(got[0]), np.arange(n, dtype=np.int32))


def test_noalloc_rejects_sizes_beyond_its_scratch(device):
    lib = L.load()
    cap = int(lib.jdem_noalloc_scratch_bytes())
    assert cap == 64 << 20
    n = 4 << 20  # 4 M particles: more scratch than the fixed array holds
    print(f"[noalloc] oversize n={n}", flush=True)
    st = _state(n, 160.0, 3, device)
    g, _ = grid_struct(np.full(3, 160.0, np.float32), np.float32(1.0), True, 1)
    outs = [torch.empty(n, dtype=torch.int32, device=device) for _ in range(2)]
    outs += [torch.empty(int(g.n_cells), dtype=torch.int32, device=device) for _ in range(2)]
    sy = jd.System.create(st.shape, domain_type="periodic",
                          domain_kw=dict(box_size=np.full(3, 160.0, np.float32), anchor=np.zeros(3, np.float32)))
    keep: list = []
    S = jd.system._system_struct(st, sy, keep)
    dom = sy.domain.to_c(3)
    ctx = L.context_for(device)
    rc = ctx.lib.jdem_cell_build_noalloc(ctx.ptr, ctypes.byref(S.state), ctypes.byref(dom), ctypes.byref(g),
                                         *[o.data_ptr() for o in outs])
    assert L.ERRORS.get(rc) == "JDEM_EINVAL"
    assert b"more than the fixed" in ctx.lib.jdem_last_error(ctx.ptr)
