"""Hypothesis fuzz of small systems on the GPU step against the oracle, bit
for bit: 1 to 24 spheres, 2D and 3D, every collider (naive, CellList,
NeighborList over a CellList), periodic / reflective / free domains (a
CellList in a free domain hashes on the box fitted every step), with
and without history, volume fractions from dilute to dense (small boxes,
grids of few cells), 1 to 6 fused steps.  The examples are derandomized, so
every run draws the same cases."""

from __future__ import annotations

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as hs

import jaxdem_amd as jd
from oracle import jdem_oracle as O
from tests import oracle_bridge as B
from tests.test_gpu_parity import _compare_states, make_system, packed

pytestmark = pytest.mark.gpu

CASES = hs.tuples(
    hs.integers(1, 24),                                   # N
    hs.sampled_from([2, 3]),                              # dim
    hs.sampled_from(["naive", "CellList", "NeighborList"]),
    hs.sampled_from(["periodic", "reflect", "free"]),
    hs.sampled_from(["spring", "hertz", "cundallstrack", "springfriction"]),
    hs.integers(1, 6),                                    # steps
    hs.integers(0, 2**16),                                # seed
    hs.sampled_from([0.1, 0.3, 0.5, 0.65]),               # volume fraction
)


@settings(max_examples=60, deadline=None, derandomize=True, database=None,
          suppress_health_check=[HealthCheck.function_scoped_fixture, HealthCheck.too_slow])
@given(case=CASES)
def test_fuzz_small_systems(device, case):
    n, dim, collider, domain, law, steps, seed, phi = case
    if collider == "NeighborList" and domain == "free":
        domain = "reflect"  # a NeighborList over a CellList needs a fixed grid
    if law == "springfriction" and collider != "NeighborList":
        law = "cundallstrack"  # history laws take the NeighborList (as in the reference)
    tag = f"N={n} dim={dim} {collider} {domain} {law} steps={steps} seed={seed} phi={phi}"
    print(f"[fuzz] {tag}", flush=True)
    st, box = packed(n, dim, seed=seed, domain=domain, phi=phi, device=device)
    sy = make_system(st, box, domain, law, collider, dt=1e-3)
    nls = B.nl_state_of(st, sy) if collider == "NeighborList" else None
    ost = O.step(B.state_of(st), B.params_of(st, sy), steps, nls)
    st2, sy2 = jd.System.step(st, sy, n=steps)
    _compare_states(st2, ost, tag)
    assert sy2.step_count == steps
    if nls is not None:
        assert sy2.collider.n_build_times == nls.n_build
        assert np.array_equal(sy2.collider.neighbor_list.cpu().numpy().astype(np.int64), nls.nl), tag


@pytest.mark.parametrize("law", ["hertz", "cundallstrack"])
@pytest.mark.parametrize("dim", [3, 2])
def test_celllist_free_domain(device, dim, law):
    """A CellList in a free domain at a few thousand spheres: 25 steps, the
    grid rebuilt on the box fitted each step, bit for bit with the oracle,
    and the final fitted box."""
    print(f"[free] CellList dim={dim} {law}", flush=True)
    st, box = packed(3000 if dim == 3 else 2000, dim, seed=17, domain="free", phi=0.4, device=device)
    sy = make_system(st, box, "free", law, "CellList", dt=2e-4)
    ost = O.step(B.state_of(st), P := B.params_of(st, sy), 25)
    st2, sy2 = jd.System.step(st, sy, n=25)
    _compare_states(st2, ost, f"free/CellList/{dim}d/{law}")
    assert np.array_equal(np.asarray(sy2.domain.box_size, np.float32)[:dim], P.domain.box.astype(np.float32))
