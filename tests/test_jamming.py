"""Jamming by bisection (utils/jamming.py:53-267) on the GPU minimiser.

Pinned by a closed form: a triangular lattice of equal disks is the densest
packing of the plane and stays triangular under an affine compression, so it
jams at exactly pi / (2 sqrt 3) = 0.9068997 (Thue), to float32 resolution."""

from __future__ import annotations

import math

import numpy as np
import pytest
import jaxdem_amd as jd
from jaxdem_amd.utils import jamming as J


def triangular(nx: int, ny: int, phi: float, device):
    """nx x ny disks of diameter 1 on a triangular lattice (ny even) filling a
    periodic box at packing fraction phi."""
    a = math.sqrt(math.pi / (2.0 * math.sqrt(3.0) * phi))
    h = a * math.sqrt(3.0) / 2.0
    pos = np.array([[(i + 0.5 + 0.5 * (j % 2)) * a, (j + 0.5) * h] for j in range(ny) for i in range(nx)], np.float32)
    n = len(pos)
    st = jd.State.create(pos=pos, vel=np.zeros_like(pos), rad=np.full(n, 0.5, np.float32),
                         mass=np.ones(n, np.float32), device=device)
    box = np.array([nx * a, ny * h], np.float32)
    return st, box


def make(st, box, collider="naive"):
    mt = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=1.0, young=1e3, poisson=0.3)])
    ckw = {}
    if collider == "CellList":
        ckw = dict(state=st)
    return jd.System.create(st.shape, dt=1e-3, domain_type="periodic",
                            domain_kw=dict(box_size=box, anchor=np.zeros(2, np.float32)),
                            collider_type=collider, collider_kw=ckw, force_model_type="spring", mat_table=mt)


def test_scale_to_packing_fraction_cpu():
    st, box = triangular(6, 4, 0.5, "cpu")
    sy = make(st, box)
    phi0 = J.compute_packing_fraction(st, sy)
    assert abs(float(phi0) - 0.5) < 1e-6
    s2, y2 = J.scale_to_packing_fraction(st, sy, 0.7)
    assert abs(float(J.compute_packing_fraction(s2, y2)) - 0.7) < 1e-6
    b0, b1 = np.asarray(sy.domain.box_size), np.asarray(y2.domain.box_size)
    assert abs(b1[0] / b1[1] - b0[0] / b0[1]) < 1e-6  # aspect ratio kept
    scale = b1[0] / b0[0]
    np.testing.assert_allclose(s2.pos_c.numpy(), st.pos_c.numpy() * scale, rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(np.asarray(y2.domain.inv_box_size), 1.0 / b1, rtol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("collider", ["naive", "CellList"])
def test_bisection_jam_triangular_lattice(device, collider):
    print(f"[jamming] triangular lattice, {collider}", flush=True)
    st, box = triangular(8, 8, 0.85, device)
    sy = make(st, box, collider)
    res = J.bisection_jam(st, sy, n_minimization_steps=2000, packing_fraction_tolerance=1e-6,
                          packing_fraction_increment=1e-2, verbose=False)
    phi_j = math.pi / (2.0 * math.sqrt(3.0))
    assert abs(res.packing_fraction - phi_j) < 2e-5, res.packing_fraction
    assert res.potential_energy > 1e-16  # the returned jammed state overlaps
    assert float(J.compute_packing_fraction(res.unjammed_state, res.unjammed_system)) <= phi_j + 2e-6
    assert float(J.compute_packing_fraction(res.jammed_state, res.jammed_system)) >= phi_j - 2e-6


@pytest.mark.gpu
def test_pe_band_jam_triangular_lattice(device):
    """On the compressed lattice every disk keeps six equal overlaps
    delta = 1 - a(phi), so PE/N = 3 * k delta^2 / 2 with a = sqrt(phi_J / phi)
    (disk diameter 1): the returned state must sit in the band and on that
    curve (tolerance 2 %: float32 resolution of delta ~ 3e-4)."""
    print("[jamming] pe band, triangular lattice", flush=True)
    st, box = triangular(8, 8, 0.85, device)
    sy = make(st, box)
    pe_tol, k = 1e-4, 1e3
    res = J.pe_band_jam(st, sy, n_minimization_steps=2000, pe_tol=pe_tol, pe_band_factor=2.0,
                        packing_fraction_increment=1e-2, n_jamming_steps=200, verbose=False)
    assert pe_tol <= res.potential_energy <= 2 * pe_tol, res.potential_energy
    phi_j = math.pi / (2.0 * math.sqrt(3.0))
    delta = 1.0 - math.sqrt(phi_j / res.packing_fraction)
    want = 1.5 * k * delta**2
    assert abs(res.potential_energy - want) < 0.02 * want, (res.potential_energy, want)
    assert float(J.compute_packing_fraction(res.unjammed_state, res.unjammed_system)) <= res.packing_fraction


@pytest.mark.gpu
def test_pressure_bisection_jam_triangular_lattice(device):
    """Contact virial of the compressed lattice: three contacts per disk of
    length a carrying k delta over the area a^2 sqrt(3) / 2 per disk give
    P = sqrt(3) k delta / a (2D pressure = trace / 2).  The returned state must
    sit in the band [P_lo, 1.01 P_lo] and on that curve (tolerance 2 %)."""
    from jaxdem_amd.utils.contacts import compute_pressure

    print("[jamming] pressure band, triangular lattice", flush=True)
    st, box = triangular(8, 8, 0.85, device)
    sy = make(st, box)
    k, p_lo = 1e3, 1.0
    res = J.pressure_bisection_jam(st, sy, n_minimization_steps=2000, pressure_threshold=p_lo,
                                   growth_rate=1.01, fine_growth_rate=1.0001, n_jamming_steps=500,
                                   verbose=False)
    p = compute_pressure(res.jammed_state, res.jammed_system)
    assert p_lo * 0.999 <= p <= 1.01 * p_lo * 1.001, p
    phi_j = math.pi / (2.0 * math.sqrt(3.0))
    a = math.sqrt(phi_j / res.packing_fraction)  # lattice spacing, disk diameter 1
    want = math.sqrt(3.0) * k * (1.0 - a) / a
    assert abs(p - want) < 0.02 * want, (p, want)
    assert compute_pressure(res.unjammed_state, res.unjammed_system) < p_lo


@pytest.mark.gpu
def test_quasistatic_compression_triangular_lattice(device):
    """Compressing the lattice below its jamming point leaves it free of
    overlaps; past it, PE/N = 3 k delta^2 / 2 and P = sqrt(3) k delta / a
    (float32 delta resolution: 2 %); expanding back removes every overlap."""
    from jaxdem_amd.utils.contacts import compute_contact_pressure
    from jaxdem_amd.utils.packing_utils import quasistatic_compress_to_packing_fraction as compress

    print("[jamming] quasistatic compression, triangular lattice", flush=True)
    st, box = triangular(8, 8, 0.80, device)
    sy = make(st, box)
    k = 1e3
    phi_j = math.pi / (2.0 * math.sqrt(3.0))
    s1, y1, phi1, pe1 = compress(st, sy, 0.88, step=0.02, max_n_min_steps_per_outer=2000)
    assert abs(phi1 - 0.88) < 1e-10 and pe1 == 0.0
    assert abs(float(J.compute_packing_fraction(s1, y1)) - 0.88) < 1e-6
    s2, y2, phi2, pe2 = compress(s1, y1, 0.92, step=0.01, max_n_min_steps_per_outer=2000)
    a = math.sqrt(phi_j / phi2)
    assert abs(pe2 - 1.5 * k * (1.0 - a) ** 2) < 0.02 * 1.5 * k * (1.0 - a) ** 2, pe2
    _, _, p = compute_contact_pressure(s2, y2)
    assert abs(p - math.sqrt(3.0) * k * (1.0 - a) / a) < 0.02 * math.sqrt(3.0) * k * (1.0 - a) / a, p
    s3, y3, phi3, pe3 = compress(s2, y2, 0.85, step=0.05, max_n_min_steps_per_outer=2000)
    assert abs(phi3 - 0.85) < 1e-10 and pe3 == 0.0
