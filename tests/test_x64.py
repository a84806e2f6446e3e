"""The float64 switch on the host (no GPU): ``jax_enable_x64`` makes states,
systems, material and law tables and cell grids float64, computed with the
same operations as the reference under x64; with the switch off everything
stays float32."""

from __future__ import annotations

import contextlib

import numpy as np
import pytest
import torch

import jaxdem_amd as jd
from jaxdem_amd import _lib as L
from jaxdem_amd.colliders import grid_struct
from jaxdem_amd.forces import CundallStrackForce, HertzianForce
from oracle import jdem_oracle as O


@contextlib.contextmanager
def x64():
    old = jd.config.jax_enable_x64
    jd.config.update("jax_enable_x64", True)
    try:
        yield
    finally:
        jd.config.update("jax_enable_x64", old)


def test_config_switch():
    assert jd.config.jax_enable_x64 is False
    with pytest.raises(KeyError):
        jd.config.update("jax_platform_name", "cpu")
    with x64():
        assert jd.config.jax_enable_x64 is True
    assert jd.config.jax_enable_x64 is False


def _scene():
    pos = np.array([[0.1, 0.2, 0.3], [1.05, 0.2, 0.3], [0.1, 1.1, 0.35]])
    mats = [jd.Material.create("elasticfrict", density=1.3, young=1.1e4, poisson=0.3, mu=0.4, e=0.7, mu_r=0.05),
            jd.Material.create("elasticfrict", density=0.7, young=3.3e3, poisson=0.21, mu=0.2, e=0.5, mu_r=0.1)]
    mt = jd.MaterialTable.from_materials(mats)
    st = jd.State.create(pos=pos, rad=[0.5, 0.45, 0.55], mat_id=[0, 1, 1], mat_table=mt, device="cpu")
    sy = jd.System.create(st.shape, dt=1e-4, domain_type="periodic", domain_kw=dict(box_size=[5.3, 5.3, 5.3]),
                          mat_table=mt, force_model_type="hertz", force_manager_kw={"gravity": [0.0, -9.81, 0.0]})
    return st, sy, mt


def test_x64_states_and_systems_are_float64():
    with x64():
        st, sy, mt = _scene()
    for f in ("pos_c", "vel", "force", "ang_vel", "torque", "rad", "mass", "inertia", "volume"):
        assert getattr(st, f).dtype == torch.float64, f
    assert st.q.data.dtype == torch.float64
    assert isinstance(sy.dt, np.float64) and sy.dt == 1e-4  # 1e-4 rounded once, to float64
    assert sy.domain.box_size.dtype == np.float64 and sy.domain.box_size[0] == 5.3
    assert mt.young.dtype == np.float64 and mt.young_eff.dtype == np.float64
    assert sy.force_manager.gravity.dtype == np.float64
    # mass = density * volume in float64 (state.py:620-640), inertia 0.4 m r^2
    r = np.array([0.5, 0.45, 0.55])
    m = np.array([1.3, 0.7, 0.7]) * (4.0 / 3.0 * np.pi * r**3)
    np.testing.assert_allclose(st.mass.numpy(), m, rtol=1e-15)
    np.testing.assert_allclose(st.inertia.numpy()[:, 0], 0.4 * m * r * r, rtol=1e-15)
    # switched off: float32 again
    st32, sy32, mt32 = _scene()
    assert st32.pos_c.dtype == torch.float32 and isinstance(sy32.dt, np.float32) and mt32.young.dtype == np.float32


def test_x64_law_tables_follow_the_pair_formulas_in_float64():
    with x64():
        _, _, mt = _scene()
    E = np.array([1.1e4, 3.3e3])
    nu = np.array([0.3, 0.21])
    # Hertz: 4/3 E* with E* = 1 / ((1 - nu_i^2)/E_i + (1 - nu_j^2)/E_j) (hertz.py:107)
    t0 = HertzianForce().tables(mt, np.float64)["t0"]
    Es = 1.0 / ((1 - nu[:, None] ** 2) / E[:, None] + (1 - nu[None, :] ** 2) / E[None, :])
    assert t0.dtype == np.float64
    np.testing.assert_allclose(t0, 4.0 / 3.0 * Es, rtol=1e-15)
    # Cundall-Strack damping ratio beta(e_eff) = -ln e / sqrt(pi^2 + ln^2 e), e_eff = min
    cs = CundallStrackForce().tables(mt, np.float64)
    e = np.minimum.outer(np.array([0.7, 0.5]), np.array([0.7, 0.5]))
    np.testing.assert_allclose(cs["t0"], -np.log(e) / np.sqrt(np.pi**2 + np.log(e) ** 2), rtol=1e-15)
    np.testing.assert_allclose(cs["m1"], E / (2 * (1 + nu)), rtol=1e-15)
    # float32 tables are float32 roundings of the float32 computation
    assert HertzianForce().tables(mt).get("t0").dtype == np.float32


@pytest.mark.parametrize("periodic", [True, False])
def test_x64_grid_equals_the_float64_oracle(periodic):
    box = np.array([10.3, 7.7, 9.1])
    cs = 1.37
    g, ovf = grid_struct(box, cs, periodic, 1, f64=True)
    assert isinstance(g, L.JdemGridF64) and not ovf
    dims, strides, cell, _ = O.grid_params(box, cs, periodic)
    assert [g.dims[a] for a in range(3)] == dims.tolist()
    assert [g.strides[a] for a in range(3)] == strides.tolist()
    assert [g.cell_size[a] for a in range(3)] == cell.tolist()
    # the float32 grid rounds the box first: a different cell size
    g32, _ = grid_struct(box, cs, periodic, 1)
    if periodic:
        assert g32.cell_size[0] != g.cell_size[0]


def test_x64_domain_and_law_structs():
    with x64():
        st, sy, mt = _scene()
    d = sy.domain.to_c(3, f64=True)
    assert isinstance(d, L.JdemDomainF64) and d.box_size[0] == 5.3 and d.inv_box_size[0] == 1.0 / 5.3
    from jaxdem_amd.forces import law_struct

    law, dev = law_struct(sy.force_model, mt, torch.device("cpu"), f64=True)
    assert isinstance(law, L.JdemLawF64) and dev["t0"].dtype == torch.float64
