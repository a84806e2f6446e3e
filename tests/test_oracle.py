"""Pins of the CPU oracle (oracle/jdem_oracle.py) that do not retype it:
closed forms, finite-difference gradients, brute-force pair sets, invariants
and the reference's own known-answer tests.  CPU only.

The reference has no golden-vector files (tests/ holds only .py), so its
known-answer tests are restated here: test_excluded_pairs.py:11-61 and
test_rotation_integrators.py:25-97.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import jdem_oracle as O

F64 = np.float64
F32 = np.float32


def two(dx, r=0.5, dim=3, dtype=F64, **kw):
    pos = np.zeros((2, dim), dtype)
    pos[1, 0] = dx
    return O.make_state(pos, [r, r], dtype=dtype, **kw)


def params(kind, dim=3, dtype=F64, domain="free", box=10.0, **law):
    d = O.Domain(domain, np.full(dim, box, dtype), np.zeros(dim, dtype))
    L = O.Law(kind, **law)
    return O.Params(dim=dim, dt=1e-4, domain=d, law=L, collider="naive")


def fd_force(st, P, i, h=1e-6):
    """-dE/dx_i by central differences of the pair energy sum (float64)."""
    g = np.zeros(st.dim)
    for c in range(st.dim):
        for sgn in (1, -1):
            s2 = st.copy()
            s2.pos[i, c] += sgn * h
            e = O.potential_energy(s2, P)
            g[c] += -sgn * e / (2 * h)
    return g


# --------------------------------------------------------------------------
# pair laws
# --------------------------------------------------------------------------
def test_spring_closed_form_and_gradient():
    k = 1e4
    st = two(0.9)
    P = params("spring", young_eff=np.array([[k]]))
    O.compute_force(st, P)
    # F_0 = k (R - d) n_hat with n_hat = (x0 - x1)/d = -x
    assert st.force[0, 0] == pytest.approx(-k * 0.1, rel=1e-12)
    np.testing.assert_allclose(st.force[1], -st.force[0], rtol=1e-14)
    assert O.potential_energy(st, P) == pytest.approx(0.5 * k * 0.1**2, rel=1e-12)
    np.testing.assert_allclose(st.force[0], fd_force(st, P, 0), rtol=1e-6, atol=1e-6)


def test_spring_coincident_particles_finite_zero():
    st = two(0.0)
    P = params("spring", young_eff=np.array([[1e4]]))
    O.compute_force(st, P)
    assert np.all(np.isfinite(st.force)) and np.all(st.force == 0)


def test_hertz_closed_form_and_gradient():
    E, nu, r, d = 1e4, 0.3, 0.5, 0.92
    st = two(d, r=r)
    P = params("hertz", young=np.array([E]), poisson=np.array([nu]))
    O.compute_force(st, P)
    Es = 1.0 / (2 * (1 - nu * nu) / E)
    Rs = r * r / (2 * r)
    k = 4.0 / 3.0 * Es * math.sqrt(Rs)
    delta = 2 * r - d
    assert -st.force[0, 0] == pytest.approx(k * delta**1.5, rel=1e-12)
    assert O.potential_energy(st, P) == pytest.approx(0.4 * k * delta**2.5, rel=1e-12)
    np.testing.assert_allclose(st.force[0], fd_force(st, P, 0), rtol=1e-6)


def cs_law(E=1e4, nu=0.3, e=0.7, mu=0.4, mu_r=0.0):
    return dict(young=np.array([E]), poisson=np.array([nu]), e=np.array([e]), mu=np.array([mu]),
                mu_r=np.array([mu_r]))


def test_cundall_strack_static_is_normal_spring():
    E, r = 1e4, 0.5
    st = two(0.95, r=r)
    P = params("cundallstrack", **cs_law(E=E))
    O.compute_force(st, P)
    kn = 2 * E * r * E * r / (E * r + E * r)
    assert -st.force[0, 0] == pytest.approx(kn * 0.05, rel=1e-12)
    assert np.all(st.torque == 0)


@pytest.mark.parametrize("vt", [5.0, 50.0])
def test_cundall_strack_damping_and_coulomb_cap(vt):
    E, nu, r, m, e, mu = 1e4, 0.3, 0.5, 1.0, 0.7, 0.4
    st = two(0.95, r=r, mass=[m, m])
    st.vel[0] = [1.0, vt, 0.0]  # approaching (v_n) and sliding (v_t)
    P = params("cundallstrack", **cs_law(E=E, nu=nu, e=e, mu=mu))
    O.compute_force(st, P)
    G = E / (2 * (1 + nu))
    kn, kt = E * r, G * r
    beta = -math.log(e) / math.sqrt(math.pi**2 + math.log(e) ** 2)
    m_eff = m * m / (m + m)
    gn = 2 * beta * math.sqrt(kn * m_eff)
    gt = 2 * beta * math.sqrt(kt * m_eff)
    vn = -1.0  # v_rel . n with n = -x
    Fn = kn * 0.05 - gn * vn
    assert -st.force[0, 0] == pytest.approx(Fn, rel=1e-12)
    # tangential dashpot opposite the slip, capped by Coulomb
    assert -st.force[0, 1] == pytest.approx(min(gt * vt, mu * Fn), rel=1e-12)
    # torque r_ci x F with r_ci = -R n = +R x
    assert st.torque[0, 2] == pytest.approx(r * st.force[0, 1], rel=1e-12)


def friction_pair(dx=0.95, m=1.0):
    st = two(dx, mass=[m, m])
    P = params("springfriction", young_eff=np.array([[1e4]]), mu_eff=np.array([[0.5]]), kt_ratio=2.0 / 7.0)
    P.dt = 1e-4
    return st, P


def test_friction_spring_grows_then_saturates():
    """Constant sliding v_t: |F_t| = k_t v_t t until mu F_n, then mu F_n."""
    st, P = friction_pair()
    st.vel[0] = [0.0, 10.0, 0.0]
    i = np.array([0])
    j = np.array([1])
    xi = np.zeros((1, 3))
    kn, kt, mu = 1e4, 1e4 * 2 / 7, 0.5
    Fn = kn * 0.05
    for step in range(1, 200):
        f, t, xi, c = O.law_force(P.law, st, i, j, P.domain, P.dt, xi)
        ft = -f[0, 1]
        expect = min(kt * 10.0 * P.dt * step, mu * Fn)
        assert ft == pytest.approx(expect, rel=1e-9), step
        assert -f[0, 0] == pytest.approx(Fn, rel=1e-12)
    assert ft == pytest.approx(mu * Fn, rel=1e-12)


def test_friction_history_resets_on_separation():
    st, P = friction_pair(dx=1.2)  # not touching
    f, t, xi, c = O.law_force(P.law, st, np.array([0]), np.array([1]), P.domain, P.dt, np.ones((1, 3)))
    assert not c[0] and np.all(xi == 0) and np.all(f == 0)


def test_friction_spring_projected_on_contact_plane():
    st, P = friction_pair()
    xi_old = np.array([[0.01, 0.002, 0.0]])  # has a normal component (along x)
    f, t, xi, c = O.law_force(P.law, st, np.array([0]), np.array([1]), P.domain, P.dt, xi_old)
    assert xi[0, 0] == pytest.approx(0.0, abs=1e-15)
    assert xi[0, 1] == pytest.approx(0.002, rel=1e-12)


def test_friction_dissipates_energy():
    """A glancing frictional, damped collision loses kinetic energy."""
    pos = np.array([[0.0, 0.0, 0.0], [1.1, 0.3, 0.0]])
    st = O.make_state(pos, [0.5, 0.5], vel=[[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]], dtype=F64)
    P = params("springfriction", young_eff=np.array([[1e4]]), mu_eff=np.array([[0.5]]), damping=0.1)
    P.dt = 1e-4
    P.collider = "neighborlist"
    P.nl_cutoff, P.nl_skin, P.nl_max_neighbors, P.nl_secondary = 1.0, 0.5, 1, "naive"
    nls = O.nl_init(st, P, history=True)
    e0 = O.kinetic_energy(st)
    for _ in range(2000):
        O.step_once(st, P, nls)
    e1 = O.kinetic_energy(st) + O.potential_energy(st, P)
    assert e1 < e0
    assert np.linalg.norm(st.ang_vel) > 0  # friction spun the spheres up


# --------------------------------------------------------------------------
# integrators
# --------------------------------------------------------------------------
def test_velocity_verlet_free_flight_and_gravity_exact():
    st = O.make_state([[0.0, 0.0, 0.0]], [0.1], vel=[[1.0, -2.0, 0.5]], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]), box=100.0)
    P.gravity = np.array([0.0, 0.0, -9.81])
    P.dt = 1e-3
    st.force[0] = st.mass[0] * P.gravity  # the force of the previous step
    n = 500
    for _ in range(n):
        O.step_once(st, P)
    t = n * P.dt
    np.testing.assert_allclose(st.pos[0], [t, -2 * t, 0.5 * t - 0.5 * 9.81 * t * t], rtol=1e-10)
    np.testing.assert_allclose(st.vel[0], [1.0, -2.0, 0.5 - 9.81 * t], rtol=1e-10)


def _energy_drift(dt, T=0.05):
    pos = np.array([[0.0, 0.0, 0.0], [1.05, 0.1, 0.0]])
    st = O.make_state(pos, [0.5, 0.5], vel=[[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]))
    P.dt = dt
    O.compute_force(st, P)
    es = []
    for _ in range(int(round(T / dt))):
        O.step_once(st, P)
        es.append(O.kinetic_energy(st) + O.potential_energy(st, P))
    return np.std(es)


def test_velocity_verlet_second_order_energy():
    """test_energy_conservation.py:34-89 restated: std(E) ~ dt^2."""
    dts = np.array([4e-4, 2e-4, 1e-4])
    s = np.array([_energy_drift(dt) for dt in dts])
    slope = np.polyfit(np.log(dts), np.log(s), 1)[0]
    assert slope > 1.8, slope


@pytest.mark.parametrize("rot", ["spiral", "verletspiral"])
def test_aspherical_rotation_analytic(rot):
    """test_rotation_integrators.py:25-97: constant body-frame torque on an
    aspherical body; closed-form omega(t) (dt 2e-5 here, 35000 steps)."""
    rho, Rad, H = 7750.0, 0.05, 0.15
    M = rho * math.pi * Rad * Rad * H
    Ix = 0.5 * M * Rad * Rad
    Iy = M * H * H / 12.0 + 0.25 * M * Rad * Rad
    Iz = Iy
    tx = Rad * 0.5
    w0 = np.array([0.3, -0.9, 0.6])
    st = O.make_state([[0.0, 0.0, 0.0]], [Rad], mass=[M], inertia=[[Ix, Iy, Iz]], ang_vel=[w0], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]))
    P.rot = rot
    P.dt = 2e-5
    steps = 35000
    tb = np.array([[tx * 0.5 * 0 + 0.025, 0.0, 0.0]])
    for _ in range(steps):
        O.domain_apply(st, P)
        O.lin_before(st, P)
        O.rot_before(st, P)
        st.force = np.zeros((1, 3))
        st.torque = O.q_rotate(st.q, tb)
        O.lin_after(st, P)
        O.rot_after(st, P)
    t = P.dt * steps
    tx = 0.025
    A = (Ix - Iy) * (Iz - Ix) / (Iy * Iz)
    B = Iy / (Iz - Ix)
    E = 2.0 * tx * B / Ix
    wx = w0[0] + tx * t / Ix
    eta = 0.5 * Ix * math.sqrt(-A) / tx
    C = E * eta
    K1 = (C * w0[1] * math.cos(eta * w0[0] ** 2) - w0[2] * math.sin(eta * w0[0] ** 2)) / C
    K2 = (C * w0[1] * math.sin(eta * w0[0] ** 2) + w0[2] * math.cos(eta * w0[0] ** 2)) / C
    Dd = eta * wx**2
    ww = np.array([wx, K1 * math.cos(Dd) + K2 * math.sin(Dd), C * (K2 * math.cos(Dd) - K1 * math.sin(Dd))])
    wb = O.q_rotate_back(st.q, st.ang_vel)[0]
    err = math.log10(np.linalg.norm(wb - ww) / np.linalg.norm(ww))
    assert err < -4.0, err
    assert abs(np.linalg.norm(st.q[0]) - 1.0) < 1e-12


def test_torque_free_sphere_keeps_omega():
    st = O.make_state([[0.0, 0.0, 0.0]], [0.5], ang_vel=[[0.3, -1.0, 2.0]], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]))
    for _ in range(200):
        O.step_once(st, P)
    np.testing.assert_allclose(st.ang_vel[0], [0.3, -1.0, 2.0], rtol=1e-12)


# --------------------------------------------------------------------------
# domains
# --------------------------------------------------------------------------
def test_reflect_wall_exact_bounce():
    """A free sphere crossing a wall with e = 1 is mirrored: the overshoot is
    folded back and the normal velocity reversed, energy conserved."""
    st = O.make_state([[0.15, 5.0, 5.0]], [0.1], vel=[[-1.0, 0.0, 0.0]], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]), domain="reflect")
    P.dt = 0.1
    O.step_once(st, P)  # 0.15 -> 0.05: 0.05 past the wall at x = 0.1
    O.step_once(st, P)  # apply folds it back: x = 0.15 (then moves +0.1)
    assert st.vel[0, 0] == pytest.approx(1.0, rel=1e-12)
    assert st.pos[0, 0] == pytest.approx(0.25, rel=1e-12)
    assert O.kinetic_energy(st) == pytest.approx(0.5, rel=1e-12)


def test_periodic_min_image_invariants():
    rng = np.random.default_rng(0)
    B = np.array([3.0, 4.0, 5.0])
    d = O.Domain("periodic", B, np.zeros(3))
    a = rng.uniform(-20, 20, (500, 3))
    b = rng.uniform(-20, 20, (500, 3))
    r = O.displacement(a, b, d)
    assert np.all(np.abs(r) <= B / 2 + 1e-12)
    r2 = O.displacement(a + B * rng.integers(-3, 4, (500, 3)), b, d)
    np.testing.assert_allclose(r, r2, atol=1e-9)
    # round half to even, as jnp.round: 0.5 -> 0, 1.5 -> 2, 2.5 -> 2, -2.5 -> -2
    # (half away from zero would give -0.5, -0.5, -0.5, 0.5)
    unit = O.Domain("periodic", np.array([1.0]), np.zeros(1))
    for x, want in ((0.5, 0.5), (1.5, -0.5), (2.5, 0.5), (-2.5, -0.5)):
        assert O.displacement(np.array([x]), np.array([0.0]), unit)[0] == want, x
        assert O.displacement_(np.array([x]), np.array([0.0]), unit)[0] == want, x


# --------------------------------------------------------------------------
# pair sets
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("box", [2.2, 9.0])
def test_cell_list_pairs_equal_brute_force(dim, domain, box):
    """The candidate set of the cell list restatement equals brute force --
    including boxes of fewer than 3 cells per axis (stencil dedup), positions
    on cell faces, and unwrapped periodic positions."""
    rng = np.random.default_rng(dim * 7 + int(box))
    N = 300
    pos = rng.uniform(0, box, (N, dim))
    pos[:20] = np.round(pos[:20] / 1.1) * 1.1  # exactly on cell faces
    if domain == "periodic":
        pos[20:40] += box * rng.integers(-2, 3, (20, dim))
    st = O.make_state(pos, np.full(N, 0.5), dtype=F64)
    d = O.Domain(domain, np.full(dim, box), np.zeros(dim))
    cutoff = 1.0
    nl, ovf = O.neighbor_list_cells(st, d, 1.1, 1, cutoff, N)
    got = {(i, int(j)) for i in range(N) for j in nl[i] if j >= 0}
    assert got == O.brute_force_pairs(st, d, cutoff)
    assert not ovf


def test_cell_list_pairs_independent_of_cell_size():
    rng = np.random.default_rng(3)
    st = O.make_state(rng.uniform(0, 10, (400, 3)), np.full(400, 0.5), dtype=F64)
    d = O.Domain("periodic", np.full(3, 10.0), np.zeros(3))
    sets = []
    for cs, sr in [(1.0, 1), (0.5, 2), (2.5, 1), (0.34, 3)]:
        nl, _ = O.neighbor_list_cells(st, d, cs, sr, 1.0, 128)
        sets.append({(i, int(j)) for i in range(400) for j in nl[i] if j >= 0})
    assert all(s == sets[0] for s in sets)


def test_neighbor_list_overflow_and_row_order():
    rng = np.random.default_rng(4)
    st = O.make_state(rng.uniform(0, 5, (200, 3)), np.full(200, 0.5), dtype=F64)
    d = O.Domain("periodic", np.full(3, 5.0), np.zeros(3))
    nl, ovf = O.neighbor_list_cells(st, d, 1.0, 1, 1.5, 3)
    assert ovf and nl.shape == (200, 3)
    nl2, _ = O.neighbor_list_naive(st, d, 1.5, 64)
    for r in range(200):
        row = nl2[r][nl2[r] >= 0]
        assert np.all(np.diff(row) < 0)  # naive rows: descending ids (naive.py:174-176)


def test_drop_far_candidates_keeps_sums():
    rng = np.random.default_rng(5)
    N = 150
    st = O.make_state(rng.uniform(0, 4, (N, 3)), rng.uniform(0.3, 0.5, N), dtype=F32)
    P = params("hertz", dtype=F32, domain="periodic", box=4.0, young=np.array([1e4], F32), poisson=np.array([0.3], F32))
    P.domain = O.Domain("periodic", np.full(3, 4.0, F32), np.zeros(3, F32))
    F_full, T_full, _ = O.accumulate(st, P, O.naive_candidates(N))
    F_c, T_c, _ = O.accumulate(st, P, O.drop_far_candidates(st, P, O.naive_candidates(N)))
    assert np.array_equal(F_full, F_c)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("domain", ["periodic", "reflect"])
@pytest.mark.parametrize("box", [1.7, 3.1, 12.0])
def test_naive_rows_through_cells_equal_all_pairs(dim, domain, box):
    """naive_near_candidates (the all-pairs rows found through a cell list)
    equals drop_far_candidates over all pairs, row for row -- with boxes of
    one, two and many cells per axis, touching and far pairs, polydisperse
    radii, positions outside a reflecting box and unwrapped periodic ones."""
    rng = np.random.default_rng(int(10 * box) + dim)
    N = 700
    pos = rng.uniform(0, box, (N, dim)).astype(F32)
    pos[:30] = pos[30:60] + rng.uniform(-0.6, 0.6, (30, dim)).astype(F32)  # touching pairs
    if domain == "periodic":
        pos[60:90] += (box * rng.integers(-2, 3, (30, dim))).astype(F32)
    else:
        pos[60:90] += rng.uniform(-1.5, 1.5, (30, dim)).astype(F32)
    st = O.make_state(pos, rng.uniform(0.1, 0.5, N).astype(F32), dtype=F32)
    P = params("spring", dtype=F32, domain=domain, box=box)
    P.domain = O.Domain(domain, np.full(dim, box, F32), np.zeros(dim, F32))
    P.dim = dim
    want = O.drop_far_candidates(st, P, O.naive_candidates(N))
    got = O.naive_near_candidates(st, P)
    K = max(want.shape[1], got.shape[1])
    pad = lambda a: np.pad(a, ((0, 0), (0, K - a.shape[1])), constant_values=-1)  # noqa: E731
    assert np.array_equal(pad(got), pad(want))
    assert (want >= 0).sum() > N  # more than the self pairs


def test_excluded_pairs_known_answer():
    """test_excluded_pairs.py:11-61 with the oracle (float64)."""
    st = O.make_state([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0]], [1.1, 1.1, 1.1], bond_id=[[1, -1], [0, 2], [1, -1]],
                      dtype=F64)
    P = params("spring", dim=2, young_eff=np.array([[1000.0]]))
    O.compute_force(st, P)
    assert np.allclose(st.force[1], 0.0)
    assert st.force[0, 0] == pytest.approx(-1000.0 * 0.2, rel=1e-12)
    np.testing.assert_allclose(st.force[0], -st.force[2])


def test_neighborlist_rebuild_criterion():
    """Rebuild iff some particle moved more than skin/2 (neighbor_list.py:78-85)."""
    st = O.make_state(np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]), [0.5, 0.5], dtype=F64)
    P = params("spring", young_eff=np.array([[1e4]]))
    P.collider, P.nl_cutoff, P.nl_skin, P.nl_max_neighbors = "neighborlist", 1.0, 0.4, 4
    nls = O.nl_init(st, P, history=False)
    assert O.nl_check_and_rebuild(st, P, nls)  # first call builds
    st.pos[0, 0] += 0.19
    assert not O.nl_check_and_rebuild(st, P, nls)
    st.pos[0, 0] += 0.02
    assert O.nl_check_and_rebuild(st, P, nls)
    assert nls.n_build == 2


def test_cpu_port_matches_numpy_oracle():
    """The timed C++/OpenMP port equals the numpy oracle bit for bit."""
    from oracle.port import PortRun

    rng = np.random.default_rng(11)
    N = 1000
    L = 11.0
    pos = rng.uniform(0, L, (N, 3)).astype(F32)
    rad = np.where(rng.random(N) < 0.5, 0.5, 0.7).astype(F32)
    vel = rng.standard_normal((N, 3)).astype(F32)
    mass = ((4 / 3) * np.pi * rad.astype(F64) ** 3).astype(F32)
    st = O.make_state(pos, rad, vel=vel, mass=mass, dtype=F32)
    P = O.Params(dim=3, dt=2e-4, domain=O.Domain("periodic", np.full(3, L, F32), np.zeros(3, F32)),
                 law=O.Law("springfriction", young_eff=np.array([[1e4]], F32), mu_eff=np.array([[0.5]], F32),
                           kt_ratio=2 / 7, damping=0.1),
                 collider="neighborlist", nl_cutoff=1.4, nl_skin=0.05, nl_max_neighbors=40,
                 nl_secondary="celllist", nl_secondary_cell_size=1.4, nl_secondary_search_range=1)
    nls = O.nl_init(st, P, history=True)
    ref = O.step(st, P, 30, nls)
    run = PortRun(pos, vel, rad, mass, np.full(3, L, F32), np.zeros(3, F32), domain="periodic", dt=2e-4,
                  law="springfriction", kn=1e4, kt=F32(2 / 7) * F32(1e4), mu=0.5, damping=0.1, cutoff=1.4, skin=0.05,
                  M=40, cell_size=1.4, sr=1)
    run.steps(30, threads=4)
    for name, a, b in [("pos", run.pos, ref.pos), ("vel", run.vel, ref.vel), ("q", run.q, ref.q),
                       ("ang", run.angv, ref.ang_vel), ("force", run.force, ref.force)]:
        assert np.array_equal(a, b), name
    assert run.n_build.value == nls.n_build


# --------------------------------------------------------------------------
# FIRE minimisation (minimizers/optimizers.py:150-339, routines.py:150-383)
# --------------------------------------------------------------------------
def _packing(N=64, phi=0.7, seed=0, dtype=F64):
    rng = np.random.default_rng(seed)
    rad = np.where(np.arange(N) % 2 == 0, 0.5, 0.7)
    L = float((np.sum(4.0 / 3.0 * np.pi * rad**3) / phi) ** (1.0 / 3.0))
    st = O.make_state(rng.uniform(0, L, (N, 3)), rad, dtype=dtype)
    d = O.Domain("periodic", np.full(3, L, dtype), np.zeros(3, dtype))
    P = O.Params(dim=3, dt=0.05, domain=d, law=O.Law("spring", young_eff=np.ones((1, 1), dtype)), collider="naive")
    return st, P


def test_fire_separates_an_overlapping_pair_to_zero_energy():
    """The minimum of a harmonic overlap is any separation >= r_i + r_j: FIRE
    reaches pe = 0 (the pe_tol stop) with the pair out of contact."""
    st = two(0.6, r=0.5)
    P = params("spring", young_eff=np.array([[1.0]]))
    steps, pe = O.minimize(st, P, O.FireParams(dt=0.1), max_steps=10000, pe_tol=0.0, pe_diff_tol=-1.0)
    assert pe == 0.0 and 0 < steps < 10000
    assert abs(st.pos[1, 0] - st.pos[0, 0]) >= 1.0
    # momentum-free: both spheres moved symmetrically
    assert st.pos[0, 0] + st.pos[1, 0] == pytest.approx(0.6, abs=1e-12)


def test_fire_fixed_particle_stays_put():
    st = two(0.6, r=0.5, fixed=[True, False])
    P = params("spring", young_eff=np.array([[1.0]]))
    O.minimize(st, P, O.FireParams(dt=0.1), max_steps=10000, pe_tol=0.0, pe_diff_tol=-1.0)
    assert st.pos[0, 0] == 0.0 and st.pos[1, 0] >= 1.0


def test_fire_uphill_and_downhill_rules():
    """Bitzek et al. (2006) with the reference's uphill-counter reset: P > 0
    counts downhill steps and, past N_min, grows dt by f_inc (capped at
    dt_max) and shrinks alpha by f_alpha; P <= 0 shrinks dt by f_dec (not
    below dt_min), resets alpha and zeroes the velocity; more than N_bad_max
    uphill steps in a row restore dt0."""
    fp = O.FireParams(dt=0.1)
    F = np.array([[1.0, 0.0, 0.0]])
    Z = np.zeros((1, 3))
    fs = O.FireState(vel_pos=np.array([[0.5, 0.0, 0.0]]), vel_rot=Z.copy(), dt=F64(0.1), alpha=F64(0.1))
    for k in range(1, 8):
        _, _, fs = O.fire_update(F, Z, fs, fp, power=1.0)
        assert fs.N_good == k and fs.N_bad == 0
        if k <= fp.N_min:
            assert fs.dt == pytest.approx(0.1) and fs.alpha == pytest.approx(0.1)
    assert fs.dt == pytest.approx(0.1 * 1.1**2) and fs.alpha == pytest.approx(0.1 * 0.99**2)
    for _ in range(40):
        _, _, fs = O.fire_update(F, Z, fs, fp, power=1.0)
    assert fs.dt == pytest.approx(1.0)  # dt_max = 10 dt0
    dts = []
    for k in range(1, 13):
        up, _, fs = O.fire_update(F, Z, fs, fp, power=-1.0)
        assert fs.N_good == 0 and fs.alpha == pytest.approx(0.1)
        np.testing.assert_array_equal(fs.vel_pos, 0.0)
        dts.append(float(fs.dt))
    assert dts[:10] == pytest.approx([max(1.0 * 0.5**k, 1e-4) for k in range(1, 11)])
    assert dts[10] == pytest.approx(0.1)  # the 11th uphill step in a row restores dt0
    assert dts[11] == pytest.approx(0.05)


def test_fire_velocity_mixing_is_bitzek():
    """Downhill, v <- (1 - alpha) v + alpha |v| F_hat with v the half-kicked
    velocity (Bitzek et al. 2006, eq. 3)."""
    fp = O.FireParams(dt=0.2)
    F = np.array([[0.0, 2.0, 0.0]])
    v0 = np.array([[1.0, 0.0, 0.0]])
    fs = O.FireState(vel_pos=v0, vel_rot=np.zeros((1, 3)), dt=F64(0.2), alpha=F64(0.1))
    _, _, ns = O.fire_update(F, np.zeros((1, 3)), fs, fp, power=1.0)
    v_half = v0 + F * 0.2  # two half kicks of dt/2 each
    want = 0.9 * v_half + 0.1 * np.linalg.norm(v_half) * F / 2.0
    np.testing.assert_allclose(ns.vel_pos, want, rtol=1e-14)


def test_fire_relaxes_a_jammed_packing():
    """phi = 0.7 is above jamming: FIRE ends in mechanical equilibrium
    (max |F| below 1e-10) at positive energy lower than the start."""
    st, P = _packing()
    O.compute_force(st, P)
    e0 = O.potential_energy(st, P)
    f0 = np.abs(st.force).max()
    steps, pe = O.minimize(st, P, O.FireParams(dt=0.05), max_steps=20000, pe_tol=0.0, pe_diff_tol=-1.0,
                           force_tol=1e-10)
    assert np.abs(st.force).max() <= 1e-10 < f0
    assert 0.0 < pe * st.N < e0
    # equilibrium: the net force on the packing vanishes
    assert np.abs(st.force.sum(axis=0)).max() < 1e-9


def test_fire_float32_tracks_float64():
    s32, P32 = _packing(N=32, dtype=F32)
    s64, P64 = _packing(N=32, dtype=F64)
    _, pe32 = O.minimize(s32, P32, O.FireParams(dt=0.05), max_steps=60, pe_tol=0.0, pe_diff_tol=-1.0)
    _, pe64 = O.minimize(s64, P64, O.FireParams(dt=0.05), max_steps=60, pe_tol=0.0, pe_diff_tol=-1.0)
    assert pe32 == pytest.approx(pe64, rel=1e-4)
    np.testing.assert_allclose(s32.pos, s64.pos, atol=1e-4)


def test_fire64_reference_matches_numpy_oracle():
    """The C++ float64 FIRE (the precision form's full-FP64 reference) takes
    the same iterates as the numpy oracle in float64."""
    from oracle import build_port, port

    build_port.build()
    rng = np.random.default_rng(0)
    N = 64
    rad = np.where(np.arange(N) % 2 == 0, 0.5, 0.7)
    L = float((np.sum(4.0 / 3.0 * np.pi * rad**3) / 0.7) ** (1.0 / 3.0))
    pos = rng.uniform(0, L, (N, 3))
    r = port.fire64(pos, rad, L, k=1.0, dt=0.05, max_steps=20000, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=1e-12)
    st = O.make_state(pos, rad, dtype=F64)
    d = O.Domain("periodic", np.full(3, L), np.zeros(3))
    P = O.Params(dim=3, dt=0.05, domain=d, law=O.Law("spring", young_eff=np.ones((1, 1))), collider="naive")
    steps, pe = O.minimize(st, P, O.FireParams(dt=0.05), max_steps=20000, pe_tol=0.0, pe_diff_tol=-1.0,
                           force_tol=1e-12)
    assert r["steps"] == steps and r["max_grad"] <= 1e-12
    assert r["pe"] == pytest.approx(pe, rel=1e-12)
    np.testing.assert_allclose(r["pos"], st.pos, atol=1e-9)


# ---------------------------------------------------------------------------
# float64 evaluation and rattler rule of the precision-form comparison
# ---------------------------------------------------------------------------
def _fcc(cells, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(cells)] * 3, indexing="ij"), -1).reshape(-1, 1, 3)
    return ((g + base[None]) * a).reshape(-1, 3)


def test_evaluate64_fcc_closed_form():
    """A compressed FCC crystal: every sphere has 12 contacts at distance
    d = a/sqrt(2), the forces cancel by symmetry, E/N = 6 (k/2) delta^2 and
    P = (N 12/2) d k delta / (3 V) with delta = 2r - d."""
    from tests.precision_vs_fp64 import evaluate64, rattlers

    cells, a = 4, 1.0
    d = a / math.sqrt(2.0)
    r = 0.5 * d * 1.01
    pos = _fcc(cells, a) + 0.1
    N = len(pos)
    rad = np.full(N, r)
    L = cells * a
    e = evaluate64(pos, rad, L, k=2.0)
    delta = 2 * r - d
    assert len(e["contacts"][0]) == N * 6
    assert e["max_grad"] < 1e-13
    assert e["pe"] == pytest.approx(6 * 0.5 * 2.0 * delta * delta, rel=1e-12)
    assert e["pressure"] == pytest.approx(N * 6 * d * 2.0 * delta / (3 * L**3), rel=1e-12)
    assert not rattlers(pos, rad, L).any()


def test_rattler_rule_iterates():
    """A sphere touching four others that touch nothing else: the four have
    one contact each and go first, which leaves the centre with none."""
    from tests.precision_vs_fp64 import rattlers

    L = 20.0
    c = np.array([10.0, 10.0, 10.0])
    dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / math.sqrt(3.0)
    pos = np.vstack([c, c + 0.95 * dirs, [[2.0, 2.0, 2.0]]])
    rad = np.full(len(pos), 0.5)
    assert rattlers(pos, rad, L).all()


def test_jamming_fixture_is_a_float64_minimum():
    """The fixture's float64 minimum, re-evaluated independently in numpy:
    max|F| < 1e-12, and E/N and P as the C++ FIRE reported them."""
    from tests.precision_vs_fp64 import evaluate64, rattlers

    fx = np.load(Path(__file__).parent / "golden" / "jamming_n1024.npz")
    e = evaluate64(fx["ref"], fx["rad"], float(fx["L"]))
    assert e["max_grad"] < 1e-12
    assert e["pe"] == pytest.approx(float(fx["ref_pe"]), rel=1e-12)
    assert e["pressure"] == pytest.approx(float(fx["ref_pressure"]), rel=1e-12)
    assert 0 < rattlers(fx["ref"], fx["rad"], float(fx["L"])).sum() < 100


def test_fp64_minimisation_is_chaotic_near_jamming():
    """The reason positions are compared only at the minimum: from the
    fixture's start (float64-relaxed to max|F| <= 1e-4), a 1e-9 perturbation
    sends the float64 FIRE to a different minimum, several percent away in
    energy."""
    from tests.precision_vs_fp64 import evaluate64, fp64_run

    fx = np.load(Path(__file__).parent / "golden" / "jamming_n1024.npz")
    rad, L = fx["rad"], float(fx["L"])
    rng = np.random.default_rng(5)
    r = fp64_run(fx["start"] + 1e-9 * rng.standard_normal(fx["start"].shape), rad, L, max_steps=200000)
    assert r["max_grad"] < 1e-12
    e = evaluate64(r["pos"], rad, L)
    assert abs(e["pe"] - float(fx["ref_pe"])) > 0.02 * float(fx["ref_pe"])
