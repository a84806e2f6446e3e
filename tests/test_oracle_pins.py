"""Closed-form and invariant pins of five oracle functions that the other
oracle tests leave open (CPU only):

* the history remap across neighbour-list rebuilds (neighbor_list.py:29-54);
* the Hertz-Mindlin friction law ``hertzfriction`` (no reference law; the
  normal force is Hertz's, hertz.py:98-117, the tangential stiffness
  Mindlin's k_t = 8 G* sqrt(R* delta));
* the time-of-contact quadratic of the reflecting wall with a nonzero
  acceleration (domains/_toc.py:66-94, reflect.py:165-297);
* symplectic Euler (integrators/direct_euler.py:62-66);
* the rolling resistance of Cundall-Strack (cundall_strack.py:187-192).

Every expected value is computed here from the mechanics (textbook
formulas, exact trajectories), not by calling the oracle's own helpers, and
each test is chosen so that a plausible slip in the oracle (a slot kept
instead of a partner, a wrong sign, a dropped term, G in place of G*) fails
it.  ``test_remap_pin_fails_for_slot_keeping_remap`` plants such a slip in
the oracle and checks that the remap pin catches it.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import jdem_oracle as O

F64 = np.float64


def _dom(kind, dim=3, box=20.0, restitution=1.0):
    return O.Domain(kind, np.full(dim, box, F64), np.zeros(dim, F64), restitution)


# --------------------------------------------------------------------------
# history remap (neighbor_list.py:29-54)
# --------------------------------------------------------------------------
KN, MU, KT_RATIO = 1e4, 0.5, 2.0 / 7.0


def _sliding_scene(secondary: str):
    """Sphere 0 (fixed, spinning about z) presses on sphere 1 (fixed, at
    rest): the contact point slides at v_t = w r_0 with a constant overlap,
    so the tangential spring grows as k_t v_t t until mu F_n.  Small
    spectator spheres fly past sphere 0 inside its list radius but never
    touch anything; as they enter and leave its row, the slot of sphere 1 in
    that row moves, and the tiny skin forces a rebuild every few steps."""
    r0, w = 0.5, 8.0
    pos = [[0.0, 0.0, 0.0], [0.95, 0.0, 0.0]]
    vel = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    rad = [r0, 0.5]
    # spectators: lines at distance 0.7 from sphere 0's centre on its -x
    # side, flying along y; spectator k crosses y = 0 at t_k and is listed
    # for |y| < sqrt(1.02^2 - 0.7^2).  Distance to sphere 0 stays >= 0.7 >
    # 0.5 + 0.1, to sphere 1 >= 0.95, to each other >= 0.35.
    speed = 40.0
    n_spec = 6
    for k in range(n_spec):
        ang = math.pi * (0.6 + 0.16 * k)  # directions on the -x half
        cx, cz = 0.7 * math.cos(ang), 0.7 * math.sin(ang)
        t_k = 0.005 + 0.008 * k
        pos.append([cx, -speed * t_k, cz])
        vel.append([0.0, speed, 0.0])
        rad.append(0.1)
    N = len(pos)
    fixed = np.ones(N, bool)  # fixed spheres keep their velocities and spins
    ang_vel = np.zeros((N, 3))
    ang_vel[0, 2] = w
    st = O.make_state(pos, rad, vel=vel, ang_vel=ang_vel, fixed=fixed, dtype=F64)
    law = O.Law("springfriction", young_eff=np.array([[KN]]), mu_eff=np.array([[MU]]), kt_ratio=KT_RATIO)
    P = O.Params(dim=3, dt=2e-4, domain=_dom("periodic", box=12.0), law=law, collider="neighborlist",
                 rot="verletspiral")
    P.nl_cutoff, P.nl_skin, P.nl_max_neighbors = 1.0, 0.02, 16
    P.nl_secondary = secondary
    P.nl_secondary_cell_size, P.nl_secondary_search_range = 1.02, 1
    nls = O.nl_init(st, P, history=True)
    return st, P, nls, w * r0


def _run_sliding(secondary: str, steps=300):
    st, P, nls, vt = _sliding_scene(secondary)
    kt = KT_RATIO * KN
    fn = KN * 0.05
    errs, slots = [], []
    for n in range(1, steps + 1):
        O.step_once(st, P, nls)
        row = list(nls.nl[0])
        slots.append(row.index(1))
        ft = math.hypot(st.force[0, 1], st.force[0, 2])
        expect = min(kt * vt * P.dt * n, MU * fn)
        errs.append(abs(ft - expect) / expect)
        # the spring opposes the slip of sphere 0's contact point (+y)
        assert st.force[0, 1] < 0
    return np.array(errs), slots, nls


@pytest.mark.parametrize("secondary", ["naive", "celllist"])
def test_remap_keeps_sliding_spring_across_slot_changes(secondary):
    errs, slots, nls = _run_sliding(secondary)
    # the rebuilds did move sphere 1 to other slots of sphere 0's row
    assert len(set(slots)) >= 3, slots
    assert nls.n_build > 20
    # |F_t| = k_t v_t t, then mu F_n, continuous through every rebuild
    assert errs.max() < 1e-9, errs.max()


def test_remap_pin_fails_for_slot_keeping_remap(monkeypatch):
    """The pin above catches a remap that keeps springs by slot index."""

    def by_slot(xi, live, old_nl, new_nl):
        keep = new_nl >= 0
        return np.where(keep[..., None], xi, 0.0), live & keep

    monkeypatch.setattr(O, "remap_history", by_slot)
    errs, slots, _ = _run_sliding("naive")
    assert len(set(slots)) >= 3
    assert errs.max() > 1e-3


def _frictional_packing(skin, steps, seed=5):
    rng = np.random.default_rng(seed)
    n_side = 5
    N = n_side**3
    box = n_side * 0.97
    g = (np.arange(n_side) + 0.5) * 0.97
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(N, 3)
    pos = pos + rng.uniform(-0.04, 0.04, (N, 3))
    vel = rng.normal(0.0, 1.0, (N, 3))
    vel[:, 0] += 0.6 * np.sin(2 * math.pi * pos[:, 1] / box)  # shear
    rad = rng.uniform(0.49, 0.52, N)
    st = O.make_state(pos, rad, vel=vel, dtype=F64)
    law = O.Law("springfriction", young_eff=np.array([[KN]]), mu_eff=np.array([[MU]]), kt_ratio=KT_RATIO,
                damping=0.1)
    P = O.Params(dim=3, dt=1e-4, domain=_dom("periodic", box=box), law=law, collider="neighborlist")
    P.nl_cutoff, P.nl_skin, P.nl_max_neighbors = 1.04, skin, 48
    P.nl_secondary = "celllist"
    P.nl_secondary_cell_size, P.nl_secondary_search_range = 1.04 + skin, 1
    nls = O.nl_init(st, P, history=True)
    builds = 0
    for _ in range(steps):
        O.step_once(st, P, nls)
    builds = nls.n_build
    return st, nls, builds


def test_friction_forces_independent_of_skin():
    """A sheared frictional packing: the skin changes how often the list is
    rebuilt (and with it every row's order and each spring's slot), not the
    contacts' forces -- up to the order of the float64 sums."""
    a, na, ba = _frictional_packing(0.01, 120)
    b, nb, bb = _frictional_packing(0.4, 120)
    assert ba > 10 * bb  # many more rebuilds, hence remaps, with the small skin
    assert int(na.live.sum()) > 50  # many sliding contacts carried the whole way
    scale = np.abs(b.force).max()
    np.testing.assert_allclose(a.force, b.force, rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(a.torque, b.torque, rtol=0, atol=1e-9 * np.abs(b.torque).max())
    np.testing.assert_allclose(a.pos, b.pos, rtol=0, atol=1e-12)


# --------------------------------------------------------------------------
# hertzfriction: Hertz normal force + Mindlin tangential stiffness
# --------------------------------------------------------------------------
def _mindlin_moduli(E1, nu1, E2, nu2):
    """Textbook effective moduli (Johnson, Contact Mechanics, ch. 4 and 7):
    1/E* = (1 - nu1^2)/E1 + (1 - nu2^2)/E2,
    1/G* = 2 (2 - nu1)(1 + nu1)/E1 + 2 (2 - nu2)(1 + nu2)/E2."""
    Es = 1.0 / ((1 - nu1 * nu1) / E1 + (1 - nu2 * nu2) / E2)
    Gs = 1.0 / (2 * (2 - nu1) * (1 + nu1) / E1 + 2 * (2 - nu2) * (1 + nu2) / E2)
    return Es, Gs


def _hf_pair(r=(0.5, 0.7), delta=0.03, E=(1e4, 2.5e4), nu=(0.3, 0.2), mu=0.4, damping=0.0):
    d = r[0] + r[1] - delta
    st = O.make_state([[0.0, 0.0, 0.0], [d, 0.0, 0.0]], list(r), mat_id=[0, 1], dtype=F64)
    mu_eff = np.full((2, 2), mu)
    law = O.Law("hertzfriction", young=np.array(E, F64), poisson=np.array(nu, F64), mu_eff=mu_eff,
                damping=damping)
    return st, law


def test_hertzfriction_normal_force_is_hertz():
    st, law = _hf_pair()
    Es, _ = _mindlin_moduli(1e4, 0.3, 2.5e4, 0.2)
    Rs = 0.5 * 0.7 / 1.2
    delta = 0.03
    f, t, xi, c = O.law_force(law, st, np.array([0, 1]), np.array([1, 0]), _dom("free"), 1e-4, np.zeros((2, 3)))
    Fn = 4.0 / 3.0 * Es * math.sqrt(Rs) * delta**1.5
    assert c.all()
    assert -f[0, 0] == pytest.approx(Fn, rel=1e-12)
    np.testing.assert_allclose(f[1], -f[0], rtol=1e-13)
    assert np.all(t == 0)


@pytest.mark.parametrize("vt", [0.05, 3.0])
def test_hertzfriction_tangential_growth_and_cap(vt):
    """Constant slip v_t with a constant overlap: |F_t| = k_t v_t t with the
    Mindlin stiffness k_t = 8 G* sqrt(R* delta), until it saturates at mu F_n."""
    st, law = _hf_pair()
    st.vel[0] = [0.0, 0.0, vt]
    Es, Gs = _mindlin_moduli(1e4, 0.3, 2.5e4, 0.2)
    Rs = 0.5 * 0.7 / 1.2
    delta = 0.03
    kt = 8.0 * Gs * math.sqrt(Rs * delta)
    Fn = 4.0 / 3.0 * Es * math.sqrt(Rs) * delta**1.5
    dt = 1e-4
    xi = np.zeros((1, 3))
    saturated = False
    for n in range(1, 400):
        f, t, xi, c = O.law_force(law, st, np.array([0]), np.array([1]), _dom("free"), dt, xi)
        ft = -f[0, 2]
        expect = min(kt * vt * dt * n, 0.4 * Fn)
        saturated |= kt * vt * dt * n > 0.4 * Fn
        assert ft == pytest.approx(expect, rel=1e-10), n
        assert -f[0, 0] == pytest.approx(Fn, rel=1e-12)
        # torque of the tangential force at the contact point r_c = +r_0 x
        assert t[0, 1] == pytest.approx(-0.5 * f[0, 2], rel=1e-12)
    assert saturated == (vt > 1.0)


def test_hertzfriction_projection_reset_and_damping():
    st, law = _hf_pair(damping=0.2)
    # a spring with a normal (x) part is projected onto the contact plane
    f, t, xi, c = O.law_force(law, st, np.array([0]), np.array([1]), _dom("free"), 1e-4,
                              np.array([[0.02, 0.001, -0.002]]))
    assert xi[0, 0] == pytest.approx(0.0, abs=1e-16)
    np.testing.assert_allclose(xi[0, 1:], [0.001, -0.002], rtol=1e-12)
    # normal damping: 2 zeta sqrt(k_n m_eff) with the tangent stiffness k_n = 4/3 E* sqrt(R* delta)
    st.vel[0] = [1.0, 0.0, 0.0]  # approaching
    Es, _ = _mindlin_moduli(1e4, 0.3, 2.5e4, 0.2)
    Rs, delta = 0.5 * 0.7 / 1.2, 0.03
    kn = 4.0 / 3.0 * Es * math.sqrt(Rs * delta)
    m_eff = 0.5
    f, *_ = O.law_force(law, st, np.array([0]), np.array([1]), _dom("free"), 1e-4, np.zeros((1, 3)))
    assert -f[0, 0] == pytest.approx(kn * delta + 2 * 0.2 * math.sqrt(kn * m_eff) * 1.0, rel=1e-12)
    # apart: no force, spring reset
    st2, _ = _hf_pair(delta=-0.01)
    f, t, xi, c = O.law_force(law, st2, np.array([0]), np.array([1]), _dom("free"), 1e-4, np.ones((1, 3)))
    assert not c[0] and np.all(f == 0) and np.all(xi == 0)


def test_hertzfriction_collision_dissipates():
    """An oblique damped frictional Hertz collision loses energy and spins
    the spheres up; undamped and frictionless it keeps it (to O(dt^2))."""

    def collide(mu, damping):
        st = O.make_state([[0.0, 0.0, 0.0], [1.15, 0.35, 0.0]], [0.6, 0.6],
                          vel=[[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]], dtype=F64)
        law = O.Law("hertzfriction", young=np.array([1e4]), poisson=np.array([0.3]),
                    mu_eff=np.array([[mu]]), damping=damping)
        P = O.Params(dim=3, dt=5e-5, domain=_dom("free"), law=law, collider="neighborlist")
        P.nl_cutoff, P.nl_skin, P.nl_max_neighbors, P.nl_secondary = 1.2, 0.3, 2, "naive"
        nls = O.nl_init(st, P, history=True)
        e0 = O.kinetic_energy(st)
        for _ in range(4000):
            O.step_once(st, P, nls)
        return e0, O.kinetic_energy(st), float(np.linalg.norm(st.ang_vel))

    e0, e1, w = collide(0.5, 0.2)
    assert e1 < 0.95 * e0 and w > 0.1
    e0, e1, w = collide(0.0, 0.0)
    assert e1 == pytest.approx(e0, rel=1e-5) and w == pytest.approx(0.0, abs=1e-12)


# --------------------------------------------------------------------------
# reflecting wall under gravity: the time-of-contact quadratic
# --------------------------------------------------------------------------
@pytest.mark.parametrize("e", [1.0, 0.5])
@pytest.mark.parametrize("h0,v0", [(2.0, 0.0), (1.3, -3.0), (0.9, 2.5)])
def test_reflect_gravity_bounce_exact(e, h0, v0):
    """A sphere falls under gravity onto the floor z = anchor + r.  Velocity
    Verlet is exact for a constant force, and the wall correction rewinds the
    step to the time of contact (alpha of the quadratic), reverses the
    normal velocity there with restitution e and replays the rest of the
    step, so every step after the bounce lies on the exact trajectory:
    z(t) = r + (-e v_c)(t - t_c) - g/2 (t - t_c)^2."""
    g, r, dt = 9.81, 0.1, 1e-3
    st = O.make_state([[5.0, 5.0, r + h0]], [r], vel=[[0.3, 0.0, v0]], dtype=F64)
    law = O.Law("spring", young_eff=np.array([[1e4]]))
    P = O.Params(dim=3, dt=dt, domain=_dom("reflect", box=10.0, restitution=e), law=law, rot="")
    P.gravity = np.array([0.0, 0.0, -g])
    st.force[0] = st.mass[0] * P.gravity
    # exact first contact: h0 + v0 t - g t^2 / 2 = 0
    t_c = (v0 + math.sqrt(v0 * v0 + 2 * g * h0)) / g
    v_c = v0 - g * t_c
    n_steps = int(t_c / dt) + 40
    # the bounce is corrected at the start of the step after the one that
    # crossed the wall; keep away from a contact on a step boundary
    assert abs(t_c / dt - round(t_c / dt)) > 1e-3
    for _ in range(n_steps):
        O.step_once(st, P)
    # the correction of the last step's crossing is applied by the next
    # step's domain.apply: compare after it
    O.domain_apply(st, P)
    t = n_steps * dt
    tau = t - t_c
    z = r + (-e * v_c) * tau - 0.5 * g * tau * tau
    vz = -e * v_c - g * tau
    assert st.pos[0, 2] == pytest.approx(z, abs=1e-9)
    assert st.vel[0, 2] == pytest.approx(vz, abs=1e-9)
    assert st.pos[0, 0] == pytest.approx(5.0 + 0.3 * t, abs=1e-12)
    assert st.vel[0, 0] == 0.3


def test_toc_fraction_known_answer():
    """alpha of the quadratic for a crossing with acceleration: the start
    of the step, the exact crossing time and the end of the step."""
    dt, a = 0.1, -4.0
    x_start, v_start, lo = 1.0, -5.0, 0.5
    # Verlet (exact for constant a): x(s) = x_start + v_start s + a s^2 / 2
    x_end = x_start + v_start * dt + 0.5 * a * dt * dt
    v_end = v_start + a * dt
    s_c = (-v_start - math.sqrt(v_start * v_start - 2 * a * (x_start - lo))) / a
    alpha = O.verlet_collision_fraction(np.array([v_end]), np.array([a]), np.array([lo - x_end]), np.array([1.0]),
                                        dt)
    assert 0 < s_c < dt
    assert alpha[0] == pytest.approx(s_c / dt, rel=1e-12)
    # a ceiling at hi (wall sign -1), moving up against gravity
    x_start, v_start, hi = 0.0, 6.0, 0.5
    x_end = x_start + v_start * dt + 0.5 * a * dt * dt
    v_end = v_start + a * dt
    s_c = (v_start - math.sqrt(v_start * v_start + 2 * a * (hi - x_start))) / (-a)
    alpha = O.verlet_collision_fraction(np.array([v_end]), np.array([a]), np.array([x_end - hi]), np.array([-1.0]),
                                        dt)
    assert alpha[0] == pytest.approx(s_c / dt, rel=1e-12)


# --------------------------------------------------------------------------
# symplectic Euler (direct_euler.py:62-66)
# --------------------------------------------------------------------------
def test_euler_closed_form_under_gravity():
    """v_n = v_0 + n g dt and x_n = x_0 + n dt v_0 + g dt^2 n (n + 1) / 2
    (velocity first, then position with the new velocity)."""
    g = np.array([0.5, -9.81, 2.0])
    v0 = np.array([1.0, 3.0, -0.5])
    x0 = np.array([0.1, 0.2, 0.3])
    st = O.make_state([x0], [0.1], vel=[v0], mass=[2.5], dtype=F64)
    law = O.Law("spring", young_eff=np.array([[1e4]]))
    P = O.Params(dim=3, dt=2e-3, domain=_dom("free"), law=law, lin="euler", rot="")
    P.gravity = g
    n = 300
    for _ in range(n):
        O.step_once(st, P)
    dt = P.dt
    np.testing.assert_allclose(st.vel[0], v0 + n * g * dt, rtol=1e-12)
    np.testing.assert_allclose(st.pos[0], x0 + n * dt * v0 + g * dt * dt * n * (n + 1) / 2, rtol=1e-11)


# --------------------------------------------------------------------------
# Cundall-Strack rolling resistance (cundall_strack.py:187-192)
# --------------------------------------------------------------------------
def _cs_pair(w_i, w_j, mu_r=0.05, r=(0.5, 0.3)):
    d = r[0] + r[1] - 0.02
    st = O.make_state([[0.0, 0.0, 0.0], [d, 0.0, 0.0]], list(r), ang_vel=[w_i, w_j], dtype=F64)
    law = O.Law("cundallstrack", young=np.array([1e4]), poisson=np.array([0.3]), e=np.array([0.7]),
                mu=np.array([0.4]), mu_r=np.array([mu_r]))
    return st, law


def test_cs_rolling_torque_along_normal_spin():
    """Relative spin about the normal: no slip at the contact point, so the
    only torque is the rolling resistance -mu_r R* F_n w_hat."""
    E, r0, r1 = 1e4, 0.5, 0.3
    st, law = _cs_pair([3.0, 0.0, 0.0], [-1.0, 0.0, 0.0])
    f, t, _, _ = O.law_force(law, st, np.array([0]), np.array([1]), _dom("free"), 1e-4, None)
    kn = 2 * E * r0 * E * r1 / (E * r0 + E * r1)
    Fn = kn * 0.02
    Rs = r0 * r1 / (r0 + r1)
    np.testing.assert_allclose(f[0], [-Fn, 0.0, 0.0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(t[0], [-0.05 * Rs * Fn, 0.0, 0.0], rtol=1e-12, atol=1e-14)


def test_cs_rolling_torque_with_sliding():
    """Spin about z: the contact point of sphere 0 slides at w r_0 along +y;
    the torque is r_c x F (from the capped tangential dashpot) plus the
    rolling term along -z."""
    E, nu, e, mu, mu_r, r0, r1 = 1e4, 0.3, 0.7, 0.4, 0.05, 0.5, 0.3
    w = 2.0
    st, law = _cs_pair([0.0, 0.0, w], [0.0, 0.0, 0.0], mu_r=mu_r)
    f, t, _, _ = O.law_force(law, st, np.array([0]), np.array([1]), _dom("free"), 1e-4, None)
    G = E / (2 * (1 + nu))
    kn = 2 * E * r0 * E * r1 / (E * r0 + E * r1)
    kt = 2 * G * r0 * G * r1 / (G * r0 + G * r1)
    beta = -math.log(e) / math.sqrt(math.pi**2 + math.log(e) ** 2)
    gt = 2 * beta * math.sqrt(kt * 0.5)
    Fn = kn * 0.02
    Ft = min(gt * w * r0, mu * Fn)
    Rs = r0 * r1 / (r0 + r1)
    np.testing.assert_allclose(f[0], [-Fn, -Ft, 0.0], rtol=1e-12, atol=1e-12)
    # r_c = +r0 x:  x cross (-Fn x - Ft y) = -Ft z
    np.testing.assert_allclose(t[0], [0.0, 0.0, -r0 * Ft - mu_r * Rs * Fn], rtol=1e-12, atol=1e-14)
    # no relative spin: no rolling term
    st2, law2 = _cs_pair([0.0, 0.0, w], [0.0, 0.0, w], mu_r=mu_r)
    _, t2, _, _ = O.law_force(law2, st2, np.array([0]), np.array([1]), _dom("free"), 1e-4, None)
    v_slip = w * r0 + w * r1  # sphere 1's contact point moves along -y
    Ft2 = min(gt * v_slip, mu * Fn)
    np.testing.assert_allclose(t2[0], [0.0, 0.0, -r0 * Ft2], rtol=1e-12, atol=1e-14)
