"""Clumps (rigid bodies of spheres) in the oracle, on the CPU: the clump
path of the collider and ForceManager (forces at the members' lab positions,
same-clump pairs masked, the lever torque R(q) pos_p x F, the segment sum
over the clump, force_manager.py:341-425; naive.py:210-233) pinned against
what the mechanics fixes, so that a dropped lever term, a wrong sign, a
missing segment sum or gravity counted per member fails:

* a clump of single spheres is the sphere path bit for bit;
* the force and torque on a dimer pressed by a sphere equal the float64
  sum of the pair forces and their lever torques, on every member;
* gravity on a clump of three is M g, not 3 M g;
* two spinning dimers colliding in free space keep the total linear and
  angular momentum (orbital + spin) while the spin changes by O(1);
* the free domain's box bounds |pos_p| + rad; the kinetic energy counts each
  clump once;
* the reflective domain: singleton clumps bounce as spheres do, and a tilted
  dimer striking the floor with one sphere leaves with its contact point's
  normal velocity reversed times e, spinning, its kinetic energy kept at
  e = 1 (a wrong effective mass or lever sign breaks both).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jdem_oracle as O

F64 = np.float64


def _quat_z(theta):
    return np.array([np.cos(theta / 2), 0.0, 0.0, np.sin(theta / 2)])


def _dimers(specs, r=0.5, m=1.0, d=0.5, dim=3):
    """Dimers: two spheres of radius r at +-d along the body x axis.  specs:
    (center, velocity, lab angular velocity, quaternion) per dimer."""
    pos, pp, vel, w, q, inertia, clump = [], [], [], [], [], [], []
    M = 2 * m
    Ib = 2 * (0.4 * m * r * r)
    Ip = Ib + 2 * m * d * d
    for k, (c, v, om, qq) in enumerate(specs):
        for s in (-1.0, 1.0):
            pos.append(c)
            pp.append([s * d, 0.0, 0.0][:dim])
            vel.append(v)
            w.append(om)
            q.append(qq)
            inertia.append([Ib, Ip, Ip] if dim == 3 else [Ip])
            clump.append(k)
    n = len(pos)
    st = O.make_state(np.array(pos), np.full(n, r), vel=np.array(vel), mass=np.full(n, M),
                      inertia=np.array(inertia), ang_vel=np.array(w), q=np.array(q), dtype=F64)
    st.clump = np.array(clump)
    st.pos_p = np.array(pp, F64)
    return st


def _params(dim, law="spring", k=1e3, domain="free", box=20.0, collider="naive", dt=1e-5):
    dom = O.Domain(domain, np.full(dim, box, F64), np.zeros(dim, F64), 1.0)
    lw = O.Law(law, young_eff=np.full((1, 1), k, F64))
    return O.Params(dim=dim, dt=dt, domain=dom, law=lw, collider=collider, cell_size=1.1, search_range=1)


def test_singleton_clumps_equal_spheres_bitwise():
    rng = np.random.Generator(np.random.PCG64(3))
    N, D = 300, 3
    pos = rng.uniform(0, 9.0, (N, D)).astype(np.float32)
    st = O.make_state(pos, rng.uniform(0.4, 0.6, N).astype(np.float32),
                      vel=rng.standard_normal((N, D)).astype(np.float32),
                      ang_vel=rng.standard_normal((N, D)).astype(np.float32))
    P = O.Params(dim=D, dt=1e-3, domain=O.Domain("periodic", np.full(D, 9.0, np.float32), np.zeros(D, np.float32)),
                 law=O.Law("cundallstrack", young=np.array([1e4], np.float32), poisson=np.array([0.3], np.float32),
                           e=np.array([0.7], np.float32), mu=np.array([0.4], np.float32),
                           mu_r=np.array([0.05], np.float32)),
                 gravity=np.array([0.0, -1.0, 0.0], np.float32))
    c = st.copy()
    c.clump = np.arange(N)
    c.pos_p = np.zeros((N, D), np.float32)
    a = O.step(st, P, 10)
    b = O.step(c, P, 10)
    for f in ("pos", "vel", "ang_vel", "q", "force", "torque"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("dim", [3, 2])
def test_dimer_force_and_torque_are_the_lever_sums(dim):
    """A sphere pressed against one end of a rotated dimer; spring law."""
    th = 0.3
    qd = _quat_z(th)
    c0 = np.zeros(dim)
    st = _dimers([(c0, np.zeros(dim), np.zeros(3 if dim == 3 else 1), qd)], dim=dim)
    # the sphere overlaps the +x member (lab position R(th) (0.5, 0))
    tip = np.array([0.5 * np.cos(th), 0.5 * np.sin(th), 0.0][:dim])
    ball = tip + np.array([0.8, 0.3, 0.0][:dim])
    sp = O.make_state(ball[None], [0.5], dtype=F64)
    n = 3
    A = 3 if dim == 3 else 1
    full = O.make_state(np.vstack([st.pos, ball[None]]), np.full(n, 0.5), mass=np.r_[st.mass, 1.0],
                        inertia=np.vstack([st.inertia, sp.inertia]), q=np.vstack([st.q, sp.q]), dtype=F64)
    full.clump = np.array([0, 0, 1])
    full.pos_p = np.vstack([st.pos_p, np.zeros((1, dim))])
    P = _params(dim)
    O.compute_force(full, P)
    O.force_manager_apply(full, P)
    # float64 by hand: spring k d n between the +x member and the ball
    k = 1e3
    rvec = tip - ball
    dist = np.linalg.norm(rvec)
    f_on_member = k * (1.0 - dist) * rvec / dist
    lever = tip - c0
    tq = lever[0] * f_on_member[1] - lever[1] * f_on_member[0]
    for m in (0, 1):  # both members carry the clump's totals
        np.testing.assert_allclose(full.force[m], f_on_member, rtol=1e-12)
        if dim == 3:
            np.testing.assert_allclose(full.torque[m], [0.0, 0.0, tq], rtol=1e-12, atol=1e-12 * abs(tq))
        else:
            np.testing.assert_allclose(full.torque[m], [tq], rtol=1e-12)
    np.testing.assert_allclose(full.force[2], -f_on_member, rtol=1e-12)
    assert abs(tq) > 1.0
    assert full.torque.shape[1] == A


def test_gravity_on_a_clump_is_total_mass_times_g():
    rng = np.random.Generator(np.random.PCG64(1))
    st = O.make_state(np.zeros((3, 3)), [0.3, 0.3, 0.3], mass=np.full(3, 7.0), dtype=F64)
    st.clump = np.zeros(3, np.int64)
    st.pos_p = rng.uniform(-0.2, 0.2, (3, 3))
    P = _params(3)
    P.gravity = np.array([0.0, -9.81, 0.0])
    O.compute_force(st, P)
    O.force_manager_apply(st, P)
    np.testing.assert_allclose(st.force, np.tile([0.0, -9.81 * 7.0, 0.0], (3, 1)), rtol=1e-14)
    # gravity acts at the centre of mass: no torque
    np.testing.assert_allclose(st.torque, 0.0, atol=1e-14)


def _momenta(st):
    """Total linear momentum and angular momentum about the origin, one term
    per clump: M X x V + R I_b R^T w."""
    P = np.zeros(3)
    L = np.zeros(3)
    for c in np.unique(st.clump):
        i = np.nonzero(st.clump == c)[0][0]
        M, X, V, w = st.mass[i], st.pos[i], st.vel[i], st.ang_vel[i]
        wb = O.q_rotate_back(st.q[i:i + 1], w[None])[0]
        Lspin = O.q_rotate(st.q[i:i + 1], (st.inertia[i] * wb)[None])[0]
        P += M * V
        L += M * np.cross(X, V) + Lspin
    return P, L


def _spin(st):
    out = np.zeros(3)
    for c in np.unique(st.clump):
        i = np.nonzero(st.clump == c)[0][0]
        wb = O.q_rotate_back(st.q[i:i + 1], st.ang_vel[i][None])[0]
        out += O.q_rotate(st.q[i:i + 1], (st.inertia[i] * wb)[None])[0]
    return out


def test_colliding_dimers_conserve_momentum_and_angular_momentum():
    """Off-centre collision of two dimers (velocity Verlet + VV-SPIRAL,
    float64): linear momentum to rounding, angular momentum to O(dt^2),
    while the spins change by O(1) of the total -- a missing or wrong-sign
    lever torque breaks the balance."""
    st = _dimers([
        (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 0.5]), _quat_z(0.4)),
        (np.array([2.4, 0.6, 0.1]), np.array([-1.0, 0.0, 0.0]), np.array([0.3, 0.0, -0.2]), _quat_z(-0.9)),
    ])
    P = _params(3, k=2e3, dt=1e-4)
    P0, L0 = _momenta(st)
    S0 = _spin(st)
    out = O.step(st, P, 12000)
    P1, L1 = _momenta(out)
    S1 = _spin(out)
    # the bodies did collide and spin up / down
    assert np.linalg.norm(S1 - S0) > 0.1 * np.linalg.norm(L0)
    np.testing.assert_allclose(P1, P0, atol=1e-12)
    assert np.linalg.norm(L1 - L0) < 1e-5 * np.linalg.norm(L0)
    # the members stay rigid: lab distance of the two members = 2 d
    lab = O.lab_pos(out)
    for c in (0, 1):
        a, b = lab[out.clump == c]
        assert abs(np.linalg.norm(a - b) - 1.0) < 1e-12


def test_free_box_and_kinetic_energy_of_clumps():
    st = _dimers([(np.array([1.0, 2.0, 3.0]), np.array([0.5, 0.0, 0.0]), np.array([0.0, 0.0, 2.0]),
                   _quat_z(0.0))])
    P = _params(3)
    O.domain_apply(st, P)
    # bounding radius |pos_p| + r = 1.0 around the common centre
    np.testing.assert_allclose(P.domain.anchor, [0.0, 1.0, 2.0])
    np.testing.assert_allclose(P.domain.box, [2.0, 2.0, 2.0])
    # 1/2 M V^2 + 1/2 I_zz w^2, counted once per clump
    M, Izz = 2.0, 2 * (0.4 * 0.25) + 2 * 0.25
    assert O.kinetic_energy(st) == pytest.approx(0.5 * M * 0.25 + 0.5 * Izz * 4.0, rel=1e-14)


def test_singleton_clumps_reflect_like_spheres():
    """The clump form of the reflective domain reduces to the sphere form."""
    rng = np.random.Generator(np.random.PCG64(8))
    N, D = 200, 3
    L = 8.0
    pos = rng.uniform(-0.3, L + 0.3, (N, D)).astype(np.float32)  # some past the walls
    st = O.make_state(pos, np.full(N, 0.4, np.float32), vel=rng.standard_normal((N, D)).astype(np.float32),
                      ang_vel=rng.standard_normal((N, D)).astype(np.float32))
    st.force = rng.standard_normal((N, D)).astype(np.float32)
    dom = O.Domain("reflect", np.full(D, L, np.float32), np.zeros(D, np.float32), 0.8)
    P = O.Params(dim=D, dt=1e-2, domain=dom, law=O.Law("spring", young_eff=np.ones((1, 1), np.float32)))
    c = st.copy()
    c.clump = np.arange(N)
    c.pos_p = np.zeros((N, D), np.float32)
    O.domain_apply(st, P)
    O.domain_apply(c, P)
    for f in ("pos", "vel", "ang_vel", "q"):
        assert np.array_equal(getattr(st, f), getattr(c, f)), f


@pytest.mark.parametrize("dim", [3, 2])
def test_tilted_dimer_bounce_is_elastic_at_e1(dim):
    """A tilted dimer whose lower sphere reaches the floor: with e = 1 the
    impulse keeps the kinetic energy (translation + rotation), reverses the
    contact point's normal velocity and spins the body; with e = 0.5 the
    contact point leaves at half its approach speed."""
    for e in (1.0, 0.5):
        th = 0.6
        c = np.array([3.0, 0.5 + 0.45 * np.sin(th) - 0.01, 3.0][:dim])  # lower sphere 0.01 into the floor
        w0 = np.zeros(3 if dim == 3 else 1)
        st = _dimers([(c, np.array([0.3, -1.0, 0.0][:dim]), w0, _quat_z(th))], d=0.45, dim=dim)
        st.pos = np.tile(c, (2, 1))
        dom = O.Domain("reflect", np.full(dim, 6.0), np.zeros(dim), e)
        P = O.Params(dim=dim, dt=1e-3, domain=dom, law=O.Law("spring", young_eff=np.ones((1, 1))))
        ke0 = O.kinetic_energy(st)
        lab0 = O.lab_pos(st)
        low = int(np.argmin(lab0[:, 1]))
        r0 = O.pos_p_rot(st)[low:low + 1]  # the lever at the impulse (before the orientation correction)
        vp0 = st.vel[low] + O.cross_w_r(st.ang_vel[low:low + 1], r0)[0]
        O.domain_apply(st, P)
        vp1 = st.vel[low] + O.cross_w_r(st.ang_vel[low:low + 1], r0)[0]
        # the members stay one body
        assert np.array_equal(st.vel[0], st.vel[1]) and np.array_equal(st.ang_vel[0], st.ang_vel[1])
        assert np.abs(st.ang_vel).max() > 0.1
        assert vp1[1] == pytest.approx(-e * vp0[1], rel=1e-12)
        if e == 1.0:
            assert O.kinetic_energy(st) == pytest.approx(ke0, rel=1e-12)
