"""Routed laws -- ForceRouter (forces/router.py:26-339) and LawCombiner
(forces/law_combiner.py:22-191) -- on the CPU: the host's routing layout
against the reference's rules, and the oracle's routed force and energy
against values computed here from the textbook pair laws (float64 brute
force over all pairs), so that a swapped species index, a lower-triangle
entry, a missing combiner term or a spring kept on a spring-free entry
fails a test.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import jaxdem_amd as jd
from jaxdem_amd.forces import ForceRouter, LawCombiner, route_layout
from oracle import jdem_oracle as O

F32 = np.float32


def _names(terms):
    return [None if t is None else t.type_name for t in terms]


def test_route_layout_follows_router_rules():
    spring, hertz = jd.ForceModel.create("spring"), jd.ForceModel.create("hertz")
    fr = jd.ForceModel.create("springfriction")
    # from_dict symmetrises; unnamed pairs get an empty combiner
    r = ForceRouter.from_dict(3, {(0, 0): spring, (2, 1): hertz})
    S, off, terms = route_layout(r)
    assert S == 3 and off.tolist() == list(range(10))
    assert _names(terms) == ["spring", None, None, None, None, "hertz", None, "hertz", None]
    # a table built by hand: the upper triangle serves both orders (router.py:113-131)
    t = ((spring, hertz), (fr, spring))
    S, off, terms = route_layout(ForceRouter(table=t))
    assert _names(terms) == ["spring", "hertz", "hertz", "spring"]
    # a combiner starts from its zeros; nested combiners flatten
    c = LawCombiner(laws=(spring, LawCombiner(laws=(hertz, fr))))
    S, off, terms = route_layout(c)
    assert S == 0 and off.tolist() == [0, 5]
    assert _names(terms) == [None, "spring", None, "hertz", "springfriction"]
    assert c.requires_history and not LawCombiner(laws=(spring, hertz)).requires_history
    assert c.required_material_properties == ("mu_eff", "poisson", "young", "young_eff")
    with pytest.raises(NotImplementedError):
        route_layout(LawCombiner(laws=(fr, jd.ForceModel.create("hertzfriction"))))
    with pytest.raises(NotImplementedError):
        route_layout(LawCombiner(laws=(r,)))
    # registry keys and the reference's constructor path
    assert jd.ForceModel.create("forcerouter", table=r.table).table == r.table
    assert type(jd.ForceModel.create("LawCombiner", laws=(spring,))) is LawCombiner


def _scene(N=220, dim=3, seed=4, n_species=2):
    rng = np.random.Generator(np.random.PCG64(seed))
    rad = rng.uniform(0.4, 0.6, N)
    vol = (4.0 / 3.0 * math.pi * rad**3) if dim == 3 else (math.pi * rad**2)
    L = (vol.sum() / 0.5) ** (1.0 / dim)
    pos = rng.uniform(0.0, L, (N, dim))
    st = O.make_state(pos.astype(F32), rad.astype(F32), vel=rng.standard_normal((N, dim)).astype(F32))
    st.species = rng.integers(0, n_species, N)
    dom = O.Domain("periodic", np.full(dim, L, F32), np.zeros(dim, F32), 1.0)
    return st, dom


E, NU, KS = 1.0e4, 0.3, 2.0e3


def _law(kind):
    one = np.ones(1, F32)
    if kind == "spring":
        return O.Law("spring", young_eff=np.full((1, 1), KS, F32))
    if kind == "hertz":
        return O.Law("hertz", young=E * one, poisson=NU * one)
    if kind == "springfriction":
        return O.Law("springfriction", young_eff=np.full((1, 1), KS, F32), mu_eff=np.full((1, 1), 0.4, F32))
    return O.Law("none")


def _routed(n_species, entries):
    off, terms = [0], []
    for e in entries:
        terms += [_law(k) for k in e]
        off.append(len(terms))
    return O.Law("routed", n_species=n_species, route_off=np.asarray(off), terms=tuple(terms))


def _params(st, law, dom):
    return O.Params(dim=st.dim, dt=1e-4, domain=dom, law=law, collider="naive")


def _textbook(st, dom, kind_of_pair):
    """Float64 forces and energy summed over all pairs, the law of each
    pair from ``kind_of_pair(s_i, s_j)``: spring k d n (energy k d^2 / 2),
    Hertz 4/3 E* sqrt(R*) d^1.5 n (energy 2/5 of k d^2.5)."""
    pos = st.pos.astype(np.float64)
    N, D = pos.shape
    Lb = dom.box.astype(np.float64)
    Estar = 1.0 / (2.0 * (1.0 - NU * NU) / E)
    F = np.zeros((N, D))
    energy = 0.0
    for i in range(N):
        d = pos[i] - pos
        d -= Lb * np.round(d / Lb)
        r = np.linalg.norm(d, axis=1)
        R = st.rad[i].astype(np.float64) + st.rad.astype(np.float64)
        delta = np.maximum(R - r, 0.0)
        delta[i] = 0.0
        for j in np.nonzero(delta > 0)[0]:
            kind = kind_of_pair(st.species[i], st.species[j])
            n = d[j] / r[j]
            if kind == "spring":
                F[i] += KS * delta[j] * n
                energy += 0.5 * 0.5 * KS * delta[j] ** 2
            elif kind == "hertz":
                Rs = st.rad[i] * st.rad[j] / (st.rad[i] + st.rad[j])
                k = 4.0 / 3.0 * Estar * math.sqrt(Rs)
                F[i] += k * delta[j] ** 1.5 * n
                energy += 0.5 * 0.4 * k * delta[j] ** 2.5
            elif kind == "spring+hertz":
                Rs = st.rad[i] * st.rad[j] / (st.rad[i] + st.rad[j])
                k = 4.0 / 3.0 * Estar * math.sqrt(Rs)
                F[i] += (KS * delta[j] + k * delta[j] ** 1.5) * n
                energy += 0.5 * (0.5 * KS * delta[j] ** 2 + 0.4 * k * delta[j] ** 2.5)
    return F, energy


@pytest.mark.parametrize("dim", [3, 2])
def test_router_selects_the_law_of_the_species_pair(dim):
    """S = 2: (0, 0) spring, (0, 1) nothing, (1, 1) spring + Hertz."""
    st, dom = _scene(dim=dim)
    law = _routed(2, [["spring"], ["none"], ["none"], ["none", "spring", "hertz"]])
    P = _params(st, law, dom)
    O.compute_force(st, P)

    def kind(a, b):
        return {(0, 0): "spring", (1, 1): "spring+hertz"}.get((a, b), "none")

    F, energy = _textbook(st, dom, kind)
    assert np.abs(F).max() > 1.0
    np.testing.assert_allclose(st.force, F, rtol=2e-4, atol=2e-3 * np.abs(F).max())
    e = O.potential_energy(st, P, O.naive_candidates(st.N))
    assert e == pytest.approx(energy, rel=1e-4)
    # the species pairs that get no law feel nothing: a particle of species 1
    # whose contacts are all of species 0 is force-free
    assert np.count_nonzero(~np.any(st.force != 0, axis=1)) > 0


def test_routed_index_is_row_species_times_S_plus_column_species():
    """An ordered table that differs between (0, 1) and (1, 0) (the ABI's
    p = s_i * S + s_j): the force on i from j follows entry (s_i, s_j)."""
    st, dom = _scene(dim=3, seed=9)
    law = _routed(2, [["none"], ["spring"], ["none"], ["none"]])
    O.compute_force(st, _params(st, law, dom))
    F, _ = _textbook(st, dom, lambda a, b: "spring" if (a, b) == (0, 1) else "none")
    assert np.abs(F).max() > 1.0
    np.testing.assert_allclose(st.force, F, rtol=2e-4, atol=2e-3 * np.abs(F).max())
    assert not np.any(st.force[st.species == 1])


def test_router_of_one_law_equals_the_law_bitwise():
    st, dom = _scene()
    st2 = st.copy()
    O.compute_force(st, _params(st, _routed(2, [["hertz"]] * 4), dom))
    O.compute_force(st2, _params(st2, _law("hertz"), dom))
    assert np.array_equal(st.force, st2.force) and np.array_equal(st.torque, st2.torque)


def test_combiner_of_two_springs_doubles_the_force_bitwise():
    """zeros + f + f = 2 f exactly in float32 (law_combiner.py:72-86)."""
    st, dom = _scene()
    st2 = st.copy()
    O.compute_force(st, _params(st, _routed(0, [["none", "spring", "spring"]]), dom))
    O.compute_force(st2, _params(st2, _law("spring"), dom))
    assert np.array_equal(st.force, F32(2.0) * st2.force)


def test_empty_combiner_is_no_interaction():
    st, dom = _scene()
    P = _params(st, _routed(0, [["none"]]), dom)
    O.compute_force(st, P)
    assert not np.any(st.force) and not np.any(st.torque)
    assert O.potential_energy(st, P, O.naive_candidates(st.N)) == 0.0


def test_springs_only_on_the_history_entry():
    """(0, 0) spring friction, (0, 1) spring, (1, 1) nothing: after steps on a
    neighbour list, live springs sit exactly on the (0, 0) pairs in contact."""
    st, dom = _scene(N=300)
    law = _routed(2, [["springfriction"], ["spring"], ["spring"], ["none"]])
    P = _params(st, law, dom)
    P.collider = "neighborlist"
    P.nl_cutoff = float(2 * st.rad.max())
    P.nl_skin = 0.1
    P.nl_max_neighbors = 32
    P.nl_secondary = "naive"
    nls = O.nl_init(st, P, history=True)
    st = O.step(st, P, 5, nls)
    live = nls.live
    i, k = np.nonzero(live)
    j = nls.nl[i, k]
    assert i.size > 0
    assert np.all(st.species[i] == 0) and np.all(st.species[j] == 0)
    d = O.displacement_(st.pos[i], st.pos[j], dom)
    # every live spring belongs to a touching pair (the spring law's contact test)
    assert np.all(np.sqrt(O.dot(d, d).astype(np.float64)) < (st.rad[i] + st.rad[j]).astype(np.float64) + 1e-6)
    # and every touching (0, 0) pair of the list carries one
    jj = np.maximum(nls.nl, 0)
    dd = O.displacement_(st.pos[:, None, :], st.pos[jj], dom)
    touch = (nls.nl >= 0) & (np.sqrt(O.dot(dd, dd).astype(np.float64)) < (st.rad[:, None] + st.rad[jj]) - 1e-6)
    same0 = (st.species[:, None] == 0) & (st.species[jj] == 0)
    assert np.all(live[touch & same0])
    assert not np.any(live[touch & ~same0])
