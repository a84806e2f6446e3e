"""Conversions between jaxdem_amd objects and the CPU oracle (tests only)."""

from __future__ import annotations

import numpy as np

from oracle import jdem_oracle as O


def law_of(system, dtype=np.float32) -> O.Law:
    return _law(system.force_model, system.mat_table, dtype)


def _law(fm, mt, dtype=np.float32) -> O.Law:
    from jaxdem_amd import _lib as L
    from jaxdem_amd.forces import route_layout

    if fm.kind == L.LAW_ROUTED:
        S, off, terms = route_layout(fm)
        return O.Law(kind="routed", n_species=S, route_off=np.asarray(off, np.int64),
                     terms=tuple(O.Law(kind="none") if t is None else _law(t, mt, dtype) for t in terms))
    kind = fm.type_name
    get = lambda k: (np.asarray(getattr(mt, k), dtype) if mt.has(k) else None)  # noqa: E731
    return O.Law(
        kind=kind,
        young_eff=get("young_eff"),
        young=get("young"),
        poisson=get("poisson"),
        e=get("e"),
        mu=get("mu"),
        mu_r=get("mu_r"),
        mu_eff=get("mu_eff"),
        kt_ratio=float(getattr(fm, "kt_ratio", 2.0 / 7.0)),
        damping=float(getattr(fm, "damping", 0.0)),
    )


def params_of(state, system, dtype=np.float32) -> O.Params:
    dom = system.domain
    odom = O.Domain(
        kind={"free": "free", "periodic": "periodic", "reflect": "reflect"}.get(dom.type_name, "free"),
        box=np.asarray(dom.box_size, dtype).copy(),
        anchor=np.asarray(dom.anchor, dtype).copy(),
        restitution=float(getattr(dom, "restitution_coefficient", 1.0)),
    )
    law = law_of(system, dtype)
    col = system.collider
    cname = col.type_name.lower()
    P = O.Params(
        dim=state.dim,
        dt=float(dtype(system.dt)),
        domain=odom,
        law=law,
        lin=system.linear_integrator.type_name,
        rot=system.rotation_integrator.type_name,
        collider={"naive": "naive", "celllist": "celllist", "neighborlist": "neighborlist", "": "none"}[cname],
        gravity=np.asarray(system.force_manager.gravity, dtype),
        interact_same_bond_id=system.interact_same_bond_id,
    )
    if cname == "celllist":
        P.cell_size = float(col.cell_size)
        P.search_range = int(col.search_range)
    if cname == "neighborlist":
        P.nl_cutoff = float(col.cutoff)
        P.nl_skin = float(col.skin)
        P.nl_max_neighbors = int(col.max_neighbors)
        sec = col.secondary_collider
        P.nl_secondary = "celllist" if sec.type_name.lower() == "celllist" else "naive"
        if P.nl_secondary == "celllist":
            P.nl_secondary_cell_size = float(sec.cell_size)
            P.nl_secondary_search_range = int(sec.search_range)
    return P


def state_of(state, dtype=np.float32) -> O.State:
    d = state.numpy()
    bond = d["bond_id"].astype(np.int64)
    if not state.has_bonds:
        bond = np.full((state.N, 0), -1, np.int64)
    return O.State(
        pos=d["pos_c"].astype(dtype),
        vel=d["vel"].astype(dtype),
        ang_vel=d["ang_vel"].astype(dtype),
        q=d["q"].astype(dtype),
        force=d["force"].astype(dtype),
        torque=d["torque"].astype(dtype),
        rad=d["rad"].astype(dtype),
        mass=d["mass"].astype(dtype),
        inertia=d["inertia"].astype(dtype),
        fixed=d["fixed"].astype(bool),
        mat_id=d["mat_id"].astype(np.int64),
        bond_id=bond,
        species=d["species_id"].astype(np.int64),
        clump=None if state.is_sphere_state else d["clump_id"].astype(np.int64),
        pos_p=None if state.is_sphere_state else d["pos_p"].astype(dtype),
    )


def nl_state_of(state, system, dtype=np.float32) -> O.NLState:
    col = system.collider
    nls = O.NLState(
        nl=col.neighbor_list.cpu().numpy().astype(np.int64),
        old_pos=col.old_pos.cpu().numpy().astype(dtype),
        n_build=int(col.n_build_times),
        overflow=bool(col.overflow),
    )
    if col.history is not None:
        nls.xi, nls.live = history_of(col)
    return nls


def history_of(col):
    xi = col.history["xi"].cpu().numpy().astype(np.float32)
    words = col.history["live"].cpu().numpy().view(np.uint32)
    N, M = xi.shape[0], xi.shape[1]
    live = np.zeros((N, M), bool)
    for s in range(M):
        live[:, s] = ((words[:, s // 32] >> np.uint32(s % 32)) & np.uint32(1)).astype(bool)
    xi = np.where(live[..., None], xi, np.float32(0.0))
    return xi, live
