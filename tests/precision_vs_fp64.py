"""Precision form against float64 on the 3D bidisperse jamming minimisation
(harmonic spheres, k = 1, FIRE): test infrastructure, run by the GPU test of
the precision form and, staged, from the command line at full size.  The
float64 runs are oracle/jdem_fire64.cpp on the CPU.

Minimisation from random positions is chaotic: near jamming the landscape
is flat, and float64 runs whose starts differ by 1e-9, even after a float64
relaxation to max|F| <= 1e-4, end in different minima (energy 10-20 %
apart, positions O(0.1) apart).  The comparison therefore has two parts:

  same seed      -- float64 and the float32 precision form from the same
                    random start; they reach different minima of one
                    ensemble, so E/N and P are compared, not positions;
  at the minimum -- both start from the float64 minimum (max|F| < 1e-12);
                    the float32 run must stay there, and its positions
                    (rattlers excluded, mean shift removed), E/N and P are
                    compared with the minimum's.

Every final configuration is also re-evaluated on the host in float64
(``evaluate64``), so each run's own E/N and P and the configuration's true
ones are both reported.

    python tests/precision_vs_fp64.py relax64 --n 262144 --force-tol 0 --out seed.npz
    python tests/precision_vs_fp64.py fp64 --start seed.npz --out ref.npz
    python tests/precision_vs_fp64.py gpu --start seed.npz --out gpu_seed.npz
    python tests/precision_vs_fp64.py gpu --start ref.npz --out gpu_min.npz
    python tests/precision_vs_fp64.py compare ref.npz gpu_seed.npz gpu_min.npz
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

K = 1.0
DT = 1e-2


def _wrap_into(pos, L):
    x = np.mod(np.asarray(pos, np.float64), L)
    return np.where(x >= L, x - L, x)


def min_image(d, L):
    return d - L * np.rint(d / L)


def contact_pairs(pos, rad, L):
    """Overlapping pairs (i < j), their min-image r_ij = x_i - x_j and
    distance, in float64 (scipy's periodic k-d tree)."""
    from scipy.spatial import cKDTree

    x = _wrap_into(pos, L)
    tree = cKDTree(x, boxsize=L)
    pr = tree.query_pairs(2.0 * float(np.max(rad)), output_type="ndarray")
    if len(pr) == 0:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros((0, 3)), np.zeros(0)
    i, j = pr[:, 0], pr[:, 1]
    r = min_image(x[i] - x[j], L)
    d = np.sqrt(np.einsum("ij,ij->i", r, r))
    keep = (d < rad[i] + rad[j]) & (d > 0)
    return i[keep], j[keep], r[keep], d[keep]


def evaluate64(pos, rad, L, k: float = K) -> dict:
    """float64 forces, energy per particle, contact virial sum r_ij (x) F_ij,
    pressure tr(W)/(3V) and max |F| component of a configuration."""
    N = len(rad)
    i, j, r, d = contact_pairs(pos, rad, L)
    delta = rad[i] + rad[j] - d
    s = k * delta / d
    fij = s[:, None] * r
    F = np.zeros((N, 3))
    np.add.at(F, i, fij)
    np.add.at(F, j, -fij)
    W = np.einsum("pi,pj->ij", r, fij)
    return {"pe": float(np.sum(0.5 * k * delta * delta) / N), "pressure": float(np.trace(W) / (3 * L**3)),
            "max_grad": float(np.abs(F).max()) if N else 0.0, "virial": W, "contacts": (i, j)}


def rattlers(pos, rad, L) -> np.ndarray:
    """Spheres with fewer than D + 1 = 4 contacts, removed iteratively (the
    usual rattler rule of jammed packings); their positions are free in
    their cages and carry no information."""
    N = len(rad)
    i, j, _, _ = contact_pairs(pos, rad, L)
    alive = np.ones(N, bool)
    while True:
        m = alive[i] & alive[j]
        z = np.bincount(i[m], minlength=N) + np.bincount(j[m], minlength=N)
        drop = alive & (z < 4)
        if not drop.any():
            return ~alive
        alive &= ~drop


def fp64_run(pos, rad, L, force_tol=1e-12, max_steps=400000, threads=None) -> dict:
    """float64 FIRE to max|F| <= force_tol; one thread below 20000 spheres
    (where OpenMP barriers would cost more than they save), else all."""
    from oracle import build_port, port

    build_port.build()
    if threads is None:
        threads = 1 if len(rad) < 20000 else 0
    r = port.fire64(pos, rad, L, k=K, dt=DT, max_steps=max_steps, pe_tol=0.0, pe_diff_tol=-1.0,
                    force_tol=force_tol, threads=threads)
    r["pressure"] = float(np.trace(r["virial"]) / (3 * L**3))
    return r


def gpu_precise(pos, rad, L, max_steps=400000, stall_steps=20000, device="cuda", skin=0.0) -> dict:
    import torch

    from jaxdem_amd import precision as PR

    torch.cuda.synchronize()
    t = time.perf_counter()
    r = PR.minimize_precise(pos, rad, L, k=K, dt=DT, max_steps=max_steps, pe_tol=0.0, pe_diff_tol=-1.0,
                            force_tol=1e-12, stall_steps=stall_steps, skin=skin, device=device)
    torch.cuda.synchronize()
    r["seconds"] = time.perf_counter() - t
    return r


def gpu_absolute(pos, rad, L, max_steps=400000, device="cuda") -> dict:
    """The same FIRE on absolute float32 coordinates (System minimize over a
    periodic CellList), for the residual floor it reaches."""
    import torch

    import jaxdem_amd as jd

    st = jd.State.create(pos=_wrap_into(pos, L).astype(np.float32), rad=np.asarray(rad, np.float32), device=device)
    mt = jd.MaterialTable.from_materials([jd.Material.create("elastic", density=1.0, young=K, poisson=0.3)])
    sy = jd.System.create(st.shape, dt=DT, domain_type="periodic",
                          domain_kw=dict(box_size=np.full(3, L, np.float32), anchor=np.zeros(3, np.float32)),
                          collider_type="CellList", collider_kw=dict(state=st), mat_table=mt)
    torch.cuda.synchronize()
    t = time.perf_counter()
    st2, sy2, steps, pe = jd.minimize(st, sy, max_steps=max_steps, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=1e-12)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    return {"pos": st2.pos.detach().cpu().numpy().astype(np.float64), "pe": pe, "steps": steps,
            "max_grad": float(sy2.__dict__["_last_minimize"].max_grad), "seconds": sec}


def compare(ref: dict, got: dict, rad, L) -> dict:
    """Errors of ``got`` against the float64 ``ref``: E/N and P as the run
    reported them and as float64 re-evaluates its configuration, and the
    positions of the non-rattlers after removing the mean shift."""
    e_ref = evaluate64(ref["pos"], rad, L)
    e_got = evaluate64(got["pos"], rad, L)
    keep = ~rattlers(ref["pos"], rad, L)
    dx = min_image(np.asarray(got["pos"]) - np.asarray(ref["pos"]), L)
    shift = dx[keep].mean(axis=0)
    err = np.abs(dx[keep] - shift)
    out = {
        "pe_rel_err": abs(e_got["pe"] - e_ref["pe"]) / e_ref["pe"],
        "pressure_rel_err": abs(e_got["pressure"] - e_ref["pressure"]) / e_ref["pressure"],
        "pos_max_err": float(err.max()),
        "pos_rms_err": float(np.sqrt(np.mean(err**2))),
        "non_rattlers": int(keep.sum()),
        "max_grad_fp64_eval": e_got["max_grad"],
        "ref_pe": e_ref["pe"],
        "ref_pressure": e_ref["pressure"],
        "ref_max_grad": e_ref["max_grad"],
    }
    if "pressure" in got:  # the run's own float64 reductions
        out["reported_pe_rel_err"] = abs(got["pe"] - e_ref["pe"]) / e_ref["pe"]
        out["reported_pressure_rel_err"] = abs(got["pressure"] - e_ref["pressure"]) / e_ref["pressure"]
    return out


def _save(path, d):
    np.savez(path, **{k: np.asarray(v) for k, v in d.items() if not isinstance(v, (dict, tuple))})


def _load(path) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("relax64", help="float64 FIRE from the random start to --force-tol")
    a.add_argument("--n", type=int, default=1 << 18)
    a.add_argument("--seed", type=int, default=0)
    a.add_argument("--phi", type=float, default=0.66)
    a.add_argument("--force-tol", type=float, default=1e-4)
    a.add_argument("--out", required=True)
    for name in ("fp64", "gpu"):
        b = sub.add_parser(name)
        b.add_argument("--start", required=True)
        b.add_argument("--out", required=True)
        b.add_argument("--max-steps", type=int, default=400000)
    sub.choices["gpu"].add_argument("--stall-steps", type=int, default=20000)
    sub.choices["gpu"].add_argument("--chunk-steps", type=int, default=0,
                                    help="restart FIRE every this many steps, printing progress (0: one run)")
    sub.choices["gpu"].add_argument("--no-absolute", action="store_true")
    c = sub.add_parser("compare")
    c.add_argument("ref")
    c.add_argument("got", nargs="+")
    c.add_argument("--start", help="start file (radii and box); default: the ref file's")
    args = ap.parse_args(argv)

    if args.cmd == "relax64":
        from jaxdem_amd.workloads import bidisperse_jamming

        pos, rad, L = bidisperse_jamming(args.n, args.seed, args.phi)
        if args.force_tol > 0:
            r = fp64_run(pos, rad, L, force_tol=args.force_tol)
            print(f"relax64: {r['steps']} steps, max|F| {r['max_grad']:.3e}, {r['seconds']:.1f} s", flush=True)
            pos = r["pos"]
        _save(args.out, {"pos": pos, "rad": rad, "L": L, "seed": args.seed, "phi": args.phi,
                         "force_tol": args.force_tol})
    elif args.cmd == "fp64":
        s = _load(args.start)
        r = fp64_run(s["pos"], s["rad"], s["L"], max_steps=args.max_steps)
        print(f"fp64: {r['steps']} steps, max|F| {r['max_grad']:.3e}, E/N {r['pe']:.17g}, "
              f"P {r['pressure']:.17g}, {r['seconds']:.1f} s", flush=True)
        _save(args.out, {**r, "rad": s["rad"], "L": s["L"]})
    elif args.cmd == "gpu":
        s = _load(args.start)
        if args.chunk_steps > 0:
            pos, done, secs = s["pos"], 0, 0.0
            while done < args.max_steps:
                r = gpu_precise(pos, s["rad"], s["L"], max_steps=min(args.chunk_steps, args.max_steps - done),
                                stall_steps=args.stall_steps)
                pos, done, secs = r["pos"], done + r["steps"], secs + r["seconds"]
                print(f"  {done} steps, max|F| {r['max_grad']:.3e}, E/N {r['pe']:.12g}", flush=True)
                if r["steps"] < args.chunk_steps:
                    break  # stalled or converged within the chunk
            r["steps"], r["seconds"] = done, secs
        else:
            r = gpu_precise(s["pos"], s["rad"], s["L"], max_steps=args.max_steps, stall_steps=args.stall_steps)
        print(f"precise: {r['steps']} steps, max|F| {r['max_grad']:.3e}, E/N {r['pe']:.17g}, "
              f"P {r['pressure']:.17g}, {r['seconds']:.2f} s", flush=True)
        _save(args.out, {"pos": r["pos"], "pe": r["pe"], "pressure": r["pressure"], "max_grad": r["max_grad"],
                         "steps": r["steps"], "seconds": r["seconds"], "rad": s["rad"], "L": s["L"]})
        if not args.no_absolute:
            q = gpu_absolute(s["pos"], s["rad"], s["L"], max_steps=min(args.max_steps, 2 * max(r["steps"], 1)))
            print(f"absolute: {q['steps']} steps, max|F| {q['max_grad']:.3e}, E/N {q['pe']:.17g}, "
                  f"{q['seconds']:.2f} s", flush=True)
            out = Path(args.out)
            _save(out.with_name(out.stem + "_absolute.npz"), {**q, "rad": s["rad"], "L": s["L"]})
    else:
        ref = _load(args.ref)
        src = _load(args.start) if args.start else ref
        for g in args.got:
            got = _load(g)
            res = compare(ref, got, src["rad"], float(src["L"]))
            res.update({"file": g, "steps": int(got.get("steps", -1)), "max_grad": float(got.get("max_grad", -1))})
            print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
