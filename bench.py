#!/usr/bin/env python3
"""Throughput of the DEM time step on one MI355X.

Workload (default): 2^20 bidisperse frictional spheres in a periodic box at
volume fraction 0.50 -- spring k_n = 1e4, tangential spring with Coulomb
mu = 0.5, normal damping ratio 0.1, NeighborList over a CellList of cell size
1.4, velocity Verlet + VV-SPIRAL, float32, dt = t_c/50 (see
jaxdem_amd/workloads.py).  One "step" is one DEM time step of the whole
system; ``--steps`` of them are timed after ``--warmup`` untimed ones, inputs
already resident in HBM.

Prints one JSON line: particle-steps/s, the roofline of the fused step kernel
(algorithmic bytes per launch / its mean HIP-event duration against the
8 TB/s HBM3E peak), and the CPU baseline (the C++/OpenMP port of the oracle
on the host cores, on a bounded sample of the same workload).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s at 1M spheres; p50/p99 rollout latency, batched vs plain"
HBM_PEAK_GBS = 8000.0


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--skin", type=float, default=None,
                    help="NeighborList skin of the plain workload (default: autotuned over --plain-skins)")
    ap.add_argument("--plain-skins", type=str, default="0.1,0.15,0.2",
                    help="candidate skins of the plain-workload autotuner")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=4096, help="systems of the batching workload")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--plain-sample", type=int, default=16, help="systems stepped one at a time for the comparison")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-skins", type=str, default="0.1,0.2,0.3,0.45,0.6",
                    help="candidate skins of the batch autotuner (empty: keep the workload's 0.1)")
    ap.add_argument("--batch-cell-factors", type=str, default="1.0,0.5,1.5",
                    help="cell sizes of the batch list build, in list radii, tried by the autotuner")
    ap.add_argument("--batch-probe-every", type=int, default=4, help="the autotuner probes every k-th system")
    ap.add_argument("--batch-mode", type=int, default=-1,
                    help="batch launch mode 0/1/2 (with --batch-tile: no launch autotuning)")
    ap.add_argument("--batch-tile", type=int, default=0, help="batch step-kernel tile width 128/256")
    ap.add_argument("--batch-probe-steps", type=int, default=200)
    ap.add_argument("--batch-probe-warm", type=int, default=300,
                    help="untimed steps before each autotuner probe (the lattice start melts first)")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-max-active", type=int, default=8, help="requests in the serving batch at most")
    ap.add_argument("--serve-policy", type=str, default="fcfs", help="serving admission: fcfs or sjf")
    ap.add_argument("--serve-loads", type=str, default="0.3,0.6,0.9",
                    help="also sweep offered loads, as fractions of the one-at-a-time capacity (e.g. 0.3,0.6,0.9)")
    ap.add_argument("--serve-sweep-requests", type=int, default=16)
    ap.add_argument("--precision-n", type=int, default=1 << 18, help="spheres of the precision-form minimisation")
    ap.add_argument("--precision-steps", type=int, default=20000,
                    help="FIRE steps timed (0: run to the float32 floor, which stalls after ~1e5 steps)")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-skin", type=float, default=0.05,
                    help="Verlet-list skin of the precision form (0: the cell stencil every step)")
    ap.add_argument("--rollout-stride", type=int, default=50, help="steps between saved frames of the rollout line")
    ap.add_argument("--rollout-steps", type=int, default=1000, help="steps of the rollout line")
    ap.add_argument("--no-rollout", action="store_true")
    ap.add_argument(
        "--copy-probe-mib",
        type=int,
        default=0,
        help="also run three plain device-to-device copies of this many MiB before the timed steps "
        "(a kernel of known traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)",
    )
    return ap.parse_args()


def live_contacts(sysm) -> int:
    import torch

    live = sysm.collider.history["live"].view(torch.int32)
    # popcount of the 32-bit live words
    x = live.to(torch.int64) & 0xFFFFFFFF
    cnt = torch.zeros_like(x)
    for b in range(32):
        cnt += (x >> b) & 1
    return int(cnt.sum().item())


def listed_pairs(sysm) -> int:
    """Entries of the neighbour list (i-side candidates a step tests)."""
    return int((sysm.collider.neighbor_list >= 0).sum().item())


def traffic_from_profiles(name: str = "fused_step_spheres_1m_friction_pmc.txt") -> tuple:
    """HBM bytes per launch of a step kernel from a committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled)."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return (float(m.group(1)), str(p.relative_to(ROOT))) if m else (None, None)


def host_cpu() -> dict:
    """The host's CPUs as the OS reports them: logical CPUs, the ones this
    process may run on, and the model name."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"logical_cpus": os.cpu_count(), "usable_cpus": usable, "model": model,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(w, col, budget_s: float) -> dict:
    import numpy as np

    from oracle.port import PortRun

    host = host_cpu()
    # every usable CPU, within the per-GPU CPU share the environment sets
    # (OMP_NUM_THREADS; 16 threads per GPU on the benchmark hosts)
    threads = host["usable_cpus"] or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        threads = min(threads, env)
    run = PortRun(
        w.pos, w.vel, w.rad, w.mass, w.box, w.anchor, domain=w.domain, dt=w.dt, law=w.law, kn=w.young,
        kt=np.float32(w.kt_ratio) * np.float32(w.young), mu=w.mu, damping=w.damping, cutoff=col.cutoff,
        skin=col.skin, M=col.max_neighbors, cell_size=col.secondary_collider.cell_size,
        sr=col.secondary_collider.search_range,
    )
    run.steps(1, threads)  # first neighbour-list build, not timed
    steps, secs = 0, 0.0
    chunk = 2
    while secs < budget_s and steps < 2000:
        secs += run.steps(chunk, threads)
        steps += chunk
        chunk = min(chunk * 2, 64)
    return {
        "value": w.N * steps / secs,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the same {w.N}-sphere workload after its first list build "
        f"(C++/OpenMP port of the oracle, bit-identical to it; {secs:.1f} s)",
        "host": host,
    }


def batch_bench(args, dev) -> dict:
    """Batching workload: independent 2D Hertz-friction disk systems of
    log-uniform size in one segmented layout, against the same systems
    stepped one at a time with System.step (a size-stratified sample)."""
    import numpy as np
    import torch

    import jaxdem_amd as jd
    from jaxdem_amd import _lib as L
    from jaxdem_amd import workloads as W

    skin, cell_factor, tune = 0.1, 1.0, []
    skins = [float(x) for x in args.batch_skins.split(",") if x]
    factors = [float(x) for x in args.batch_cell_factors.split(",") if x] or [1.0]
    if len(skins) == 1 and len(factors) == 1:
        skin, cell_factor = skins[0], factors[0]
    elif skins:
        from jaxdem_amd.batching import autotune

        def sample(sk):
            return [W.make_gpu_system(w, device="cpu")
                    for w in W.batch_disks(args.batch_systems, seed=args.seed, skin=sk, every=args.batch_probe_every)]

        cfg, tune = autotune(sample, skins, cell_factors=factors,
                             tiles=(), modes=(), probe_steps=args.batch_probe_steps,
                             warm_steps=args.batch_probe_warm, device=dev)
        skin, cell_factor = cfg["skin"], cfg["cell_factor"]
        print(f"[bench] batch autotuner: skin {skin}, cell factor {cell_factor} from {tune}", file=sys.stderr,
              flush=True)
    from jaxdem_amd.batching import tune_launch, with_cell_factor

    t0 = time.perf_counter()
    ws = W.batch_disks(args.batch_systems, seed=args.seed, skin=skin)
    pairs = with_cell_factor([W.make_gpu_system(w, device="cpu") for w in ws], cell_factor)
    b = jd.Batch.pack(pairs).to(dev)
    del pairs
    prep = time.perf_counter() - t0
    N = b.n_particles
    print(f"[bench] batch: {len(ws)} systems, {N} disks, packed in {prep:.1f} s", file=sys.stderr, flush=True)
    if args.batch_mode >= 0 and args.batch_tile > 0:
        b.launch_mode, b.tile_width = args.batch_mode, args.batch_tile
    else:
        launch_cfg, launch_tune = tune_launch(b, probe_steps=args.batch_probe_steps, warm_steps=args.batch_probe_warm)
        tune = tune + launch_tune
        print(f"[bench] batch launch autotuner: {launch_cfg} from {launch_tune}", file=sys.stderr, flush=True)
    b = b.step(20)  # first list builds, not timed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b2 = b.step(args.batch_steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    batched = N * args.batch_steps / secs
    stats = b2.system.__dict__.get("_last_stats")
    # roofline pass: the same steps again with HIP events around every step
    # kernel and every rebuild chain (kernels launched one by one)
    ctx = L.context_for(dev)
    ctx.set_timing(True)
    n_rl = args.batch_steps
    b3 = b.step(n_rl)
    timing = ctx.get_timing()
    ctx.set_timing(False)
    torch.cuda.synchronize()
    zh = live_contacts(b3.system) / N
    bytes_per_disk = 96.0 + 24.0 * zh
    avg_ms = timing["step_kernel_ms"] / max(timing["step_kernels"], 1)
    achieved = N * bytes_per_disk / (avg_ms * 1e-3) / 1e9
    del b3
    # one at a time: a sample stratified by size, the same number of steps
    order = np.argsort([w.N for w in ws])
    pick = order[np.linspace(0, len(ws) - 1, min(args.plain_sample, len(ws))).astype(int)]
    n_tot, t_tot = 0, 0.0
    for k in pick:
        st, sy = W.make_gpu_system(ws[int(k)], device=dev)
        st, sy = jd.System.step(st, sy, n=20)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        jd.System.step(st, sy, n=args.batch_steps)
        torch.cuda.synchronize()
        t_tot += time.perf_counter() - t1
        n_tot += ws[int(k)].N * args.batch_steps
    plain = n_tot / t_tot
    print(f"[bench] batch: {batched:.3e} particle-steps/s batched, {plain:.3e} one at a time", file=sys.stderr,
          flush=True)
    return {
        "workload": f"batch_disks: {len(ws)} 2D systems, N log-uniform in [256, 16384] (total {N}), "
        f"area fraction U[0.3, 0.8], Hertz + tangential spring, mu 0.3, NeighborList over CellList, skin {skin}"
        + (" (autotuned)" if tune else ""),
        "n_systems": len(ws),
        "n_particles": N,
        "steps": args.batch_steps,
        "value": batched,
        "unit": "particle-steps/s",
        "ms_per_step": 1e3 * secs / args.batch_steps,
        "plain_one_at_a_time": plain,
        "plain_sample": f"{len(pick)} systems stratified by size, {args.batch_steps} steps each with System.step",
        "batched_over_plain": batched / plain,
        "rebuilds": int(stats.rebuilds) if stats is not None else None,
        "skin": skin,
        "cell_factor": cell_factor,
        "tile_width": b.tile_width,
        "launch_mode": b.launch_mode,
        "autotuner": {"knob_value_rate_rebuilds": tune,
                      "probe": f"skin and cell factor on every {args.batch_probe_every}th system, tile width and "
                      f"launch mode on the whole batch; {args.batch_probe_steps} steps after {args.batch_probe_warm}"}
        if tune else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("batch_disks_4096_pmc.txt")[0],
            "traffic_source": traffic_from_profiles("batch_disks_4096_pmc.txt")[1],
            "kernel": "k_ws_step<2, HERTZ_FRICTION, segmented>",
            "kernel_avg_us": 1e3 * avg_ms,
            "algorithmic_bytes_per_disk": bytes_per_disk,
            "mean_history_contacts": zh,
            "step_kernels_ms_per_step": timing["step_kernel_ms"] / n_rl,
            "rebuild_chain_ms_per_step": timing["rebuild_ms"] / n_rl,
        },
    }


def serving_bench(args, dev) -> dict:
    """Serving workload: Poisson-arriving rollout requests (N log-uniform in
    [1e3, 2e5], 100-10000 steps, polydisperse spheres with friction, periodic
    or reflect) served by continuous batching, against the same requests run
    one at a time with System.step in arrival order."""
    import torch

    from jaxdem_amd import serving as SV
    from jaxdem_amd import workloads as W

    spec = W.serving_requests(args.serve_requests, rate=args.serve_rate, seed=args.seed)
    reqs = []
    for k, (t, steps, w) in enumerate(spec):
        st, sy = W.make_gpu_system(w, device=dev)
        reqs.append(SV.Request(k, st, sy, steps, arrival=t))
    torch.cuda.synchronize()
    work = sum(r.n * r.steps for r in reqs)
    print(f"[bench] serving: {len(reqs)} requests, {work:.3e} particle-steps, rate {args.serve_rate}/s",
          file=sys.stderr, flush=True)
    server = SV.Server(dev, chunk=50, max_active=args.serve_max_active, policy=args.serve_policy)
    batched = SV.latency_stats(server.run(reqs, realtime=True))
    print(f"[bench] serving batched: {batched}", file=sys.stderr, flush=True)
    plain = SV.latency_stats(SV.run_one_at_a_time(reqs, dev, realtime=True))
    print(f"[bench] serving one at a time: {plain}", file=sys.stderr, flush=True)
    out = {
        "workload": f"serving_requests: {len(reqs)} Poisson arrivals at {args.serve_rate}/s, N log-uniform in "
        "[1e3, 2e5], steps U[100, 10000], log-normal radii (mean 0.5, sigma 0.2), periodic/reflect, "
        "phi U[0.3, 0.55], spring + friction mu U[0, 0.5]",
        "server": {"policy": args.serve_policy, "max_active": args.serve_max_active, "chunk_ms": 20},
        "continuous_batching": batched,
        "one_at_a_time": plain,
        "throughput_ratio": batched["particle_steps_per_s"] / plain["particle_steps_per_s"],
        "p99_ratio": plain["p99_latency_s"] / batched["p99_latency_s"],
    }
    if args.serve_loads:
        # offered load = arrival rate x mean work per request / one-at-a-time throughput
        mean_work = work / len(reqs)
        sweep = []
        for load in [float(x) for x in args.serve_loads.split(",") if x]:
            rate = load * plain["particle_steps_per_s"] / mean_work
            spec = W.serving_requests(args.serve_sweep_requests, rate=rate, seed=args.seed + 1)
            rs = []
            for k, (t, steps, w) in enumerate(spec):
                st, sy = W.make_gpu_system(w, device=dev)
                rs.append(SV.Request(k, st, sy, steps, arrival=t))
            torch.cuda.synchronize()
            b = SV.latency_stats(SV.Server(dev, chunk=50, max_active=args.serve_max_active,
                                           policy=args.serve_policy).run(rs, realtime=True))
            p = SV.latency_stats(SV.run_one_at_a_time(rs, dev, realtime=True))
            print(f"[bench] serving load {load}: rate {rate:.2f}/s batched {b} one at a time {p}", file=sys.stderr,
                  flush=True)
            sweep.append({"offered_load": load, "rate_per_s": rate, "continuous_batching": b, "one_at_a_time": p})
            del rs
            torch.cuda.empty_cache()
        out["load_sweep"] = sweep
    return out


def precision_bench(args, dev) -> dict:
    """Precision form: FIRE energy minimisation of the 3D bidisperse jamming
    start (radii 0.5 / 0.7, phi 0.66, harmonic k = 1, dt0 = 1e-2) with
    positions as integer cell + float32 offset and float64 energy, virial,
    power and residual, for --precision-steps steps from the random start."""
    import torch

    from jaxdem_amd import precision as PR
    from jaxdem_amd import workloads as W

    pos, rad, L = W.bidisperse_jamming(args.precision_n, seed=args.seed, phi=0.66)
    steps = args.precision_steps if args.precision_steps > 0 else 400000
    stall = 0 if args.precision_steps > 0 else 20000
    # warm-up: a few steps, so the timed run does not include the first allocations
    sk = max(args.precision_skin, 0.0)
    PR.minimize_precise(pos, rad, L, k=1.0, dt=1e-2, max_steps=10, pe_tol=0.0, pe_diff_tol=-1.0, force_tol=0.0,
                        skin=sk, device=dev)
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = PR.minimize_precise(pos, rad, L, k=1.0, dt=1e-2, max_steps=steps, pe_tol=0.0, pe_diff_tol=-1.0,
                            force_tol=1e-12, stall_steps=stall, skin=sk, device=dev)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t
    print(f"[bench] precision: {r['steps']} steps in {secs:.2f} s, max|F| {r['max_grad']:.3e}", file=sys.stderr,
          flush=True)
    # roofline of the force kernel: the same steps again with HIP events
    # around every k_prec_force launch
    from jaxdem_amd import _lib as Lb

    ctx = Lb.context_for(dev)
    ctx.set_timing(True)
    PR.minimize_precise(pos, rad, L, k=1.0, dt=1e-2, max_steps=int(r["steps"]), pe_tol=0.0, pe_diff_tol=-1.0,
                        force_tol=0.0, skin=sk, device=dev)
    tm = ctx.get_timing()
    ctx.set_timing(False)
    avg_ms = tm["step_kernel_ms"] / max(tm["step_kernels"], 1)
    bps = 48.0  # per sphere: cell (12 B) + offset (12) + radius (4) read, force (12) + energy (8, float64) written
    achieved = args.precision_n * bps / (avg_ms * 1e-3) / 1e9
    return {
        "workload": f"bidisperse_jamming: {args.precision_n} spheres (r 0.5/0.7, 50:50), periodic cube L = {L:.2f} "
        "at phi 0.66, harmonic k = 1, FIRE dt0 = 1e-2, integer cell + float32 offset positions, float64 "
        "energy / virial / power / residual, " + (f"Verlet list skin {sk}" if sk > 0 else "27-cell stencil"),
        "steps": r["steps"],
        "seconds": secs,
        "ms_per_step": secs * 1e3 / max(r["steps"], 1),
        "particle_steps_per_s": args.precision_n * r["steps"] / secs,
        "max_grad": r["max_grad"],
        "pe_per_particle": r["pe"],
        "pressure": r["pressure"],
        "roofline": {"bound": "hbm", "kernel": "k_prec_vforce" if sk > 0 else "k_prec_force", "kernel_avg_us": 1e3 * avg_ms,
                     "algorithmic_bytes_per_sphere": bps, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS},
        "fp64_error": precision_fp64_error(dev),
    }


def _rattlers(pos, rad, L):
    """Spheres with fewer than dim + 1 contacts, removed until none is left
    (the contact network that carries the float64 minimum's forces)."""
    import numpy as np

    N, D = pos.shape
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.rint(d / L)
    touch = np.sum(d * d, axis=-1) < (rad[:, None] + rad[None, :]) ** 2
    np.fill_diagonal(touch, False)
    keep = np.ones(N, bool)
    while True:
        z = (touch & keep[None, :]).sum(axis=1)
        drop = keep & (z < D + 1)
        if not drop.any():
            return ~keep
        keep &= ~drop


def precision_fp64_error(dev) -> dict:
    """The error of the precision form against a float64 result: FIRE started
    at the committed float64 minimum of 1024 bidisperse spheres at phi 0.66
    (tests/golden/jamming_n1024.npz, max|F| < 1e-12, made by float64 FIRE)
    must stay there; reported are the run's own float64 E/N and pressure
    against the minimum's, and the non-rattler positions after the mean
    shift."""
    import numpy as np

    from jaxdem_amd import precision as PR

    fx = np.load(ROOT / "tests" / "golden" / "jamming_n1024.npz")
    rad, L, ref = fx["rad"], float(fx["L"]), fx["ref"]
    r = PR.minimize_precise(ref, rad, L, k=1.0, dt=1e-2, max_steps=20000, pe_tol=0.0, pe_diff_tol=-1.0,
                            force_tol=1e-12, stall_steps=5000, device=dev)
    keep = ~_rattlers(ref, rad, L)
    dx = PR.min_image_diff(r["pos"], ref, L)
    err = np.abs(dx[keep] - dx[keep].mean(axis=0))
    return {
        "reference": "float64 FIRE minimum of 1024 spheres, tests/golden/jamming_n1024.npz",
        "steps": r["steps"],
        "max_grad": r["max_grad"],
        "pe_per_particle_rel_err": abs(r["pe"] - float(fx["ref_pe"])) / float(fx["ref_pe"]),
        "pressure_rel_err": abs(r["pressure"] - float(fx["ref_pressure"])) / float(fx["ref_pressure"]),
        "pos_max_err": float(err.max()),
        "pos_rms_err": float(np.sqrt(np.mean(err ** 2))),
        "non_rattlers": int(keep.sum()),
    }


def rollout_bench(args, st, sy, n_particles: int, plain_value: float) -> dict:
    """trajectory_rollout of the same workload with a frame saved every
    --rollout-stride steps: frames written on the device by one jdem_rollout
    call (save_fn=save_state), and the same frames through one System.step
    call per frame."""
    import torch

    import jaxdem_amd as jd

    k = max(1, args.rollout_stride)
    nf = max(1, args.rollout_steps // k)
    out = {"stride": k, "frames": nf, "steps": nf * k}
    for name, fn in (("device_frames", jd.save_state), ("step_per_frame", lambda s, y: (s, y))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, traj = jd.System.trajectory_rollout(st, sy, n=nf, stride=k, save_fn=fn)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        out[name] = {"particle_steps_per_s": n_particles * nf * k / secs, "ms_per_frame": 1e3 * secs / nf}
        del traj
    out["device_frames_vs_plain_step"] = out["device_frames"]["particle_steps_per_s"] / plain_value
    print(f"[bench] rollout: {out}", file=sys.stderr, flush=True)
    return out


def main():
    args = parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the step runs on one GPU; "
            "there is no multi-GPU path)",
            file=sys.stderr,
        )
        sys.exit(2)
    import torch

    import jaxdem_amd as jd
    from jaxdem_amd import _lib as L
    from jaxdem_amd import workloads as W

    dev = torch.device("cuda:0")
    skin, skin_tune = args.skin, None
    if skin is None:
        from jaxdem_amd.tuning import autotune_skin_system

        def make_pair(sk):
            return W.make_gpu_system(W.spheres_friction(N=args.n, seed=args.seed, skin=sk), device=dev)

        skin, skin_tune = autotune_skin_system(make_pair, [float(x) for x in args.plain_skins.split(",") if x])
        print(f"[bench] plain autotuner: skin {skin} from {skin_tune}", file=sys.stderr, flush=True)
    w = W.spheres_friction(N=args.n, seed=args.seed, skin=skin)
    st, sy = W.make_gpu_system(w, device=dev)
    print(f"[bench] N={w.N} dt={w.dt:.4e} M={sy.collider.max_neighbors} warmup={args.warmup}", file=sys.stderr, flush=True)
    if args.warmup > 0:
        st, sy = jd.System.step(st, sy, n=args.warmup)
    if args.copy_probe_mib > 0:
        nel = args.copy_probe_mib * (1 << 20) // 4
        src = torch.rand(nel, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        print(f"[bench] copy probe: 3 x {nel * 4} bytes read and written", file=sys.stderr, flush=True)
        del src, dst
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st1, sy1 = jd.System.step(st, sy, n=args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = w.N * args.steps / secs
    stats = sy1.__dict__.get("_last_stats")
    print(f"[bench] timed {args.steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)

    # roofline pass: the same steps with HIP events around every kernel
    ctx = L.context_for(dev)
    ctx.set_timing(True)
    st2, sy2 = jd.System.step(st, sy, n=args.steps)
    timing = ctx.get_timing()
    ctx.set_timing(False)
    torch.cuda.synchronize()
    zh = live_contacts(sy2) / w.N
    bytes_per_particle = 148.0 + 32.0 * zh
    launches = max(timing["step_kernels"], 1)
    avg_ms = timing["step_kernel_ms"] / launches
    algo_bytes = w.N * bytes_per_particle
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = traffic_from_profiles()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"spheres_{w.N}_friction: 2^20 bidisperse frictional spheres (r 0.5/0.7), periodic, phi 0.50",
            "n_particles": w.N,
            "collider": f"NeighborList(CellList cell 1.4, skin {skin:g})",
            "law": "springfriction (k_n 1e4, k_t 2/7 k_n, mu 0.5, damping ratio 0.1)",
            "integrators": "verlet + verletspiral",
            "dt": w.dt,
            "skin": float(sy.collider.skin),
            "skin_autotuner": {"candidates_skin_value_rebuilds_M": skin_tune,
                               "probe": "200 steps after 20, System.step"} if skin_tune else None,
            "max_neighbors": int(sy.collider.max_neighbors),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_ws_step<3, SPRING_FRICTION, periodic>",
            "kernel_avg_us": 1e3 * avg_ms,
            "algorithmic_bytes_per_particle": bytes_per_particle,
            "mean_history_contacts": zh,
        },
        "breakdown_ms_per_step": {
            "step_kernel": timing["step_kernel_ms"] / args.steps,
            "idle_step_kernels": timing["idle_kernel_ms"] / args.steps,
            "rebuilds": timing["rebuild_ms"] / args.steps,
            "n_rebuilds": timing["rebuilds"],
            "host_syncs": int(stats.host_syncs) if stats is not None else None,
            "list_overflow": bool(stats.overflow) if stats is not None else None,
        },
        "contact_evals_per_s": live_contacts(sy1) * args.steps / secs,
        "candidate_tests_per_s": listed_pairs(sy1) * args.steps / secs,
        "traffic_source": traffic_src,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, sy.collider, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    del st1, sy1, st2, sy2
    if not args.no_rollout:
        out["rollout"] = rollout_bench(args, st, sy, w.N, value)
        torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched"] = batch_bench(args, dev)
        torch.cuda.empty_cache()
    if not args.no_serve:
        out["serving"] = serving_bench(args, dev)
        torch.cuda.empty_cache()
    if not args.no_precision:
        out["precision"] = precision_bench(args, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
