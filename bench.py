"""Benchmark: particles/s of SEREN integration steps on one MI355X.

Workload: N = 1,048,576 equal-mass particles, rho ~ r^-2 truncated at r = 1, with a
divergence-free P(k) ~ k^-4 velocity field on a 128^3 grid at rms Mach 2
(c_s = 0.2), octupole moments and the GADGET2 multipole MAC (abserror 1e-4),
FP64 throughout.  One "step" (--mode step, the default) is one kick-drift-
kick leapfrog step with SEREN's timestep criteria (timestep_size.F90,
sph_timesteps.F90, one time-step level): two device-resident sph_integrate
calls, the second holding the force evaluation of every particle -- tree
build (every 8 steps, as nbuildstep) or restock, grad-h h/rho solve with
neighbour lists, isothermal thermal update, grad-h hydro pair forces,
Barnes-Hut gravity -- and the particle advance.  --mode eval times bare force
evaluations (tree build every time) instead.  Inputs are resident in HBM;
the timed region is K steps after W untimed ones and the setup evaluation.

Prints one JSON line.  --cpu-baseline times the reference Fortran/OpenMP path
(the flang build of the reference sources in oracle/_ref) on the same cloud.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s (tree grav+grad-h hydro), FP32-mixed vs FP64 @1e-4 rms; p99 lat"

# Algorithmic flop counts per interaction (counted from the reference source,
# + - * / sqrt = 1, compares/table loads = 0):
#   octupole cell term, BHgrav_accel.F90:146-226        190
#   quadrupole cell term                                  69
#   kernel-softened pair, gravity_gradh.F90:42-86        29 + 8 [r < 2h_p] + 9 [r < 2h_j]
F_CELL = {"octupole": 190.0, "octupole_exact": 190.0, "quadrupole": 69.0}
F_PAIR = 29.0
MI355X_FP64_TFLOPS = 78.6
MI355X_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="step", choices=["step", "eval", "forest", "serve"])
    ap.add_argument("--clouds", type=int, default=256, help="forest mode: number of clouds")
    ap.add_argument("--requests", type=int, default=2000, help="serve mode: requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve mode: requests per second")
    ap.add_argument("--slots", type=int, default=8, help="serve mode: concurrent requests")
    ap.add_argument("--pool", type=int, default=4, help="serve mode: resident clouds per size")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--profile", default="r2")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--multipole", default="octupole")
    ap.add_argument("--mac", default="gadget2")
    ap.add_argument("--abserror", type=float, default=1e-4)
    ap.add_argument("--mach", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--direct-targets", type=int, default=16384)
    ap.add_argument("--cpu-baseline", dest="cpu", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB device buffer once (HBM counter calibration)")
    return ap.parse_args()


def calibrate_copy():
    """One device-to-device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of rocprofv3 to bytes."""
    import torch
    n = 32 * 1024 * 1024                      # 256 MiB of float64
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n}), flush=True)


def measured_traffic(workload):
    """HBM bytes per gravity-walk dispatch from the committed rocprofv3 counter
    record of this workload (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel
    calibration), or None when no record matches."""
    import glob
    for f in glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")):
        try:
            rec = json.load(open(f))
        except (OSError, ValueError):
            continue
        if rec.get("workload") == workload:
            return rec.get("traffic_bytes_per_dispatch")
    return None


def cpu_baseline(pos, vel, mass, h, args):
    """Time the reference Fortran/OpenMP force evaluation on the same cloud."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from oracle_lib import RefSeren
        ref = RefSeren(args.multipole != "quadrupole", args.mac)   # same walk cost
    except Exception as e:  # reference build absent
        return {"value": None, "unit": "particles/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {e}"}
    n = len(mass)
    ref.setup(pos, mass, h, vel, thetamaxsqd=0.25, abserror=args.abserror)
    ref.force_eval(0.2, hguess=True)          # setup evaluation (untimed)
    t0 = time.perf_counter()
    ref.tree_update()
    ref.sph_update()
    ref.thermal_iso(0.2)
    ref.hydro_forces()
    ref.grav_forces()
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n / dt, "unit": "particles/s", "cores": cores, "kind": "reference",
            "sample": f"one force evaluation of the full {n}-particle cloud (after one "
                      f"untimed setup evaluation), {dt:.2f} s wall"}


def bench_forest(args):
    """Batched ensemble: independent uniform clouds, N log-uniform in
    [4096, 262144], each with its own P(k) ~ k^-4 turbulent seed at Mach
    U(1, 5), loaded as one forest; a step is one force evaluation of the
    whole forest (every phase one launch over all clouds)."""
    from seren_amd import ics
    from seren_amd.engine import ForceEngine
    rng = np.random.default_rng(args.seed)
    ns = np.round(np.exp(rng.uniform(np.log(4096), np.log(262144), args.clouds))).astype(int)
    machs = rng.uniform(1.0, 5.0, args.clouds)
    print(f"# generating {args.clouds} clouds, {ns.sum()} particles", file=sys.stderr, flush=True)
    clouds = [ics.make_cloud(int(n), "uniform", seed=args.seed + 1 + k, mach=float(m), cs=0.2,
                             ngrid=32) for k, (n, m) in enumerate(zip(ns, machs))]
    eng = ForceEngine(multipole=args.multipole, mac=args.mac, precision=args.precision,
                      thetamaxsqd=0.25, abserror=args.abserror, sound_speed=0.2)
    off = eng.load_forest(clouds)
    eng.force_eval(hguess=True)
    for _ in range(max(1, args.warmup)):
        eng.force_eval()
    eng.sync()
    phases = np.zeros(4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.force_eval()
        t = eng.timings()
        phases += np.array([t["tree"], t["sph"], t["hydro"], t["grav"]])
    eng.sync()
    dt = time.perf_counter() - t0
    ntot = int(off[-1])
    tg = np.sort(rng.choice(ntot, min(ntot, args.direct_targets), replace=False))
    eng.force_eval(hydro=False)
    a_tree = eng.fetch(neighbours=False)["a"]
    a_dir, _ = eng.direct_gravity(tg.astype(np.int32))
    err = np.linalg.norm(a_tree[tg] - a_dir, axis=1) / np.linalg.norm(a_dir, axis=1)
    eng.close()
    out = {"metric": METRIC, "value": ntot * args.steps / dt, "unit": "particles/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "fp64" if args.precision == "fp64" else "fp32-mixed",
           "data": "synthetic (seeded uniform spheres, P(k)~k^-4 turbulent velocities)",
           "config": {"workload": f"forest of {args.clouds} clouds, N log-uniform "
                                  f"[4096, 262144] (total {ntot}), Mach U(1,5), "
                                  f"{args.multipole} + {args.mac} MAC, one force evaluation "
                                  f"of the forest per step, {args.precision}",
                      "clouds": args.clouds, "n_total": ntot, "precision": args.precision},
           "phases_ms": {k: float(v / args.steps) for k, v in
                         zip(("tree", "sph", "hydro", "grav"), phases)},
           "accuracy_vs_fp64_direct_sum": {"targets": int(len(tg)),
                                            "rms_rel": float(np.sqrt(np.mean(err * err))),
                                            "max_rel": float(err.max())}}
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Serving: Poisson requests (one KDK leapfrog step each) on resident
    clouds of N in {2k, 16k, 128k, 1M} drawn with p = {0.4, 0.3, 0.2, 0.1},
    profiles uniform / r^-2 / Plummer, mixed precision with an FP64 guard
    step every 4, admission by HBM budget; p50 / p99 latency and throughput."""
    from seren_amd import ics
    from seren_amd.serve import CloudServer, poisson_run
    sizes, probs = [2048, 16384, 131072, 1048576], [0.4, 0.3, 0.2, 0.1]
    profiles = ["uniform", "r2", "plummer"]
    srv = CloudServer(hbm_budget=200e9, slots=args.slots, precision=args.precision,
                      guard_every=4, engine_kw=dict(multipole=args.multipole, mac=args.mac,
                                                    abserror=args.abserror, sound_speed=0.2))
    rng = np.random.default_rng(args.seed)
    ids, p_cloud = [], []
    print(f"# admitting {len(sizes) * args.pool} clouds", file=sys.stderr, flush=True)
    for si, (n, pr) in enumerate(zip(sizes, probs)):
        for j in range(args.pool):
            cid = len(ids)
            prof = profiles[int(rng.integers(3))]
            pos, vel, mass, h = ics.make_cloud(n, prof, seed=args.seed + 100 + cid, mach=2.0,
                                               cs=0.2, ngrid=32 if n < 200000 else 64)
            srv.admit(cid, pos, vel, mass, h)
            ids.append(cid)
            p_cloud.append(pr / args.pool)
    srv.start()
    # warm every cloud once
    for f in [srv.submit(c) for c in ids]:
        f.result()
    print("# warm", file=sys.stderr, flush=True)
    res = poisson_run(srv, args.rate, args.requests, ids, p_cloud, seed=args.seed)
    used = srv.used
    srv.close()
    out = {"metric": METRIC, "value": res["particles_per_s"], "unit": "particles/s",
           "n_gpus": 1, "steps": args.requests, "warmup": len(ids),
           "ms_per_step": res["mean_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "fp64" if args.precision == "fp64" else "fp32-mixed",
           "data": "synthetic (seeded uniform / r^-2 / Plummer clouds, Mach 2)",
           "config": {"workload": f"serving: Poisson {args.rate:g} req/s, {args.requests} "
                                  f"requests (one KDK step each) on {len(ids)} resident "
                                  f"clouds, N in {sizes} with p {probs}, {args.slots} slots, "
                                  f"{args.precision} with FP64 guard step every 4",
                      "precision": args.precision, "slots": args.slots},
           "latency_ms": {"p50": res["p50_ms"], "p99": res["p99_ms"], "mean": res["mean_ms"]},
           "offered_particles_per_s": res["offered_particles_per_s"],
           "resident_bytes": used, "wall_s": res["wall_s"]}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this force path runs on one GPU; sharding over "
                                   "several GPUs is not built, use --gpus 1"}))
        sys.exit(2)
    if args.mode == "forest":
        return bench_forest(args)
    if args.mode == "serve":
        return bench_serve(args)
    from seren_amd import ics
    from seren_amd.engine import ForceEngine
    print(f"# generating {args.n} particles ({args.profile}, Mach {args.mach})", file=sys.stderr,
          flush=True)
    pos, vel, mass, h = ics.make_cloud(args.n, args.profile, seed=args.seed, mach=args.mach,
                                       cs=0.2, ngrid=128)
    eng = ForceEngine(multipole=args.multipole, mac=args.mac, precision=args.precision,
                      thetamaxsqd=0.25, abserror=args.abserror, sound_speed=0.2)
    eng.load(pos, vel, mass, h)
    eng.force_eval(hguess=True)                  # setup evaluation (sph_setup)
    eng.sync()
    counts = np.zeros(4)
    grav_ms = []
    phases = np.zeros(4)
    wstats = None

    def unit():
        """One timed unit: a leapfrog step (two sph_integrate calls) or a
        bare force evaluation; returns per-phase ms and interaction counts."""
        nonlocal wstats
        ph, cnt, glaunch = np.zeros(4), np.zeros(4), []
        for _ in range(2 if args.mode == "step" else 1):
            if args.mode == "step":
                eng.step(1)
            else:
                eng.force_eval()
            t = eng.timings()
            ph += np.array([t["tree"], t["sph"], t["hydro"], t["grav"]])
            c = np.array(eng.interaction_counts(), np.float64)
            if c[0] > 0:                          # the call that evaluated forces
                wstats = eng.walk_stats()
                glaunch.append(t["grav"])
            cnt += c
        return ph, cnt, glaunch

    if args.mode == "step":
        eng.step_begin("lfkdk", nlevels=1, nbuildstep=8, accel_mult=0.5, courant_mult=0.2)
    else:
        eng.force_eval()                         # provides |a| for the multipole MAC
    for _ in range(args.warmup):
        unit()
    eng.sync()
    print("# warm", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ph, cnt, gl = unit()
        phases += ph
        counts += cnt
        grav_ms.extend(gl)
    eng.sync()
    dt = time.perf_counter() - t0
    step_info = eng.step_info() if args.mode == "step" else None
    ncell, npp, npp_in_p, npp_in_j = counts
    # accuracy of the tree gravity against the FP64 direct sum on a seeded
    # subsample of targets (one more, gravity-only, evaluation; untimed)
    eng.force_eval(hydro=False)
    a_tree = eng.fetch(neighbours=False)["a"]
    rng = np.random.default_rng(args.seed + 1)
    tg = np.sort(rng.choice(args.n, min(args.n, args.direct_targets), replace=False))
    a_dir, _ = eng.direct_gravity(tg.astype(np.int32))
    err = np.linalg.norm(a_tree[tg] - a_dir, axis=1) / np.linalg.norm(a_dir, axis=1)
    n = args.n
    value = n * args.steps / dt
    ms_step = 1e3 * dt / args.steps
    # gravity-walk roofline: algorithmic flops of the force-evaluating
    # launches / their measured duration (HIP events on the library's stream)
    flops = ncell * F_CELL[args.multipole] + npp * F_PAIR + 8.0 * npp_in_p + 9.0 * npp_in_j
    gms = float(np.sum(grav_ms))
    achieved = flops / (gms * 1e-3) / 1e12
    peak = MI355X_FP64_TFLOPS if args.precision == "fp64" else MI355X_FP32_TFLOPS
    what = ("KDK leapfrog steps (timestep_size criteria, 1 level, tree build every 8)"
            if args.mode == "step" else "force evaluation")
    workload = (f"{what}, N={n} {args.profile} sphere, Mach {args.mach}, "
                f"{args.multipole} + {args.mac} MAC (abserror {args.abserror}), {args.precision}")
    traffic = measured_traffic(workload)
    out = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64" if args.precision == "fp64" else "fp32-mixed",
        "data": "synthetic (seeded r^-2 sphere, P(k)~k^-4 turbulent velocities)",
        "config": {"workload": workload,
                   "n": n, "profile": args.profile, "multipole": args.multipole,
                   "mac": args.mac, "abserror": args.abserror, "precision": args.precision},
        "phases_ms": {k: float(v / args.steps) for k, v in
                      zip(("tree", "sph", "hydro", "grav"), phases)},
        "interactions_per_particle": {"cell": ncell / n / args.steps,
                                      "pair": npp / n / args.steps},
        "walk_stats": wstats,
        "step_state": step_info,
        "accuracy_vs_fp64_direct_sum": {"targets": int(len(tg)),
                                         "rms_rel": float(np.sqrt(np.mean(err * err))),
                                         "max_rel": float(err.max())},
        "roofline": {"kernel": "gravity walk", "bound": "valu", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic},
    }
    eng.close()
    if args.cpu:
        print("# cpu baseline (reference Fortran/OpenMP)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(pos, vel, mass, h, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
