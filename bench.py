"""Benchmark: particles/s of SEREN integration steps on one MI355X.

Workload (the GPU FP64 configuration of the reference's benchmark suite):
N = 1,048,576 equal-mass particles, rho ~ r^-2 truncated at r = 1, with a
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
F_CELL = {"octupole": 190.0, "quadrupole": 69.0}
F_PAIR = 29.0
MI355X_FP64_TFLOPS = 78.6
MI355X_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="step", choices=["step", "eval"])
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
        ref = RefSeren(args.multipole == "octupole", args.mac)
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


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this force path runs on one GPU; sharding over "
                                   "several GPUs is not built, use --gpus 1"}))
        sys.exit(2)
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
