"""Benchmark: particles/s of SEREN integration steps on one MI355X.

Workload: N = 1,048,576 equal-mass particles, rho ~ r^-2 truncated at r = 1, with a
divergence-free P(k) ~ k^-4 velocity field on a 128^3 grid at rms Mach 2
(c_s = 0.2), octupole moments and the GADGET2 multipole MAC (abserror 1e-4),
an FP64 leg and a mixed-precision leg (the headline value; FP64 value and
ratio beside it), then a bounded serving leg.  One "step" (--mode step, the default) is one kick-drift-
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

# The serving leg steps up to 7 pools at once, each on its own stream; HIP
# maps a process's streams onto at most GPU_MAX_HW_QUEUES hardware queues (4
# by default), and pools that share a queue run one after the other (at 200
# requests/s: p99 1.6-1.8 s with 4 queues, 0.5-0.7 s with 16).  Raised to 16
# before the runtime initialises; a larger value already set is kept.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

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
    ap.add_argument("--nbuild", type=int, default=8,
                    help="forest mode: tree built every nbuild evaluations, restocked between "
                         "(tree_update.F90's nbuildstep)")
    ap.add_argument("--no-graph", dest="graph", action="store_false", default=True,
                    help="forest mode: run the restock evaluations eagerly, not as HIP graphs")
    ap.add_argument("--requests", type=int, default=2000, help="serve mode: requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve mode: requests per second")
    ap.add_argument("--pool", type=int, default=4, help="serve mode: resident clouds per size")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--profile", default="r2")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"],
                    help="step/eval modes: time FP64, mixed, or both (the headline: mixed "
                         "value beside the FP64 value)")
    ap.add_argument("--no-tune", dest="tune", action="store_false", default=True,
                    help="mixed leg: keep the reference MAC tolerance (no autotuning)")
    ap.add_argument("--mixed-abserror", type=float, default=0.0,
                    help="MAC tolerance of the mixed walk (0: the autotuned one, or abserror)")
    ap.add_argument("--walk-queue", type=int, default=-1,
                    help="gravity walk blocks per compute unit (-1: autotuned, or the library's "
                         "default with --no-tune-walk)")
    ap.add_argument("--pair-blocks", type=int, default=-1,
                    help="blocks per CU of the walk's pair phase (0: the walk's; -1: the library's "
                         "default, 0)")
    ap.add_argument("--no-tune-walk", dest="tune_walk", action="store_false", default=True,
                    help="keep the library's walk blocks per CU (no timing of candidates)")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="step/eval modes: requests of the bounded serving leg (0: none)")
    ap.add_argument("--precision-serve", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--serve-walk-stream", action="store_true",
                    help="serving: each pool's gravity walk on a second stream beside its SPH "
                         "kernels (default: one stream per pool)")
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


def measured_traffic(kernel, n):
    """HBM bytes per dispatch of the walk kernel from the committed rocprofv3
    counter record of this workload (FETCH_SIZE + WRITE_SIZE, each scaled by
    a copy-kernel calibration), or None when no record matches."""
    import glob
    for f in glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")):
        try:
            rec = json.load(open(f))
        except (OSError, ValueError):
            continue
        if rec.get("kernel") == kernel and rec.get("n") == n:
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
    whole forest (every phase one launch over all clouds): the tree is built
    every nbuild steps and restocked in between, the restock evaluations
    replayed from recorded HIP graphs (recorded again after each build)."""
    from seren_amd import ics
    from seren_amd.engine import ForceEngine
    rng = np.random.default_rng(args.seed)
    ns = np.round(np.exp(rng.uniform(np.log(4096), np.log(262144), args.clouds))).astype(int)
    machs = rng.uniform(1.0, 5.0, args.clouds)
    print(f"# generating {args.clouds} clouds, {ns.sum()} particles", file=sys.stderr, flush=True)
    clouds = [ics.make_cloud(int(n), "uniform", seed=args.seed + 1 + k, mach=float(m), cs=0.2,
                             ngrid=32) for k, (n, m) in enumerate(zip(ns, machs))]
    prec = "mixed" if args.precision == "both" else args.precision
    eng = ForceEngine(multipole=args.multipole, mac=args.mac, precision=prec,
                      thetamaxsqd=0.25, abserror=args.abserror, sound_speed=0.2)
    off = eng.load_forest(clouds)
    tuned = None
    if prec == "mixed" and args.tune:
        from seren_amd.autotune import tune_mixed
        eng.force_eval(hguess=True)
        tuned = tune_mixed(eng, target_rms=1e-4, safety=0.5, seed=args.seed)
        print(f"# autotuned mixed MAC tolerance {tuned['abserror']:g}", file=sys.stderr,
              flush=True)
        off = eng.load_forest(clouds)
    eng.force_eval(hguess=True)
    walk_tuned = None
    if args.walk_queue >= 0:
        from seren_amd.engine import TUNE_WALK_QUEUE
        eng.set_tuning(TUNE_WALK_QUEUE, args.walk_queue)
    elif args.tune_walk:
        from seren_amd.autotune import tune_walk_blocks
        walk_tuned = tune_walk_blocks(
            eng, lambda: eng.force_eval(restock=True, graph=args.graph))
        print(f"# walk blocks per CU {walk_tuned['blocks_per_cu']}", file=sys.stderr, flush=True)

    def one(k):
        if k % max(args.nbuild, 1) == 0:
            eng.force_eval()
        else:
            eng.force_eval(restock=True, graph=args.graph)

    for k in range(max(1, args.warmup)):
        one(k)
    eng.sync()
    phases = np.zeros(4)
    t0 = time.perf_counter()
    for k in range(args.steps):
        one(k)
        t = eng.timings()
        phases += np.array([t["tree"], t["sph"], t["hydro"], t["grav"]])
    eng.sync()
    dt = time.perf_counter() - t0
    ginfo = eng.graph_info()
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
           "dtype": "fp64" if prec == "fp64" else "fp32-mixed",
           "data": "synthetic (seeded uniform spheres, P(k)~k^-4 turbulent velocities)",
           "config": {"workload": f"forest of {args.clouds} clouds, N log-uniform "
                                  f"[4096, 262144] (total {ntot}), Mach U(1,5), "
                                  f"{args.multipole} + {args.mac} MAC, one force evaluation "
                                  f"of the forest per step (tree built every {args.nbuild}, "
                                  f"restocked between{', as HIP graphs' if args.graph else ''}), "
                                  f"{prec}",
                      "clouds": args.clouds, "n_total": ntot, "precision": prec,
                      "nbuild": args.nbuild, "graph": bool(args.graph)},
           "graphs": ginfo,
           "autotune": tuned,
           "walk_blocks": walk_tuned,
           "phases_ms": {k: float(v / args.steps) for k, v in
                         zip(("tree", "sph", "hydro", "grav"), phases)},
           "accuracy_vs_fp64_direct_sum": {"targets": int(len(tg)),
                                            "rms_rel": float(np.sqrt(np.mean(err * err))),
                                            "max_rel": float(err.max())}}
    print(json.dumps(out), flush=True)


def step_leg(args, cloud, precision, tune=False):
    """Times K force-evaluating units (leapfrog steps, or bare evaluations)
    of one precision on the cloud; returns the measured figures."""
    from seren_amd.engine import ForceEngine
    pos, vel, mass, h = cloud
    n = len(mass)
    eng = ForceEngine(multipole=args.multipole, mac=args.mac, precision=precision,
                      thetamaxsqd=0.25, abserror=args.abserror, sound_speed=0.2)
    out = {"precision": precision}
    try:
        if precision == "mixed" and not tune and args.mixed_abserror > 0:
            from seren_amd.engine import TUNE_MIXED_ABSERROR
            eng.set_tuning(TUNE_MIXED_ABSERROR, args.mixed_abserror)
            out["mixed_abserror"] = args.mixed_abserror
        if args.walk_queue >= 0:
            from seren_amd.engine import TUNE_WALK_QUEUE
            eng.set_tuning(TUNE_WALK_QUEUE, args.walk_queue)
            out["walk_queue"] = args.walk_queue
        if args.pair_blocks >= 0:
            from seren_amd.engine import TUNE_PAIR_BLOCKS
            eng.set_tuning(TUNE_PAIR_BLOCKS, args.pair_blocks)
            out["pair_blocks"] = args.pair_blocks
        eng.load(pos, vel, mass, h)
        eng.force_eval(hguess=True)                  # setup evaluation (sph_setup)
        eng.sync()
        if tune:
            from seren_amd.autotune import tune_mixed
            t = tune_mixed(eng, target_rms=1e-4, safety=0.5, seed=args.seed)
            out["autotune"] = t
            print(f"# autotuned mixed MAC tolerance {t['abserror']:g}", file=sys.stderr,
                  flush=True)
            eng.load(pos, vel, mass, h)              # the timed run starts from the ICs
            eng.force_eval(hguess=True)
            eng.sync()
        counts = np.zeros(4)
        grav_ms = []
        phases = np.zeros(4)
        wstats = None

        def unit(instrument=False):
            """One unit: a leapfrog step (two sph_integrate calls) or a bare
            force evaluation; returns per-phase ms and the walk's (cell
            phase, pair phase) ms, read from the events the library records
            (no device work).  With instrument also the interaction counts
            and walk statistics, which copy device counters to the host:
            they are read on one more unit after the timed ones."""
            nonlocal wstats
            ph, cnt, glaunch = np.zeros(4), np.zeros(4), []
            for _ in range(2 if args.mode == "step" else 1):
                if args.mode == "step":
                    eng.step(1)
                else:
                    eng.force_eval()
                t = eng.timings()
                ph += np.array([t["tree"], t["sph"], t["hydro"], t["grav"]])
                w = eng.walk_timings()
                if w["cells"] > 0:                    # the call that evaluated forces
                    glaunch.append((w["cells"], w["pairs"]))
                    if instrument:
                        wstats = eng.walk_stats()
                if instrument:
                    cnt += np.array(eng.interaction_counts(), np.float64)
            return ph, cnt, glaunch

        if args.mode == "step":
            eng.step_begin("lfkdk", nlevels=1, nbuildstep=8, accel_mult=0.5, courant_mult=0.2)
        else:
            eng.force_eval()                         # provides |a| for the multipole MAC
        if args.walk_queue < 0 and args.tune_walk:
            # the walk's blocks per CU, timed on the unit itself (untimed here)
            from seren_amd.autotune import tune_walk_blocks
            if args.mode == "step":
                # mean over 4 steps, which hold exactly one tree rebuild (every
                # 8 sph_integrate calls); all that fit never wins a step
                out["walk_blocks"] = tune_walk_blocks(eng, unit, blocks=(3, 4, 6), reps=4,
                                                      mean=True)
            else:
                out["walk_blocks"] = tune_walk_blocks(eng, unit)
            print(f"# walk blocks per CU {out['walk_blocks']['blocks_per_cu']}",
                  file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            unit()
        eng.sync()
        print(f"# {precision} warm", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ph, _, gl = unit()
            phases += ph
            grav_ms.extend(gl)
        eng.sync()
        dt = time.perf_counter() - t0
        # interaction counts and walk statistics: one more unit (untimed),
        # counted for each timed unit (the counts of consecutive steps agree
        # to a fraction of a percent)
        counts = unit(instrument=True)[1] * args.steps
        out["step_info"] = eng.step_info() if args.mode == "step" else None
        # accuracy of the tree gravity against the FP64 direct sum on a seeded
        # subsample of targets (one more, gravity-only, evaluation; untimed)
        eng.force_eval(hydro=False)
        a_tree = eng.fetch(neighbours=False)["a"]
        rng = np.random.default_rng(args.seed + 1)
        tg = np.sort(rng.choice(n, min(n, args.direct_targets), replace=False))
        a_dir, _ = eng.direct_gravity(tg.astype(np.int32))
        err = np.linalg.norm(a_tree[tg] - a_dir, axis=1) / np.linalg.norm(a_dir, axis=1)
    finally:
        eng.close()
    ncell, npp, npp_in_p, npp_in_j = counts
    cell_flops = ncell * F_CELL[args.multipole]
    pair_flops = npp * F_PAIR + 8.0 * npp_in_p + 9.0 * npp_in_j
    cms = float(np.sum([g[0] for g in grav_ms]))
    pms = float(np.sum([g[1] for g in grav_ms]))
    nwalk = max(len(grav_ms), 1)
    peak = MI355X_FP64_TFLOPS if precision == "fp64" else MI355X_FP32_TFLOPS
    achieved = cell_flops / (cms * 1e-3) / 1e12
    kname = "k_grav_cells_fp64" if precision == "fp64" else "k_grav_cells_mixed"
    pname = "k_grav_pairs_fp64" if precision == "fp64" else "k_grav_pairs_mixed"
    out.update({
        "value": n * args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
        "phases_ms": {k: float(v / args.steps) for k, v in
                      zip(("tree", "sph", "hydro", "grav"), phases)},
        "interactions_per_particle": {"cell": ncell / n / args.steps,
                                      "pair": npp / n / args.steps},
        "walk_stats": wstats,
        "accuracy_vs_fp64_direct_sum": {"targets": int(len(tg)),
                                         "rms_rel": float(np.sqrt(np.mean(err * err))),
                                         "max_rel": float(err.max())},
        # the walk's cell phase is the dominant kernel: its algorithmic flops
        # (cell terms x 190, BHgrav_accel.F90:146-226) over its event-timed
        # duration; the pair phase beside it
        "roofline": {"kernel": kname, "bound": "valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "ms_per_launch": cms / nwalk},
        "roofline_pairs": {"kernel": pname, "bound": "valu",
                           "achieved": pair_flops / max(pms * 1e-3, 1e-12) / 1e12, "peak": peak,
                           "unit": "TFLOP/s",
                           "frac": pair_flops / max(pms * 1e-3, 1e-12) / 1e12 / peak,
                           "ms_per_launch": pms / nwalk},
    })
    return out


def serve_leg(args, nrequests, pool=4):
    """Bounded serving run: Poisson requests at args.rate on resident clouds
    of N in {2k, 16k, 128k, 1M} (p = 0.4, 0.3, 0.2, 0.1), mixed precision
    with an FP64 guard step every 4 steps of a cloud, continuous batching in
    pools; returns latency percentiles and throughput."""
    from seren_amd import ics
    from seren_amd.serve import CloudServer, poisson_run
    sizes, probs = [2048, 16384, 131072, 1048576], [0.4, 0.3, 0.2, 0.1]
    profiles = ["uniform", "r2", "plummer"]
    srv = CloudServer(hbm_budget=200e9, precision=args.precision_serve, guard_every=4,
                      engine_kw=dict(multipole=args.multipole, mac=args.mac,
                                     abserror=args.abserror, sound_speed=0.2),
                      serial=not args.serve_walk_stream)
    rng = np.random.default_rng(args.seed)
    ids, p_cloud, data = [], [], {}
    for si, (n, pr) in enumerate(zip(sizes, probs)):
        for j in range(pool):
            cid = len(ids)
            prof = profiles[int(rng.integers(3))]
            data[cid] = ics.make_cloud(n, prof, seed=args.seed + 100 + cid, mach=2.0, cs=0.2,
                                       ngrid=32 if n < 200000 else 64)
            ids.append(cid)
            p_cloud.append(pr / pool)
    print(f"# admitting {len(ids)} clouds", file=sys.stderr, flush=True)
    try:
        srv.admit_many(data)
        if args.mixed_abserror > 0:
            from seren_amd.engine import TUNE_MIXED_ABSERROR
            for p in srv.pools:
                p.eng.set_tuning(TUNE_MIXED_ABSERROR, args.mixed_abserror)
        srv.start()
        for f in [srv.submit(c) for c in ids]:   # warm every cloud once
            f.result()
        print("# serving warm", file=sys.stderr, flush=True)
        res = poisson_run(srv, args.rate, nrequests, ids, p_cloud, seed=args.seed)
        res["pools"] = len(srv.pools)
        res["resident_bytes"] = srv.used
    finally:
        srv.close()
    res.update({"sizes": sizes, "probs": probs, "clouds": len(ids),
                "precision": args.precision_serve, "guard_every": 4,
                "mixed_abserror": args.mixed_abserror or args.abserror})
    return res


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
        res = serve_leg(args, args.requests, pool=args.pool)
        out = {"metric": METRIC, "value": res["particles_per_s"], "unit": "particles/s",
               "n_gpus": 1, "steps": args.requests, "warmup": 0, "ms_per_step": res["mean_ms"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "fp32-mixed", "data": "synthetic (seeded clouds, Mach 2)",
               "config": {"workload": f"serving: Poisson {args.rate:g} req/s, {args.requests} "
                                      "requests (one KDK step each), N in {2k,16k,128k,1M}"},
               "serving": res}
        print(json.dumps(out), flush=True)
        return
    from seren_amd import ics
    print(f"# generating {args.n} particles ({args.profile}, Mach {args.mach})", file=sys.stderr,
          flush=True)
    cloud = ics.make_cloud(args.n, args.profile, seed=args.seed, mach=args.mach, cs=0.2,
                           ngrid=128)
    n = args.n
    legs = {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        legs[prec] = step_leg(args, cloud, prec, tune=(prec == "mixed" and args.tune))
        if prec == "mixed" and "autotune" in legs[prec]:
            args.mixed_abserror = legs[prec]["autotune"]["abserror"]
    head = legs.get("mixed") or legs["fp64"]
    what = ("KDK leapfrog steps (timestep_size criteria, 1 level, tree build every 8)"
            if args.mode == "step" else "force evaluation")
    workload = (f"{what}, N={n} {args.profile} sphere, Mach {args.mach}, "
                f"{args.multipole} + {args.mac} MAC (abserror {args.abserror}), "
                f"{'+'.join(legs)}")
    roof = dict(head["roofline"])
    roof["traffic"] = measured_traffic(roof["kernel"], n)
    out = {
        "metric": METRIC, "value": head["value"], "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64" if head["precision"] == "fp64" else "fp32-mixed",
        "data": "synthetic (seeded r^-2 sphere, P(k)~k^-4 turbulent velocities)",
        "config": {"workload": workload, "n": n, "profile": args.profile,
                   "multipole": args.multipole, "mac": args.mac, "abserror": args.abserror,
                   "precision": head["precision"]},
        "roofline": roof,
    }
    if "fp64" in legs and "mixed" in legs:
        out["fp64_value"] = legs["fp64"]["value"]
        out["mixed_over_fp64"] = legs["mixed"]["value"] / legs["fp64"]["value"]
        r64 = dict(legs["fp64"]["roofline"])
        r64["traffic"] = measured_traffic(r64["kernel"], n)
        out["roofline_fp64"] = r64
    out["legs"] = legs
    if args.serve_requests > 0:
        print("# serving leg", file=sys.stderr, flush=True)
        out["serving"] = serve_leg(args, args.serve_requests)
    if args.cpu:
        print("# cpu baseline (reference Fortran/OpenMP)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(*cloud, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
