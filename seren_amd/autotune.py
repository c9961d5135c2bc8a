"""Autotuner of the mixed-precision force path.

The mixed walk's accuracy bar is stated against the FP64 direct sum
(direct_sph_gravity.F90:63-79, serenx_direct_gravity): rms |da|/|a| of the
gravitational acceleration at most 1e-4.  Its multipole-MAC tolerance
(SERENX_TUNE_MIXED_ABSERROR; the FP64 walk keeps the reference's abserror)
trades that accuracy for walk time, and the h-solve list length
(SERENX_TUNE_HSOLVE_SLOTS) trades list memory traffic against fallback
walks without changing any result.

tune_mixed() measures, on the cloud loaded in an engine, the evaluation time
and the error against the direct sum on a seeded sample of targets for each
candidate tolerance, and keeps the fastest whose rms error stays within
safety * target (a margin for clouds and steps it has not seen).

tune_walk_blocks() chooses the gravity walk's blocks per CU
(SERENX_TUNE_WALK_QUEUE) by timing the caller's unit of work (a force
evaluation or a step) for each candidate: the walk shares the device's
vector ALUs with the SPH kernels beside it, and the split that finishes the
unit first depends on the cloud and the precision.  Results do not depend
on it.  Everything runs on the device through the library; nothing here is
a CPU fallback.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .engine import TUNE_HSOLVE_SLOTS, TUNE_MIXED_ABSERROR, TUNE_WALK_QUEUE, ForceEngine

DEFAULT_TOLERANCES = (1e-4, 2e-4, 4e-4, 8e-4, 1.6e-3, 3.2e-3)
DEFAULT_SLOTS = (32, 48, 64, 96)
DEFAULT_WALK_BLOCKS = (3, 4, 6, 64)


def _grav_error(eng: ForceEngine, targets: np.ndarray) -> Dict[str, float]:
    """One gravity-only evaluation (hydro skipped) and its relative error
    against the FP64 direct sum at the targets (indices in tree order)."""
    eng.force_eval(hydro=False)
    a = eng.fetch(neighbours=False)["a"][targets]
    a_dir, _ = eng.direct_gravity(targets.astype(np.int32))
    err = np.linalg.norm(a - a_dir, axis=1) / np.linalg.norm(a_dir, axis=1)
    return {"rms": float(np.sqrt(np.mean(err * err))), "max": float(err.max())}


def _eval_ms(eng: ForceEngine, reps: int) -> float:
    best = np.inf
    for _ in range(reps):
        eng.sync()
        t0 = time.perf_counter()
        eng.force_eval()
        eng.sync()
        best = min(best, time.perf_counter() - t0)
    return 1e3 * best


def tune_mixed(eng: ForceEngine, target_rms: float = 1e-4, safety: float = 0.5,
               tolerances: Sequence[float] = DEFAULT_TOLERANCES, ntargets: int = 2048,
               reps: int = 2, seed: int = 0) -> Dict[str, object]:
    """Chooses the mixed walk's MAC tolerance for the loaded cloud (the engine
    must be in mixed precision with a multipole MAC) and sets it.  Returns
    the choice and the measured table."""
    if eng.opt.precision != 1:
        raise ValueError("tune_mixed needs an engine in mixed precision")
    rng = np.random.default_rng(seed)
    eng.force_eval()                       # |a| for the multipole MAC
    table: List[Dict[str, float]] = []
    best = None
    for tol in sorted(tolerances):
        eng.set_tuning(TUNE_MIXED_ABSERROR, tol)
        n = eng.n
        targets = np.sort(rng.choice(n, min(n, ntargets), replace=False))
        err = _grav_error(eng, targets)
        ms = _eval_ms(eng, reps)
        row = {"abserror": tol, "rms": err["rms"], "max": err["max"], "ms": ms}
        table.append(row)
        if err["rms"] <= safety * target_rms and (best is None or ms < best["ms"]):
            best = row
        if err["rms"] > target_rms:        # looser tolerances only grow the error
            break
    if best is None:
        best = table[0]
    eng.set_tuning(TUNE_MIXED_ABSERROR, best["abserror"])
    return {"abserror": best["abserror"], "rms": best["rms"], "max": best["max"],
            "ms": best["ms"], "table": table}


def tune_slots(eng: ForceEngine, slots: Sequence[int] = DEFAULT_SLOTS,
               reps: int = 2) -> Dict[str, object]:
    """Chooses the h-solve list length by evaluation time (results do not
    depend on it) and sets it."""
    table = []
    for s in slots:
        eng.set_tuning(TUNE_HSOLVE_SLOTS, s)
        table.append({"slots": int(s), "ms": _eval_ms(eng, reps)})
    best = min(table, key=lambda r: r["ms"])
    eng.set_tuning(TUNE_HSOLVE_SLOTS, best["slots"])
    return {"slots": best["slots"], "table": table}


def tune_walk_blocks(eng: ForceEngine, unit: Optional[Callable[[], None]] = None,
                     blocks: Sequence[int] = DEFAULT_WALK_BLOCKS,
                     reps: int = 3, mean: bool = False) -> Dict[str, object]:
    """Chooses the walk's blocks per CU by the median (mean: the mean) time
    of reps calls of unit (default: a force evaluation) after one untimed
    call per candidate, and sets it.  A step as the unit advances the
    engine's state; the mean of a whole tree-rebuild cycle of steps weighs
    the rebuild step as a run does, where a median leaves it out."""
    if unit is None:
        unit = eng.force_eval
    table = []
    for b in blocks:
        eng.set_tuning(TUNE_WALK_QUEUE, b)
        unit()
        ts = []
        for _ in range(reps):
            eng.sync()
            t0 = time.perf_counter()
            unit()
            eng.sync()
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.mean(ts) if mean else np.median(ts))
        table.append({"blocks_per_cu": int(b), "ms": ms})
    best = min(table, key=lambda r: r["ms"])
    eng.set_tuning(TUNE_WALK_QUEUE, best["blocks_per_cu"])
    return {"blocks_per_cu": best["blocks_per_cu"], "table": table}
