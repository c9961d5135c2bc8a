"""GPU: the gravity walk's own stream changes no result.

By default the walk runs on a second stream beside the neighbour-list,
thermal and hydro kernels and its acceleration is added to a afterwards;
with serial=True it runs on the context's one stream.  Both must give the
same state bit for bit, for force evaluations and leapfrog steps, in FP64
and in mixed precision, with the energy equation's du/dt too: any missing
dependency between the two streams would show up as a difference.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("porig", "pos", "v", "h", "zeta", "rho", "omega", "div_v", "a", "gpot", "agravmag",
        "pptot")


def _run(serial, precision, thermal, pos, vel, mass, walk_queue=None, near_cap=None,
         neib_leaves=None, pair_blocks=None):
    from seren_amd.engine import (TUNE_NEAR_CAP, TUNE_NEIB_LEAVES, TUNE_PAIR_BLOCKS,
                                  TUNE_WALK_QUEUE, ForceEngine)
    eng = ForceEngine(multipole="octupole", mac="gadget2", precision=precision,
                      thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS, thermal=thermal,
                      serial=serial)
    try:
        if walk_queue is not None:
            eng.set_tuning(TUNE_WALK_QUEUE, walk_queue)
        if pair_blocks is not None:
            eng.set_tuning(TUNE_PAIR_BLOCKS, pair_blocks)
        if near_cap is not None:
            eng.set_tuning(TUNE_NEAR_CAP, near_cap)
            assert eng.get_tuning(TUNE_NEAR_CAP) == near_cap
        if neib_leaves is not None:
            eng.set_tuning(TUNE_NEIB_LEAVES, neib_leaves)
            assert eng.get_tuning(TUNE_NEIB_LEAVES) == neib_leaves
        n = len(mass)
        eng.load(pos, vel, mass, np.zeros(n))
        if thermal == "energy":
            eng.set_energy(np.full(n, CS * CS / 0.4))
        eng.force_eval(hguess=True)
        eng.force_eval()
        ev = eng.fetch()
        eng.step_begin("lfkdk", nlevels=3, nbuildstep=4)
        eng.step(6)
        st = eng.fetch()
        out = {"eval_" + k: ev[k] for k in KEYS}
        out.update({"step_" + k: st[k] for k in KEYS})
        out["pp_limit"] = eng.pp_limit
        if thermal == "energy":
            out["u"], out["du_dt"] = eng.fetch_energy()
        return out
    finally:
        eng.close()


@pytest.mark.parametrize("precision,thermal", [("fp64", "isothermal"), ("mixed", "isothermal"),
                                               ("fp64", "energy")])
def test_overlapped_walk_matches_serial(precision, thermal):
    print(f"case {precision} {thermal}", flush=True)
    pos, vel, mass, h = cloud(30000, "r2", 41)
    a = _run(False, precision, thermal, pos, vel, mass)
    b = _run(True, precision, thermal, pos, vel, mass)
    for k in a:
        ne = int(np.sum(np.asarray(a[k]) != np.asarray(b[k])))
        if ne:
            print(f"  {k}: {ne} differ", flush=True)
        assert ne == 0, k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_walk_occupancy_and_queue_change_nothing(precision):
    """The walk's blocks per CU (all that fit by default, 1, 3; the pair
    phase's own 5 beside 3) and its queue (off: one wave per group in tree
    order) only change which wave takes which group; the queue order sorted
    at the end of one walk for the next is used from the second evaluation
    on."""
    print(f"case walk queue settings {precision}", flush=True)
    pos, vel, mass, h = cloud(30000, "r2", 42)
    ref = _run(False, precision, "isothermal", pos, vel, mass)
    for wq, pb in ((0, None), (1, None), (3, None), (3, 5)):
        b = _run(False, precision, "isothermal", pos, vel, mass, walk_queue=wq, pair_blocks=pb)
        for k in ref:
            ne = int(np.sum(np.asarray(ref[k]) != np.asarray(b[k])))
            if ne:
                print(f"  walk queue {wq} pair blocks {pb} {k}: {ne} differ", flush=True)
            assert ne == 0, (wq, pb, k)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_near_list_fallback_changes_nothing(precision):
    """Groups whose near-field list does not fit walk the tree again in the
    pair phase with the same MAC decisions: with room for 4 leaves per group
    (nearly every group falls back) the state is the default's bit for bit,
    through evaluations and steps."""
    print(f"case near-list fallback {precision}", flush=True)
    pos, vel, mass, h = cloud(30000, "uniform", 43)
    ref = _run(False, precision, "isothermal", pos, vel, mass)
    b = _run(False, precision, "isothermal", pos, vel, mass, near_cap=4)
    for k in ref:
        ne = int(np.sum(np.asarray(ref[k]) != np.asarray(b[k])))
        if ne:
            print(f"  near cap 4 {k}: {ne} differ", flush=True)
        assert ne == 0, k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_overflow_rows_from_leaf_lists_change_nothing(precision):
    """Rows above pp_limit are formed from the leaves the particle's
    neighbour-walk wave recorded, by the whole wave; with no recorded leaves
    (0) every such particle walks the tree alone.  The rows -- and so every
    force, evaluation and step -- are the same bit for bit."""
    print(f"case overflow rows from leaf lists {precision}", flush=True)
    pos, vel, mass, h = cloud(20000, "plummer", 47)
    ref = _run(False, precision, "isothermal", pos, vel, mass)
    over = int(np.sum(ref["eval_pptot"] > ref["pp_limit"]))
    print(f"  rows above pp_limit: {over}", flush=True)
    assert over > 0
    b = _run(False, precision, "isothermal", pos, vel, mass, neib_leaves=0)
    for k in ref:
        ne = int(np.sum(np.asarray(ref[k]) != np.asarray(b[k])))
        if ne:
            print(f"  leaf lists off {k}: {ne} differ", flush=True)
        assert ne == 0, k
