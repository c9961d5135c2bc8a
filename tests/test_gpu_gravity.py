"""GPU gravity: the FP64 direct-sum kernel against the oracle's direct sum
(direct_sph_gravity.F90 semantics), tree gravity against that direct sum,
and the mixed-precision walk against the FP64 walk.

Tolerances, stated here and asserted below:
  * GPU direct sum vs oracle direct sum: |da|/|a| <= 1e-12 per target (the
    pair terms are evaluated with FMA contraction; same summation order);
  * FP64 tree vs FP64 direct sum, octupole + GADGET2 MAC (abserror 1e-4)
    after one evaluation that supplies |a|: rms |da|/|a| <= 2e-3 (the MAC
    bounds the error per cell relative to |a|);
  * mixed walk (same tree; group MAC, so every cell it accepts satisfies
    each particle's own MAC): mixed tree vs direct sum within 1.05x of the
    FP64 tree's rms error, and the mixed tree within 2x that error of the
    FP64 tree (the two differ by the cells the group opens).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud, vec_rel_err  # noqa: E402

pytestmark = pytest.mark.gpu


def _eng(precision, octupole=True, mac="gadget2"):
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="octupole" if octupole else "quadrupole", mac=mac,
                       precision=precision, thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS)


def test_direct_sum_matches_oracle():
    from oracle_lib import CSeren
    n = 4000
    pos, vel, mass, h = cloud(n, "plummer", 31)
    o = CSeren(False, "geometric")
    o.setup(pos, mass, h, vel, thetamaxsqd=0.25)
    o.force_eval(CS, hguess=True)
    r = o.particles()
    tg = np.arange(0, n, 7, dtype=np.int32)
    ra, rp = o.direct_gravity(tg + 1)
    e = _eng("fp64", octupole=False, mac="geometric")
    e.load(pos, vel, mass, h)
    e.force_eval(hguess=True)
    g = e.fetch()
    assert np.array_equal(g["porig"], r["porig"])
    ga, gp = e.direct_gravity(tg)
    err = vec_rel_err(ga, ra)
    print(f"direct sum max |da|/|a| {err.max():.3e}")
    assert err.max() <= 1e-12
    assert np.max(np.abs(gp - rp) / np.abs(rp)) <= 1e-12
    # momentum conservation of the symmetric pair force: sum m a = 0
    ta = np.arange(n, dtype=np.int32)
    fa, _ = e.direct_gravity(ta)
    mom = np.abs((g["m"][:, None] * fa).sum(0)).max() / np.abs(g["m"][:, None] * fa).sum(0).max()
    print(f"direct-sum momentum residual {mom:.3e}")
    assert mom <= 1e-12
    e.close()


@pytest.mark.parametrize("profile", ["plummer", "r2"])
def test_tree_and_mixed_vs_direct_sum(profile):
    n = 30000
    pos, vel, mass, h = cloud(n, profile, 41)
    res = {}
    for prec in ("fp64", "mixed"):
        e = _eng(prec)
        e.load(pos, vel, mass, h)
        e.force_eval(hguess=True)
        e.force_eval(hydro=False)          # |a| from the first evaluation feeds the MAC
        d = e.fetch(neighbours=False)
        res[prec] = d
        res[prec + "_counts"] = e.interaction_counts()
        if prec == "fp64":
            rng = np.random.default_rng(5)
            tg = np.sort(rng.choice(n, 2048, replace=False)).astype(np.int32)
            res["direct"], _ = e.direct_gravity(tg)
            res["targets"] = tg
        e.close()
    assert np.array_equal(res["fp64"]["porig"], res["mixed"]["porig"])
    tg = res["targets"]
    e64 = vec_rel_err(res["fp64"]["a"][tg], res["direct"])
    e32 = vec_rel_err(res["mixed"]["a"][tg], res["direct"])
    emx = vec_rel_err(res["mixed"]["a"], res["fp64"]["a"])
    rms = lambda v: float(np.sqrt(np.mean(v * v)))  # noqa: E731
    print(f"{profile}: fp64 tree vs direct rms {rms(e64):.3e} max {e64.max():.3e}; "
          f"mixed vs direct rms {rms(e32):.3e} max {e32.max():.3e}; "
          f"mixed vs fp64 rms {rms(emx):.3e} max {emx.max():.3e}; "
          f"counts fp64 {res['fp64_counts']} mixed {res['mixed_counts']}")
    assert rms(e64) <= 2e-3
    assert rms(e32) <= 1.05 * rms(e64) + 1e-6
    assert rms(emx) <= 2.0 * rms(e64)
