"""GPU: FP64 conservation diagnostics (serenx_diagnostics) against the
reference's own diagnostics.F90.

The fixtures hold the line diagnostics.F90 writes with DEBUG_DIAGNOSTICS
(tests/golden/make_golden.py, the reference built with its diagnostics):
mtot, rcom, vcom, linear and angular momentum, net force, ketot, gpetot,
utot (3/2 sum m P/rho isothermal; sum m u with the energy equation) and
etot, in E18.10 format (10-11 significant digits), after the setup force
evaluation and after six RK integration steps.  The device sums the same
quantities in FP64 in tree order; the bar is the fixture's own precision:
|d| <= 1e-9 |value| + 1e-12 for every entry.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["diag_r2_2k_quad_geo", "diag_uniform2k_quad_geo_steps", "diag_r2_2k_quad_geo_energy"]
NAMES = ["mtot", "rcom.x", "rcom.y", "rcom.z", "vcom.x", "vcom.y", "vcom.z", "mom.x", "mom.y",
         "mom.z", "ang.x", "ang.y", "ang.z", "force.x", "force.y", "force.z", "ketot", "gpetot",
         "utot", "etot"]


def _vector(d):
    return np.concatenate([[d["mtot"]], d["rcom"], d["vcom"], d["mom"], d["angmom"], d["force"],
                           [d["ketot"], d["gpetot"], d["utot"], d["etot"]]])


@pytest.mark.parametrize("case", CASES)
def test_diagnostics_match_the_reference(case):
    from seren_amd.engine import ForceEngine
    print(f"case {case}", flush=True)
    g = np.load(os.path.join(GOLD, case + ".npz"))
    n, seed, nsteps = (int(v) for v in g["spec"])
    energy = str(g["thermal"]) == "energy"
    eng = ForceEngine(multipole="quadrupole", mac="geometric", precision="fp64",
                      thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS,
                      thermal="energy" if energy else "isothermal")
    try:
        eng.load(g["pos"], g["vel"], g["mass"], np.zeros(n))
        if energy:
            eng.set_energy(g["u0"])
        eng.force_eval(hguess=True)
        if nsteps:
            eng.step_begin("rk", nlevels=3, nbuildstep=8, accel_mult=0.5, courant_mult=0.2,
                           levels="adaptive", neib_check=0)
            eng.step(nsteps)
        got = _vector(eng.diagnostics())
    finally:
        eng.close()
    want = g["diag"]
    err = np.abs(got - want)
    bar = 1e-9 * np.abs(want) + 1e-12
    for k, name in enumerate(NAMES):
        print(f"  {name:8s} {got[k]: .12e} ref {want[k]: .10e}  |d| {err[k]:.1e}", flush=True)
    assert np.all(err <= bar), [NAMES[k] for k in np.nonzero(err > bar)[0]]
