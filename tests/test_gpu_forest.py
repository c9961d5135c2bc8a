"""GPU: a forest of independent clouds evaluated in one pass gives every cloud
exactly the results of evaluating it alone (serenx_load_forest).

Each cloud has its own tree from its own bounding box, and every walk stays
inside its own cloud's tree, visiting cells and particles in the same order
as a standalone walk, so the bar is bit-exact: tree order, h, rho, Omega,
zeta, div v, neighbour counts and lists, accelerations, potentials and the
FP64 direct sum, for two evaluations (the second uses |a| from the first in
the multipole MAC).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu

CLOUDS = [(3000, "uniform", 31), (5000, "r2", 32), (1200, "plummer", 33), (4096, "uniform", 34),
          (2500, "r2", 35)]
KEYS = ("h", "rho", "omega", "zeta", "div_v", "a", "gpot", "agravmag", "pptot")


def _engine(octupole, mac):
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="octupole" if octupole else "quadrupole", mac=mac,
                       precision="fp64", thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS)


@pytest.mark.parametrize("octupole,mac", [(True, "gadget2"), (False, "geometric")])
def test_forest_equals_clouds_alone(octupole, mac):
    print(f"case forest of {len(CLOUDS)} clouds, octupole={octupole} mac={mac}", flush=True)
    data = [cloud(n, prof, seed) for n, prof, seed in CLOUDS]
    alone = []
    eng = _engine(octupole, mac)
    try:
        for pos, vel, mass, h in data:
            eng.load(pos, vel, mass, h)
            eng.force_eval(hguess=True)
            eng.force_eval()
            d = eng.fetch()
            tg = np.arange(0, len(mass), 97, dtype=np.int32)
            d["direct"] = eng.direct_gravity(tg)
            alone.append(d)
        off = eng.load_forest(data)
        eng.force_eval(hguess=True)
        eng.force_eval()
        f = eng.fetch()
        fdirect = []
        for k, (pos, vel, mass, h) in enumerate(data):
            tg = (np.arange(0, len(mass), 97) + off[k]).astype(np.int32)
            fdirect.append(eng.direct_gravity(tg))
    finally:
        eng.close()
    assert list(np.diff(off)) == [c[0] for c in CLOUDS]
    for k, d in enumerate(alone):
        lo, hi = off[k], off[k + 1]
        assert np.array_equal(f["porig"][lo:hi] - lo, d["porig"]), k
        for key in KEYS:
            assert np.array_equal(f[key][lo:hi], d[key]), (k, key)
        lim = d["pplist"].shape[1]
        rows = d["pptot"] <= lim
        cols = np.arange(lim)[None, :] < d["pptot"][:, None]
        mine = np.where(cols, f["pplist"][lo:hi] - lo, 0)
        ref = np.where(cols, d["pplist"], 0)
        assert np.array_equal(mine[rows], ref[rows]), k
        for q in range(2):
            assert np.array_equal(fdirect[k][q], d["direct"][q]), (k, "direct")
