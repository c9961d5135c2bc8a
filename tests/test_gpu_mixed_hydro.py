"""GPU: the mixed-precision hydro forces against the FP64 ones.

In mixed precision the grad-h hydro pair forces (hydro_gradh.F90:87-297)
run in FP32 on offsets formed in FP64 from the FP64 positions, with 64-byte
records.  The h solve and the neighbour lists stay FP64 and bit-exact, so on
the same cloud both precisions see the same h, rho, Omega and lists, and only
the pair arithmetic differs.  The bar: the hydro acceleration within 1e-5
rms and 1e-3 at most of the cloud's rms hydro acceleration, and du/dt (energy
equation) within the same bounds of its rms; the lists, h and rho identical.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu

RMS_BAR, MAX_BAR = 1e-5, 1e-3


def _hydro(precision, thermal, pos, vel, mass):
    from seren_amd.engine import ForceEngine
    eng = ForceEngine(multipole="octupole", mac="gadget2", precision=precision,
                      thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS, thermal=thermal)
    try:
        n = len(mass)
        eng.load(pos, vel, mass, np.zeros(n))
        if thermal == "energy":
            eng.set_energy(np.full(n, CS * CS / 0.4))
        eng.force_eval(hguess=True)
        eng.force_eval(grav=False)
        out = eng.fetch()
        if thermal == "energy":
            out["u"], out["du_dt"] = eng.fetch_energy()
        return out
    finally:
        eng.close()


def _rel(d, ref):
    scale = np.sqrt(np.mean(np.sum(ref * ref, axis=-1) if ref.ndim > 1 else ref * ref))
    e = np.linalg.norm(d - ref, axis=-1) if ref.ndim > 1 else np.abs(d - ref)
    return float(np.sqrt(np.mean(e * e)) / scale), float(e.max() / scale)


@pytest.mark.parametrize("thermal", ["isothermal", "energy"])
def test_mixed_hydro_within_the_bar_of_fp64(thermal):
    print(f"case mixed hydro vs fp64, {thermal}", flush=True)
    pos, vel, mass, h = cloud(30000, "r2", 81, mach=3.0)
    a = _hydro("fp64", thermal, pos, vel, mass)
    b = _hydro("mixed", thermal, pos, vel, mass)
    for k in ("porig", "h", "rho", "omega", "pptot"):
        assert np.array_equal(a[k], b[k]), k
    ok = a["pptot"] <= a["pplist"].shape[1]
    assert np.array_equal(a["pplist"][ok], b["pplist"][ok])
    rms, mx = _rel(b["a"], a["a"])
    print(f"  hydro a: rms {rms:.2e} max {mx:.2e}", flush=True)
    assert rms <= RMS_BAR and mx <= MAX_BAR
    assert rms > 0.0   # the FP32 path ran
    if thermal == "energy":
        rms, mx = _rel(b["du_dt"], a["du_dt"])
        print(f"  du/dt: rms {rms:.2e} max {mx:.2e}", flush=True)
        assert rms <= RMS_BAR and mx <= MAX_BAR
