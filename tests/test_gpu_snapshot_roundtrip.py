"""GPU: a device state written as a SEREN binary snapshot
(write_data_seren_unform.F90:47-137), read back
(read_data_seren_unform.F90:46-210) and loaded into a fresh engine gives the
same force evaluation, bit for bit, as the state it was taken from.

The state is a cloud after a setup evaluation and two device-resident
kick-drift-kick steps (h, positions and velocities all moved on the device).
The snapshot holds the particles in tree order with porig, as the reference
writes them; the binary format stores real*8 values, so the round trip must
be lossless.  The formatted (E18.10) and column ASCII (E15.7) formats round
values to their printed digits: for them the reloaded state agrees to that
precision and the evaluation with it is a valid one (every particle finds
its neighbours; the rms force difference from the lossless run stays below
1e-3, since a perturbed input can end the h iteration one step earlier or
later, which moves h within the convergence tolerance).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud, vec_rel_err  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("h", "rho", "omega", "zeta", "div_v", "a", "gpot", "pptot", "porig")


def _engine():
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="octupole", mac="gadget2", precision="fp64",
                       thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS)


def _eval_state(pos, vel, mass, h):
    eng = _engine()
    try:
        eng.load(pos, vel, mass, h)
        eng.force_eval()          # the snapshot's h is the starting h (no h guess)
        return eng.fetch()
    finally:
        eng.close()


def test_unform_snapshot_round_trip_is_bit_exact(tmp_path):
    from seren_amd import snapshot
    print("case snapshot round trip: r2 cloud of 6000, 2 KDK steps", flush=True)
    pos, vel, mass, h = cloud(6000, "r2", 41)
    eng = _engine()
    try:
        eng.load(pos, vel, mass, h)
        eng.force_eval(hguess=True)
        eng.step_begin("lfkdk", nlevels=1, nbuildstep=8, accel_mult=0.5, courant_mult=0.2)
        eng.step(4)               # two kick-drift-kick steps (two calls each)
        snap = snapshot.from_engine(eng, time=0.5, nsteps=2, snapshot=1)
    finally:
        eng.close()
    # the state as held in memory, evaluated directly
    direct = _eval_state(snap.r, snap.v, snap.m, snap.h)

    path = str(tmp_path / "snap.su")
    snapshot.write(path, snap, "seren_unform")
    back = snapshot.read(path, "seren_unform")
    for name in ("r", "v", "m", "h", "rho", "porig"):
        assert np.array_equal(getattr(back, name), getattr(snap, name)), name
    assert back.nsteps == 2 and back.snapshot == 1 and back.time == 0.5

    fresh = _engine()
    try:
        snapshot.load_into(fresh, back)
        fresh.force_eval()
        got = fresh.fetch()
    finally:
        fresh.close()
    for k in KEYS:
        assert np.array_equal(got[k], direct[k]), k
    assert np.array_equal(got["pplist"], direct["pplist"])


@pytest.mark.parametrize("fmt,digits", [("seren_form", 1e-9), ("ascii", 1e-6)])
def test_text_snapshots_reload_to_their_printed_precision(tmp_path, fmt, digits):
    from seren_amd import snapshot
    print(f"case snapshot {fmt}: uniform cloud of 3000", flush=True)
    pos, vel, mass, h = cloud(3000, "uniform", 42)
    eng = _engine()
    try:
        eng.load(pos, vel, mass, h)
        eng.force_eval(hguess=True)
        snap = snapshot.from_engine(eng)
    finally:
        eng.close()
    path = str(tmp_path / f"snap.{fmt}")
    snapshot.write(path, snap, fmt)
    back = snapshot.read(path, fmt)
    scale = np.max(np.abs(snap.r))
    assert np.max(np.abs(back.r - snap.r)) <= digits * scale
    assert np.max(np.abs(back.h - snap.h) / snap.h) <= digits
    ref = _eval_state(snap.r, snap.v, snap.m, snap.h)
    got = _eval_state(back.r, back.v, back.m, back.h)
    assert np.all(got["pptot"] > 0)
    err = vec_rel_err(got["a"], ref["a"])
    assert np.sqrt(np.mean(err * err)) < 1e-3, fmt
