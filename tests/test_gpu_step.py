"""GPU parity of the device-resident integration step (sph_integrate.F90:
tree_update on the reference cadence, sph_update, thermal, forces,
sph_timesteps, sph_advance) against golden vectors written by the reference
itself (tests/golden/make_golden.py: the flang build of sph_timesteps.F90,
timestep_size.F90, sph_advance.F90 and advance_runge_kutta.F90).

Bar:
  * integer-time state after every step (n, nsteps, nresync, level_max,
    level_step, next build / stock), the number of active particles, and
    per particle nlevel, nlast and accdo: exact;
  * time, timestep, dt_max: relative 1e-12;
  * positions, velocities, r_old, v_old, v_half: |d| <= 1e-12 (max |x|);
    laststep relative 1e-12; h and rho relative 1e-11; accelerations
    |da| <= 1e-10 |a|.  Forces agree with the reference to ~1e-15 relative
    (see test_gpu_parity), and a few steps carry that into the state.

The time-step options of the reference Makefile are covered by fixtures from
builds with them: TIMESTEP=RESTRICTED with CHECK_NEIB_TIMESTEP=1 (the
Makefile's defaults) and TIMESTEP=FIXED with CHECK_NEIB_TIMESTEP=2
(check_neighbour_timesteps.F90, reduce_timesteps.F90); nminneib is compared
exactly as well.  The other fixtures come from a build with neither option
(adaptive levels, no neighbour check); one of them with THERMAL=ENERGY_EQN
(u advanced with du/dt, the energy time-step criterion), where u is bound
to 1e-11 relative and du/dt to 1e-10 of its largest value.

The leapfrog KDK integrator has no reference counterpart (the reference's
advance_leapfrog_kdk.F90 uses dt_half before setting it); it is checked
against a numpy restatement of its update with the fix, driven by the
device's own accelerations, bit for bit.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, rel_err  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEP_CASES = ["step_rk_uniform2k_quad_geo_nlev3", "step_rk_r2_2k_oct_gadget2_nlev1",
              "step_rk_r2_2k_quad_geo_nlev6_restricted_1", "step_rk_r2_2k_quad_geo_nlev6_fixed_2",
              "step_rk_uniform2k_quad_geo_energy_nlev3", "step_rk_r2_2k_quad_geo_quintic_nlev3",
              "step_rk_r2_2k_std_quad_geo_energy_nlev3",
              "step_rk_uniform2k_std_quad_geo_consth_nlev3"]
# reference build -> (levels, neib_check)
TIMESTEP_BUILDS = {"": ("adaptive", 0), "restricted_1": ("restricted", 1),
                   "fixed_2": ("fixed", 2)}
STEP_KEYS = ["n", "nsteps", "nresync", "level_max", "level_step", "nbuild", "nstock"]


def _engine(octupole, mac, thermal="isothermal", kernel="m4", h_fac=1.2, sph="gradh",
            hmin=0.0):
    from seren_amd.engine import ForceEngine
    kw = dict(sph=sph, hmin=hmin)   # hmin > 0: HFIND=CONSTANT
    return ForceEngine(multipole="octupole" if octupole else "quadrupole", mac=mac,
                       precision="fp64", thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS,
                       thermal=thermal, kernel=kernel, h_fac=h_fac, **kw)


@pytest.mark.parametrize("name", STEP_CASES)
def test_rk_steps_match_reference_golden(name):
    print(f"case {name}", flush=True)
    g = np.load(os.path.join(GOLD, name + ".npz"))
    n, seed, octu, nlevels, nsteps = (int(v) for v in g["spec"])
    levels, check = TIMESTEP_BUILDS[str(g["timestep"])]
    energy = str(g["thermal"]) == "energy"
    kernel = str(g["kernel"]) if "kernel" in g.files else "m4"
    h_fac = float(g["h_fac"]) if "h_fac" in g.files else 1.2
    sph = str(g["sph"]) if "sph" in g.files else "gradh"
    hmin = float(g["hmin"]) if "hmin" in g.files else 0.0
    eng = _engine(bool(octu), str(g["mac"]), "energy" if energy else "isothermal", kernel, h_fac,
                  sph, hmin)
    try:
        h0 = np.zeros(n)
        eng.load(g["pos0"], g["vel0"], g["mass"], h0)
        if energy:
            eng.set_energy(g["u0"])
        eng.force_eval(hguess=hmin == 0)
        eng.step_begin("rk", nlevels=nlevels, nbuildstep=8, accel_mult=0.5, courant_mult=0.2,
                       levels=levels, dt_fixed=float(g["dt_fixed"]), neib_check=check)
        for k in range(nsteps):
            eng.step(1)
            info = eng.step_info()
            got = [info[key] for key in STEP_KEYS]
            assert got == list(g["step_ints"][k]), (k, got, g["step_ints"][k])
            assert info["nactive"] == g["step_nactive"][k]
            gr = g["step_reals"][k]
            for j, key in enumerate(("time", "timestep", "dt_max")):
                assert rel_err(info[key], gr[j]) <= 1e-12, (k, key, info[key], gr[j])
        d = eng.fetch(neighbours=False)
        st = eng.fetch_step()
        en = eng.fetch_energy() if energy else None
    finally:
        eng.close()
    if energy:   # u advanced with du/dt (advance_runge_kutta.F90:79-81)
        assert rel_err(en[0], g["u"]).max() <= 1e-11
        assert np.abs(en[1] - g["du_dt"]).max() <= 1e-10 * np.abs(g["du_dt"]).max()
    # tree order after the last build equals the reference's
    assert np.array_equal(d["porig"], g["porig"])
    scale = np.abs(g["pos"]).max()
    assert np.abs(d["pos"] - g["pos"]).max() <= 1e-12 * scale
    for key in ("r_old",):
        assert np.abs(st[key] - g[key]).max() <= 1e-12 * scale
    vscale = max(np.abs(g["v"]).max(), 1e-300)
    assert np.abs(d["v"] - g["v"]).max() <= 1e-12 * vscale
    for key in ("v_old", "v_half"):
        assert np.abs(st[key] - g[key]).max() <= 1e-12 * vscale
    assert np.array_equal(st["nlevel"], g["nlevel"])
    assert np.array_equal(st["nlast"], g["nlast"])
    if check:
        assert np.array_equal(st["nminneib"], g["nminneib"])
    assert np.array_equal(st["active"], g["active"])
    assert rel_err(st["laststep"], g["laststep"]).max() <= 1e-12
    assert rel_err(d["h"], g["h"]).max() <= 1e-11
    assert rel_err(d["rho"], g["rho"]).max() <= 1e-11
    da = np.linalg.norm(d["a"] - g["a"], axis=1)
    assert (da <= 1e-10 * np.linalg.norm(g["a"], axis=1) + 1e-300).all()


def _nlast_after_timesteps(step_prev, info_prev, info):
    """nlast as sph_timesteps leaves it: zero on a resynchronisation
    (sph_timesteps.F90:62), else rescaled with level_max (:261-277)."""
    if info_prev["n"] == info_prev["nresync"]:
        return np.zeros_like(step_prev["nlast"])
    shift = info["level_step"] - info_prev["level_step"]
    nl = step_prev["nlast"]
    return nl * 2 ** shift if shift >= 0 else nl // 2 ** (-shift)


def _kdk_restated(cur, step_prev, step_cur, info_prev, info):
    """advance_leapfrog_kdk.F90:32-116 with dt_half = timestep (dn - nhalf),
    from the step state before this step (step_prev), the levels this step's
    sph_timesteps assigned (step_cur) and this step's accelerations."""
    a = cur["a"]
    act = step_prev["active"].astype(bool)
    v_half = step_prev["v_half"].copy()
    v_old = step_prev["v_old"].copy()
    v_old[act] = v_half[act] + 0.5 * a[act] * step_prev["laststep"][act, None]
    n_int, ts, lstep = info["n"], info["timestep"], info["level_step"]
    dn = n_int - _nlast_after_timesteps(step_prev, info_prev, info)
    nfull = 2 ** (lstep - step_cur["nlevel"].astype(np.int64))
    nhalf = nfull // 2
    dt = ts * dn.astype(np.float64)
    first = dn <= nhalf
    rp = np.where(first[:, None], step_prev["r_old"] + v_old * dt[:, None],
                  step_prev["r_old"] + v_half * dt[:, None])
    dth = ts * (dn - nhalf).astype(np.float64)
    vp = np.where(first[:, None], v_old + a * dt[:, None], v_half + a * dth[:, None])
    v_half = np.where((dn == nhalf)[:, None], vp, v_half)
    return rp, vp, v_half, dn == nfull


def test_kdk_advance_matches_restatement():
    from seren_amd import ics
    print("case kdk uniform 3000 nlevels 2", flush=True)
    pos, vel, mass, h = ics.make_cloud(3000, "uniform", seed=21, mach=1.0, cs=CS, ngrid=32)
    eng = _engine(False, "geometric")
    try:
        eng.load(pos, vel, mass, h)
        eng.force_eval(hguess=True)
        eng.step_begin("lfkdk", nlevels=2, nbuildstep=8, levels="adaptive", neib_check=0)
        prev = eng.fetch(neighbours=False)
        sprev = eng.fetch_step()
        iprev = eng.step_info()
        for k in range(8):
            eng.step(1)
            cur = eng.fetch(neighbours=False)
            scur = eng.fetch_step()
            info = eng.step_info()
            if not np.array_equal(cur["porig"], prev["porig"]):
                # the tree was rebuilt: bring the previous state to the new order
                inv = np.argsort(prev["porig"])[cur["porig"] - 1]
                prev = {key: v[inv] for key, v in prev.items() if isinstance(v, np.ndarray)}
                sprev = {key: v[inv] for key, v in sprev.items()}
            rp, vp, vh, full = _kdk_restated(cur, sprev, scur, iprev, info)
            assert np.array_equal(cur["pos"], rp), k
            assert np.array_equal(cur["v"], vp), k
            assert np.array_equal(scur["v_half"], vh), k
            assert np.array_equal(scur["active"].astype(bool), full), k
            prev, sprev, iprev = cur, scur, info
    finally:
        eng.close()
