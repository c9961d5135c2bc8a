"""GPU: the kick-drift-kick leapfrog of the device step, pinned by what the
mathematics fixes rather than by a restatement of its update.

advance_leapfrog_kdk.F90:32-116 cannot serve as the checker: it uses dt_half
before setting it (line 27, 67), so the device integrator fixes that as
dt_half = timestep (dn - nhalf).  Three properties of the textbook KDK map

    v(1/2) = v(0) + a(x(0)) dt/2,  x(1) = x(0) + v(1/2) dt,
    v(1)   = v(1/2) + a(x(1)) dt/2

hold for any correct implementation and fail for the plausible mistakes (a
kick of dt instead of dt/2, a drift with v(0), the second kick with the old
acceleration):

* second-order convergence: on a circular Kepler orbit the position error at
  a fixed time falls by 4 when dt halves;
* time reversibility: N steps, velocities negated, N steps return the
  initial state to rounding;
* the orbit itself: after N steps the positions lie within 1e-3 of the
  analytic circle at dt = 1/16 (a first-order error would be ~dt/4);
* the textbook map stepped in numpy from the Newtonian pair force agrees to
  rounding.

The system is two equal point masses (total M = 1, G = 1) one length unit
apart on a circular orbit (angular frequency 1).  Standard SPH with a
constant h = 0.01 (HFIND=CONSTANT) keeps them out of each other's kernel
range, so there is no hydro force and the gravity pair term is exactly
Newtonian (gravity_gradh.F90:61-82 beyond 2h on both sides); the tree's root
is one leaf, so the pair is summed directly.  Fixed time steps (TIMESTEP=
FIXED, one level) with dt a power of two keep time exact.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def _state0():
    pos = np.array([[0.5, 0.0, 0.0], [-0.5, 0.0, 0.0]])
    vel = np.array([[0.0, 0.5, 0.0], [0.0, -0.5, 0.0]])
    mass = np.array([0.5, 0.5])
    return pos, vel, mass


def _run(pos, vel, mass, dt, nsteps):
    """nsteps KDK steps of size dt; returns (x_N, v_N) in input order.  One
    step is two sph_integrate calls (half steps); one more call evaluates
    a(x_N) and applies the second half kick, leaving x_N in r_old and v_N in
    v_old."""
    from seren_amd.engine import ForceEngine
    eng = ForceEngine(multipole="quadrupole", mac="geometric", precision="fp64",
                      thetamaxsqd=0.25, sound_speed=0.2, sph="standard", hmin=0.01)
    try:
        eng.load(pos, vel, mass, np.full(len(mass), 0.01))
        eng.force_eval()
        eng.step_begin("lfkdk", nlevels=1, nbuildstep=1, levels="fixed", dt_fixed=dt,
                       neib_check=0)
        eng.step(2 * nsteps + 1)
        info = eng.step_info()
        st = eng.fetch_step()
        d = eng.fetch(neighbours=False)
    finally:
        eng.close()
    order = np.argsort(d["porig"])
    assert info["dt_max"] == dt
    return st["r_old"][order], st["v_old"][order]


def _circle(t):
    c, s = np.cos(t), np.sin(t)
    return np.array([[0.5 * c, 0.5 * s, 0.0], [-0.5 * c, -0.5 * s, 0.0]])


def test_kdk_is_second_order_on_a_kepler_orbit():
    pos, vel, mass = _state0()
    errs = []
    for k in (4, 5, 6):
        dt = 2.0 ** -k
        x, v = _run(pos, vel, mass, dt, 2 ** k)          # to t = 1
        e = np.abs(x - _circle(1.0)).max()
        print(f"  dt 2^-{k}: |x - circle| = {e:.3e}", flush=True)
        errs.append(e)
        # momentum: equal masses, opposite velocities
        assert np.abs(mass @ v).max() < 1e-14
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print(f"  ratios {r1:.3f} {r2:.3f}", flush=True)
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4
    # the orbit is followed: a first-order kick error would leave ~dt/4
    assert errs[0] < 1e-3


def test_kdk_is_time_reversible():
    pos, vel, mass = _state0()
    dt, nsteps = 2.0 ** -5, 40
    x, v = _run(pos, vel, mass, dt, nsteps)
    assert np.abs(x - pos).max() > 0.1                   # it moved
    x2, v2 = _run(x, -v, mass, dt, nsteps)
    dx, dv = np.abs(x2 - pos).max(), np.abs(v2 + vel).max()
    print(f"  back to start: |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-13 and dv < 1e-13


def test_kdk_matches_the_textbook_map():
    """The same orbit stepped with the textbook KDK map in numpy from the
    Newtonian pair force: agreement to rounding."""
    pos, vel, mass = _state0()
    dt, nsteps = 2.0 ** -5, 32
    x, v = _run(pos, vel, mass, dt, nsteps)

    def acc(p):
        d = p[1] - p[0]
        r = np.sqrt(d @ d)
        f = d / r ** 3
        return np.array([mass[1] * f, -mass[0] * f])

    xr, vr = pos.copy(), vel.copy()
    a = acc(xr)
    for _ in range(nsteps):
        vh = vr + 0.5 * dt * a
        xr = xr + dt * vh
        a = acc(xr)
        vr = vh + 0.5 * dt * a
    print(f"  |dx| {np.abs(x - xr).max():.2e} |dv| {np.abs(v - vr).max():.2e}", flush=True)
    assert np.abs(x - xr).max() < 1e-13 and np.abs(v - vr).max() < 1e-13
