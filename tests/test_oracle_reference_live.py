"""Oracle against the live reference build on fresh seeded inputs.

Runs only where oracle/_ref has been built from the reference sources
(make -f oracle/Makefile.ref); elsewhere the committed golden vectors pin the
oracle (test_oracle_golden.py).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
pytestmark = pytest.mark.skipif(
    not os.path.exists(os.path.join(REF_DIR, "libseren_ref_quad_geo.so")),
    reason="reference not built (oracle/_ref)")

CASES = [
    (1500, "uniform", 11, False, "geometric", 1),
    (2500, "plummer", 12, False, "geometric", 1),
    (3000, "r2", 13, True, "gadget2", 2),
    (700, "uniform", 14, True, "gadget2", 3),
    (2000, "r2", 15, True, "gadget", 2),
    (2000, "plummer", 16, True, "eigen", 2),
    (1500, "uniform", 17, False, "geometric", 1, "m4tc"),
    (2000, "r2", 18, False, "geometric", 2, "quintic", 1.0),
    (1500, "plummer", 19, True, "gadget2", 2, "quintictc", 0.8),
    # SPH=STANDARD, HFIND=NUMBER
    (2000, "r2", 20, False, "geometric", 2, "m4", 1.2, "standard"),
    (1500, "plummer", 21, True, "gadget2", 3, "m4", 1.2, "standard"),
    # SPH=STANDARD with THERMAL=ENERGY_EQN
    (2000, "plummer", 22, False, "geometric", 2, "m4", 1.2, "standard", "energy"),
]


@pytest.mark.parametrize("case", CASES)
def test_oracle_bit_exact_with_reference(case):
    from oracle_lib import CSeren, RefSeren
    from seren_amd import ics
    n, prof, seed, octu, mac, nev = case[:6]
    kernel = case[6] if len(case) > 6 else "m4"
    h_fac = case[7] if len(case) > 7 else 1.2
    sph = case[8] if len(case) > 8 else "gradh"
    thermal = case[9] if len(case) > 9 else "isothermal"
    pos, vel, mass, h = ics.make_cloud(n, prof, seed=seed, mach=1.5, ngrid=16)
    u0 = np.random.default_rng(seed).uniform(0.05, 0.15, n)
    outs = []
    for cls in (RefSeren, CSeren):
        # the reference selects its THERMAL build; the oracle switches on set_energy
        kw = {"thermal": thermal} if cls is RefSeren else {}
        o = cls(octu, mac, kernel=kernel, sph=sph, **kw)
        o.setup(pos, mass, h, vel, h_fac=h_fac, thetamaxsqd=0.25, abserror=1e-4)
        if thermal == "energy":
            o.set_energy(u0, 5.0 / 3.0)
        for e in range(nev):
            o.tree_update()
            if e == 0:
                o.hguess()
            o.sph_update()
            o.thermal_iso(0.2)
            o.hydro_forces()
            o.grav_forces()
        outs.append((o.particles(), o.neib(), o.grav_tree(),
                     o.energy() if thermal == "energy" else None))
    (pr, nr, tr, er), (pc, nc, tc, ec) = outs
    if thermal == "energy":
        assert np.array_equal(er[0], ec[0]) and np.array_equal(er[1], ec[1]), "u, du_dt"
        assert np.any(er[1] != 0.0)
    for k in ("porig", "h", "rho", "omega", "zeta", "div_v", "a", "gpot", "agravmag"):
        assert np.array_equal(pr[k], pc[k]), k
    assert np.array_equal(nr[0], nc[0])
    ok = (nr[0] <= nr[1].shape[1])[:, None] & (np.arange(nr[1].shape[1])[None, :] < nr[0][:, None])
    assert np.array_equal(np.where(ok, nr[1], 0), np.where(ok, nc[1], 0))
    for k in ("leaf", "nextcell", "m", "r", "q", "s", "dminsqd", "mac"):
        assert np.array_equal(tr[k], tc[k]), k


def test_step_golden_reproduced_by_reference():
    """The committed integration-step fixtures are what the reference's own
    sph_timesteps / sph_advance (RUNGE_KUTTA) produce, bit for bit."""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_golden
    for name, spec in make_golden.STEP_CASES.items():
        g = np.load(os.path.join(HERE, "golden", name + ".npz"))
        d = make_golden.step_case_data(spec)
        for key in g.files:
            assert np.array_equal(np.asarray(d[key]), g[key]), (name, key)


@pytest.mark.parametrize("case", [(3000, "plummer", 31, False, "geometric"),
                                  (2500, "r2", 32, True, "gadget2")])
def test_oracle_direct_sum_equals_reference(case):
    """The oracle's kernel-softened direct sum (the ground truth of the
    gravity error metric) against the reference's direct_sph_gravity.F90 on
    the same state: bit for bit, every target."""
    from oracle_lib import CSeren, RefSeren
    from seren_amd import ics
    n, prof, seed, octu, mac = case
    pos, vel, mass, h = ics.make_cloud(n, prof, seed=seed, mach=1.0, ngrid=16)
    targets = np.arange(1, n + 1, 7, dtype=np.int32)   # 1-based particle ids
    res = []
    for cls in (RefSeren, CSeren):
        o = cls(octu, mac)
        o.setup(pos, mass, h, vel, thetamaxsqd=0.25, abserror=1e-4)
        o.force_eval(0.2, hguess=True)
        res.append(o.direct_gravity(targets))
    (ar, pr), (ac, pc) = res
    assert np.array_equal(ar, ac) and np.array_equal(pr, pc)
    assert np.all(np.linalg.norm(ar, axis=1) > 0)
