"""GPU parity: the gfx950 force path against the CPU restatement oracle on
identical seeded inputs.

Bar (per quantity; the reference-vector test adds u, du/dt, P and c for
the energy equation, bar in that test):
  * tree topology (BFS ids, leaf counts, plist, nextcell, ifopen, levels),
    tree order (porig), neighbour counts and lists: bit-exact;
  * stocked moments, COM, opening distances, hydro-cell centres, radii,
    hrangemax: bit-exact;
  * h, rho, Omega, zeta, div v, drho/dt: relative error <= 1e-12 and at most
    2% of particles not bit-exact.  The h solve calls pow(m/rho, 0.333333333)
    and pow(1/h, 3.0): the device evaluates them correctly rounded, host libm
    pow misrounds about 1e-3 of its arguments; all other arithmetic is the
    reference's, bit for bit;
  * hydro acceleration: |da| / |a| <= 1e-12;
  * gravity: the reference adds all cell terms before the pair terms; here
    they are summed in two accumulators, so |da| / |a_grav| <= 1e-12 and the
    potential to 1e-12.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, checker_eval, cloud, rel_err, vec_rel_err  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [
    # n, profile, seed, octupole, mac, theta^2, abserror, evaluations
    (2000, "uniform", 3, False, "geometric", 0.25, 1e-4, 1),
    (6000, "plummer", 5, False, "geometric", 0.25, 1e-4, 1),
    (5000, "r2", 7, True, "gadget2", 0.25, 1e-4, 2),
    # exact branch octupole shift (no reference counterpart; the oracle's
    # moments are pinned against direct per-cell sums)
    (4000, "r2", 8, "exact", "gadget2", 0.25, 1e-4, 2),
]

_MULTI = {False: "quadrupole", True: "octupole", "exact": "octupole_exact"}


def _ids(c):
    return f"n{c[0]}-{c[1]}-{_MULTI[c[3]]}-{c[4]}-x{c[7]}"


def _engine(octupole, mac, theta2, abserr, kernel="m4", thermal="isothermal", h_fac=1.2,
            sph="gradh", hmin=0.0):
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole=_MULTI[octupole], mac=mac,
                       precision="fp64", thetamaxsqd=theta2, abserror=abserr, sound_speed=CS,
                       kernel=kernel, thermal=thermal, h_fac=h_fac, sph=sph, hmin=hmin)


def _report(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    ne = int(np.sum(a != b))
    print(f"  {name}: {ne} of {a.size} differ", flush=True)
    return ne


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_force_eval_matches_oracle(case):
    from oracle_lib import CSeren
    n, prof, seed, octu, mac, th2, ae, nev = case
    print(f"case {_ids(case)}", flush=True)
    pos, vel, mass, h = cloud(n, prof, seed)
    ref = checker_eval(CSeren(octu, mac), pos, vel, mass, h, octupole=octu, mac=mac,
                       theta2=th2, abserr=ae, evals=nev)
    eng = _engine(octu, mac, th2, ae)
    # hydro-only twin of the first evaluation
    eng.load(pos, vel, mass, h)
    eng.force_eval(hguess=True, grav=False)
    a_hydro = eng.fetch(neighbours=False)["a"]
    eng.load(pos, vel, mass, h)
    for e in range(nev):
        eng.force_eval(hguess=(e == 0))
        g = eng.fetch()
        t = eng.export_tree()
        r = ref[e]
        print(f" evaluation {e}", flush=True)
        # tree order and topology
        assert _report("porig", g["porig"], r["porig"]) == 0
        rt = r["tree"]
        assert len(t["leaf"]) == len(rt["leaf"])
        assert _report("leaf", t["leaf"], rt["leaf"]) == 0
        assert _report("nextcell", t["nextcell"], rt["nextcell"]) == 0
        br = rt["leaf"] == 0
        assert _report("ifopen", t["ifopen"][br], rt["ifopen"][br]) == 0
        lm = np.arange(8)[None, :] < rt["leaf"][:, None]
        assert _report("plist", np.where(lm, t["plist"], 0), np.where(lm, rt["plist"], 0)) == 0
        f, l = r["levels"]
        assert _report("first_cell", t["first_cell"], f) == 0
        assert _report("last_cell", t["last_cell"], l) == 0
        for k in ("r", "m", "q", "s") + (("mac",) if mac != "geometric" else ()):
            assert _report("tree " + k, t[k], rt[k]) == 0
        # opening distances use hmax of the h stocked with the tree (h of the
        # previous evaluation); from the second evaluation on, a last-bit h
        # difference can reach a few cells
        nd = _report("tree dminsqd", t["dminsqd"], rt["dminsqd"])
        if e == 0:
            assert nd == 0
        else:
            assert rel_err(t["dminsqd"], rt["dminsqd"]).max() <= 1e-12
        assert _report("hydro r", t["hr"], r["htree"]["r"]) == 0
        assert _report("rmax", t["rmax"], r["htree"]["rmax"]) == 0
        # h solve
        for k in ("h", "rho", "omega", "zeta", "div_v", "drhodt"):
            err = rel_err(g[k], r[k]) if k not in ("div_v", "drhodt", "zeta") else \
                np.abs(g[k] - r[k]) / np.maximum(np.abs(r[k]).max(), 1e-300)
            nbit = int(np.sum(g[k] != r[k]))
            print(f"  {k}: max rel {err.max():.3e}, {nbit} not bit-exact", flush=True)
            assert err.max() <= 1e-12, k
            # host libm pow misrounds ~1e-3 of its arguments; the device pow
            # is correctly rounded, so a few particles differ in the last bit
            assert nbit <= max(3, 0.02 * n), k
        eh = rel_err(t["hrangemax"], r["htree"]["hrangemax"])
        print(f"  hrangemax: max rel {eh.max():.3e}", flush=True)
        assert eh.max() <= 1e-12
        # neighbour lists
        assert _report("pptot", g["pptot"], r["pptot"]) == 0
        ok = r["pptot"] <= r["pplist"].shape[1]
        assert _report("pplist", g["pplist"][ok], r["pplist"][ok]) == 0
        # forces
        if e == 0:
            eh = vec_rel_err(a_hydro, r["a_hydro"])
            print(f"  a_hydro: max rel {eh.max():.3e}", flush=True)
            assert eh.max() <= 1e-12
        ag_ref = r["a"] - r["a_hydro"]
        ea = np.linalg.norm(g["a"] - r["a"], axis=1) / np.maximum(
            np.linalg.norm(ag_ref, axis=1), 1e-300)
        ep = rel_err(g["gpot"], r["gpot"])
        print(f"  a: max |da|/|a_grav| {ea.max():.3e}; gpot max rel {ep.max():.3e}", flush=True)
        assert ea.max() <= 1e-12
        assert ep.max() <= 1e-12
        if mac != "geometric":
            em = rel_err(g["agravmag"], r["agravmag"])
            print(f"  agravmag max rel {em.max():.3e}", flush=True)
            assert em.max() <= 1e-12
    eng.close()


def test_host_state_dropin_matches_oracle():
    """Drive the library through the SEREN-shaped host-state entry points in
    the sph_setup order and compare with the oracle."""
    from oracle_lib import CSeren
    from seren_amd import engine as E
    n = 3000
    pos, vel, mass, h = cloud(n, "uniform", 21)
    r = checker_eval(CSeren(False, "geometric"), pos, vel, mass, h, octupole=False,
                     mac="geometric", theta2=0.25, abserr=1e-4)[0]
    eng = _engine(False, "geometric", 0.25, 1e-4)
    st = E.SerenState(pos, vel, mass, h, eng.pp_limit)
    E.tree_update(eng, st)
    E.BHhydro_hguess(eng, st)
    E.sph_update(eng, st)
    E.update_thermal_properties_isothermal(st, CS)
    E.sph_hydro_forces(eng, st)
    E.sph_grav_forces(eng, st)
    assert np.array_equal(st.porig, r["porig"])
    assert rel_err(st.parray[:, 4], r["h"]).max() <= 1e-12
    assert np.array_equal(st.pptot, r["pptot"])
    ok = r["pptot"] <= eng.pp_limit
    assert np.array_equal(st.pplist[ok], r["pplist"][ok])
    ea = np.linalg.norm(st.a - r["a"], axis=1) / np.maximum(np.linalg.norm(r["a"], axis=1), 1e-300)
    assert ea.max() <= 1e-10
    assert rel_err(st.gpot, r["gpot"]).max() <= 1e-12
    eng.close()


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLDEN_CASES = sorted(f[:-4] for f in os.listdir(GOLD)
                      if f.endswith(".npz") and not f.startswith("kernel_tables")
                      and not f.startswith("step_") and not f.startswith("diag_"))


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_gpu_matches_reference_golden_vectors(case):
    """The HIP path against vectors written by the reference routines."""
    _golden_check(case)


# golden cases whose lists go through the fallback path when the h-solve
# list holds 2 leaves per particle: grad-h (all_sph_gradh) and h_gather
LIST_OVERFLOW_CASES = ["r2_2k_oct_gadget2_x2", "plummer2k_quad_geo", "r2_2k_std_quad_geo_x2",
                       "r2_2k_quad_geo_quintic_x2"]


@pytest.mark.parametrize("case", LIST_OVERFLOW_CASES)
def test_list_overflow_walk_path_matches_reference(case):
    """A particle whose potential-neighbour list does not fit the recorded
    slots is summed over the tree walk at the same radius, the counterpart
    of the reference's pp_pot = -1 retry (all_sph_gradh.F90:171-184,
    BHhydrowalk_hgather.F90:46-49).  With 2 slots every wave takes that
    path, and the results must still be the reference's."""
    stats = _golden_check(case, slots=2)
    print(f"  fallback passes {stats['hsolve_overflow_passes']}", flush=True)
    assert stats["hsolve_overflow_passes"] > 0


def test_golden_cases_reach_the_overflow_neighbour_rows():
    """Rows with more than pp_limit neighbours get complete CSR lists (the
    reference recomputes them on the fly, hydro_gradh.F90:126-140,
    get_neib.F90:115-118); the golden comparisons above include such rows."""
    over = {}
    for case in GOLDEN_CASES:
        g = np.load(os.path.join(GOLD, case + ".npz"))
        lim = g["pplist"].shape[1] if "pplist" in g.files else 114
        over[case] = int(np.sum(g["pptot"] > lim))
    print(over, flush=True)
    assert over["plummer2k_quad_geo"] > 100 and over["r2_2k_oct_gadget2_x2"] > 50
    assert over["c1_uniform16k_quad_geo"] >= 1


def _golden_check(case, slots=None):
    import hashlib
    g = np.load(os.path.join(GOLD, case + ".npz"))
    n, seed, octu, nev = (int(v) for v in g["spec"])
    mac = str(g["mac"])
    kernel = str(g["kernel"]) if "kernel" in g.files else "m4"
    energy = "thermal" in g.files and str(g["thermal"]) == "energy"
    print(f"golden {case}", flush=True)
    h_fac = float(g["h_fac"]) if "h_fac" in g.files else 1.2
    sph = str(g["sph"]) if "sph" in g.files else "gradh"
    hmin = float(g["hmin"]) if "hmin" in g.files else 0.0
    eng = _engine(bool(octu), mac, float(g["thetamaxsqd"]), float(g["abserror"]), kernel,
                  "energy" if energy else "isothermal", h_fac, sph, hmin)
    if slots is not None:
        from seren_amd.engine import TUNE_HSOLVE_SLOTS
        eng.set_tuning(TUNE_HSOLVE_SLOTS, slots)
    eng.load(g["pos"], g["vel"], g["mass"], np.zeros(n))
    if energy:
        assert float(g["gamma"]) == eng.opt.gamma
        eng.set_energy(g["u0"])
    for e in range(nev):
        eng.force_eval(hguess=(e == 0 and hmin == 0))
    stats = eng.walk_stats()
    d = eng.fetch()
    t = eng.export_tree()
    assert np.array_equal(d["porig"], g["porig"])
    assert np.array_equal(d["pptot"], g["pptot"])
    lim = d["pplist"].shape[1]
    cols = np.arange(lim)[None, :] < d["pptot"][:, None]
    pl = np.where((d["pptot"] <= lim)[:, None] & cols, d["pplist"], 0).astype(np.int32)
    exact = ("leaf", "nextcell", "ifopen", "plist", "r", "m", "q", "s", "hr", "rmax")
    if nev == 1:
        exact = exact + ("dminsqd",) + (("mac_",) if mac != "geometric" else ())
    mine = dict(leaf=t["leaf"], nextcell=t["nextcell"], ifopen=t["ifopen"], plist=t["plist"],
                r=t["r"], m=t["m"], q=t["q"], s=t["s"], hr=t["hr"], rmax=t["rmax"],
                dminsqd=t["dminsqd"], mac_=t["mac"])
    if "pplist" in g:
        assert np.array_equal(pl, g["pplist"].astype(np.int32))
        for k in exact:
            assert np.array_equal(mine[k], g["tree_" + k]), k
    else:
        dig = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa
        assert dig(pl) == str(g["pplist_sha256"])
        for k in exact:
            assert dig(mine[k]) == str(g["tree_" + k + "_sha256"]), k
    for k in ("h", "rho", "omega"):
        err = rel_err(d[k], g[k])
        nbit = int(np.sum(d[k] != g[k]))
        print(f"  {k}: max rel {err.max():.3e}, {nbit} not bit-exact", flush=True)
        assert err.max() <= 1e-12 and nbit <= max(3, 0.02 * n)
    ag = g["a"] - g["a_hydro"]
    ea = np.linalg.norm(d["a"] - g["a"], axis=1) / np.maximum(np.linalg.norm(ag, axis=1), 1e-300)
    print(f"  a: max |da|/|a_grav| {ea.max():.3e}", flush=True)
    assert ea.max() <= 1e-12
    assert rel_err(d["gpot"], g["gpot"]).max() <= 1e-12
    if energy:
        # u moves with the tree order only; du/dt = viscous heating - P div v /
        # (Omega rho) carries the ~1e-12 relative h-solve differences of div v
        u, du = eng.fetch_energy()
        assert np.array_equal(u, g["u"])
        edu = np.abs(du - g["du_dt"]).max() / np.abs(g["du_dt"]).max()
        print(f"  du/dt: max |d(du/dt)| / max |du/dt| {edu:.3e}, "
              f"{int(np.sum(du != g['du_dt']))} not bit-exact", flush=True)
        assert edu <= 1e-11
        for k in ("press", "sound"):
            assert rel_err(d[k], g[k]).max() <= 1e-12, k
    eng.close()
    return stats


@pytest.mark.parametrize("profile,octupole,mac", [("r2", True, "gadget2"),
                                                   ("plummer", False, "geometric")])
def test_neighbour_walk_modes_agree(profile, octupole, mac):
    """The wave-uniform neighbour walk (any superset of the reference's
    leaves gives the reference's lists after the pair test) against the
    per-particle walks of BHhydro_walk.F90: identical pptot and pplist, and
    so identical forces, on a 40k cloud (two evaluations: stale and fresh
    boxes)."""
    from seren_amd.engine import TUNE_NEIB_GROUP
    pos, vel, mass, h = cloud(40000, profile, 77)
    out = []
    for mode in (0, 1):
        eng = _engine(octupole, mac, 0.25, 1e-4)
        try:
            eng.set_tuning(TUNE_NEIB_GROUP, mode)
            eng.load(pos, vel, mass, h)
            eng.force_eval(hguess=True)
            eng.force_eval()
            out.append(eng.fetch())
        finally:
            eng.close()
    a, b = out
    assert np.array_equal(a["pptot"], b["pptot"])
    lim = a["pplist"].shape[1]
    ok = a["pptot"] <= lim
    assert np.array_equal(a["pplist"][ok], b["pplist"][ok])
    assert int(np.sum(~ok)) > 0 or profile == "plummer"
    for k in ("a", "rho", "h"):
        assert np.array_equal(a[k], b[k]), k


def test_hsolve_occupancy_modes_agree():
    """The grad-h iteration kernel held to 8 waves per SIMD (registers capped,
    a few spills) and at the compiler's register budget: the same h, rho,
    Omega and neighbour lists bit for bit, in both precisions (the default
    picks by precision)."""
    from seren_amd.engine import TUNE_HSOLVE_WAVES
    pos, vel, mass, h = cloud(30000, "r2", 78)
    for precision in ("fp64", "mixed"):
        out = []
        for waves in (0, 8, -1):
            eng = _engine(True, "gadget2", 0.25, 1e-4)
            try:
                eng.set_precision(precision)
                eng.set_tuning(TUNE_HSOLVE_WAVES, waves)
                assert eng.get_tuning(TUNE_HSOLVE_WAVES) == waves
                eng.load(pos, vel, mass, h)
                eng.force_eval(hguess=True)
                eng.force_eval()
                out.append(eng.fetch())
            finally:
                eng.close()
        for d in out[1:]:
            for k in ("h", "rho", "omega", "zeta", "pptot", "a", "gpot"):
                assert np.array_equal(out[0][k], d[k]), (precision, k)
