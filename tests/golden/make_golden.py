"""Write the golden vectors of the force path from the reference itself.

Runs the SEREN reference routines -- the flang build of the reference's own
Fortran sources, oracle/_ref/libseren_ref_*.so (make -f oracle/Makefile.ref)
-- on small seeded clouds in the sph_setup order and stores inputs and
outputs as compressed .npz fixtures next to this script.  Large arrays are
stored as SHA-256 digests of their raw bytes; small ones in full.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle_lib import RefSeren  # noqa: E402
from seren_amd import ics  # noqa: E402

CS = 0.2
GAMMA = 1.4


def energy0(n, seed):
    """Seeded specific internal energies around c_s^2 / (gamma - 1) for the
    energy-equation cases (THERMAL=ENERGY_EQN)."""
    rng = np.random.default_rng(seed + 1000)
    return CS * CS / (GAMMA - 1.0) * (1.0 + 0.5 * rng.random(n))


# name: (n, profile, seed, mach, octupole, mac, thetamaxsqd, abserror, evaluations, full
#        [, kernel [, thermal [, h_fac [, sph]]]])
CASES = {
    "uniform1k_quad_geo": (1000, "uniform", 101, 1.0, False, "geometric", 0.25, 1e-4, 1, True),
    "plummer2k_quad_geo": (2048, "plummer", 202, 0.0, False, "geometric", 0.25, 1e-4, 1, True),
    "r2_2k_oct_gadget2_x2": (2048, "r2", 303, 2.0, True, "gadget2", 0.25, 1e-4, 2, True),
    "r2_2k_oct_gadget_x2": (2048, "r2", 304, 2.0, True, "gadget", 0.25, 1e-4, 2, True),
    "plummer2k_oct_eigen_x2": (2048, "plummer", 305, 1.0, True, "eigen", 0.25, 1e-4, 2, True),
    # M4 with the Thomas-Couchman gradient (TC_KERNEL)
    "uniform1k_quad_geo_tc": (1000, "uniform", 106, 1.0, False, "geometric", 0.25, 1e-4, 1, True,
                              "m4tc"),
    # the reference's CPU-sized case: N = 16384 uniform sphere, quadrupole, geometric
    # MAC theta = 0.5, one force evaluation (digests for big arrays)
    "c1_uniform16k_quad_geo": (16384, "uniform", 1, 0.0, False, "geometric", 0.25, 1e-4, 1,
                               False),
    # THERMAL=ENERGY_EQN: ideal gas from u, du/dt
    "r2_2k_quad_geo_energy_x2": (2048, "r2", 107, 2.0, False, "geometric", 0.25, 1e-4, 2, True,
                                 "m4", "energy"),
    # KERNEL=QUINTIC (range 3h): h_fac 1.0, pp_gather 113
    "r2_2k_quad_geo_quintic_x2": (2048, "r2", 109, 2.0, False, "geometric", 0.25, 1e-4, 2, True,
                                  "quintic", "isothermal", 1.0),
    # KERNEL=QUINTICTC, octupole + GADGET2: h_fac 0.8, pp_gather 57
    "plummer2k_oct_gadget2_quintictc_x2": (2048, "plummer", 110, 1.0, True, "gadget2", 0.25,
                                           1e-4, 2, True, "quintictc", "isothermal", 0.8),
    # SPH=STANDARD, HFIND=NUMBER (h_gather, all_sph, hydro, gravity_sph)
    "r2_2k_std_quad_geo_x2": (2048, "r2", 111, 2.0, False, "geometric", 0.25, 1e-4, 2, True,
                              "m4", "isothermal", 1.2, "standard"),
    "plummer2k_std_oct_gadget2_x3": (2048, "plummer", 112, 1.0, True, "gadget2", 0.25, 1e-4, 3,
                                     True, "m4", "isothermal", 1.2, "standard"),
    # SPH=STANDARD with HFIND=CONSTANT: every h = hmin = 0.15 (about 55 neighbours)
    "uniform2k_std_quad_geo_consth_x2": (2048, "uniform", 114, 1.0, False, "geometric", 0.25,
                                         1e-4, 2, True, "m4", "isothermal", 1.2, "standard",
                                         0.15),
    # SPH=STANDARD with THERMAL=ENERGY_EQN (hydro.F90's du/dt)
    "r2_2k_std_quad_geo_energy_x2": (2048, "r2", 113, 2.0, False, "geometric", 0.25, 1e-4, 2,
                                     True, "m4", "energy", 1.2, "standard"),
}


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_case(name, spec):
    n, prof, seed, mach, octu, mac, th2, ae, nev, full = spec[:10]
    kernel = spec[10] if len(spec) > 10 else "m4"
    thermal = spec[11] if len(spec) > 11 else "isothermal"
    h_fac = spec[12] if len(spec) > 12 else 1.2
    sph = spec[13] if len(spec) > 13 else "gradh"
    hmin = spec[14] if len(spec) > 14 else 0.0
    pos, vel, mass, h = ics.make_cloud(n, prof, seed=seed, mach=mach, cs=CS, ngrid=32)
    ref = RefSeren(octu, mac, kernel=kernel, thermal=thermal, sph=sph,
                   hfind="constant" if hmin > 0 else "number")
    ref.setup(pos, mass, h, vel, h_fac=h_fac, thetamaxsqd=th2, abserror=ae)
    if hmin > 0:
        ref.set_hmin(hmin)
    if thermal == "energy":
        ref.set_energy(energy0(n, seed), GAMMA)
    for e in range(nev):
        ref.tree_update()
        if e == 0 and hmin == 0:   # HFIND=CONSTANT: no h guess (sph_setup.F90:45-49)
            ref.hguess()
        ref.sph_update()
        ref.thermal_iso(CS)
        ref.hydro_forces()
        a_hydro = ref.particles()["a"].copy()
        ref.grav_forces()
    p = ref.particles()
    pptot, pplist = ref.neib()
    t = ref.grav_tree()
    ht = ref.hydro_tree()
    first, last = ref.levels()
    par = ref.params()
    # only pplist(1:pptot, p) is defined (get_neib.F90:115-118); later slots
    # keep stale ids from earlier evaluations in the reference
    valid = pptot <= pplist.shape[1]
    pplist = np.where(np.arange(pplist.shape[1])[None, :] < pptot[:, None], pplist, 0)
    lm = np.arange(8)[None, :] < t["leaf"][:, None]
    plist = np.where(lm, t["plist"], 0).astype(np.int32)
    ifopen = np.where(t["leaf"] == 0, t["ifopen"], 0).astype(np.int32)
    pl_valid = np.where(valid[:, None], pplist, 0).astype(np.int32)
    out = dict(
        spec=np.array([n, seed, int(octu), nev], np.int64),
        profile=np.array(prof), mac=np.array(mac), thetamaxsqd=np.array(th2),
        kernel=np.array(kernel),
        abserror=np.array(ae), mach=np.array(mach),
        pos=pos, vel=vel, mass=mass,
        porig=p["porig"], h=p["h"], rho=p["rho"], omega=p["omega"], zeta=p["zeta"],
        div_v=p["div_v"], drhodt=p["drhodt"], a=p["a"], a_hydro=a_hydro, gpot=p["gpot"],
        agravmag=p["agravmag"], pptot=pptot, first_cell=first, last_cell=last,
        ctot=np.array(par["ctot_grav"]), ltot=np.array(par["ltot_grav"]),
        rextent=np.array(par["rextent"]),
    )
    if h_fac != 1.2:
        out["h_fac"] = np.array(h_fac)
    if sph != "gradh":
        out["sph"] = np.array(sph)
    if hmin > 0:
        out["hmin"] = np.array(hmin)
    if thermal == "energy":
        out["thermal"] = np.array(thermal)
        out["gamma"] = np.array(GAMMA)
        out["u0"] = energy0(n, seed)
        out["u"], out["du_dt"] = ref.energy()
        out["press"], out["sound"] = p["press"], p["sound"]
    tree = dict(leaf=t["leaf"], nextcell=t["nextcell"], ifopen=ifopen, plist=plist,
                r=t["r"], m=t["m"], dminsqd=t["dminsqd"], mac_=t["mac"], q=t["q"], s=t["s"],
                hmax=t["hmax"], hr=ht["r"], rmax=ht["rmax"], hrangemax=ht["hrangemax"])
    if full:
        out["pplist"] = pl_valid.astype(np.int16 if n < 32768 else np.int32)
        for k, v in tree.items():
            out["tree_" + k] = v
    else:
        out["pplist_sha256"] = np.array(digest(pl_valid))
        for k, v in tree.items():
            out["tree_" + k + "_sha256"] = np.array(digest(v))
        out["tree_leaf"] = t["leaf"]          # small, kept for diagnostics
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "ctot", par["ctot_grav"], "ltot", par["ltot_grav"], "pptot max", pptot.max())


# Integration steps (sph_integrate.F90 with RUNGE_KUTTA) after the setup
# evaluation: name: (n, profile, seed, mach, octupole, mac, nlevels, steps)
STEP_CASES = {
    # n, profile, seed, Mach, octupole, MAC, nlevels, steps[, time-step build, dt_fixed
    # [, thermal [, kernel, h_fac]]]
    "step_rk_uniform2k_quad_geo_nlev3": (2000, "uniform", 11, 1.0, False, "geometric", 3, 6),
    "step_rk_r2_2k_oct_gadget2_nlev1": (2048, "r2", 12, 2.0, True, "gadget2", 1, 4),
    # Makefile TIMESTEP=RESTRICTED CHECK_NEIB_TIMESTEP=1 (the Makefile's defaults)
    "step_rk_r2_2k_quad_geo_nlev6_restricted_1": (2048, "r2", 13, 3.0, False, "geometric", 6,
                                                  24, "restricted_1", 1.0),
    # TIMESTEP=FIXED CHECK_NEIB_TIMESTEP=2 (IMMEDIATE_TIMESTEP_REDUCTION)
    "step_rk_r2_2k_quad_geo_nlev6_fixed_2": (2048, "r2", 13, 3.0, False, "geometric", 6, 24,
                                             "fixed_2", 0.02),
    # THERMAL=ENERGY_EQN: u advanced, the energy time-step criterion
    "step_rk_uniform2k_quad_geo_energy_nlev3": (2000, "uniform", 15, 2.0, False, "geometric", 3,
                                                8, "", 1.0, "energy"),
    # KERNEL=QUINTIC, h_fac 1.0
    "step_rk_r2_2k_quad_geo_quintic_nlev3": (2048, "r2", 16, 2.0, False, "geometric", 3, 8, "",
                                             1.0, "isothermal", "quintic", 1.0),
    # SPH=STANDARD with THERMAL=ENERGY_EQN: h_gather from the advanced h,
    # hydro.F90's du/dt in the u advance, the energy time-step criterion
    "step_rk_r2_2k_std_quad_geo_energy_nlev3": (2048, "r2", 17, 2.0, False, "geometric", 3, 8,
                                                "", 1.0, "energy", "m4", 1.2, "standard"),
    # SPH=STANDARD with HFIND=CONSTANT (every h = hmin = 0.15)
    "step_rk_uniform2k_std_quad_geo_consth_nlev3": (2048, "uniform", 18, 1.0, False,
                                                    "geometric", 3, 8, "", 1.0, "isothermal",
                                                    "m4", 1.2, "standard", 0.15),
}
STEP_KEYS = ["n", "nsteps", "nresync", "level_max", "level_step", "nbuild", "nstock"]


def step_case_data(spec):
    """Runs one step case on the reference and returns the fixture arrays."""
    n, prof, seed, mach, octu, mac, nlevels, nsteps = spec[:8]
    timestep, dt_fixed = spec[8:10] if len(spec) > 8 else ("", 1.0)
    thermal = spec[10] if len(spec) > 10 else "isothermal"
    kernel, h_fac = spec[11:13] if len(spec) > 11 else ("m4", 1.2)
    sph = spec[13] if len(spec) > 13 else "gradh"
    hmin = spec[14] if len(spec) > 14 else 0.0
    pos, vel, mass, h = ics.make_cloud(n, prof, seed=seed, mach=mach, cs=CS, ngrid=32)
    ref = RefSeren(octu, mac, timestep=timestep, thermal=thermal, kernel=kernel, sph=sph,
                   hfind="constant" if hmin > 0 else "number")
    ref.setup(pos, mass, h, vel, h_fac=h_fac, thetamaxsqd=0.25, abserror=1e-4)
    if hmin > 0:
        ref.set_hmin(hmin)
    if thermal == "energy":
        ref.set_energy(energy0(n, seed), GAMMA)
    ref.force_eval(CS, hguess=hmin == 0)
    ref.step_setup(nlevels=nlevels, nbuildstep=8, accel_mult=0.5, courant_mult=0.2,
                   dt_fixed=dt_fixed)
    ints, reals, nact = [], [], []
    for _ in range(nsteps):
        ref.integrate(CS)
        info = ref.step_info()
        ints.append([info[k] for k in STEP_KEYS])
        reals.append([info["time"], info["timestep"], info["dt_max"]])
        nact.append(int(ref.step_particles()["active"].sum()))
    p = ref.particles()
    sp = ref.step_particles()
    return dict(
        spec=np.array([n, seed, int(octu), nlevels, nsteps], np.int64), profile=np.array(prof),
        mac=np.array(mac), mach=np.array(mach), pos0=pos, vel0=vel, mass=mass,
        timestep=np.array(timestep), dt_fixed=np.array(dt_fixed), nminneib=sp["nminneib"],
        thermal=np.array(thermal), **energy_out(ref, n, seed, thermal),
        **({} if kernel == "m4" else dict(kernel=np.array(kernel), h_fac=np.array(h_fac))),
        **({} if sph == "gradh" else dict(sph=np.array(sph))),
        **({} if hmin == 0 else dict(hmin=np.array(hmin))),
        step_ints=np.array(ints, np.int64), step_reals=np.array(reals),
        step_nactive=np.array(nact, np.int64),
        porig=p["porig"], pos=p["pos"], v=p["v"], a=p["a"], h=p["h"], rho=p["rho"],
        nlevel=sp["nlevel"], nlast=sp["nlast"], laststep=sp["laststep"], active=sp["active"],
        r_old=sp["r_old"], v_old=sp["v_old"], v_half=sp["v_half"])


def energy_out(ref, n, seed, thermal):
    if thermal != "energy":
        return {}
    u, du = ref.energy()
    return dict(gamma=np.array(GAMMA), u0=energy0(n, seed), u=u, du_dt=du)


def run_step_case(name, spec):
    d = step_case_data(spec)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    print(name, "levels", np.bincount(d["nlevel"]), "time", d["step_reals"][-1][0])


# diagnostics.F90 (DEBUG_DIAGNOSTICS builds) after the setup evaluation, and
# after RK steps: name: (n, profile, seed, mach, thermal, steps)
DIAG_CASES = {
    "diag_r2_2k_quad_geo": (2048, "r2", 21, 2.0, "isothermal", 0),
    "diag_uniform2k_quad_geo_steps": (2048, "uniform", 22, 1.0, "isothermal", 6),
    "diag_r2_2k_quad_geo_energy": (2048, "r2", 23, 2.0, "energy", 0),
}


def run_diag_case(name, spec):
    n, prof, seed, mach, thermal, nsteps = spec
    pos, vel, mass, h = ics.make_cloud(n, prof, seed=seed, mach=mach, cs=CS, ngrid=32)
    ref = RefSeren(False, "geometric", thermal=thermal, diag=True)
    ref.setup(pos, mass, h, vel, thetamaxsqd=0.25, abserror=1e-4)
    if thermal == "energy":
        ref.set_energy(energy0(n, seed), GAMMA)
    ref.force_eval(CS, hguess=True)
    if nsteps:
        ref.step_setup(nlevels=3, nbuildstep=8, accel_mult=0.5, courant_mult=0.2, dt_fixed=1.0)
        for _ in range(nsteps):
            ref.integrate(CS)
    d = ref.diagnostics()
    vec = np.concatenate([[d["mtot"]], d["rcom"], d["vcom"], d["mom"], d["angmom"], d["force"],
                          [d["ketot"], d["gpetot"], d["utot"], d["etot"]]])
    out = dict(spec=np.array([n, seed, nsteps], np.int64), profile=np.array(prof),
               mach=np.array(mach), thermal=np.array(thermal), pos=pos, vel=vel, mass=mass,
               diag=vec)
    if thermal == "energy":
        out["u0"] = energy0(n, seed)
        out["gamma"] = np.array(GAMMA)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "etot", d["etot"], "mom", d["mom"])


def kernel_tables(kernel):
    """The reference's kernel tables (kernel.F90) for one KERNEL option."""
    octu, mac = (True, "gadget2") if kernel == "quintictc" else (False, "geometric")
    ref = RefSeren(octu, mac, kernel=kernel)
    pos, vel, mass, h = ics.make_cloud(100, "uniform", seed=9, mach=0.0)
    ref.setup(pos, mass, h, vel)
    return ref.kernel()


# kernel-table fixtures: full, they pin the tabulated functions
TABLES = {"kernel_tables": "m4", "kernel_tables_quintic": "quintic"}


def main(names=None):
    """Writes every fixture, or only the named ones."""
    for name, spec in CASES.items():
        if not names or name in names:
            run_case(name, spec)
    for name, spec in STEP_CASES.items():
        if not names or name in names:
            run_step_case(name, spec)
    for name, spec in DIAG_CASES.items():
        if not names or name in names:
            run_diag_case(name, spec)
    for name, kernel in TABLES.items():
        if not names or name in names:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), tables=kernel_tables(kernel))
            print(name, "written")


if __name__ == "__main__":
    main(sys.argv[1:])
