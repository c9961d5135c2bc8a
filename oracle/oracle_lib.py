"""ctypes front-ends for the two CPU checkers of the force path.

TEST INFRASTRUCTURE ONLY (imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; never by the product package).

* ``RefSeren``  -- the SEREN reference routines compiled from the reference's
  Fortran sources with flang (oracle/Makefile.ref) plus the project's dump
  harness oracle/ref_driver.F90.  Module state is global, so one live
  instance per variant per process.
* ``CSeren``    -- the C restatement oracle/seren_oracle.c.

Both expose the same calls with the same array shapes, so tests can run the
same sequence on each and compare:

    setup(pos, mass, h, vel)   tree_update()   hguess()   sph_update()
    thermal_iso(cs)   hydro_forces()   grav_forces()   direct_gravity(t)
    params() levels() particles() neib() grav_tree() hydro_tree() kernel()

Particle outputs are in the reordered (tree) order; ``porig`` maps each row
to its 1-based input index, as SEREN's own ``porig`` array does.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KERNTOT = 10000
LMAX = 30

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# kernel name -> the kernel option of the restatement and of the GPU library
KERNELS = {"m4": 0, "m4tc": 1, "quintic": 2, "quintictc": 3}

VARIANTS = {
    # (octupole, mac[, kernel]) -> reference library variant
    (False, "geometric", "m4tc"): "quad_geo_tc",
    # Makefile KERNEL=QUINTIC and QUINTICTC
    (False, "geometric", "quintic"): "quad_geo_quintic",
    (True, "gadget2", "quintictc"): "oct_gadget2_quintictc",
    (False, "geometric"): "quad_geo",
    (True, "gadget2"): "oct_gadget2",
    (True, "gadget"): "oct_gadget",
    (True, "eigen"): "oct_eigen",
    # time-step options (Makefile TIMESTEP, CHECK_NEIB_TIMESTEP)
    (False, "geometric", "m4", "restricted_1"): "quad_geo_tsr1",
    (False, "geometric", "m4", "fixed_2"): "quad_geo_tsf2",
    # THERMAL=ENERGY_EQN
    (False, "geometric", "m4", "energy"): "quad_geo_ee",
    # SPH=STANDARD, HFIND=NUMBER
    (False, "geometric", "m4", "standard"): "std_quad_geo",
    (True, "gadget2", "m4", "standard"): "std_oct_gadget2",
    # SPH=STANDARD with THERMAL=ENERGY_EQN
    (False, "geometric", "m4", "standard_energy"): "std_quad_geo_ee",
    # SPH=STANDARD, HFIND=CONSTANT
    (False, "geometric", "m4", "standard_consth"): "std_quad_geo_consth",
    # diagnostics.F90 builds (DEBUG_DIAGNOSTICS)
    (False, "geometric", "m4", "diag"): "quad_geo_diag",
    (False, "geometric", "m4", "energy_diag"): "quad_geo_ee_diag",
}


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class _Common:
    n: int
    pp_limit: int

    def force_eval(self, cs: float, hguess: bool = True) -> None:
        """The sph_setup.F90:42-66 order: tree, [hguess], sph, thermal, hydro, grav."""
        self.tree_update()
        if hguess:
            self.hguess()
        self.sph_update()
        self.thermal_iso(cs)
        self.hydro_forces()
        self.grav_forces()

    def particles(self) -> Dict[str, np.ndarray]:
        n = self.n
        porig = np.zeros(n, np.int32)
        parray = np.zeros((n, 6))
        v = np.zeros((n, 3))
        a = np.zeros((n, 3))
        outs = [np.zeros(n) for _ in range(8)]
        self._get_particles(porig, parray, v, a, *outs)
        gpot, rho, omega, divv, drhodt, press, sound, agm = outs
        return dict(porig=porig, pos=parray[:, 0:3].copy(), m=parray[:, 3].copy(),
                    h=parray[:, 4].copy(), zeta=parray[:, 5].copy(), v=v, a=a,
                    gpot=gpot, rho=rho, omega=omega, div_v=divv, drhodt=drhodt,
                    press=press, sound=sound, agravmag=agm)

    def params(self) -> Dict[str, float]:
        ip = np.zeros(8, np.int32)
        rp = np.zeros(4)
        self._get_params(ip, rp)
        keys = ["pp_gather", "pp_limit", "ctot_grav", "ltot_grav", "ctot_hydro",
                "ltot_hydro", "cmax_grav", "cmax_hydro"]
        out = {k: int(v) for k, v in zip(keys, ip)}
        out.update(rextent=float(rp[0]), thetamaxsqd=float(rp[1]),
                   abserror=float(rp[2]), beta=float(rp[3]))
        return out

    def levels(self):
        f = np.zeros(LMAX + 1, np.int32)
        l = np.zeros(LMAX + 1, np.int32)
        self._get_levels(f, l)
        return f, l

    def neib(self):
        n = self.n
        pptot = np.zeros(n, np.int32)
        pplist = np.zeros((n, self.pp_limit), np.int32)
        self._get_neib(pptot, pplist)
        return pptot, pplist

    def grav_tree(self) -> Dict[str, np.ndarray]:
        nc = self.params()["ctot_grav"] + 1
        leaf = np.zeros(nc, np.int32)
        plist = np.zeros((nc, 8), np.int32)
        nextcell = np.zeros(nc, np.int32)
        ifopen = np.zeros(nc, np.int32)
        r = np.zeros((nc, 3))
        m = np.zeros(nc)
        dminsqd = np.zeros(nc)
        mac = np.zeros(nc)
        q = np.zeros((nc, 5))
        s = np.zeros((nc, 10))
        hmax = np.zeros(nc)
        self._get_grav_tree(nc, leaf, plist, nextcell, ifopen, r, m, dminsqd, mac, q, s, hmax)
        return dict(leaf=leaf, plist=plist, nextcell=nextcell, ifopen=ifopen, r=r, m=m,
                    dminsqd=dminsqd, mac=mac, q=q, s=s, hmax=hmax)

    def hydro_tree(self) -> Dict[str, np.ndarray]:
        nc = self.params()["ctot_hydro"] + 1
        r = np.zeros((nc, 3))
        rmax = np.zeros(nc)
        hr = np.zeros(nc)
        self._get_hydro_tree(nc, r, rmax, hr)
        return dict(r=r, rmax=rmax, hrangemax=hr)

    def kernel(self) -> np.ndarray:
        tab = np.zeros((9, KERNTOT + 1))
        self._get_kernel(tab)
        return tab

    def direct_gravity(self, targets) -> tuple:
        t = np.ascontiguousarray(targets, dtype=np.int32)
        ag = np.zeros((len(t), 3))
        pot = np.zeros(len(t))
        self._direct(len(t), t, ag, pot)
        return ag, pot


class RefSeren(_Common):
    """The reference Fortran path (flang build of the reference sources)."""

    _libs: Dict[str, C.CDLL] = {}

    def __init__(self, octupole: bool, mac: str, dropin: bool = False, kernel: str = "m4",
                 timestep: str = "", thermal: str = "isothermal", sph: str = "gradh",
                 hfind: str = "number", diag: bool = False):
        """dropin=True loads the reference force path whose five drivers are
        replaced by the Fortran binding of the GPU library (make -f
        oracle/Makefile.ref dropin): the reference code as the host.
        timestep selects a build with the Makefile's TIMESTEP and
        CHECK_NEIB_TIMESTEP options ("restricted_1", "fixed_2"); the default
        builds have neither.  thermal="energy" selects the THERMAL=ENERGY_EQN
        build (set u with set_energy; thermal_iso then runs the reference's
        update_thermal_properties).  sph="standard" selects the SPH=STANDARD
        HFIND=NUMBER build (h_gather, all_sph, hydro, gravity_sph).
        diag=True selects the build with diagnostics.F90 (DEBUG_DIAGNOSTICS)
        of the quad_geo or quad_geo_ee option set (see diagnostics())."""
        if diag:
            key = (bool(octupole), mac, kernel, "energy_diag" if thermal == "energy" else "diag")
        elif sph == "standard":
            key = (bool(octupole), mac, kernel,
                   "standard_consth" if hfind == "constant" else
                   "standard_energy" if thermal == "energy" else "standard")
        elif thermal == "energy":
            key = (bool(octupole), mac, kernel, "energy")
        elif timestep:
            key = (bool(octupole), mac, kernel, timestep)
        else:
            key = (bool(octupole), mac) if kernel == "m4" else (bool(octupole), mac, kernel)
        variant = VARIANTS[key]
        kind = "dropin" if dropin else "ref"
        path = os.path.join(HERE, "_ref", f"libseren_{kind}_{variant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -f oracle/Makefile.ref)")
        variant = kind + "_" + variant
        if variant not in self._libs:
            lib = C.CDLL(path)
            lib.ref_setup.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, C.c_double]
            lib.ref_thermal_iso.argtypes = [C.c_double]
            lib.ref_direct_gravity.argtypes = [C.c_int, _ip, _dp, _dp]
            if diag:
                lib.ref_diagnostics.argtypes = [C.c_int, C.c_char_p]
            lib.ref_set_agravmag.argtypes = [C.c_int, _dp]
            lib.ref_set_h.argtypes = [C.c_int, _dp]
            lib.ref_get_params.argtypes = [_ip, _dp]
            lib.ref_get_levels.argtypes = [_ip, _ip]
            lib.ref_get_particles.argtypes = [C.c_int, _ip, _dp, _dp, _dp] + [_dp] * 8
            lib.ref_get_neib.argtypes = [C.c_int, C.c_int, _ip, _ip]
            lib.ref_get_grav_tree.argtypes = [C.c_int, _ip, _ip, _ip, _ip] + [_dp] * 7
            lib.ref_get_hydro_tree.argtypes = [C.c_int, _dp, _dp, _dp]
            lib.ref_get_kernel.argtypes = [_dp]
            lib.ref_step_setup.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double]
            lib.ref_set_energy.argtypes = [C.c_int, _dp, C.c_double]
            lib.ref_set_hmin.argtypes = [C.c_double]
            lib.ref_get_energy.argtypes = [C.c_int, _dp, _dp]
            lib.ref_integrate.argtypes = [C.c_double]
            lib.ref_get_step.argtypes = [_lp, _dp]
            lib.ref_get_step_particles.argtypes = [C.c_int, _lp, _lp, _dp, _ip, _dp, _dp, _dp,
                                                   _lp]
            self._libs[variant] = lib
        self.lib = self._libs[variant]
        self.variant = variant

    DIAG_FIELDS = ("nsteps", "time", "mtot", "rcom", "vcom", "mom", "angmom", "force", "utot",
                   "gpetot", "ketot", "etot", "Atot")

    def diagnostics(self) -> Dict[str, np.ndarray]:
        """Runs the reference's diagnostics.F90 on the current state (diag
        builds only) and reads back the line it writes: nsteps, time, mtot,
        rcom, vcom, mom, angmom, force, utot, gpetot, ketot, etot, Atot, each
        to the 11 significant digits of its E18.10 format."""
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = d.rstrip("/") + "/"
            self.lib.ref_diagnostics(len(path), path.encode())
            line = open(os.path.join(d, "ref.diag")).read().splitlines()[-1]
        vals = [float(line[8 + 18 * k:8 + 18 * (k + 1)]) for k in range((len(line) - 8) // 18)]
        out, k = {"nsteps": int(line[:8])}, 0
        for name in self.DIAG_FIELDS[1:]:
            w = 3 if name in ("rcom", "vcom", "mom", "angmom", "force") else 1
            out[name] = np.array(vals[k:k + w]) if w == 3 else vals[k]
            k += w
        return out

    def setup(self, pos, mass, h, vel, h_fac=1.2, thetamaxsqd=0.25, abserror=1e-4,
              alpha=1.0):
        pos, mass, h, vel = _f64(pos), _f64(mass), _f64(h), _f64(vel)
        self.n = len(mass)
        self.lib.ref_setup(self.n, pos, mass, h, vel, h_fac, thetamaxsqd, abserror, alpha)
        self.pp_limit = self.params()["pp_limit"]

    def tree_update(self):
        self.lib.ref_tree_update()

    def hguess(self):
        self.lib.ref_hguess()

    def sph_update(self):
        self.lib.ref_sph_update()

    def thermal_iso(self, cs):
        self.lib.ref_thermal_iso(cs)

    def hydro_forces(self):
        self.lib.ref_hydro_forces()

    def grav_forces(self):
        self.lib.ref_grav_forces()

    # ---- integration steps (sph_integrate.F90 with RUNGE_KUTTA) ---------
    def step_setup(self, nlevels=1, nbuildstep=8, accel_mult=0.5, courant_mult=0.2,
                   dt_fixed=1.0):
        """Step state after the setup evaluation (initialize_sph_variables_2)."""
        self.lib.ref_step_setup(nlevels, nbuildstep, accel_mult, courant_mult, dt_fixed)

    def integrate(self, cs):
        """One sph_integrate: tree_update, sph_update, thermal, forces,
        sph_timesteps, sph_advance."""
        self.lib.ref_integrate(cs)

    def step_info(self):
        ist = np.zeros(8, np.int64)
        rst = np.zeros(4)
        self.lib.ref_get_step(ist, rst)
        keys = ["n", "nsteps", "nresync", "level_max", "level_step", "nbuild", "nstock"]
        out = {k: int(v) for k, v in zip(keys, ist)}
        out.update(time=float(rst[0]), timestep=float(rst[1]), dt_max=float(rst[2]))
        return out

    def step_particles(self):
        n = self.n
        nlevel, nlast = np.zeros(n, np.int64), np.zeros(n, np.int64)
        laststep, accdo = np.zeros(n), np.zeros(n, np.int32)
        vh, ro, vo = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        nmn = np.zeros(n, np.int64)
        self.lib.ref_get_step_particles(n, nlevel, nlast, laststep, accdo, vh, ro, vo, nmn)
        return dict(nlevel=nlevel, nlast=nlast, laststep=laststep, active=accdo,
                    v_half=vh, r_old=ro, v_old=vo, nminneib=nmn)

    def set_agravmag(self, am):
        self.lib.ref_set_agravmag(self.n, _f64(am))

    def set_hmin(self, hmin):
        """params hmin, the h of every particle under HFIND=CONSTANT."""
        self.lib.ref_set_hmin(hmin)

    def set_energy(self, u, gamma=1.4):
        """Specific internal energy per particle (current order) and gamma."""
        self.lib.ref_set_energy(self.n, _f64(u), gamma)

    def energy(self):
        """(u, du_dt) in the current (tree) order."""
        u, du = np.zeros(self.n), np.zeros(self.n)
        self.lib.ref_get_energy(self.n, u, du)
        return u, du

    def set_h(self, h):
        self.lib.ref_set_h(self.n, _f64(h))

    def _direct(self, nt, t, ag, pot):
        self.lib.ref_direct_gravity(nt, t, ag, pot)

    def _get_params(self, ip, rp):
        self.lib.ref_get_params(ip, rp)

    def _get_levels(self, f, l):
        self.lib.ref_get_levels(f, l)

    def _get_particles(self, *a):
        self.lib.ref_get_particles(self.n, *a)

    def _get_neib(self, pptot, pplist):
        self.lib.ref_get_neib(self.n, self.pp_limit, pptot, pplist)

    def _get_grav_tree(self, *a):
        self.lib.ref_get_grav_tree(*a)

    def _get_hydro_tree(self, *a):
        self.lib.ref_get_hydro_tree(*a)

    def _get_kernel(self, tab):
        self.lib.ref_get_kernel(tab)


class CSeren(_Common):
    """The C restatement (oracle/seren_oracle.c)."""

    _lib = None

    @classmethod
    def _load(cls):
        if cls._lib is None:
            path = os.path.join(HERE, "libseren_oracle.so")
            if not os.path.exists(path):
                raise FileNotFoundError(f"{path} not built (make -f oracle/Makefile)")
            lib = C.CDLL(path)
            vp = C.c_void_p
            lib.oracle_create.restype = vp
            lib.oracle_create.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_int, C.c_int]
            lib.oracle_destroy.argtypes = [vp]
            lib.oracle_error.argtypes = [vp]
            for f in ("oracle_hguess", "oracle_sph_update", "oracle_hydro_forces",
                      "oracle_grav_forces"):
                getattr(lib, f).argtypes = [vp]
            lib.oracle_tree_update.argtypes = [vp]
            lib.oracle_thermal_iso.argtypes = [vp, C.c_double]
            lib.oracle_direct_gravity.argtypes = [vp, C.c_int, _ip, _dp, _dp]
            lib.oracle_set_agravmag.argtypes = [vp, _dp]
            lib.oracle_set_h.argtypes = [vp, _dp]
            lib.oracle_get_params.argtypes = [vp, _ip, _dp]
            lib.oracle_get_levels.argtypes = [vp, _ip, _ip]
            lib.oracle_get_particles.argtypes = [vp, _ip, _dp, _dp, _dp] + [_dp] * 8
            lib.oracle_get_neib.argtypes = [vp, _ip, _ip]
            lib.oracle_get_grav_tree.argtypes = [vp, C.c_int, _ip, _ip, _ip, _ip] + [_dp] * 7
            lib.oracle_get_hydro_tree.argtypes = [vp, C.c_int, _dp, _dp, _dp]
            lib.oracle_get_kernel.argtypes = [vp, _dp]
            lib.oracle_set_energy.argtypes = [vp, _dp, C.c_double]
            lib.oracle_set_hmin.argtypes = [vp, C.c_double]
            lib.oracle_get_energy.argtypes = [vp, _dp, _dp]
            cls._lib = lib
        return cls._lib

    def __init__(self, octupole, mac: str, kernel: str = "m4", sph: str = "gradh"):
        """octupole: False (quadrupole), True (the reference's octupole
        stocking) or "exact" (exact branch parallel-axis shift)."""
        self.lib = self._load()
        self.octupole = 2 if octupole == "exact" else int(bool(octupole))
        self.mac = {"geometric": 0, "gadget": 1, "gadget2": 2, "eigen": 3}[mac]
        self.kernel_id = KERNELS[kernel]
        self.sph_std = {"gradh": 0, "standard": 1}[sph]
        self.ctx = None

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.oracle_destroy(self.ctx)
            self.ctx = None

    def setup(self, pos, mass, h, vel, h_fac=1.2, thetamaxsqd=0.25, abserror=1e-4,
              alpha=1.0):
        if self.ctx:
            self.lib.oracle_destroy(self.ctx)
        pos, mass, h, vel = _f64(pos), _f64(mass), _f64(h), _f64(vel)
        self.n = len(mass)
        self.ctx = self.lib.oracle_create(self.n, pos, mass, h, vel, int(self.octupole),
                                          self.mac, h_fac, thetamaxsqd, abserror, alpha,
                                          self.kernel_id, self.sph_std)
        self.pp_limit = self.params()["pp_limit"]

    def tree_update(self):
        err = self.lib.oracle_tree_update(self.ctx)
        if err:
            raise RuntimeError("Reached maximum level")

    def hguess(self):
        self.lib.oracle_hguess(self.ctx)

    def sph_update(self):
        self.lib.oracle_sph_update(self.ctx)
        if self.lib.oracle_error(self.ctx) == 2:
            raise RuntimeError("Too few particles for pp_gather neighbours (h_gather)")

    def thermal_iso(self, cs):
        self.lib.oracle_thermal_iso(self.ctx, cs)

    def hydro_forces(self):
        self.lib.oracle_hydro_forces(self.ctx)

    def grav_forces(self):
        self.lib.oracle_grav_forces(self.ctx)

    def set_agravmag(self, am):
        self.lib.oracle_set_agravmag(self.ctx, _f64(am))

    def set_hmin(self, hmin):
        """HFIND=CONSTANT: every h is hmin from the next sph_update on."""
        self.lib.oracle_set_hmin(self.ctx, hmin)

    def set_energy(self, u, gamma=1.4):
        """THERMAL=ENERGY_EQN: u per particle (current order) and gamma."""
        self.lib.oracle_set_energy(self.ctx, _f64(u), gamma)

    def energy(self):
        u, du = np.zeros(self.n), np.zeros(self.n)
        self.lib.oracle_get_energy(self.ctx, u, du)
        return u, du

    def set_h(self, h):
        self.lib.oracle_set_h(self.ctx, _f64(h))

    def _direct(self, nt, t, ag, pot):
        self.lib.oracle_direct_gravity(self.ctx, nt, t, ag, pot)

    def _get_params(self, ip, rp):
        self.lib.oracle_get_params(self.ctx, ip, rp)

    def _get_levels(self, f, l):
        self.lib.oracle_get_levels(self.ctx, f, l)

    def _get_particles(self, *a):
        self.lib.oracle_get_particles(self.ctx, *a)

    def _get_neib(self, pptot, pplist):
        self.lib.oracle_get_neib(self.ctx, pptot, pplist)

    def _get_grav_tree(self, *a):
        self.lib.oracle_get_grav_tree(self.ctx, *a)

    def _get_hydro_tree(self, *a):
        self.lib.oracle_get_hydro_tree(self.ctx, *a)

    def _get_kernel(self, tab):
        self.lib.oracle_get_kernel(self.ctx, tab)
