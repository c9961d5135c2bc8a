"""Python binding of the gfx950 force-evaluation library (libserenx.so).

Two layers over the C ABI of include/serenx.h:

* ``ForceEngine`` -- device-resident use: ``load`` a cloud once, run
  ``force_eval`` repeatedly with no host transfers, ``fetch`` results.
* The drop-in drivers ``tree_update``, ``BHhydro_hguess``, ``sph_update``,
  ``sph_hydro_forces`` and ``sph_grav_forces`` operate on a ``SerenState``
  whose arrays have the memory layout of SEREN's module arrays (parray(6,N),
  v(3,N), pplist(pp_limit,N), 1-based ids), the way SEREN's own drivers
  (src/main/tree_update.F90, src/sph_sim/sph_update.F90,
  src/sph_sim/sph_hydro_forces.F90, src/sph_sim/sph_grav_forces.F90) operate
  on its module state.  ``tree_update`` permutes every state array with the
  returned tree order, as reorder_particle_arrays (src/main/reorder_array.F90)
  does.

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise ``SerenxError``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SERENX_LIB names another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("SERENX_LIB") or os.path.join(HERE, "libserenx.so")

QUADRUPOLE, OCTUPOLE, OCTUPOLE_EXACT = 2, 3, 4
MAC_GEOMETRIC, MAC_GADGET, MAC_GADGET2, MAC_EIGEN = 0, 1, 2, 3
FP64, MIXED = 0, 1
EVAL_HGUESS, EVAL_NO_GRAV, EVAL_NO_HYDRO, EVAL_RESTOCK, EVAL_GRAPH = 1, 2, 4, 8, 16

_MULTIPOLE = {"quadrupole": QUADRUPOLE, "octupole": OCTUPOLE, "octupole_exact": OCTUPOLE_EXACT}
_MAC = {"geometric": MAC_GEOMETRIC, "gadget": MAC_GADGET, "gadget2": MAC_GADGET2,
        "eigen": MAC_EIGEN}
_PREC = {"fp64": FP64, "mixed": MIXED}
KERNEL_M4, KERNEL_M4_TC, KERNEL_QUINTIC, KERNEL_QUINTIC_TC = 0, 1, 2, 3
_KERNEL = {"m4": KERNEL_M4, "m4tc": KERNEL_M4_TC, "quintic": KERNEL_QUINTIC,
           "quintictc": KERNEL_QUINTIC_TC}

EXPORTED = [
    "serenx_default_options", "serenx_create", "serenx_destroy", "serenx_last_error",
    "serenx_neighbour_params", "serenx_bind_host_state", "serenx_tree_update",
    "serenx_tree_stock",
    "serenx_hguess", "serenx_sph_update", "serenx_hydro_forces", "serenx_grav_forces",
    "serenx_load", "serenx_set_agravmag", "serenx_force_eval", "serenx_sync",
    "serenx_fetch", "serenx_direct_gravity", "serenx_tree_sizes", "serenx_export_tree",
    "serenx_last_timings", "serenx_interaction_counts", "serenx_walk_stats",
    "serenx_default_step_options", "serenx_step_begin", "serenx_step", "serenx_step_get_info",
    "serenx_fetch_step", "serenx_load_forest", "serenx_cloud_offsets", "serenx_set_precision",
    "serenx_memory", "serenx_pow_third", "serenx_set_energy", "serenx_fetch_energy",
    "serenx_set_tuning", "serenx_get_tuning", "serenx_serve_begin", "serenx_serve_step",
    "serenx_serve_info", "serenx_diagnostics", "serenx_walk_timings",
    "serenx_diagnostics_noalloc", "serenx_graph_info",
]
INT_RK, INT_LFKDK = 0, 1
_INTEGRATOR = {"rk": INT_RK, "lfkdk": INT_LFKDK}
# Makefile TIMESTEP and CHECK_NEIB_TIMESTEP options
_LEVELS = {"adaptive": 0, "restricted": 1, "fixed": 2}
_NEIB_CHECK = {0: 0, 1: 1, 2: 2, "off": 0, "end": 1, "immediate": 2}
_THERMAL = {"isothermal": 0, "energy": 1}
_SPH = {"gradh": 0, "standard": 1}


class SerenxError(RuntimeError):
    """A non-zero status of the library; code is the SERENX_E* value."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


# status codes (include/serenx.h)
EINVAL, EHIP, ELMAX, ESTATE, ENOMEM, EHNEG, EFEW = 1, 2, 3, 4, 5, 6, 7
# tuning keys (serenx_set_tuning)
TUNE_HSOLVE_SLOTS, TUNE_MIXED_ABSERROR, TUNE_WALK_QUEUE, TUNE_NEIB_GROUP = 1, 2, 3, 4
TUNE_HSOLVE_WAVES = 5
TUNE_NEAR_CAP = 6
TUNE_NEIB_LEAVES = 7
TUNE_PAIR_BLOCKS = 8


class Options(C.Structure):
    _fields_ = [("multipole", C.c_int32), ("mac", C.c_int32), ("precision", C.c_int32),
                ("kernel", C.c_int32), ("h_fac", C.c_double), ("thetamaxsqd", C.c_double),
                ("abserror", C.c_double), ("alpha", C.c_double), ("sound_speed", C.c_double),
                ("thermal", C.c_int32), ("sph", C.c_int32), ("gamma", C.c_double),
                ("pp_gather", C.c_int32), ("serial", C.c_int32),
                ("hfind", C.c_int32), ("hmin", C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class HostState(C.Structure):
    _fields_ = [("ptot", C.c_int32), ("pp_limit", C.c_int32), ("parray", _dp), ("v", _dp),
                ("a", _dp), ("gpot", _dp), ("agravmag", _dp), ("rho", _dp), ("omega", _dp),
                ("press", _dp), ("sound", _dp), ("div_v", _dp), ("drhodt", _dp),
                ("pptot", _ip), ("pplist", _ip), ("accdo", _ip), ("rextent", _dp),
                ("du_dt", _dp)]


class StepOptions(C.Structure):
    _fields_ = [("integrator", C.c_int32), ("nlevels", C.c_int32), ("nbuildstep", C.c_int32),
                ("levels", C.c_int32), ("accel_mult", C.c_double),
                ("courant_mult", C.c_double), ("dt_fixed", C.c_double),
                ("neib_check", C.c_int32), ("reserved", C.c_int32)]


class StepInfo(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("n", "nsteps", "nresync", "level_max", "level_step",
                                         "nbuild", "nstock", "nactive")] + \
               [(k, C.c_double) for k in ("time", "timestep", "dt_max", "reserved")]


_lib: Optional[C.CDLL] = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libserenx.so (built by seren_amd/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SerenxError(f"{path} is not built; run python -m seren_amd.build")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.serenx_last_error.restype = C.c_char_p
    lib.serenx_create.argtypes = [C.POINTER(Options), C.c_int, C.POINTER(vp)]
    lib.serenx_default_options.argtypes = [C.POINTER(Options)]
    lib.serenx_destroy.argtypes = [vp]
    lib.serenx_neighbour_params.argtypes = [vp, _ip, _ip]
    lib.serenx_bind_host_state.argtypes = [vp, C.POINTER(HostState)]
    lib.serenx_tree_update.argtypes = [vp, _ip]
    for f in ("serenx_tree_stock", "serenx_hguess", "serenx_sph_update", "serenx_hydro_forces",
              "serenx_grav_forces", "serenx_sync"):
        getattr(lib, f).argtypes = [vp]
    lib.serenx_load.argtypes = [vp, C.c_int32, _dp, _dp, _dp, _dp]
    lib.serenx_load_forest.argtypes = [vp, C.c_int32, _ip, _dp, _dp, _dp, _dp]
    lib.serenx_cloud_offsets.argtypes = [vp, _ip]
    lib.serenx_set_precision.argtypes = [vp, C.c_int32]
    lib.serenx_diagnostics.argtypes = [vp, _dp]
    lib.serenx_diagnostics_noalloc.argtypes = [vp, vp]   # (context, device pointer)
    lib.serenx_graph_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.serenx_walk_timings.argtypes = [vp, _dp]
    lib.serenx_serve_begin.argtypes = [vp, C.c_double, C.c_double]
    lib.serenx_serve_step.argtypes = [vp, C.c_int32, C.POINTER(C.c_uint8)]
    lib.serenx_serve_info.argtypes = [vp, C.c_int32, _dp, _dp]
    lib.serenx_set_tuning.argtypes = [vp, C.c_int32, C.c_double]
    lib.serenx_get_tuning.argtypes = [vp, C.c_int32, _dp]
    lib.serenx_memory.argtypes = [vp] + [C.POINTER(C.c_int64)] * 3
    lib.serenx_set_agravmag.argtypes = [vp, _dp]
    lib.serenx_force_eval.argtypes = [vp, C.c_int32]
    lib.serenx_fetch.argtypes = [vp, _ip] + [_dp] * 14 + [_ip, _ip]
    lib.serenx_direct_gravity.argtypes = [vp, C.c_int32, _ip, _dp, _dp]
    lib.serenx_tree_sizes.argtypes = [vp, _ip, _ip]
    lib.serenx_export_tree.argtypes = [vp, _ip, _ip, _ip, _ip] + [_dp] * 9 + [_ip, _ip]
    lib.serenx_last_timings.argtypes = [vp, _dp]
    lib.serenx_interaction_counts.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.serenx_walk_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.serenx_default_step_options.argtypes = [C.POINTER(StepOptions)]
    lib.serenx_default_step_options.restype = None
    lib.serenx_step_begin.argtypes = [vp, C.POINTER(StepOptions)]
    lib.serenx_step.argtypes = [vp, C.c_int32]
    lib.serenx_step_get_info.argtypes = [vp, C.POINTER(StepInfo)]
    lib.serenx_pow_third.argtypes = [vp, C.c_int32, _dp, _dp]
    lib.serenx_set_energy.argtypes = [vp, _dp]
    lib.serenx_fetch_energy.argtypes = [vp, _dp, _dp]
    lib.serenx_fetch_step.argtypes = [vp, _ip, C.POINTER(C.c_int64), _dp, _ip, _dp, _dp, _dp,
                                      _ip]
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_library().serenx_last_error().decode(errors="replace")
        raise SerenxError(f"serenx error {rc}: {msg}", rc)


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _iptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class ForceEngine:
    """One device context holding one resident cloud."""

    def __init__(self, multipole: str = "quadrupole", mac: str = "geometric",
                 precision: str = "fp64", h_fac: float = 1.2, thetamaxsqd: float = 0.25,
                 abserror: float = 1e-4, alpha: float = 1.0, sound_speed: float = 0.2,
                 device: int = 0, kernel: str = "m4", thermal: str = "isothermal",
                 gamma: float = 1.4, sph: str = "gradh", pp_gather: int = 0,
                 serial: bool = False, hmin: float = 0.0):
        """thermal: "isothermal" (P = c_s^2 rho with sound_speed) or "energy"
        (the energy equation: set u with set_energy).  sph: "gradh" or
        "standard" (h from the pp_gather-th neighbour; pp_gather 0 takes the
        h_fac value).  hmin > 0 with sph="standard" selects HFIND=CONSTANT:
        every h is hmin.  serial: keep the gravity walk on the context's one
        stream (many concurrent contexts)."""
        self.lib = load_library()
        opt = Options()
        _check(self.lib.serenx_default_options(C.byref(opt)))
        opt.multipole = _MULTIPOLE[multipole]
        opt.mac = _MAC[mac]
        opt.precision = _PREC[precision]
        opt.kernel = _KERNEL[kernel]
        opt.h_fac = h_fac
        opt.thetamaxsqd = thetamaxsqd
        opt.abserror = abserror
        opt.alpha = alpha
        opt.sound_speed = sound_speed
        opt.thermal = _THERMAL[thermal]
        opt.gamma = gamma
        opt.sph = _SPH[sph]
        opt.pp_gather = pp_gather
        opt.serial = int(serial)
        opt.hfind = 1 if hmin > 0.0 else 0
        opt.hmin = hmin
        self.opt = opt
        self.ctx = C.c_void_p()
        _check(self.lib.serenx_create(C.byref(opt), device, C.byref(self.ctx)))
        g, l = C.c_int32(), C.c_int32()
        _check(self.lib.serenx_neighbour_params(self.ctx, C.byref(g), C.byref(l)))
        self.pp_gather, self.pp_limit = g.value, l.value
        self.n = 0

    def close(self) -> None:
        if getattr(self, "ctx", None) and self.ctx.value:
            self.lib.serenx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device-resident mode ---------------------------------------------
    def load(self, pos, vel, mass, h=None) -> None:
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        h = None if h is None else np.ascontiguousarray(h, np.float64)
        self.n = len(mass)
        _check(self.lib.serenx_load(self.ctx, self.n, _dptr(pos), _dptr(vel), _dptr(mass),
                                    _dptr(h)))

    def load_forest(self, clouds) -> np.ndarray:
        """Loads independent clouds [(pos, vel, mass, h or None), ...] as one
        forest; returns the particle offsets of the clouds (ncloud + 1)."""
        counts = np.array([len(c[2]) for c in clouds], np.int32)
        pos = np.ascontiguousarray(np.concatenate([c[0] for c in clouds]), np.float64)
        vel = np.ascontiguousarray(np.concatenate([c[1] for c in clouds]), np.float64)
        mass = np.ascontiguousarray(np.concatenate([c[2] for c in clouds]), np.float64)
        hs = [c[3] if len(c) > 3 and c[3] is not None else np.zeros(len(c[2])) for c in clouds]
        h = np.ascontiguousarray(np.concatenate(hs), np.float64)
        _check(self.lib.serenx_load_forest(self.ctx, len(counts), _iptr(counts), _dptr(pos),
                                           _dptr(vel), _dptr(mass), _dptr(h)))
        self.n = int(counts.sum())
        return self.cloud_offsets(len(counts))

    def cloud_offsets(self, ncloud: int) -> np.ndarray:
        off = np.zeros(ncloud + 1, np.int32)
        _check(self.lib.serenx_cloud_offsets(self.ctx, _iptr(off)))
        return off

    def set_precision(self, precision: str) -> None:
        """'fp64' or 'mixed' for later evaluations (guard steps)."""
        _check(self.lib.serenx_set_precision(self.ctx, _PREC[precision]))

    DIAG_KEYS = ("mtot", "rcom", "vcom", "mom", "angmom", "force", "ketot", "gpetot", "utot",
                 "etot")

    def diagnostics(self) -> Dict[str, np.ndarray]:
        """FP64 conservation sums of the loaded particles (diagnostics.F90)."""
        v = np.zeros(20)
        _check(self.lib.serenx_diagnostics(self.ctx, _dptr(v)))
        return self._diag_dict(v)

    @classmethod
    def _diag_dict(cls, v):
        return {"mtot": v[0], "rcom": v[1:4].copy(), "vcom": v[4:7].copy(), "mom": v[7:10].copy(),
                "angmom": v[10:13].copy(), "force": v[13:16].copy(), "ketot": v[16],
                "gpetot": v[17], "utot": v[18], "etot": v[19]}

    # ---- batched serving steps (serenx_serve_*) ---------------------------
    def serve_begin(self, accel_mult: float = 0.5, courant_mult: float = 0.2) -> None:
        """After the setup evaluation: every cloud's first step size."""
        _check(self.lib.serenx_serve_begin(self.ctx, accel_mult, courant_mult))

    def serve_step(self, select) -> None:
        """One KDK step of the clouds with select[k] true, each with its own
        step size (synchronous)."""
        sel = np.ascontiguousarray(np.asarray(select, np.uint8))
        _check(self.lib.serenx_serve_step(self.ctx, len(sel),
                                          sel.ctypes.data_as(C.POINTER(C.c_uint8))))

    def serve_info(self, ncloud: int):
        """Per-cloud next step size and time."""
        dt, t = np.zeros(ncloud), np.zeros(ncloud)
        _check(self.lib.serenx_serve_info(self.ctx, ncloud, _dptr(dt), _dptr(t)))
        return dt, t

    def set_tuning(self, key: int, value: float) -> None:
        """serenx_set_tuning: TUNE_HSOLVE_SLOTS or TUNE_MIXED_ABSERROR."""
        _check(self.lib.serenx_set_tuning(self.ctx, key, float(value)))

    def get_tuning(self, key: int) -> float:
        v = C.c_double()
        _check(self.lib.serenx_get_tuning(self.ctx, key, C.byref(v)))
        return v.value

    def memory(self) -> Dict[str, int]:
        """Free / total device bytes and the bytes this context holds."""
        fr, tot, mine = C.c_int64(), C.c_int64(), C.c_int64()
        _check(self.lib.serenx_memory(self.ctx, C.byref(fr), C.byref(tot), C.byref(mine)))
        return {"free": fr.value, "total": tot.value, "context": mine.value}

    def set_agravmag(self, am) -> None:
        am = np.ascontiguousarray(am, np.float64)
        _check(self.lib.serenx_set_agravmag(self.ctx, _dptr(am)))

    def force_eval(self, hguess: bool = False, grav: bool = True, hydro: bool = True,
                   restock: bool = False, graph: bool = False) -> None:
        """One force evaluation.  restock: on the tree of the last build,
        restocked with the current positions (tree_update.F90 between
        rebuilds); graph (with restock): as recorded HIP graphs."""
        flags = (EVAL_HGUESS if hguess else 0) | (0 if grav else EVAL_NO_GRAV) | \
            (0 if hydro else EVAL_NO_HYDRO) | (EVAL_RESTOCK if restock or graph else 0) | \
            (EVAL_GRAPH if graph else 0)
        _check(self.lib.serenx_force_eval(self.ctx, flags))

    def graph_info(self) -> Dict[str, int]:
        """Recordings and replays of the restock evaluation's graphs."""
        cap, lau = C.c_int64(), C.c_int64()
        _check(self.lib.serenx_graph_info(self.ctx, C.byref(cap), C.byref(lau)))
        return {"captures": int(cap.value), "launches": int(lau.value)}

    def sync(self) -> None:
        _check(self.lib.serenx_sync(self.ctx))

    def fetch(self, neighbours: bool = True) -> Dict[str, np.ndarray]:
        n = self.n
        out = dict(porig=np.zeros(n, np.int32), pos=np.zeros((n, 3)), v=np.zeros((n, 3)),
                   m=np.zeros(n), h=np.zeros(n), zeta=np.zeros(n), rho=np.zeros(n),
                   omega=np.zeros(n), div_v=np.zeros(n), drhodt=np.zeros(n),
                   press=np.zeros(n), sound=np.zeros(n), a=np.zeros((n, 3)), gpot=np.zeros(n),
                   agravmag=np.zeros(n), pptot=np.zeros(n, np.int32))
        pl = np.zeros((n, self.pp_limit), np.int32) if neighbours else None
        _check(self.lib.serenx_fetch(
            self.ctx, _iptr(out["porig"]), _dptr(out["pos"]), _dptr(out["v"]), _dptr(out["m"]),
            _dptr(out["h"]), _dptr(out["zeta"]), _dptr(out["rho"]), _dptr(out["omega"]),
            _dptr(out["div_v"]), _dptr(out["drhodt"]), _dptr(out["press"]), _dptr(out["sound"]),
            _dptr(out["a"]), _dptr(out["gpot"]), _dptr(out["agravmag"]), _iptr(out["pptot"]),
            _iptr(pl)))
        if pl is not None:
            out["pplist"] = pl
        return out

    def direct_gravity(self, targets):
        t = np.ascontiguousarray(targets, np.int32)
        acc = np.zeros((len(t), 3))
        pot = np.zeros(len(t))
        _check(self.lib.serenx_direct_gravity(self.ctx, len(t), _iptr(t), _dptr(acc),
                                              _dptr(pot)))
        return acc, pot

    def tree_sizes(self):
        nc, lt = C.c_int32(), C.c_int32()
        _check(self.lib.serenx_tree_sizes(self.ctx, C.byref(nc), C.byref(lt)))
        return nc.value, lt.value

    def export_tree(self) -> Dict[str, np.ndarray]:
        nc, _ = self.tree_sizes()
        t = dict(leaf=np.zeros(nc, np.int32), plist=np.zeros((nc, 8), np.int32),
                 nextcell=np.zeros(nc, np.int32), ifopen=np.zeros(nc, np.int32),
                 r=np.zeros((nc, 3)), m=np.zeros(nc), dminsqd=np.zeros(nc), mac=np.zeros(nc),
                 q=np.zeros((nc, 5)), s=np.zeros((nc, 10)), hr=np.zeros((nc, 3)),
                 rmax=np.zeros(nc), hrangemax=np.zeros(nc),
                 first_cell=np.zeros(31, np.int32), last_cell=np.zeros(31, np.int32))
        _check(self.lib.serenx_export_tree(
            self.ctx, _iptr(t["leaf"]), _iptr(t["plist"]), _iptr(t["nextcell"]),
            _iptr(t["ifopen"]), _dptr(t["r"]), _dptr(t["m"]), _dptr(t["dminsqd"]),
            _dptr(t["mac"]), _dptr(t["q"]), _dptr(t["s"]), _dptr(t["hr"]), _dptr(t["rmax"]),
            _dptr(t["hrangemax"]), _iptr(t["first_cell"]), _iptr(t["last_cell"])))
        return t

    def timings(self):
        ms = np.zeros(4)
        _check(self.lib.serenx_last_timings(self.ctx, _dptr(ms)))
        return dict(tree=ms[0], sph=ms[1], hydro=ms[2], grav=ms[3])

    def walk_timings(self):
        """ms of the last walk's cell phase and pair phase."""
        ms = np.zeros(2)
        _check(self.lib.serenx_walk_timings(self.ctx, _dptr(ms)))
        return dict(cells=ms[0], pairs=ms[1])

    def walk_stats(self) -> Dict[str, float]:
        """Per-walk SIMD utilisation and h-solve pass counts of the last evaluation."""
        v = (C.c_int64 * 16)()
        _check(self.lib.serenx_walk_stats(self.ctx, v))
        v = [int(x) for x in v]
        util = lambda a, b: (b / (64.0 * a)) if a else 0.0  # noqa: E731
        return {"grav_cells_per_wave_visit_util": util(v[0], v[1]), "grav_wave_visits": v[0],
                "hsolve_util": util(v[2], v[3]), "hsolve_wave_visits": v[2],
                "hsolve_mean_passes": v[4] / max(v[7], 1), "hsolve_max_passes": v[5],
                "hsolve_mean_iterations": v[6] / max(self.n, 1),
                "neib_util": util(v[8], v[9]), "neib_wave_visits": v[8],
                "hsolve_list_walks_per_wave": v[10] / max(v[7], 1),
                "hsolve_overflow_passes": v[11], "hsolve_longest_list_leaves": v[12]}

    # ---- device-resident integration (sph_integrate.F90) ----------------
    def step_begin(self, integrator: str = "rk", nlevels: int = 1, nbuildstep: int = 8,
                   accel_mult: float = 0.5, courant_mult: float = 0.2,
                   levels: str = "restricted", dt_fixed: float = 1.0,
                   neib_check=1) -> None:
        """Start stepping after the setup evaluation (force_eval(hguess=True)).
        levels and neib_check are the Makefile's TIMESTEP (adaptive,
        restricted, fixed) and CHECK_NEIB_TIMESTEP (0, 1, 2) options, with
        the reference Makefile's defaults."""
        o = StepOptions()
        self.lib.serenx_default_step_options(C.byref(o))
        o.integrator = _INTEGRATOR[integrator]
        o.nlevels, o.nbuildstep = nlevels, nbuildstep
        o.accel_mult, o.courant_mult = accel_mult, courant_mult
        o.levels, o.dt_fixed = _LEVELS[levels], dt_fixed
        o.neib_check = _NEIB_CHECK[neib_check]
        _check(self.lib.serenx_step_begin(self.ctx, C.byref(o)))

    def step(self, count: int = 1) -> None:
        """count x sph_integrate on the device (synchronous)."""
        _check(self.lib.serenx_step(self.ctx, count))

    def step_info(self) -> Dict[str, float]:
        i = StepInfo()
        _check(self.lib.serenx_step_get_info(self.ctx, C.byref(i)))
        return {k: getattr(i, k) for k, _ in StepInfo._fields_ if k != "reserved"}

    def fetch_step(self) -> Dict[str, np.ndarray]:
        """Per-particle step state in tree order (see fetch()['porig'])."""
        n = self.n
        nlevel = np.zeros(n, np.int32)
        nlast = np.zeros(n, np.int64)
        laststep = np.zeros(n)
        active = np.zeros(n, np.int32)
        r_old, v_old, v_half = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        nminneib = np.zeros(n, np.int32)
        _check(self.lib.serenx_fetch_step(
            self.ctx, _iptr(nlevel), nlast.ctypes.data_as(C.POINTER(C.c_int64)),
            _dptr(laststep), _iptr(active), _dptr(r_old), _dptr(v_old), _dptr(v_half),
            _iptr(nminneib)))
        return dict(nlevel=nlevel, nlast=nlast, laststep=laststep, active=active,
                    r_old=r_old, v_old=v_old, v_half=v_half, nminneib=nminneib)

    def set_energy(self, u) -> None:
        """Specific internal energy per particle in the current particle order
        (load order before the first evaluation, tree order after)."""
        u = np.ascontiguousarray(u, np.float64)
        if len(u) != self.n:
            raise ValueError("u must have one value per particle")
        _check(self.lib.serenx_set_energy(self.ctx, _dptr(u)))

    def fetch_energy(self):
        """(u, du_dt) in tree order."""
        u, du = np.empty(self.n), np.empty(self.n)
        _check(self.lib.serenx_fetch_energy(self.ctx, _dptr(u), _dptr(du)))
        return u, du

    def pow_third(self, x: np.ndarray) -> np.ndarray:
        """x**0.333333333 as the device h solve evaluates it."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        _check(self.lib.serenx_pow_third(self.ctx, len(x), _dptr(x), _dptr(y)))
        return y

    def interaction_counts(self):
        """(cell terms, pair terms, pairs inside 2h_p, pairs inside 2h_j)."""
        v = (C.c_int64 * 4)()
        _check(self.lib.serenx_interaction_counts(self.ctx, v))
        return tuple(int(x) for x in v)


# ---------------------------------------------------------------------------
# Drop-in drivers over SEREN-shaped host state
# ---------------------------------------------------------------------------
class SerenState:
    """Host arrays with the memory layout of SEREN's module arrays.

    parray is (N, 6) C-order == parray(6, N) Fortran order: [x y z m h zeta];
    v, a are (N, 3); pplist is (N, pp_limit) with 1-based ids; porig tracks
    the original 1-based index as SEREN's porig array does.
    """

    def __init__(self, pos, vel, mass, h, pp_limit: int):
        n = len(mass)
        self.n = n
        self.pp_limit = pp_limit
        self.parray = np.zeros((n, 6))
        self.parray[:, 0:3] = pos
        self.parray[:, 3] = mass
        self.parray[:, 4] = h
        self.v = np.ascontiguousarray(vel, np.float64).copy()
        self.a = np.zeros((n, 3))
        self.gpot = np.zeros(n)
        self.agravmag = np.zeros(n)
        self.rho = np.ones(n)       # initialize_sph_variables_1.F90:68-72
        self.omega = np.ones(n)
        self.press = np.zeros(n)
        self.sound = np.zeros(n)
        self.div_v = np.zeros(n)
        self.drhodt = np.zeros(n)
        self.pptot = np.zeros(n, np.int32)
        self.pplist = np.zeros((n, pp_limit), np.int32)
        self.accdo = np.ones(n, np.int32)
        self.porig = np.arange(1, n + 1, dtype=np.int32)
        self.rextent = np.zeros(1)

    def reorder(self, dummylist: np.ndarray) -> None:
        """reorder_particle_arrays (reorder_array.F90:9-270): new p <- old dummylist(p)."""
        idx = dummylist - 1
        for name in ("parray", "v", "a", "gpot", "agravmag", "rho", "omega", "press", "sound",
                     "div_v", "drhodt", "accdo", "porig"):
            arr = getattr(self, name)
            arr[...] = arr[idx]   # in place: the bound pointers stay valid


def bind(engine: ForceEngine, st: SerenState) -> None:
    hs = HostState()
    hs.ptot = st.n
    hs.pp_limit = st.pp_limit
    hs.parray, hs.v, hs.a = _dptr(st.parray), _dptr(st.v), _dptr(st.a)
    hs.gpot, hs.agravmag = _dptr(st.gpot), _dptr(st.agravmag)
    hs.rho, hs.omega, hs.press, hs.sound = (_dptr(st.rho), _dptr(st.omega), _dptr(st.press),
                                            _dptr(st.sound))
    hs.div_v, hs.drhodt = _dptr(st.div_v), _dptr(st.drhodt)
    hs.pptot, hs.pplist, hs.accdo = _iptr(st.pptot), _iptr(st.pplist), _iptr(st.accdo)
    hs.rextent = _dptr(st.rextent)
    engine._hs = hs   # keep the struct alive with the engine
    engine._hs_state = st
    _check(engine.lib.serenx_bind_host_state(engine.ctx, C.byref(hs)))


def tree_update(engine: ForceEngine, st: SerenState) -> None:
    """tree_update.F90:9-92 incl. BH_reorder_particles' permutation of all arrays."""
    if getattr(engine, "_hs_state", None) is not st:
        bind(engine, st)
    dl = np.zeros(st.n, np.int32)
    _check(engine.lib.serenx_tree_update(engine.ctx, _iptr(dl)))
    st.reorder(dl)


def tree_stock(engine: ForceEngine, st: SerenState) -> None:
    """The stock-only branch of tree_update.F90 (nstock == nsteps, no rebuild)."""
    _check(engine.lib.serenx_tree_stock(engine.ctx))


def BHhydro_hguess(engine: ForceEngine, st: SerenState) -> None:
    _check(engine.lib.serenx_hguess(engine.ctx))


def sph_update(engine: ForceEngine, st: SerenState) -> None:
    _check(engine.lib.serenx_sph_update(engine.ctx))


def update_thermal_properties_isothermal(st: SerenState, cs: float) -> None:
    """thermal.F90:52-60 with the sound speed given (P = cs^2 rho, c = cs)."""
    st.press[:] = cs * cs * st.rho
    st.sound[:] = cs


def sph_hydro_forces(engine: ForceEngine, st: SerenState) -> None:
    _check(engine.lib.serenx_hydro_forces(engine.ctx))


def sph_grav_forces(engine: ForceEngine, st: SerenState) -> None:
    _check(engine.lib.serenx_grav_forces(engine.ctx))
