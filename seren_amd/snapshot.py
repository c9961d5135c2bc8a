"""SEREN snapshot interchange: the binary and formatted SEREN dumps
(SERENBINARYDUMPV2 / SERENASCIIDUMPV2) and the column ASCII format, so that
SEREN snapshots can feed the device-resident engine and results can be read
by SEREN tooling.

Layouts follow the reference's readers and writers:

* ``seren_unform``  write_data_seren_unform.F90:47-137 /
  read_data_seren_unform.F90:46-210: Fortran sequential unformatted records
  (4-byte little-endian length markers around each record) -- format id
  (character*20), PR, NDIM, VDIM, BDIM, idata(50) int32, ilpdata(50) int64,
  rdata(50) and dpdata(50) real*8, [unit names], data ids (character*20),
  typedata(5, ndata) int32, then one record per array in data-id order;
* ``seren_form``    write_data_seren_form.F90:128-250 /
  read_data_seren_form.F90:56-250: the same content as text, one value (or
  one particle's vector) per line, in (I2), (I10), (E18.10), (5I10) fields;
* ``ascii``         write_data_ascii.F90:38-92 / read_data_ascii.F90:40-150:
  one particle per line, '(I3,100E15.7)', columns named by a column list
  (the reference reads it from asciicolumns.dat).

Only gas particles and the arrays of the isothermal grad-h option set are
handled (porig, r, m, h, v, rho, temp, and u when present); files carrying
sinks or MHD fields are rejected with an error.  Values are stored in code
units (DIMENSIONLESS: every scale factor is 1, units.F90:37-60).
"""
from __future__ import annotations

import dataclasses
import math
import re
import struct
from typing import List, Optional, Sequence

import numpy as np

NDIM = VDIM = BDIM = 3
PR = 8                      # kind value of real(kind=PR) with DOUBLE_PRECISION
DMDT_RANGE = 1000           # macros.h:159
FORMAT_UNFORM = "SERENBINARYDUMPV2"
FORMAT_FORM = "SERENASCIIDUMPV2"
# datafiles/asciicolumns.dat of the reference: the default column list
ASCII_COLUMNS = ("ptype", "x", "y", "z", "vx", "vy", "vz", "m", "h", "u")

# typedata rows (write_data_seren_unform.F90:96-115): (ncomponents, first,
# last, type id, unit id) with type 2 = integer, 4 = real(PR)
_TYPES = {"porig": (1, 2, 0), "r": (NDIM, 4, 1), "m": (1, 4, 2), "h": (1, 4, 1),
          "v": (VDIM, 4, 4), "rho": (1, 4, 6), "temp": (1, 4, 21), "u": (1, 4, 20)}


class SnapshotError(ValueError):
    pass


@dataclasses.dataclass
class Snapshot:
    """One SEREN snapshot of gas particles (arrays in file order)."""
    r: np.ndarray                      # (n, 3)
    m: np.ndarray                      # (n,)
    h: np.ndarray                      # (n,)
    v: np.ndarray                      # (n, 3)
    rho: np.ndarray                    # (n,)
    porig: Optional[np.ndarray] = None   # (n,) int32, 1-based
    temp: Optional[np.ndarray] = None
    u: Optional[np.ndarray] = None
    time: float = 0.0
    lastsnap: float = 0.0
    mgas_orig: float = 0.0
    snapshot: int = 0
    nsteps: int = 0
    ntempnext: int = 0
    ndiagnext: int = 0
    nsnapnext: int = 0
    nsinknext: int = 0
    h_fac: float = 1.2
    gamma: float = 1.4
    mu_bar: float = 2.35
    pp_gather: int = 0
    pgas_orig: int = 0
    units: List[str] = dataclasses.field(default_factory=list)

    @property
    def n(self) -> int:
        return len(self.m)

    def __post_init__(self):
        n = len(self.m)
        self.r = np.asarray(self.r, np.float64).reshape(n, NDIM)
        self.v = np.asarray(self.v, np.float64).reshape(n, VDIM)
        for k in ("m", "h", "rho", "temp", "u"):
            a = getattr(self, k)
            if a is not None:
                setattr(self, k, np.asarray(a, np.float64).reshape(n))
        if self.porig is None:
            self.porig = np.arange(1, n + 1, dtype=np.int32)
        self.porig = np.asarray(self.porig, np.int32).reshape(n)

    def _data_ids(self) -> List[str]:
        ids = ["porig", "r", "m", "h", "v", "rho"]
        if self.temp is not None:
            ids.append("temp")
        if self.u is not None:
            ids.append("u")
        return ids

    def _headers(self, ndata: int):
        n = self.n
        idata = np.zeros(50, np.int32)
        idata[0], idata[4] = n, n           # ptot, pgas (gas only)
        idata[19], idata[20] = len(self.units), ndata
        idata[29] = DMDT_RANGE
        idata[30] = self.pgas_orig
        idata[31] = self.pp_gather
        ilp = np.zeros(50, np.int64)
        ilp[:6] = [self.snapshot, self.nsteps, self.ntempnext, self.ndiagnext, self.nsnapnext,
                   self.nsinknext]
        rdata = np.zeros(50)
        rdata[:3] = [self.h_fac, self.gamma, self.mu_bar]
        dpdata = np.zeros(50)
        dpdata[:3] = [self.time, self.lastsnap, self.mgas_orig]
        return idata, ilp, rdata, dpdata

    def _set_headers(self, idata, ilp, rdata, dpdata):
        (self.snapshot, self.nsteps, self.ntempnext, self.ndiagnext, self.nsnapnext,
         self.nsinknext) = (int(x) for x in ilp[:6])
        self.h_fac, self.gamma, self.mu_bar = (float(x) for x in rdata[:3])
        self.time, self.lastsnap, self.mgas_orig = (float(x) for x in dpdata[:3])
        self.pgas_orig = int(idata[30])
        self.pp_gather = int(idata[31])


def _check_counts(idata):
    ptot, stot = int(idata[0]), int(idata[1])
    others = [int(idata[k]) for k in (2, 3, 5, 6, 7)]   # boundary icm cdm dust ion
    if stot != 0:
        raise SnapshotError("snapshot contains sink particles (not supported)")
    if any(others) or int(idata[4]) != ptot:
        raise SnapshotError("snapshot contains non-gas particle types (not supported)")
    return ptot


def _char20(s: str) -> bytes:
    b = s.encode("ascii")
    if len(b) > 20:
        raise SnapshotError(f"identifier longer than 20 characters: {s!r}")
    return b.ljust(20, b" ")


# ---------------------------------------------------------------------------
# Binary SEREN dump
# ---------------------------------------------------------------------------
def _rec(payload: bytes) -> bytes:
    if len(payload) >= 2 ** 31:
        raise SnapshotError("record of 2 GiB or more (split records are not supported)")
    m = struct.pack("<i", len(payload))
    return m + payload + m


def write_seren_unform(path: str, snap: Snapshot) -> None:
    """write_data_seren_unform.F90:47-137."""
    ids = snap._data_ids()
    idata, ilp, rdata, dpdata = snap._headers(len(ids))
    out = [_rec(_char20(FORMAT_UNFORM)), _rec(struct.pack("<i", PR)),
           _rec(struct.pack("<i", NDIM)), _rec(struct.pack("<i", VDIM)),
           _rec(struct.pack("<i", BDIM)), _rec(idata.astype("<i4").tobytes()),
           _rec(ilp.astype("<i8").tobytes()), _rec(rdata.astype("<f8").tobytes()),
           _rec(dpdata.astype("<f8").tobytes())]
    if snap.units:
        out.append(_rec(b"".join(_char20(u) for u in snap.units)))
    out.append(_rec(b"".join(_char20(k) for k in ids)))
    typedata = np.array([[_TYPES[k][0], 1, snap.n, _TYPES[k][1], _TYPES[k][2]] for k in ids],
                        np.int32)
    out.append(_rec(typedata.astype("<i4").tobytes()))
    for k in ids:
        a = getattr(snap, k)
        out.append(_rec(np.ascontiguousarray(a, "<i4" if k == "porig" else "<f8").tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))


class _Records:
    def __init__(self, data: bytes):
        self.data, self.pos = data, 0

    def next(self) -> bytes:
        if self.pos + 4 > len(self.data):
            raise SnapshotError("truncated unformatted file")
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        if n < 0:
            raise SnapshotError("split records are not supported")
        end = self.pos + 4 + n
        if end + 4 > len(self.data) or struct.unpack_from("<i", self.data, end)[0] != n:
            raise SnapshotError("corrupt record markers")
        rec = self.data[self.pos + 4:end]
        self.pos = end + 4
        return rec


def read_seren_unform(path: str) -> Snapshot:
    """read_data_seren_unform.F90:46-210 for gas-only files."""
    with open(path, "rb") as f:
        rs = _Records(f.read())
    fmt = rs.next().decode("ascii").strip()
    if fmt != FORMAT_UNFORM:
        raise SnapshotError("Incorrect format of IC file")
    pr = struct.unpack("<i", rs.next())[0]
    if pr not in (2, 8):
        raise SnapshotError("Incorrect PR of IC file (double precision expected)")
    for want, what in ((NDIM, "NDIM"), (VDIM, "VDIM"), (BDIM, "BDIM")):
        if struct.unpack("<i", rs.next())[0] != want:
            raise SnapshotError(f"Incorrect {what} of IC file")
    idata = np.frombuffer(rs.next(), "<i4")
    ilp = np.frombuffer(rs.next(), "<i8")
    rdata = np.frombuffer(rs.next(), "<f8")
    dpdata = np.frombuffer(rs.next(), "<f8")
    n = _check_counts(idata)
    nunit, ndata = int(idata[19]), int(idata[20])
    units = []
    if nunit > 0:
        u = rs.next()
        units = [u[20 * i:20 * i + 20].decode("ascii").rstrip() for i in range(nunit)]
    ids = []
    typedata = np.zeros((0, 5), np.int32)
    if ndata > 0:
        d = rs.next()
        ids = [d[20 * i:20 * i + 20].decode("ascii").strip() for i in range(ndata)]
        typedata = np.frombuffer(rs.next(), "<i4").reshape(ndata, 5)
    arrays = {}
    for k, t in zip(ids, typedata):
        ncomp, first, last, tid = int(t[0]), int(t[1]), int(t[2]), int(t[3])
        rec = rs.next()
        if k == "B":
            raise SnapshotError("snapshot contains MHD fields (not supported)")
        if tid == 2:
            a = np.frombuffer(rec, "<i4")
        elif tid == 3:
            a = np.frombuffer(rec, "<i8")
        elif tid in (4, 5):
            a = np.frombuffer(rec, "<f8")
        else:
            continue                           # logical or unknown: skipped as the reader does
        if ncomp > 1:
            a = a.reshape(last - first + 1, ncomp)
        arrays[k] = a.copy()
    for k in ("r", "m", "h", "v", "rho"):
        if k not in arrays:
            raise SnapshotError(f"snapshot has no {k!r} array")
    snap = Snapshot(r=arrays["r"], m=arrays["m"], h=arrays["h"], v=arrays["v"],
                    rho=arrays["rho"], porig=arrays.get("porig"), temp=arrays.get("temp"),
                    u=arrays.get("u"), units=units)
    if snap.n != n:
        raise SnapshotError("particle arrays do not match ptot")
    snap._set_headers(idata, ilp, rdata, dpdata)
    return snap


# ---------------------------------------------------------------------------
# Formatted SEREN dump
# ---------------------------------------------------------------------------
def fortran_e(x: float, w: int = 18, d: int = 10) -> str:
    """Fortran Ew.d output editing: 0.dddE+ee (or 0.ddd+eee for three-digit
    exponents), rounded to nearest, right-justified in w columns."""
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(w)
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    ax = abs(x)
    if ax == 0.0:
        digits, e = "0" * d, 0
    else:
        s = f"{ax:.{d - 1}e}"                   # d significant digits, correctly rounded
        mant, ex = s.split("e")
        digits, e = mant.replace(".", ""), int(ex) + 1
    exp = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
    s = f"{sign}0.{digits}{exp}"
    return s.rjust(w) if len(s) <= w else "*" * w


_E3 = re.compile(r"(\d)([+-]\d{3})$")


def _parse_real(field: str) -> float:
    f = field.strip().replace("D", "E").replace("d", "E")
    return float(_E3.sub(r"\1E\2", f))


def _reals(line: str, k: int, w: int = 18) -> List[float]:
    return [_parse_real(line[i * w:(i + 1) * w]) for i in range(k)]


def write_seren_form(path: str, snap: Snapshot) -> None:
    """write_data_seren_form.F90:128-250."""
    ids = snap._data_ids()
    idata, ilp, rdata, dpdata = snap._headers(len(ids))
    L = [FORMAT_FORM, f"{PR:2d}", f"{NDIM:2d}", f"{VDIM:2d}", f"{BDIM:2d}"]
    L += [f"{int(x):10d}" for x in idata]
    L += [f"{int(x):10d}" for x in ilp]
    L += [fortran_e(float(x)) for x in rdata]
    L += [fortran_e(float(x)) for x in dpdata]
    L += [u.ljust(20) for u in snap.units]
    L += [k.ljust(20) for k in ids]
    L += ["".join(f"{v:10d}" for v in (_TYPES[k][0], 1, snap.n, _TYPES[k][1], _TYPES[k][2]))
          for k in ids]
    for k in ids:
        a = getattr(snap, k)
        if k == "porig":
            L += [f"{int(v):10d}" for v in a]
        elif a.ndim == 2:
            L += ["".join(fortran_e(float(v)) for v in row) for row in a]
        else:
            L += [fortran_e(float(v)) for v in a]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def read_seren_form(path: str) -> Snapshot:
    """read_data_seren_form.F90:56-250 for gas-only files."""
    with open(path) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    if next(it).strip() != FORMAT_FORM:
        raise SnapshotError("Incorrect format of IC file")
    if int(next(it)) not in (2, 8):
        raise SnapshotError("Incorrect PR of IC file (double precision expected)")
    for want, what in ((NDIM, "NDIM"), (VDIM, "VDIM"), (BDIM, "BDIM")):
        if int(next(it)) != want:
            raise SnapshotError(f"Incorrect {what} of IC file")
    idata = np.array([int(next(it)[:10]) for _ in range(50)], np.int64)
    ilp = np.array([int(next(it)[:10]) for _ in range(50)], np.int64)
    rdata = np.array([_reals(next(it), 1)[0] for _ in range(50)])
    dpdata = np.array([_reals(next(it), 1)[0] for _ in range(50)])
    n = _check_counts(idata)
    nunit, ndata = int(idata[19]), int(idata[20])
    units = [next(it)[:20].rstrip() for _ in range(nunit)]
    ids = [next(it).strip() for _ in range(ndata)]
    typedata = [[int(x) for x in next(it).split()] for _ in range(ndata)]
    arrays = {}
    for k, t in zip(ids, typedata):
        ncomp, first, last, tid = t[0], t[1], t[2], t[3]
        cnt = last - first + 1
        if k == "B":
            raise SnapshotError("snapshot contains MHD fields (not supported)")
        if tid in (2, 3):
            arrays[k] = np.array([int(next(it).split()[0]) for _ in range(cnt)], np.int64)
        elif tid in (4, 5):
            rows = [_reals(next(it), ncomp) for _ in range(cnt)]
            a = np.array(rows)
            arrays[k] = a if ncomp > 1 else a[:, 0]
        else:
            for _ in range(cnt):
                next(it)
    snap = Snapshot(r=arrays["r"], m=arrays["m"], h=arrays["h"], v=arrays["v"],
                    rho=arrays["rho"], porig=arrays.get("porig"), temp=arrays.get("temp"),
                    u=arrays.get("u"), units=units)
    if snap.n != n:
        raise SnapshotError("particle arrays do not match ptot")
    snap._set_headers(idata, ilp, rdata, dpdata)
    return snap


# ---------------------------------------------------------------------------
# Column ASCII
# ---------------------------------------------------------------------------
_ASCII_GETTERS = {
    "x": lambda s: s.r[:, 0], "y": lambda s: s.r[:, 1], "z": lambda s: s.r[:, 2],
    "vx": lambda s: s.v[:, 0], "vy": lambda s: s.v[:, 1], "vz": lambda s: s.v[:, 2],
    "m": lambda s: s.m, "h": lambda s: s.h, "rho": lambda s: s.rho,
    "u": lambda s: s.u,
}


def write_ascii(path: str, snap: Snapshot, columns: Sequence[str] = ASCII_COLUMNS) -> None:
    """write_data_ascii.F90:38-92: '(I3,100E15.7)' per gas particle (ptype 1);
    columns the reference does not fill (u without INTERNAL_ENERGY, unknown
    names) are written as zero."""
    columns = list(columns)
    if len(columns) <= 1 or columns[0] != "ptype":
        raise SnapshotError("first column must be ptype")
    cols = []
    for c in columns[1:]:
        g = _ASCII_GETTERS.get(c)
        a = g(snap) if g is not None else None
        cols.append(np.zeros(snap.n) if a is None else a)
    with open(path, "w") as f:
        for p in range(snap.n):
            f.write(f"{1:3d}" + "".join(fortran_e(float(a[p]), 15, 7) for a in cols) + "\n")


def read_ascii(path: str, columns: Sequence[str] = ASCII_COLUMNS) -> Snapshot:
    """read_data_ascii.F90:40-150 for gas particles (ptype 1); rows of other
    types are rejected."""
    columns = list(columns)
    if len(columns) <= 1 or columns[0] != "ptype":
        raise SnapshotError("First column is not particle type id")
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                break
            ptype = int(t[0])
            if ptype != 1:
                raise SnapshotError(f"particle type {ptype} (only gas, type 1, is supported)")
            rows.append([_parse_real(x) for x in t[1:len(columns)]])
    a = np.array(rows, np.float64).reshape(len(rows), len(columns) - 1)
    n = len(rows)
    val = {c: a[:, i] for i, c in enumerate(columns[1:])}
    z = np.zeros(n)
    r = np.stack([val.get("x", z), val.get("y", z), val.get("z", z)], 1)
    v = np.stack([val.get("vx", z), val.get("vy", z), val.get("vz", z)], 1)
    return Snapshot(r=r, m=val.get("m", z), h=val.get("h", z), v=v, rho=val.get("rho", z),
                    u=val.get("u"))


def from_engine(engine, time: float = 0.0, nsteps: int = 0, snapshot: int = 0,
                temp: float = 1.0) -> Snapshot:
    """The device state as SEREN would dump it: particles in the current
    (tree) order with porig, as write_data_seren_unform.F90:153-186 writes
    them; isothermal runs carry temp = 1 (DIMENSIONLESS)."""
    d = engine.fetch(neighbours=False)
    n = len(d["m"])
    return Snapshot(r=d["pos"], m=d["m"], h=d["h"], v=d["v"], rho=d["rho"], porig=d["porig"],
                    temp=np.full(n, temp), time=time, lastsnap=time,
                    mgas_orig=float(np.sum(d["m"])), snapshot=snapshot, nsteps=nsteps,
                    h_fac=engine.opt.h_fac, pp_gather=engine.pp_gather, pgas_orig=n)


def load_into(engine, snap: Snapshot) -> None:
    """Loads a snapshot's particles (file order) into a ForceEngine."""
    engine.load(snap.r, snap.v, snap.m, snap.h)


def read(path: str, fmt: str, **kw) -> Snapshot:
    return {"seren_unform": read_seren_unform, "su": read_seren_unform,
            "seren_form": read_seren_form, "sf": read_seren_form,
            "ascii": read_ascii}[fmt](path, **kw)


def write(path: str, snap: Snapshot, fmt: str, **kw) -> None:
    {"seren_unform": write_seren_unform, "su": write_seren_unform,
     "seren_form": write_seren_form, "sf": write_seren_form,
     "ascii": write_ascii}[fmt](path, snap, **kw)
