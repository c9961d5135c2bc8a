"""SEREN snapshot interchange (seren_amd.snapshot) against the reference's
own readers and writers (oracle/_ref/libseren_ref_io.so, built from
io/read_data_*.F90 and io/write_data_*.F90 by make -f oracle/Makefile.ref io).

* files written here and by the reference from the same state are byte-for-
  byte identical (binary dump, formatted dump, column ASCII);
* the reference reads files written here back to the same values, and this
  module reads the reference's files back to the same values;
* without the reference build: round trips (exact for the binary dump, to the
  printed digits for the text formats) and the Fortran E-editing rules.
"""
import ctypes as C
import os

import numpy as np
import pytest

from seren_amd import snapshot as S

HERE = os.path.dirname(os.path.abspath(__file__))
IOLIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libseren_ref_io.so")
KINDS = {"seren_unform": 0, "seren_form": 1, "ascii": 2}
ASCII_COLS = ("ptype", "x", "y", "z", "vx", "vy", "vz", "m", "h", "rho", "u")


def _state(n=257, seed=5):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * 0.3
    vel = rng.normal(size=(n, 3)) * 0.2
    mass = np.full(n, 1.0 / n)
    h = rng.uniform(0.01, 0.1, n)
    rho = rng.lognormal(size=n)
    temp = np.ones(n)
    porig = rng.permutation(n).astype(np.int32) + 1
    return pos, vel, mass, h, rho, temp, porig


def _snap(state, mgas_orig, time=0.75, nsteps=123, snapshot=4, h_fac=1.2, pp_gather=57):
    pos, vel, mass, h, rho, temp, porig = state
    n = len(mass)
    return S.Snapshot(r=pos, m=mass, h=h, v=vel, rho=rho, porig=porig, temp=temp, time=time,
                      lastsnap=time, mgas_orig=mgas_orig, snapshot=snapshot, nsteps=nsteps,
                      ntempnext=nsteps + 1, ndiagnext=nsteps + 2, nsnapnext=nsteps + 3,
                      nsinknext=nsteps + 4, h_fac=h_fac, gamma=1.4, mu_bar=2.35,
                      pp_gather=pp_gather, pgas_orig=n)


def test_fortran_e_editing():
    assert S.fortran_e(1.2) == "  0.1200000000E+01"
    assert S.fortran_e(0.0) == "  0.0000000000E+00"
    assert S.fortran_e(-2.5e-7) == " -0.2500000000E-06"
    assert S.fortran_e(9.99999999996) == "  0.1000000000E+02"     # carries into the exponent
    assert S.fortran_e(1.5e120) == "  0.1500000000+121"        # three-digit exponent
    assert S.fortran_e(-0.0) == " -0.0000000000E+00"
    assert S.fortran_e(1.0e-310) == "  0.1000000000-309"
    assert S._parse_real("0.1500000000+121") == 1.5e120
    assert S.fortran_e(0.125, 15, 7) == "  0.1250000E+00"


@pytest.mark.parametrize("fmt", ["seren_unform", "seren_form", "ascii"])
def test_round_trip(tmp_path, fmt):
    st = _state()
    snap = _snap(st, float(np.sum(st[2])))
    path = str(tmp_path / "snap")
    kw = {"columns": ASCII_COLS} if fmt == "ascii" else {}
    S.write(path, snap, fmt, **kw)
    back = S.read(path, fmt, **kw)
    if fmt == "seren_unform":
        for k in ("r", "m", "h", "v", "rho", "porig", "temp"):
            assert np.array_equal(getattr(back, k), getattr(snap, k)), k
        assert (back.time, back.nsteps, back.snapshot, back.pp_gather) == (0.75, 123, 4, 57)
    else:
        digits = 10 if fmt == "seren_form" else 7
        for k in ("r", "m", "h", "v", "rho"):
            a, b = getattr(back, k), getattr(snap, k)
            assert np.all(np.abs(a - b) <= 10.0 ** (1 - digits) * np.abs(b)), k
        if fmt == "seren_form":
            assert np.array_equal(back.porig, snap.porig)
    # a second write of what was read reproduces the file exactly
    path2 = str(tmp_path / "snap2")
    S.write(path2, back, fmt, **kw)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_rejects_foreign_files(tmp_path):
    p = tmp_path / "bad"
    p.write_bytes(b"\x14\x00\x00\x00" + b"DRAGON".ljust(20) + b"\x14\x00\x00\x00")
    with pytest.raises(S.SnapshotError):
        S.read_seren_unform(str(p))


# ---------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not os.path.exists(IOLIB),
                               reason="reference I/O harness not built (make ... io)")


class _RefIO:
    def __init__(self):
        lib = C.CDLL(IOLIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        lib.ref_io_set.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, ip, C.c_double,
                                   C.c_longlong, C.c_longlong, C.c_double, C.c_int]
        lib.ref_io_write.argtypes = [C.c_int, C.c_char_p, C.c_int]
        lib.ref_io_read.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_int]
        lib.ref_io_get.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, ip, ip, dp, lp]
        lib.ref_io_ptot.argtypes = [C.POINTER(C.c_int)]
        self.lib = lib

    def set(self, st, time=0.75, nsteps=123, snapshot=4):
        pos, vel, mass, h, rho, temp, porig = (np.ascontiguousarray(a) for a in st)
        self.lib.ref_io_set(len(mass), pos, mass, h, vel, rho, temp, porig, time, nsteps,
                            snapshot, 1.2, 57)

    def write(self, kind, path):
        b = path.encode()
        self.lib.ref_io_write(KINDS[kind], b, len(b))

    def read(self, kind, path):
        b = path.encode()
        self.lib.ref_io_read(KINDS[kind], b, len(b), 57)

    def get(self):
        n = C.c_int()
        self.lib.ref_io_ptot(C.byref(n))
        n = n.value
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        mass, h, rho, temp = (np.zeros(n) for _ in range(4))
        porig, counts = np.zeros(n, np.int32), np.zeros(8, np.int32)
        tm, steps = np.zeros(3), np.zeros(6, np.int64)
        self.lib.ref_io_get(n, pos, mass, h, vel, rho, temp, porig, counts, tm, steps)
        return dict(r=pos, v=vel, m=mass, h=h, rho=rho, temp=temp, porig=porig,
                    counts=counts, times=tm, steps=steps)


@pytest.fixture(scope="module")
def refio():
    return _RefIO()


@needs_ref
@pytest.mark.parametrize("fmt", ["seren_unform", "seren_form", "ascii"])
def test_files_identical_to_reference(tmp_path, monkeypatch, refio, fmt):
    monkeypatch.chdir(tmp_path)        # the reference's ASCII routines read ./asciicolumns.dat
    (tmp_path / "asciicolumns.dat").write_text("\n".join(ASCII_COLS) + "\n")
    st = _state()
    refio.set(st)
    ref_path = str(tmp_path / "ref_out")
    refio.write(fmt, ref_path)
    mgas = refio.get()["times"][2]
    mine = str(tmp_path / "mine_out")
    kw = {"columns": ASCII_COLS} if fmt == "ascii" else {}
    S.write(mine, _snap(st, mgas), fmt, **kw)
    assert open(mine, "rb").read() == open(ref_path, "rb").read()


@needs_ref
@pytest.mark.parametrize("fmt", ["seren_unform", "seren_form", "ascii"])
def test_cross_read(tmp_path, monkeypatch, refio, fmt):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "asciicolumns.dat").write_text("\n".join(ASCII_COLS) + "\n")
    st = _state(n=300, seed=9)
    kw = {"columns": ASCII_COLS} if fmt == "ascii" else {}
    # reference writes, this module reads
    refio.set(st, time=1.5, nsteps=9, snapshot=2)
    ref_path = str(tmp_path / "ref_out")
    refio.write(fmt, ref_path)
    mine = S.read(ref_path, fmt, **kw)
    # this module writes, the reference reads
    my_path = str(tmp_path / "my_out")
    S.write(my_path, mine, fmt, **kw)
    refio.read(fmt, my_path)
    got = refio.get()
    for k in ("r", "m", "h", "v", "rho"):
        assert np.array_equal(got[k], getattr(mine, k)), k
    if fmt != "ascii":
        assert np.array_equal(got["porig"], mine.porig)
        assert np.array_equal(got["temp"], mine.temp)
        assert got["times"][0] == mine.time == 1.5
        assert list(got["steps"][:2]) == [2, 9]
    if fmt == "seren_unform":
        for k, i in (("r", 0), ("v", 1), ("m", 2), ("h", 3), ("rho", 4)):
            assert np.array_equal(getattr(mine, k), st[i]), k
        assert np.array_equal(mine.porig, st[6])
