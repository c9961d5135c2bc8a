"""GPU: standard SPH (SPH=STANDARD, HFIND=NUMBER) beyond the golden sizes.

h_gather.F90 brackets h and sorts the r^2 of a band of candidates; its
result is h = INVKERNRANGE sqrt(r^2 of the pp_gather-th nearest other
particle), whatever path the bracketing took.  At sizes the oracle does not
reach in seconds, every h from the device must equal that order statistic
computed here by brute force (r^2 formed as distance2.F90 forms it), bit for
bit; with a large pp_gather the band exceeds the kernel's buffer and is
narrowed by bisection first.  The density is checked as a sum over the
particle's own gather neighbours, and a cloud with no more particles than
pp_gather is an error rather than a search that never ends.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud, rel_err  # noqa: E402

pytestmark = pytest.mark.gpu


def _engine(pp_gather=0, kernel="m4"):
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="quadrupole", mac="geometric", precision="fp64",
                       thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS, kernel=kernel,
                       sph="standard", pp_gather=pp_gather)


def _kth_h(pos, idx, k, invrange):
    out = np.empty(len(idx))
    for t, i in enumerate(idx):
        d = pos - pos[i]
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        r2[i] = np.inf
        out[t] = invrange * np.sqrt(np.partition(r2, k - 1)[k - 1])
    return out


@pytest.mark.parametrize("n,profile,pp_gather,kernel", [
    (40000, "plummer", 0, "m4"),        # pp_gather 57 from h_fac 1.2
    (30000, "r2", 300, "m4"),           # bands above the 64-value buffer
    (30000, "uniform", 0, "quintic"),   # range 3h: pp_gather 195
])
def test_h_is_the_pp_gather_th_neighbour(n, profile, pp_gather, kernel):
    print(f"case n={n} {profile} pp_gather={pp_gather} {kernel}", flush=True)
    pos, vel, mass, h = cloud(n, profile, 31)
    eng = _engine(pp_gather, kernel)
    try:
        eng.load(pos, vel, mass, np.zeros(n))
        eng.force_eval(hguess=True)
        d = eng.fetch()
        k = eng.pp_gather
        invrange = 0.3333333333333 if kernel == "quintic" else 0.5
        rng = np.random.default_rng(5)
        idx = rng.choice(n, 1500, replace=False)
        want = _kth_h(d["pos"], idx, k, invrange)
        got = d["h"][idx]
        print(f"  pp_gather {k}: {int(np.sum(got != want))} of {len(idx)} differ", flush=True)
        assert np.array_equal(got, want)
        # rho = m_p w0(0) / h^3 + sum over gather neighbours (all_sph.F90):
        # positive, and at least the self term
        assert np.all(d["rho"] > 0)
        assert np.all(d["omega"] == 1.0) and np.all(d["zeta"] == 0.0)
    finally:
        eng.close()


def test_h_in_a_lattice_shell_of_more_than_64_points():
    """pp_gather inside a cubic-lattice shell of 72 points at one r^2 (26
    spacings^2; 514 points lie closer): the band of r^2 values cannot be
    narrowed below 64 values by bisection, and the pp_gather-th neighbour is
    that shell, exactly as an insertion sort of the whole band finds it."""
    g = np.arange(24, dtype=np.float64) / 16.0          # exact coordinates and r^2
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    eng = _engine(550)
    try:
        eng.load(pos, np.zeros_like(pos), mass, np.zeros(n))
        eng.force_eval(hguess=True)
        d = eng.fetch()
        rng = np.random.default_rng(9)
        idx = rng.choice(n, 1200, replace=False)
        want = _kth_h(d["pos"], idx, 550, 0.5)
        got = d["h"][idx]
        print(f"  {int(np.sum(got != want))} of {len(idx)} differ", flush=True)
        assert np.array_equal(got, want)
        inner = np.all((d["pos"] > 6 / 16) & (d["pos"] < 17 / 16), axis=1)
        assert inner.sum() > 500
        assert np.all(d["h"][inner] == 0.5 * np.sqrt(26.0 / 256.0))
    finally:
        eng.close()


def test_density_of_a_lattice_particle():
    """A particle in a cubic lattice with its neighbours at known distances:
    rho = sum m W(r, h) over r < 2h with the tabulated M4 kernel, within the
    table's resolution."""
    g = np.arange(-6, 7, dtype=np.float64) * 0.1
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    eng = _engine(0)
    try:
        eng.load(pos, np.zeros_like(pos), mass, np.full(n, 0.1))
        eng.force_eval(hguess=True)
        d = eng.fetch()
        c = int(np.argmin(np.sum(d["pos"] ** 2, 1)))   # the lattice centre
        hc = d["h"][c]
        r = np.sqrt(np.sum((d["pos"] - d["pos"][c]) ** 2, 1))
        s = r / hc
        w = np.where(s < 1, 1 - 1.5 * s**2 + 0.75 * s**3,
                     np.where(s < 2, 0.25 * (2 - s) ** 3, 0.0)) / np.pi
        want = np.sum(mass * w) / hc**3
        print(f"  rho {d['rho'][c]:.12e} want {want:.12e}", flush=True)
        assert rel_err(np.array([d["rho"][c]]), np.array([want]))[0] < 2e-3
    finally:
        eng.close()


def test_too_few_particles_is_an_error():
    from seren_amd.engine import SerenxError
    pos, vel, mass, h = cloud(40, "uniform", 3)
    eng = _engine(57)
    try:
        eng.load(pos, vel, mass, np.zeros(40))
        with pytest.raises(SerenxError):
            eng.force_eval(hguess=True)
    finally:
        eng.close()
