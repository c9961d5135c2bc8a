"""Seeded synthetic initial conditions for the force-evaluation workloads.

All generators return float64 numpy arrays: ``pos`` (N, 3), ``vel`` (N, 3),
``mass`` (N,) and ``h`` (N,), with total mass 1 (m = 1/N) and G = 1.
Randomness comes from ``numpy.random.Generator(PCG64(seed))`` only.

Profiles
--------
* ``uniform_sphere``: uniform density inside r < 1 (r = U^(1/3)).
* ``r2_sphere``: rho ~ r^-2 truncated at r = 1, so M(r) ~ r and r = U.
* ``plummer``: Plummer sphere of scale length ``a`` truncated at ``rtrunc``,
  by inverse-CDF sampling of M(r) = r^3 / (r^2 + a^2)^(3/2).

Velocities
----------
``turbulent_velocity`` builds a divergence-free Gaussian random field with
power spectrum P(k) ~ k^-4 on an ``ngrid``^3 periodic grid by FFT and
interpolates it trilinearly onto the particles. It follows the construction
of SEREN's IC tools (src/ic/velfield.F90:42-112 builds the same spectrum by a
direct mode sum; src/ic/ic_subroutines.F90:419-530 interpolates trilinearly)
but uses an FFT so that 128^3 grids are practical.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _isotropic(rng: np.random.Generator, n: int) -> np.ndarray:
    cost = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    return np.stack([sint * np.cos(phi), sint * np.sin(phi), cost], axis=1)


def uniform_sphere(n: int, seed: int) -> np.ndarray:
    rng = _rng(seed)
    r = rng.uniform(0.0, 1.0, n) ** (1.0 / 3.0)
    return r[:, None] * _isotropic(rng, n)


def r2_sphere(n: int, seed: int) -> np.ndarray:
    rng = _rng(seed)
    r = rng.uniform(0.0, 1.0, n)
    return r[:, None] * _isotropic(rng, n)


def plummer(n: int, seed: int, a: float = 0.1, rtrunc: float = 1.0) -> np.ndarray:
    rng = _rng(seed)
    mmax = rtrunc ** 3 / (rtrunc ** 2 + a ** 2) ** 1.5
    x = rng.uniform(0.0, mmax, n)
    x = np.maximum(x, 1e-300)
    r = a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    return r[:, None] * _isotropic(rng, n)


def turbulent_velocity(pos: np.ndarray, seed: int, vrms: float,
                       ngrid: int = 128, box: float = 2.0) -> np.ndarray:
    """Divergence-free P(k) ~ k^-4 velocity field sampled at ``pos``.

    The grid spans [-box/2, box/2)^3 periodically; the field is rescaled so
    that the particle rms speed equals ``vrms``.
    """
    rng = _rng(seed)
    k1 = np.fft.fftfreq(ngrid) * ngrid
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    # |v_k|^2 ~ P(k) / (4 pi k^2) in 3-D for P(k) ~ k^-4  ->  amplitude k^-3
    amp = k2 ** (-1.5)
    amp[0, 0, 0] = 0.0
    fields = []
    for _ in range(3):
        re = rng.standard_normal((ngrid, ngrid, ngrid))
        im = rng.standard_normal((ngrid, ngrid, ngrid))
        fields.append((re + 1j * im) * amp)
    vk = np.stack(fields)  # (3, n, n, n)
    # project out the compressive part: v_k -> v_k - k (k.v_k)/k^2
    kdotv = kx * vk[0] + ky * vk[1] + kz * vk[2]
    vk[0] -= kx * kdotv / k2
    vk[1] -= ky * kdotv / k2
    vk[2] -= kz * kdotv / k2
    vgrid = np.real(np.fft.ifftn(vk, axes=(1, 2, 3)))
    # trilinear interpolation onto particles
    u = (pos + 0.5 * box) / box * ngrid
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    out = np.zeros_like(pos)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1.0 - f[:, 0]
        ix = (i0[:, 0] + dx) % ngrid
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1.0 - f[:, 1]
            iy = (i0[:, 1] + dy) % ngrid
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1.0 - f[:, 2]
                iz = (i0[:, 2] + dz) % ngrid
                w = wx * wy * wz
                for c in range(3):
                    out[:, c] += w * vgrid[c, ix, iy, iz]
    rms = np.sqrt(np.mean(np.sum(out * out, axis=1)))
    if rms > 0:
        out *= vrms / rms
    return out


def make_cloud(n: int, profile: str, seed: int, mach: float = 0.0,
               cs: float = 0.2, ngrid: int = 128):
    """Return (pos, vel, mass, h) for one cloud; h = 0 (guessed by the tree)."""
    if profile == "uniform":
        pos = uniform_sphere(n, seed)
    elif profile == "r2":
        pos = r2_sphere(n, seed)
    elif profile == "plummer":
        pos = plummer(n, seed)
    else:
        raise ValueError(f"unknown profile {profile!r}")
    if mach > 0.0:
        vel = turbulent_velocity(pos, seed + 7919, mach * cs, ngrid=ngrid)
    else:
        vel = np.zeros_like(pos)
    mass = np.full(n, 1.0 / n)
    h = np.zeros(n)
    return np.ascontiguousarray(pos), np.ascontiguousarray(vel), mass, h
