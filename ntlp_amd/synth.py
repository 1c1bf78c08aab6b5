"""Synthetic inputs of the droplet-step workloads.

Fields: u, v, w are sums of random-phase Fourier modes on the periodic box
with amplitude |k|^(-11/6) (a k^(-5/3) energy spectrum), scaled to the rms
asked for; T and q per workload, with q from relative humidity by the
reference's conversion q = RH/100 * Mw/Ru/T * e_s(T)/rho_a (les.F:10632).
Droplets: uniform positions, lognormal or log-uniform radii, salt mass per
workload.  Everything is seeded (numpy PCG64).

Field arrays are returned in the CORE layout (0:nnz+1, 1:nny, 1:nnx) with z
fastest, i.e. numpy arrays of shape (nnx, nny, nnz+2); the ghost levels 0 and
nnz+1 hold the periodic images of levels nnz and 1.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import COL, Consts, Flags, Particles
from .grid import Grid


def mod_magnus(T):
    """Saturation vapour pressure (les.F:10764-10774)."""
    return 610.94 * np.exp((17.6257 * (T - 273.15)) / (243.04 + (T - 273.15)))


def spectral_field(rng: np.random.Generator, nnx: int, nny: int, nnz: int, rms: float) -> np.ndarray:
    noise = rng.standard_normal((nnx, nny, nnz))
    fk = np.fft.rfftn(noise)
    kx = np.fft.fftfreq(nnx) * nnx
    ky = np.fft.fftfreq(nny) * nny
    kz = np.fft.rfftfreq(nnz) * nnz
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (-11.0 / 12.0), 0.0)
    f = np.fft.irfftn(fk * amp, s=(nnx, nny, nnz), axes=(0, 1, 2))
    f *= rms / np.sqrt(np.mean(f * f))
    return f


def with_ghosts(f: np.ndarray) -> np.ndarray:
    """(nnx, nny, nnz) periodic box -> (nnx, nny, nnz+2) with z ghosts."""
    return np.concatenate([f[:, :, -1:], f, f[:, :, :1]], axis=2)


def humidity(T: np.ndarray, rh_percent: np.ndarray, c: Consts) -> np.ndarray:
    return rh_percent / 100.0 * c.Mw / c.Ru / T * mod_magnus(T) / c.surf_rho


def make_fields(grid: Grid, c: Consts, rng: np.random.Generator, T0: float = 300.0, T_sd: float = 0.1,
                T_grad: float = 0.0, rh_mean: float = 100.0, rh_sd: float = 0.5, rh_uniform=None,
                urms: float = 0.1):
    nx, ny, nz = grid.nnx, grid.nny, grid.nnz
    u = with_ghosts(spectral_field(rng, nx, ny, nz, urms))
    v = with_ghosts(spectral_field(rng, nx, ny, nz, urms))
    w = with_ghosts(spectral_field(rng, nx, ny, nz, urms))
    zz = grid.zz[None, None, :]
    T = T0 + T_grad * np.maximum(zz, 0.0) + T_sd * rng.standard_normal((nx, ny, nz + 2))
    if rh_uniform is not None:
        rh = rng.uniform(rh_uniform[0], rh_uniform[1], size=(nx, ny, nz + 2))
    else:
        rh = rh_mean + rh_sd * rng.standard_normal((nx, ny, nz + 2))
    q = humidity(T, rh, c)
    return u, v, w, T, q


def make_droplets(grid: Grid, rng: np.random.Generator, n: int, radius: np.ndarray, m_s: np.ndarray,
                  kappa_s: float = 1.2, mult: int = 1) -> Particles:
    """Positions uniform over [0,xl) x [0,yl) x [r, z(nnz)-r]; vp, Tp and rc
    are left zero for the caller to initialise (interpolation, then
    crit_radius) on the device."""
    p = Particles.empty(n)
    p["x"] = rng.uniform(0.0, grid.xl, n)
    p["y"] = rng.uniform(0.0, grid.yl, n)
    zl = grid.zl
    p["z"] = radius + rng.uniform(0.0, 1.0, n) * (zl - 2 * radius)
    p["radius"] = radius
    p["m_s"] = m_s
    p["kappa_s"] = kappa_s
    p.mult[:] = mult
    return p


def lognormal(rng, n, median, sigma_ln):
    return median * np.exp(sigma_ln * rng.standard_normal(n))


def dry_mass(r_dry, rhos=2000.0):
    """m_s = rhos * (4/3) pi r_d^3 (particles.f90:2498-2499)."""
    return rhos * (4.0 / 3.0) * np.pi * r_dry ** 3


def ensemble(k: int, seed: int = 0):
    """The droplet ensemble of workload k as drawn on the device
    (ntlp_gpu_particles_generate): the same distributions as workload()'s
    host draws, from the counter-based generator instead of numpy."""
    from . import (DIST_LOGNORMAL, DIST_LOGUNIFORM, SALT_CHOICE3, SALT_DRY_LOGNORMAL, SALT_FIXED, Ensemble)

    if k == 1:
        return Ensemble(seed=seed, radius_kind=DIST_LOGNORMAL, radius_p=(10e-6, 0.3), salt_kind=SALT_FIXED,
                        salt_p=(1e-18,))
    if k in (2, 4):
        return Ensemble(seed=seed, radius_kind=DIST_LOGNORMAL, radius_p=(8e-6, 0.4), salt_kind=SALT_DRY_LOGNORMAL,
                        salt_p=(0.05e-6, 0.5))
    if k == 3:
        return Ensemble(seed=seed, radius_kind=DIST_LOGUNIFORM, radius_p=(5e-6, 20e-6), radius2_kind=DIST_LOGUNIFORM,
                        radius2_p=(1e-6, 2e-6), frac2=0.6, salt_kind=SALT_FIXED, salt_p=(5e-18,))
    if k == 5:
        return Ensemble(seed=seed, radius_kind=DIST_LOGNORMAL, radius_p=(10e-6, 0.35), salt_kind=SALT_CHOICE3,
                        salt_p=(1e-19, 1e-17, 1e-15))
    raise ValueError(f"unknown workload {k}")


@dataclass
class Workload:
    name: str
    grid: Grid
    consts: Consts
    flags: Flags
    fields: tuple
    droplets: Particles
    nsteps: int
    mode: int = 2


def workload(k: int, n: int | None = None, seed: int = 0, make_fields_: bool = True,
             droplets: bool = True) -> Workload:
    """The five workloads: 1 = 64^3 / 100k / 1 step; 2 = 128^3 / 2M / 100
    steps; 3 = 256x256x128 tanh-stretched / 20M (60% haze) / 50 steps;
    4 = 128^3 serving fields; 5 = 512x512x256 / 100M / 20 steps.  `n`
    overrides the droplet count (parity tests at reduced size); droplets=False
    leaves them to be drawn on the device (ensemble(k))."""
    rng = np.random.default_rng(seed)
    # the fields have their own stream, so they are the same with or without
    # host-drawn droplets
    rng_f = np.random.default_rng([seed, 1])
    flags = Flags()
    if k == 1:
        grid = Grid.uniform(64, 64, 64, 0.0156, 0.0156, 0.0156)
        c = Consts(dt=1.0e-3, radius_init=10e-6)
        n = 100_000 if n is None else n
        r = lognormal(rng, n, 10e-6, 0.3)
        ms = np.full(n, 1e-18)
        fl = make_fields(grid, c, rng_f) if make_fields_ else None
        d = make_droplets(grid, rng, n, r, ms) if droplets else None
        return Workload("64^3 periodic, 100k droplets, 1 step", grid, c, flags, fl, d, 1)
    if k in (2, 4):
        grid = Grid.uniform(128, 128, 128, 0.0156, 0.0156, 0.0156)
        c = Consts(dt=5.0e-4, radius_init=8e-6)
        n = 2_000_000 if n is None else n
        r = lognormal(rng, n, 8e-6, 0.4) if droplets else None
        ms = dry_mass(lognormal(rng, n, 0.05e-6, 0.5)) if droplets else None
        fl = make_fields(grid, c, rng_f) if make_fields_ else None
        d = make_droplets(grid, rng, n, r, ms) if droplets else None
        return Workload("128^3 periodic, 2M droplets, 100 steps", grid, c, flags, fl, d, 100)
    if k == 3:
        grid = Grid.tanh_stretched(256, 256, 128, 0.05, 0.05, 0.01, 0.10)
        c = Consts(dt=5.0e-4, radius_init=1.5e-6)
        n = 20_000_000 if n is None else n
        if droplets:
            nh = int(round(0.6 * n))
            r = np.empty(n)
            r[:nh] = np.exp(rng.uniform(np.log(1e-6), np.log(2e-6), nh))
            r[nh:] = np.exp(rng.uniform(np.log(5e-6), np.log(20e-6), n - nh))
            r = r[rng.permutation(n)]
        ms = np.full(n, 5e-18)  # critical radius ~1.4 um: the haze sits at activation
        fl = (make_fields(grid, c, rng_f, T0=295.0, T_sd=0.0, T_grad=0.02, rh_uniform=(99.8, 100.2))
              if make_fields_ else None)
        d = make_droplets(grid, rng, n, r, ms) if droplets else None
        return Workload("256x256x128 tanh-stretched, 20M droplets (60% haze), 50 steps", grid, c, flags, fl, d, 50)
    if k == 5:
        grid = Grid.uniform(512, 512, 256, 0.0156, 0.0156, 0.0156)
        c = Consts(dt=5.0e-4, radius_init=10e-6)
        n = 100_000_000 if n is None else n
        r = lognormal(rng, n, 10e-6, 0.35) if droplets else None
        fl = make_fields(grid, c, rng_f) if make_fields_ else None
        d = make_droplets(grid, rng, n, r, rng.choice(np.array([1e-19, 1e-17, 1e-15]), size=n)) if droplets else None
        return Workload("512x512x256 periodic, 100M droplets, 20 steps", grid, c, flags, fl, d, 20)
    raise ValueError(f"unknown workload {k}")
