"""Shared test inputs: small seeded droplet ensembles on small grids, with the
initial state (vp = uf, Tp = Tf, rc) set by the oracle so the device and the
oracle start from identical bits."""
from __future__ import annotations

import numpy as np

from ntlp_amd import COL, Consts, Flags, Particles, synth
from ntlp_amd.grid import Grid
from oracle import Oracle


def small_grid(kind: str = "uniform", nnx: int = 16, nny: int = 12, nnz: int = 20) -> Grid:
    if kind == "uniform":
        return Grid.uniform(nnx, nny, nnz, 0.02, 0.025, 0.015)
    return Grid.tanh_stretched(nnx, nny, nnz, 0.02, 0.025, 0.005, 0.03)


def mixed_droplets(grid: Grid, rng: np.random.Generator, n: int) -> Particles:
    """40% haze near activation (1-2 um, m_s = 5e-18 kg), 60% cloud droplets
    (lognormal 8 um, dry radius lognormal 0.05 um)."""
    nh = int(0.4 * n)
    r = np.empty(n)
    r[:nh] = np.exp(rng.uniform(np.log(1e-6), np.log(2e-6), nh))
    r[nh:] = synth.lognormal(rng, n - nh, 8e-6, 0.5)
    ms = np.empty(n)
    ms[:nh] = 5e-18
    ms[nh:] = synth.dry_mass(synth.lognormal(rng, n - nh, 0.05e-6, 0.5))
    perm = rng.permutation(n)
    return synth.make_droplets(grid, rng, n, r[perm], ms[perm], kappa_s=1.2, mult=1)


def init_state(ora: Oracle, p: Particles) -> None:
    """vp = uf(xp), Tp = Tf(xp), rc = crit_radius(m_s, kappa_s, Tp)."""
    ora.interp(p.soa)
    for d in "xyz":
        p["v" + d] = p["uf" + d]
    p["Tp"] = p["Tf"]
    p["rc"] = ora.crit_radius(p.soa, COL)
    p.mult[:] = np.random.default_rng(7).integers(1, 2000, p.n)


def small_case(seed: int = 0, n: int = 3000, kind: str = "uniform", flags: Flags | None = None,
               dt: float = 5e-4, rh=(99.7, 100.5)):
    rng = np.random.default_rng(seed)
    grid = small_grid(kind)
    consts = Consts(dt=dt, radius_init=5e-6)
    flags = flags or Flags()
    fields = synth.make_fields(grid, consts, rng, T0=300.0, T_sd=0.1, rh_uniform=rh)
    p = mixed_droplets(grid, rng, n)
    ora = Oracle(grid, consts, flags)
    ora.set_fields(*fields)
    init_state(ora, p)
    return grid, consts, flags, fields, p, ora


def edge_positions(grid: Grid) -> np.ndarray:
    """Positions on the cell-index edges: nodes, x = 0, x = xl, z on zz(k)
    and z(k), below zz(1), above zz(nnz)."""
    pts = []
    for k in range(grid.nnz + 1):
        pts.append((grid.dx * 3, grid.dy * 2, grid.zz[k]))
        pts.append((grid.dx * 3.5, grid.dy * 2.5, grid.z[k]))
    pts += [(0.0, 0.0, grid.zz[1] * 0.5), (grid.xl, grid.yl, grid.z[grid.nnz]),
            (grid.xl, 0.5 * grid.yl, grid.zz[grid.nnz] + 1e-9), (grid.dx * 5, grid.dy * 7, grid.z[1]),
            (np.nextafter(grid.dx * 4, 0), np.nextafter(grid.dy * 4, 1), grid.zz[2]),
            (grid.xl * 0.999999, grid.yl * 1e-9, grid.zz[grid.nnz])]
    return np.array(pts)


def rel_err(a: np.ndarray, b: np.ndarray, floor: float) -> float:
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def radius_outliers(fp64: np.ndarray, mixed: np.ndarray, fp64_aj: np.ndarray, bound: float = 1e-5):
    """Mixed-precision radii against the FP64 reference path, with the FP64
    exact-Jacobian path (PREC_FP64_AJ) as the arbiter.  The mixed solve uses
    the exact Jacobian; where the reference's forward-difference Jacobian
    makes its solve fail (Gauss-Newton, then Levenberg-Marquardt, then the
    reset to the pre-step radius, particles.f90:3127-3153) the exact-Jacobian
    solve converges, and the two FP64 paths differ by that step's growth.
    Returns (relative errors mixed vs fp64, mixed vs fp64_aj, indices above
    the bound that the FP64 paths' own disagreement does not explain)."""
    e_ref = np.abs(mixed / fp64 - 1.0)
    e_aj = np.abs(mixed / fp64_aj - 1.0)
    ref_aj = np.abs(fp64 / fp64_aj - 1.0)
    unexplained = np.flatnonzero((e_ref > bound) & ~((ref_aj > bound) & (e_aj <= bound)))
    return e_ref, e_aj, unexplained


def oracle_spread(ora, before, idx, its, m: int = 24):
    """(lo, hi), each (2, len(idx)), of the radius and Tp that the oracle's BE
    steps `its` give for the droplets idx of `before` when each droplet's
    radius or temperature is moved by 1..m ulps either way: how far last-ulp
    differences (those of exp/log between the device and the host) move an
    ill-conditioned solve."""
    k = np.arange(1, m + 1, dtype=float) * 2.0 ** -52
    scale = np.concatenate([1 + k, 1 - k])
    nv = 2 * scale.size
    idx = np.asarray(idx)
    q = before.copy()
    q.soa = np.ascontiguousarray(np.repeat(before.soa[:, idx], nv, axis=1))
    q.mult = np.repeat(before.mult[idx], nv)
    rad = q["radius"].reshape(idx.size, nv)
    tp = q["Tp"].reshape(idx.size, nv)
    rad[:, : scale.size] *= scale
    tp[:, scale.size:] *= scale
    q["radius"] = rad.ravel()
    q["Tp"] = tp.ravel()
    for it in its:
        ora.update_be(q.soa, it)
    r = q["radius"].reshape(idx.size, nv)
    t = q["Tp"].reshape(idx.size, nv)
    return np.stack([r.min(1), t.min(1)]), np.stack([r.max(1), t.max(1)])
