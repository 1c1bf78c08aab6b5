"""CPU pins of the oracle's coupling amounts, RK3 microphysics, wall
reflection and flux statistics, each against the physics written down
independently here (numpy or 50-digit mpmath), not against the oracle's own
expressions:

- coupling (particles.f90:1187-1264): the node sums of each source field times
  the cell volume equal the droplets' budgets -- vapour rho_w 4 pi r^2 rdot,
  sensible heat m_w Cpp dTs/dt, enthalpy, and the drag on the air of a droplet
  released in still air, whose backward-Euler velocity is dt g/(1 + dt/tau);
- one RK3 stage (particles.f90:2916-2966) with zetas(1) = 0: y + dt gama(1)
  y' with y' from the diffusional-growth, heat and drag laws;
- walls (particles.f90:2718-2723, 3099): the mirrored height, the negated
  velocity, x advanced with the pre-step velocity, to the bit;
- flux statistics (particles.f90:3208-3238): crossings of the w levels from
  numpy's searchsorted, masses from m_s + (4/3) pi r^3 rho_w.
"""
from __future__ import annotations

import mpmath as mp
import numpy as np
import pytest

from ntlp_amd import Consts, Flags, Particles
from ntlp_amd.grid import Grid
from oracle import Oracle


def uniform_case(n=400, seed=0, rh=(99.5, 100.6), flags=None, dt=5e-4, grav=(0.0, 0.0, -9.81)):
    """A uniform grid (constant cell volume), smooth fields and a population of
    cloud and haze droplets at their fluid state (vp = uf, Tp = Tf + noise)."""
    rng = np.random.default_rng(seed)
    g = Grid.uniform(10, 9, 12, 0.02, 0.02, 0.02)
    c = Consts(dt=dt, surf_rho=1.2, radius_init=8e-6, part_grav=grav)
    f = flags or Flags()
    o = Oracle(g, c, f)
    shape = (g.nnx, g.nny, g.nnz + 2)
    u, v, w = (0.1 * rng.standard_normal(shape) for _ in range(3))
    T = 300.0 + 0.1 * rng.standard_normal(shape)
    RH = rng.uniform(rh[0], rh[1], shape)
    q = RH / 100.0 * c.Mw / c.Ru / T * np.array([o.L.or_mod_magnus(t) for t in T.ravel()]).reshape(shape) / c.surf_rho
    o.set_fields(u, v, w, T, q)
    p = Particles.empty(n)
    p["x"], p["y"] = rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n)
    r = np.exp(rng.uniform(np.log(1e-6), np.log(2e-5), n))
    p["z"] = rng.uniform(2 * r, g.zl - 2 * r)
    p["radius"], p["m_s"], p["kappa_s"] = r, 10.0 ** rng.uniform(-18.5, -16.5, n), 1.2
    o.interp(p.soa)
    for d in "xyz":
        p["v" + d] = p["uf" + d] + 0.02 * rng.standard_normal(n)
    p["Tp"] = p["Tf"] + 0.05 * rng.standard_normal(n)
    p.mult[:] = rng.integers(1, 500, n)
    return g, c, o, p


# ------------------------------------------------------------- coupling

def test_coupling_budgets_vapour_heat_enthalpy():
    """After one BE step, for each field f: sum over nodes of src_f * dV equals
    the sum over droplets of its budget (the trilinear weights sum to one, dV
    is one constant on a uniform grid):
      S_H:  -rho_w/rho_a * 4 pi r^2 * rdot * mult
      S_T:  -m_w Cpp / (rho_a Cpa) * dTs/dt * mult,  m_w = (4/3) pi r^3 rho_w,
            dTs/dt = (Tp - Tp_old)/dt - 3 Lv rdot / (Cpp r)
      S_TE: rho_w/rho_a * 4 pi r^2 rdot (Cpv/Cpa) (Tf - Tp) * mult
    with rdot = (r - r_old)/dt.  A wrong factor in the heat amount
    Tprhs_s*6*rhow/rhop/CpaCpp/taup_i*(pi2/2)*r*nuf, a missing mult or a
    wrong sign fails."""
    g, c, o, p = uniform_case(seed=1)
    o.update_be(p.soa, 2)
    src = o.coupling(p.soa, p.mult)
    dV = g.dx * g.dy * (g.zz[2] - g.zz[1])
    r, r0, Tp, Tp0, Tf = p["radius"], p["radiusold"], p["Tp"], p["Tpold"], p["Tf"]
    m = p.mult.astype(float)
    rdot = (r - r0) / c.dt
    assert np.abs(rdot).max() > 0  # droplets grow and shrink
    area = 4.0 * np.pi * r ** 2
    vap = c.rhow / c.surf_rho * area * rdot
    dTs = (Tp - Tp0) / c.dt - 3.0 * c.Lv * rdot / (c.Cpp * r)
    m_w = 4.0 / 3.0 * np.pi * r ** 3 * c.rhow
    want = {3: np.sum(-m_w * c.Cpp / (c.surf_rho * c.Cpa) * dTs * m),
            4: np.sum(-vap * m),
            5: np.sum(vap * (c.Cpv / c.Cpa) * (Tf - Tp) * m)}
    for f, w in want.items():
        got = src[f].sum() * dV
        # the enthalpy amount is -A*Tp + A*Tf (particles.f90:1258-1260): two
        # terms ~Tp/(Tp-Tf) times larger than their sum
        tol = 1e-11 if f < 5 else 1e-8
        scale = np.sum(np.abs(-m_w * c.Cpp / (c.surf_rho * c.Cpa) * dTs * m)) if f == 3 else \
            np.sum(np.abs(vap * m)) * (1.0 if f == 4 else c.Cpv / c.Cpa * np.abs(Tf - Tp).max())
        assert abs(got - w) <= tol * scale, (f, got, w)


def test_coupling_budget_momentum_of_a_droplet_released_in_still_air():
    """Still air, ievap = 0: one BE step from rest gives v = dt g/(1 + dt/tau)
    (tau = rho_p (2r)^2/(18 rho_a nu), Stokes drag, corrfac = 1), so the drag
    force on the air is S_m = -m_p/rho_a (dv/dt - g) mult =
    m_p/rho_a g (dt/tau)/(1 + dt/tau) mult, m_p = m_s + (4/3) pi r^3 rho_w."""
    g = Grid.uniform(8, 8, 8, 0.02, 0.02, 0.02)
    c = Consts(dt=1e-3, surf_rho=1.2, radius_init=1e-5, part_grav=(0.0, 0.0, -9.81))
    o = Oracle(g, c, Flags(ievap=0))
    zero = np.zeros((g.nnx, g.nny, g.nnz + 2))
    o.set_fields(zero, zero, zero, zero + 300.0, zero + 0.01)
    n = 5
    p = Particles.empty(n)
    p["x"], p["y"] = np.linspace(0.03, 0.12, n), np.linspace(0.14, 0.02, n)
    p["z"] = np.linspace(0.03, 0.13, n)
    r = np.array([1e-6, 4e-6, 1e-5, 2.5e-5, 6e-5])
    p["radius"], p["m_s"], p["kappa_s"], p["Tp"] = r, 1e-17, 1.2, 300.0
    p.mult[:] = [1, 3, 7, 11, 400]
    o.update_be(p.soa, 2)
    src = o.coupling(p.soa, p.mult)
    dV = g.dx * g.dy * (g.zz[2] - g.zz[1])
    m_p = p["m_s"] + 4.0 / 3.0 * np.pi * r ** 3 * c.rhow
    tau = m_p / (4.0 / 3.0 * np.pi * r ** 3) * (2 * r) ** 2 / (18 * c.surf_rho * c.nuf)
    a = c.dt / tau
    assert np.allclose(p["vz"], c.dt * c.part_grav[2] / (1 + a), rtol=1e-13, atol=0)
    force = m_p / c.surf_rho * c.part_grav[2] * a / (1 + a) * p.mult
    assert abs(src[2].sum() * dV / force.sum() - 1) < 1e-12
    assert np.abs(src[0]).max() == 0 and np.abs(src[1]).max() == 0
    # no evaporation: no vapour or enthalpy source
    assert np.abs(src[4]).max() == 0 and np.abs(src[5]).max() == 0


# ------------------------------------------------------------- RK3 stage

def mp_rhs(c, r, Tp, Tf, qinf, m_s, kappa, slip, rhoa):
    """Right-hand sides of the explicit droplet equations at 50 digits:
    rdot from diffusional growth with Ranz-Marshall ventilation and the Koehler
    curvature and solute terms, dTs/dt and dTL/dt from sensible and latent
    heat, and the drag factor f tau^-1 with f = 1 + 0.15 Re^0.687."""
    mp.mp.dps = 50
    r, Tp, Tf, qinf, m_s, kappa, slip = (mp.mpf(x) for x in (r, Tp, Tf, qinf, m_s, kappa, slip))
    rhow, rhos, nu, Sc, Pr = (mp.mpf(x) for x in (c.rhow, c.rhos, c.nuf, c.Sc, c.Pra))
    Mw, Ru, Gam, Lv, Cpa, Cpp = (mp.mpf(x) for x in (c.Mw, c.Ru, c.Gam, c.Lv, c.Cpa, c.Cpp))
    rhoa = mp.mpf(rhoa)
    V = mp.mpf(4) / 3 * mp.pi * r ** 3
    rhop = rhow + m_s / V
    tau = rhop * (2 * r) ** 2 / (18 * rhoa * nu)
    Re = 2 * r * slip / nu
    Sh = 2 + mp.mpf("0.6") * mp.sqrt(Re) * mp.cbrt(Sc)
    Nu = 2 + mp.mpf("0.6") * mp.sqrt(Re) * mp.cbrt(Pr)
    es = mp.mpf("610.94") * mp.exp(mp.mpf("17.6257") * (Tf - mp.mpf("273.15")) / (mp.mpf("243.04") + Tf - mp.mpf("273.15")))
    qs = Mw * es / (Ru * Tp * rhoa) * mp.exp(Lv * Mw / Ru * (1 / Tf - 1 / Tp) + 2 * Mw * Gam / (Ru * rhow * r * Tp)
                                             - kappa * m_s / (rhos * V))
    rdot = Sh / (9 * Sc) * (rhop / rhow) * (r / tau) * (qinf - qs)
    dTs = -Nu / (3 * Pr) * (Cpa / Cpp) * (rhop / rhow) * (Tp - Tf) / tau
    dTL = 3 * Lv * rdot / (Cpp * r)
    drag = (1 + mp.mpf("0.15") * Re ** mp.mpf("0.687")) / tau
    return rdot, dTs, dTL, drag, qs


def test_rk3_first_stage_matches_the_growth_and_heat_laws():
    """Stage 1 of particle_update_rk3 (zetas(1) = 0, so stale right-hand sides
    do not enter): r1 = r + dt gama(1) rdot, Tp1 = Tp + dt gama(1) (dTs + dTL),
    v1 = v + dt gama(1) (f/tau (u_f - v) + g), with the right-hand sides at the
    pre-stage state; q* at Tp is stored.  ievap = 1, super- and
    subsaturated droplets, haze to drizzle."""
    g, c, o, p = uniform_case(n=60, seed=2)
    # stale right-hand sides that zetas(1) = 0 must discard
    for col in ("xrhsx", "vrhsz", "Tprhs_s", "Tprhs_L", "radrhs"):
        p[col] = 123.0
    p0 = p.copy()
    o.update_rk3(p.soa, 1, 2)
    gam = c.gama[0]
    for i in range(p.n):
        slip = np.sqrt(sum((p0["v" + d][i] - p["uf" + d][i]) ** 2 for d in "xyz"))
        rdot, dTs, dTL, drag, qs = mp_rhs(c, p0["radius"][i], p0["Tp"][i], p["Tf"][i], p["qinf"][i], p0["m_s"][i],
                                          p0["kappa_s"][i], slip, c.surf_rho)
        dr_want = float(c.dt * gam * rdot)
        dr_got = p["radius"][i] - p0["radius"][i]
        assert abs(dr_got - dr_want) <= 1e-9 * abs(dr_want) + 1e-22, (i, dr_got, dr_want)
        dT_want = float(c.dt * gam * (dTs + dTL))
        dT_got = p["Tp"][i] - p0["Tp"][i]
        assert abs(dT_got - dT_want) <= 1e-8 * abs(dT_want) + 1e-12, (i, dT_got, dT_want)
        assert abs(p["qstar"][i] / float(qs) - 1) < 1e-12, i
        for d, gd in zip("xyz", c.part_grav):
            dv_want = float(c.dt * gam * (drag * (mp.mpf(p["uf" + d][i]) - mp.mpf(p0["v" + d][i])) + gd))
            dv_got = p["v" + d][i] - p0["v" + d][i]
            assert abs(dv_got - dv_want) <= 1e-9 * abs(dv_want) + 1e-15, (i, d)


# ----------------------------------------------------------------- walls

def test_walls_mirror_and_periodic_wrap_to_the_bit():
    """particle_bcs_nonperiodic (icase = 0, 2718-2723) after the BE advance,
    which moves x with the pre-step velocity (3099 before 3100):
      z_adv = z + dt vz_old, top = z(nnz) - r:  z = top - (z_adv - top),
      bottom bot = r:                           z = bot + (bot - z_adv),
    vz is the negated velocity of the same droplet away from the wall, and
    x = (x + dt vx_old) - xl through x = xl (particle_bcs_periodic)."""
    g = Grid.uniform(8, 8, 8, 0.02, 0.02, 0.02)
    c = Consts(dt=1e-3, radius_init=1e-5)
    o = Oracle(g, c, Flags(ievap=0))
    zero = np.zeros((g.nnx, g.nny, g.nnz + 2))
    o.set_fields(zero, zero, zero, zero + 300.0, zero + 0.01)
    p = Particles.empty(5)
    r = 1e-5
    p["radius"], p["m_s"], p["Tp"] = r, 1e-18, 300.0
    p["x"] = [0.05, g.xl - 1e-4, 1e-4, g.xl, 0.05]
    p["y"] = [0.05, 0.05, 0.05, 0.05, 0.05]
    p["z"] = [g.z[g.nnz] - 2 * r, 1.5 * r, 0.08, 0.08, 0.08]
    p["vx"] = [0.0, 1.0, -1.0, 0.0, 0.0]
    p["vz"] = [3.0, -3.0, 0.0, 0.0, 3.0]
    p0 = p.copy()
    # droplet 4 has droplet 0's velocity away from the walls; droplet 1's
    # velocity is the same with the opposite sign
    o.update_be(p.soa, 2)
    top = g.z[g.nnz] - r
    z_adv = p0["z"] + c.dt * p0["vz"]
    assert p["z"][0] == top - (z_adv[0] - top)
    assert p["vz"][0] == -p["vz"][4]
    bot = 0.0 + r
    assert p["z"][1] == bot + (bot - z_adv[1])
    q = p0.copy()
    q["vz"] = [0.0, -3.0, 0.0, 0.0, 0.0]
    q["z"] = [0.08, 0.08, 0.08, 0.08, 0.08]
    o.update_be(q.soa, 2)
    assert p["vz"][1] == -q["vz"][1]
    x_adv = p0["x"] + c.dt * p0["vx"]
    assert p["x"][1] == x_adv[1] - g.xl
    assert p["x"][2] == g.xl + x_adv[2]
    # x == xl is not wrapped (strict >, particles.f90:2826)
    assert p["x"][3] == g.xl
    # the position moved with the pre-step velocity, not the updated one
    assert p["vx"][1] != p0["vx"][1]


# ------------------------------------------------------- flux statistics

def test_be_flux_counts_level_crossings():
    """pflux/pmassflux/penegflux of one BE step against numpy: the w levels
    z(iz) a droplet crosses between its pre-step height and its advanced height
    z + dt vz_old (before the walls; above zl = z(nnz) counts as level nnz+1,
    below 0 as level 0), from searchsorted; +mult, +m_p mult, +m_p Cpp Tp mult
    upward and the negatives downward, with m_p = m_s + (4/3) pi r^3 rho_w at
    the new radius and Tp the new temperature."""
    g = Grid.tanh_stretched(8, 8, 16, 0.02, 0.02, 0.004, 0.02)
    c = Consts(dt=2e-3, surf_rho=1.2, radius_init=8e-6)
    rng = np.random.default_rng(7)
    o = Oracle(g, c, Flags())
    shape = (g.nnx, g.nny, g.nnz + 2)
    T = 300.0 + 0.1 * rng.standard_normal(shape)
    q = np.full(shape, 1.001 * c.Mw / c.Ru / 300.0 * o.L.or_mod_magnus(300.0) / c.surf_rho)
    o.set_fields(*(0.1 * rng.standard_normal(shape) for _ in range(3)), T, q)
    n = 3000
    p = Particles.empty(n)
    p["x"], p["y"] = rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n)
    r = np.exp(rng.uniform(np.log(2e-6), np.log(2e-5), n))
    p["z"] = rng.uniform(r, g.zl - r)
    p["radius"], p["m_s"], p["kappa_s"], p["Tp"] = r, 1e-17, 1.2, 300.0
    # fast droplets: many cross one or several levels, some leave the domain
    p["vz"] = rng.normal(0.0, 8.0, n)
    p["vx"], p["vy"] = rng.normal(0, 0.1, n), rng.normal(0, 0.1, n)
    p.mult[:] = rng.integers(1, 1000, n)
    p0 = p.copy()
    cnt, flux = o.update_be_flux(p.soa, p.mult, 2)
    z0 = p0["z"]
    z1 = p0["z"] + c.dt * p0["vz"]
    zl = g.z[g.nnz]
    k0 = np.searchsorted(g.z, z0, side="right")
    k1 = np.where(z1 > zl, g.nnz + 1, np.where(z1 < 0, 0, np.searchsorted(g.z, z1, side="right")))
    assert np.count_nonzero(k1 != k0) > n // 4 and np.any(z1 > zl) and np.any(z1 < 0)
    m_p = p["m_s"] + 4.0 / 3.0 * np.pi * p["radius"] ** 3 * c.rhow
    want = np.zeros((3, g.nnz + 1))
    for i in range(n):
        lo, hi = min(k0[i], k1[i]), max(k0[i], k1[i])
        s = 1.0 if k1[i] > k0[i] else -1.0
        for iz in range(lo, hi):
            want[0, iz] += s * p.mult[i]
            want[1, iz] += s * m_p[i] * p.mult[i]
            want[2, iz] += s * m_p[i] * c.Cpp * p["Tp"][i] * p.mult[i]
    assert np.array_equal(flux[0], want[0])
    for k in (1, 2):
        scale = np.abs(want[k]).max()
        assert np.max(np.abs(flux[k] - want[k])) <= 1e-12 * scale, k
