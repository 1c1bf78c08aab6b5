"""CPU pins of the oracle (no GPU): each check compares the oracle with
something other than its own formula -- numpy's searchsorted/argsort,
polynomial exactness, closed forms, an mpmath root of the BE equations
(tests/golden/be_roots.json), conservation and hand-worked boundary cases."""
from __future__ import annotations

import json
import os

import mpmath as mp
import numpy as np
import pytest

from ntlp_amd import COL, NREAL, Consts, Flags, Particles
from ntlp_amd.grid import Grid
from oracle import Oracle

from .helpers import edge_positions, small_case, small_grid

HERE = os.path.dirname(os.path.abspath(__file__))


def ora_for(grid, consts=None, flags=None):
    return Oracle(grid, consts or Consts(), flags or Flags())


# ------------------------------------------------------------ cell indices

@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_cell_index_matches_searchsorted(kind):
    g = small_grid(kind)
    rng = np.random.default_rng(0)
    pts = np.concatenate([edge_positions(g),
                          np.column_stack([rng.uniform(0, g.xl, 500), rng.uniform(0, g.yl, 500),
                                           rng.uniform(-0.1 * g.zl, 1.2 * g.zl, 500)])])
    ci = ora_for(g).cell_index(pts[:, 0], pts[:, 1], pts[:, 2])
    len_ = g.nnz + 2
    # first index with zz > z == number of zz <= z
    pzz = np.searchsorted(g.zz, pts[:, 2], side="right")
    pz = np.searchsorted(g.z, pts[:, 2], side="right")
    assert np.array_equal(ci[:, 0], np.floor(pts[:, 0] / g.dx).astype(int) + 1)
    assert np.array_equal(ci[:, 1], np.floor(pts[:, 1] / g.dy).astype(int) + 1)
    assert np.array_equal(ci[:, 2], np.where(pzz == len_, -2, pzz - 1))
    assert np.array_equal(ci[:, 3], np.where(pz == len_, -2, pz - 1))
    assert np.array_equal(ci[:, 4], np.where(pz == len_, -1, pz))


def test_counting_sort_is_a_stable_argsort():
    g = small_grid("stretched")
    rng = np.random.default_rng(1)
    n = 5000
    x, y, z = rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n), rng.uniform(0, g.zl, n)
    keys, counts, offsets, perm = ora_for(g).bin(x, y, z)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    assert np.array_equal(counts, np.bincount(keys, minlength=len(counts)))
    assert offsets[0] == 0 and offsets[-1] == n and np.array_equal(np.diff(offsets), counts)


def test_wcell_counts_are_a_bincount_of_searchsorted_cells():
    """calc_tauc_min's num_array (particles.f90:3320-3339): droplets counted
    at (ipt, jpt, kc) with kc = first w level above z, 1..nnz."""
    g = small_grid("stretched")
    rng = np.random.default_rng(4)
    n = 4000
    x, y, z = rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n), rng.uniform(0, g.zl, n)
    num = ora_for(g).wcell_counts(x, y, z)  # (nnz, nny, nnx)
    ix = np.floor(x / g.dx).astype(int)
    iy = np.floor(y / g.dy).astype(int)
    kc = np.searchsorted(g.z, z, side="right")  # first z(k) > z, k counted from 0
    ok = (ix < g.nnx) & (iy < g.nny) & (kc >= 1) & (kc <= g.nnz)
    want = np.zeros((g.nnz, g.nny, g.nnx), dtype=np.int64)
    np.add.at(want, (kc[ok] - 1, iy[ok], ix[ok]), 1)
    assert np.array_equal(num, want)


# ------------------------------------------------------------------ fields

def test_fill_ext_is_a_periodic_wrap():
    """fill_ext on one domain: ext(0:nnz+1, -1:nny+3, -1:nnx+3) holds the
    core (1:nnx, 1:nny) and its periodic images."""
    g = small_grid("uniform")
    core = np.random.default_rng(5).standard_normal((g.nnx, g.nny, g.nnz + 2))
    ext = ora_for(g).fill_ext(core)
    ii = (np.arange(-1, g.nnx + 4) - 1) % g.nnx
    jj = (np.arange(-1, g.nny + 4) - 1) % g.nny
    assert np.array_equal(ext, core[ii][:, jj])


# ----------------------------------------------------------- interpolation

def node_coords(g):
    """x, y of ext nodes (i = -1..nnx+3: x = dx (i-1)) and the z levels."""
    xi = g.dx * (np.arange(-1, g.nnx + 4) - 1.0)
    yj = g.dy * (np.arange(-1, g.nny + 4) - 1.0)
    return xi, yj


def ext_from(g, fn, levels):
    xi, yj = node_coords(g)
    X, Y, Z = np.meshgrid(xi, yj, levels, indexing="ij")
    return fn(X, Y, Z)


def interp_points(o, g, pts):
    p = Particles.empty(len(pts))
    p["x"], p["y"], p["z"] = pts[:, 0], pts[:, 1], pts[:, 2]
    o.interp(p.soa)
    return p


@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_trilinear_is_exact_for_affine_fields(kind):
    g = small_grid(kind)
    aff = lambda X, Y, Z: 0.3 + 1.7 * X - 0.9 * Y + 2.3 * Z
    o = ora_for(g)
    ext_uv = ext_from(g, aff, g.zz)
    ext_w = ext_from(g, aff, g.z)
    o.set_fields(ext_uv, ext_uv, ext_w, ext_uv, ext_uv, ext=True)
    rng = np.random.default_rng(2)
    n = 400
    # interior in z: between zz(1) and zz(nnz), where no wall override applies
    pts = np.column_stack([rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n), rng.uniform(g.zz[1], g.zz[g.nnz], n)])
    p = interp_points(o, g, pts)
    want = aff(pts[:, 0], pts[:, 1], pts[:, 2])
    for c in ("ufx", "ufy", "ufz", "Tf", "qinf"):
        assert np.max(np.abs(p[c] - want)) < 1e-12, c


def test_trilinear_wall_override_takes_node_ipt1_jpt1():
    """particles.f90:831-842: between the last u level and the wall (and below
    the first), u, v, T and q are the values of node (ipt+1, jpt+1) at level
    kpt (nnz) or 1; w keeps its trilinear value."""
    g = small_grid("uniform")
    o = ora_for(g)
    rng = np.random.default_rng(3)
    exts = [rng.standard_normal((g.nnx + 5, g.nny + 5, g.nnz + 2)) for _ in range(5)]
    o.set_fields(*exts, ext=True)
    xq, yq = 4.3 * g.dx, 6.6 * g.dy
    ipt, jpt = 5, 7
    for zq, level in ((0.5 * (g.zz[g.nnz] + g.z[g.nnz]), g.nnz), (0.5 * g.zz[1], 1)):
        p = interp_points(o, g, np.array([[xq, yq, zq]]))
        ix, iy = ipt + 1 + 1, jpt + 1 + 1  # ext index offset: i = -1 is row 0
        assert p["ufx"][0] == exts[0][ix, iy, level]
        assert p["ufy"][0] == exts[1][ix, iy, level]
        assert p["Tf"][0] == exts[3][ix, iy, level]
        assert p["qinf"][0] == exts[4][ix, iy, level]
        assert p["ufz"][0] != exts[2][ix, iy, level]


def test_sixth_order_is_exact_for_quintics_in_the_interior():
    g = small_grid("stretched")
    o = ora_for(g, flags=Flags(ilin=0))
    poly = lambda X, Y, Z: 1 + X ** 5 - 3 * X ** 2 * Y ** 3 + Y ** 4 * Z + 7 * Z ** 5 - X * Y * Z ** 3
    ext_uv = ext_from(g, poly, g.zz)
    ext_w = ext_from(g, poly, g.z)
    o.set_fields(ext_uv, ext_uv, ext_w, ext_uv, ext_uv, ext=True)
    rng = np.random.default_rng(4)
    n = 200
    pts = np.column_stack([rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n),
                           rng.uniform(g.z[4], g.z[g.nnz - 4], n)])
    p = interp_points(o, g, pts)
    want = poly(pts[:, 0], pts[:, 1], pts[:, 2])
    for c in ("ufx", "ufz", "Tf"):
        assert np.max(np.abs(p[c] - want)) < 1e-9 * np.max(np.abs(want)), c


# ----------------------------------------------------------- thermodynamics

def test_magnus_at_freezing_point():
    from oracle import lib
    assert lib().or_mod_magnus(273.15) == 610.94


def test_base_pressure_is_the_dry_adiabat_and_exner_inverts_it():
    """func_p_base (les.F:10787-10795) is the hydrostatic pressure of a dry
    adiabatic column: dp/dz = -p g / (Rd T(z)) with T(z) = Ts - g z / Cpa;
    exner (les.F:10775-10784) of that pressure is T(z)/Ts."""
    c = Consts()
    o = ora_for(small_grid(), c)
    ps, Ts = 101325.0, 300.0
    assert o.p_base(ps, Ts, 0.0) == ps
    assert o.exner(ps, ps) == 1.0
    for z in (10.0, 250.0, 1500.0):
        h = 1e-3
        dpdz = (o.p_base(ps, Ts, z + h) - o.p_base(ps, Ts, z - h)) / (2 * h)
        Tz = Ts - c.grav * z / c.Cpa
        assert abs(dpdz / (-o.p_base(ps, Ts, z) * c.grav / (c.Rd * Tz)) - 1) < 1e-7
        assert abs(o.exner(ps, o.p_base(ps, Ts, z)) - Tz / Ts) < 1e-13


def test_crit_radius_within_one_table_step_of_closed_form():
    """r_c = sqrt(3B/A) maximises A/r - B/r^3 (Koehler without the exponent)."""
    c = Consts()
    o = ora_for(small_grid())
    for ms in (1e-19, 5e-18, 1e-17, 1e-15):
        p = Particles.empty(1)
        p["m_s"], p["kappa_s"], p["Tp"] = ms, 1.2, 295.0
        rc = o.crit_radius(p.soa, COL)[0]
        A = 2 * c.Mw * c.Gam / (c.Ru * c.rhow * 295.0)
        B = 1.2 * ms / (c.rhos * 4.0 / 3.0 * np.pi)
        assert abs(np.log10(rc / np.sqrt(3 * B / A))) <= 5.0 / 1024 + 1e-12


def test_koehler_guess_solves_the_equilibrium_cubic():
    """rad_solver2's guess r satisfies ln S = A/r - B/r^3 (checked at 50 digits)."""
    c = Consts(surf_rho=1.2)
    o = ora_for(small_grid(), c)
    mp.mp.dps = 50
    # below the critical supersaturation the haze branch has an equilibrium;
    # 1.0003 is above S_c = 1 + sqrt(4A^3/(27B)) ~ 1.000035 for m_s = 1e-15
    # kg, where the reference falls back to the current radius (mflag = 1)
    for rh, ms, has_eq in ((0.999, 5e-18, True), (0.9995, 1e-17, True), (0.99, 1e-18, True),
                           (1.00001, 1e-15, True), (1.0003, 1e-15, False)):
        Tf = 290.0
        es = o.L.or_mod_magnus(Tf)
        q = rh * c.Mw / c.Ru / Tf * es / c.surf_rho
        rec = np.zeros(NREAL)
        rec[COL["Tf"]], rec[COL["qinf"]], rec[COL["m_s"]], rec[COL["kappa_s"]], rec[COL["radius"]] = Tf, q, ms, 1.2, 1e-6
        g, mflag = o.rad_solver2(rec, c.surf_rho)
        if not has_eq:
            assert mflag == 1 and g == 1e-6
            continue
        assert mflag == 0 and g > 0, (rh, ms)
        A = mp.mpf(2 * c.Mw * c.Gam) / (c.Ru * c.rhow * Tf)
        B = mp.mpf(1.2 * ms) / (mp.mpf(4) / 3 * mp.pi * c.rhos)
        lnS = mp.log(mp.mpf(c.Ru) * Tf * c.surf_rho * q / (c.Mw * es))
        resid = A / g - B / mp.mpf(g) ** 3 - lnS
        assert abs(resid) <= 1e-9 * (abs(A / g) + abs(B / mp.mpf(g) ** 3)), (rh, ms, resid)


# ------------------------------------------------------- Newton vs mpmath

def golden_roots():
    with open(os.path.join(HERE, "golden", "be_roots.json")) as fh:
        return json.load(fh)["roots"]


@pytest.mark.parametrize("radius_init", [1.5e-6, 1e-5])
def test_newton_matches_mpmath_roots(radius_init):
    """The oracle's Koehler guess + Gauss-Newton (+ LM) lands on the
    50-digit root of the BE equations, for any radius_init (taup0 cancels)."""
    c = Consts(surf_rho=1.2, radius_init=radius_init)
    o = ora_for(small_grid(), c)
    ri = radius_init
    for s in golden_roots():
        rec = np.zeros(NREAL)
        rec[COL["radius"]], rec[COL["Tp"]], rec[COL["Tf"]], rec[COL["qinf"]] = s["r"], s["Tp"], s["Tf"], s["qinf"]
        rec[COL["m_s"]], rec[COL["kappa_s"]], rec[COL["ufx"]] = s["m_s"], s["kappa"], s["slip"]
        taup0 = ((s["m_s"] / ((2.0 / 3.0) * c.pi2 * ri ** 3) + c.rhow) * (ri * 2) ** 2) / (18 * c.surf_rho * c.nuf)
        h = s["dt"] / taup0
        g, mflag = o.rad_solver2(rec, c.surf_rho)
        start = (g / s["r"], s["Tf"] / s["Tp"]) if mflag == 0 else (1.0, 1.0)
        v2, flag, _ = o.gauss_newton(rec, np.zeros(3), h, c.surf_rho, start)
        if flag:
            v2, flag, _ = o.lv_solver(rec, np.zeros(3), h, c.surf_rho, start)
        assert flag == 0, s
        rn, Tn = v2[0] * s["r"], v2[1] * s["Tp"]
        assert abs(rn / float(s["r_new"]) - 1) < 1e-9, (s, rn)
        assert abs(Tn / float(s["Tp_new"]) - 1) < 1e-11, (s, Tn)
        F = o.ie_vrt_nd(rec, c.surf_rho, np.zeros(3), float(s["r_new"]) / s["r"], float(s["Tp_new"]) / s["Tp"], h)
        assert np.max(np.abs(F)) < 1e-12


def test_levenberg_marquardt_alone_reaches_the_mpmath_roots():
    """LV_solver (particles.f90:4143-4223) started from the Koehler guess
    converges to the same roots on its own (its stop test is the same 1e-8
    correction norm, so it lands within the Newton tolerance)."""
    c = Consts(surf_rho=1.2, radius_init=1e-5)
    o = ora_for(small_grid(), c)
    ri = c.radius_init
    for s in golden_roots():
        rec = np.zeros(NREAL)
        rec[COL["radius"]], rec[COL["Tp"]], rec[COL["Tf"]], rec[COL["qinf"]] = s["r"], s["Tp"], s["Tf"], s["qinf"]
        rec[COL["m_s"]], rec[COL["kappa_s"]], rec[COL["ufx"]] = s["m_s"], s["kappa"], s["slip"]
        taup0 = ((s["m_s"] / ((2.0 / 3.0) * c.pi2 * ri ** 3) + c.rhow) * (ri * 2) ** 2) / (18 * c.surf_rho * c.nuf)
        h = s["dt"] / taup0
        g, mflag = o.rad_solver2(rec, c.surf_rho)
        start = (g / s["r"], s["Tf"] / s["Tp"]) if mflag == 0 else (1.0, 1.0)
        v2, flag, iters = o.lv_solver(rec, np.zeros(3), h, c.surf_rho, start)
        assert flag == 0 and iters < 1000, s
        assert abs(v2[0] * s["r"] / float(s["r_new"]) - 1) < 1e-8, s
        assert abs(v2[1] * s["Tp"] / float(s["Tp_new"]) - 1) < 1e-10, s


def be_case(dt=5e-4, rh=1.0, r=1e-5, ms=1e-18, n=1):
    g = Grid.uniform(8, 8, 8, 0.02, 0.02, 0.02)
    c = Consts(dt=dt, surf_rho=1.2, radius_init=r)
    o = ora_for(g, c)
    shape = (g.nnx, g.nny, g.nnz + 2)
    T = np.full(shape, 300.0)
    q = np.full(shape, rh * c.Mw / c.Ru / 300.0 * o.L.or_mod_magnus(300.0) / c.surf_rho)
    zero = np.zeros(shape)
    o.set_fields(zero, zero, zero, T, q)
    p = Particles.empty(n)
    p["x"], p["y"], p["z"] = 0.05, 0.05, 0.08
    p["radius"], p["m_s"], p["kappa_s"], p["Tp"], p["Tf"] = r, ms, 1.2, 300.0, 300.0
    return g, c, o, p


def test_growth_sign_follows_supersaturation():
    for rh, sign in ((1.01, 1), (0.95, -1)):
        g, c, o, p = be_case(rh=rh)
        r0 = p["radius"][0]
        o.update_be(p.soa, 2)
        assert np.sign(p["radius"][0] - r0) == sign


def test_stokes_terminal_settling():
    """Still air: the BE velocity update's fixed point is v = g tau_p
    (Stokes settling, corrfac = 1 in the BE step)."""
    g, c, o, p = be_case(dt=1e-3, rh=1.0, r=3e-6)
    p["vz"] = 0.0
    o.f.ievap = 0
    for it in range(2, 200):
        p["z"] = 0.08  # keep it inside the domain
        o.update_be(p.soa, it)
    r = p["radius"][0]
    V = 4.0 / 3.0 * np.pi * r ** 3
    rhop = c.rhow + p["m_s"][0] / V
    tau = rhop * (2 * r) ** 2 / (18 * c.surf_rho * c.nuf)
    assert abs(p["vz"][0] / (c.part_grav[2] * tau) - 1) < 1e-10


def test_rk3_velocity_matches_third_order_taylor():
    """Linear drag (tiny slip: corrfac -> 1), no gravity: one low-storage RK3
    step multiplies the slip by 1 + z + z^2/2 + z^3/6, z = -dt/tau_p."""
    g, c, o, p = be_case(dt=2e-5, r=4e-6)
    c2 = Consts(dt=2e-5, surf_rho=1.2, radius_init=4e-6, part_grav=(0.0, 0.0, 0.0))
    o.set_consts(c2)
    o.f.ievap = 0
    slip = 1e-9
    p["vx"] = slip
    for st in (1, 2, 3):
        p["x"], p["y"], p["z"] = 0.05, 0.05, 0.08
        o.update_rk3(p.soa, st, 2)
    r = p["radius"][0]
    V = 4.0 / 3.0 * np.pi * r ** 3
    tau = (c.rhow + p["m_s"][0] / V) * (2 * r) ** 2 / (18 * c.surf_rho * c.nuf)
    z = -c2.dt / tau
    assert abs(p["vx"][0] / (slip * (1 + z + z * z / 2 + z ** 3 / 6)) - 1) < 1e-6


# ------------------------------------------------------------- coupling

def test_coupling_conserves_and_is_linear():
    """sum(src * dV) over nodes equals the deposit of the same droplet placed on
    a node (the trilinear weights sum to one); deposits are linear in mult."""
    g = Grid.uniform(10, 9, 12, 0.02, 0.02, 0.02)
    c = Consts(dt=5e-4, radius_init=8e-6)
    o = ora_for(g, c)
    rng = np.random.default_rng(5)
    p = Particles.empty(1)
    for k in ("vx", "vy", "vz", "vxold", "vyold", "vzold"):
        p[k] = rng.normal(0, 0.1)
    p["radius"], p["radiusold"], p["Tp"], p["Tpold"], p["Tf"] = 8e-6, 7.99e-6, 300.02, 300.0, 300.01
    p["m_s"], p["kappa_s"] = 1e-17, 1.2
    p.mult[:] = 7
    dV = g.dx * g.dy * (g.zz[5] - g.zz[4])
    p["x"], p["y"], p["z"] = 3 * g.dx, 4 * g.dy, g.zz[4]
    on_node = o.coupling(p.soa, p.mult)
    p["x"], p["y"], p["z"] = 3.37 * g.dx, 4.81 * g.dy, g.zz[4] + 0.4 * (g.zz[5] - g.zz[4])
    spread = o.coupling(p.soa, p.mult)
    for f in range(6):
        tol = 1e-12 if f < 5 else 1e-9  # field 5 is a difference of two ~Tp-sized terms
        assert abs(spread[f].sum() * dV - on_node[f].sum() * dV) <= tol * abs(on_node[f].sum() * dV) + 1e-300
        big = np.abs(on_node[f]).max()
        assert np.count_nonzero(np.abs(on_node[f]) > 1e-12 * big) <= 1
        assert np.count_nonzero(np.abs(spread[f]) > 1e-12 * big) == (8 if big > 0 else 0)
    p.mult[:] = 21
    triple = o.coupling(p.soa, p.mult)
    for f in range(6):
        tol = 1e-14 if f < 5 else 1e-9
        assert np.max(np.abs(triple[f] - 3 * spread[f])) <= tol * np.max(np.abs(3 * spread[f])), f


def test_coupling_exchange_folds_ghosts_onto_their_periodic_images():
    """particle_coupling_exchange on one periodic domain: a deposit on the +1
    ghost plane/row lands on its periodic image (x = nnx+1 -> 1, y = nny+1 ->
    1, the corner -> (1, 1)); every other node keeps its value; the column
    sums over the core equal the sums over all nodes before the fold."""
    g = small_grid("uniform")
    o = ora_for(g)
    shape = (6, g.nnx + 1, g.nny + 1, g.nnz + 2)
    src = np.random.default_rng(2).standard_normal(shape)
    xs = o.coupling_exchange(src.copy())
    for f in range(6):
        assert np.allclose(xs[f].sum(axis=(1, 2)), src[f].sum(axis=(0, 1)), rtol=1e-12, atol=1e-12)
    for (ix, iy), (jx, jy) in {(g.nnx + 1, g.nny + 1): (1, 1), (g.nnx + 1, 3): (1, 3), (4, g.nny + 1): (4, 1),
                               (5, 6): (5, 6), (1, 1): (1, 1)}.items():
        d = np.zeros(shape)
        d[2, ix - 1, iy - 1, 7] = 1.5
        x = o.coupling_exchange(d)
        assert x[2, 7, jy - 1, jx - 1] == 1.5 and np.count_nonzero(x) == 1


def test_apply_coupling_adds_dt_times_the_sources():
    """apply_particle_coupling: T += dt*(S_T (/ exner) + S_TE), q += dt*S_H at
    levels 1..nnz of the core columns; level 0, nnz+1 and the two x-pencil
    padding columns are untouched."""
    g = small_grid("stretched")
    c = Consts(dt=2e-3)
    rng = np.random.default_rng(3)
    xs = rng.standard_normal((6, g.nnz + 2, g.nny, g.nnx))
    for iexner in (0, 1):
        o = ora_for(g, c, Flags(iexner=iexner))
        T0 = 300 + rng.standard_normal((g.nnz + 2, g.nny, g.nnx + 2))
        q0 = 0.01 + 1e-3 * rng.standard_normal((g.nnz + 2, g.nny, g.nnx + 2))
        T, q = T0.copy(), q0.copy()
        o.apply_coupling(xs, T, q)
        core = (slice(1, g.nnz + 1), slice(None), slice(0, g.nnx))
        ex = np.array([o.exner(c.surf_p, o.p_base(c.surf_p, c.tsfcc1, z)) for z in g.zz[1:g.nnz + 1]])
        sT = xs[3][core] / (ex[:, None, None] if iexner else 1.0)
        assert np.allclose(T[core] - T0[core], c.dt * (sT + xs[5][core]), rtol=1e-9, atol=1e-12)
        assert np.allclose(q[core] - q0[core], c.dt * xs[4][core], rtol=1e-9, atol=1e-15)
        for a, a0 in ((T, T0), (q, q0)):
            assert np.array_equal(a[0], a0[0]) and np.array_equal(a[-1], a0[-1])
            assert np.array_equal(a[:, :, g.nnx:], a0[:, :, g.nnx:])


# ------------------------------------------------------------ statistics

def stats_population(g, n=3000, seed=6):
    rng = np.random.default_rng(seed)
    p = Particles.empty(n)
    p["x"], p["y"], p["z"] = rng.uniform(0, g.xl, n), rng.uniform(0, g.yl, n), rng.uniform(0, g.zl, n)
    for c in ("vx", "vy", "vz", "ufx", "ufy", "ufz"):
        p[c] = 0.1 * rng.standard_normal(n)
    p["Tp"], p["Tf"] = 300 + rng.standard_normal(n), 300 + rng.standard_normal(n)
    p["radius"] = np.exp(rng.uniform(np.log(1e-6), np.log(2e-5), n))
    p["rc"] = 1.5e-6
    p["m_s"], p["qinf"], p["qstar"] = 1e-18, 0.02 + 1e-3 * rng.standard_normal(n), 0.02
    p.mult[:] = rng.integers(1, 100, n)
    return p


def test_tauc_min_single_cell_closed_form():
    """calc_tauc_min with every droplet in one w cell: 5/(2 pi (nuf/Sc) 2 r
    N mult/(dx dy dzw) + 1e-10)."""
    g = small_grid("stretched")
    c = Consts()
    o = ora_for(g, c)
    n, r, m = 7, 4e-6, 30
    p = Particles.empty(n)
    p["x"], p["y"], p["z"], p["radius"] = 2.5 * g.dx, 3.5 * g.dy, 0.5 * (g.z[4] + g.z[5]), r
    p.mult[:] = m
    Nc = n * m / (g.dx * g.dy * g.dzw[5])
    want = 5.0 / (2 * np.pi * (c.nuf / c.Sc) * 2 * r * Nc + 1e-10)
    assert abs(o.calc_tauc_min(p.soa, p.mult) / want - 1) < 1e-13
    assert o.calc_tauc_min(p.soa[:, :0], p.mult[:0]) == 5.0e10


def test_statistics_match_numpy_bincounts():
    """calc_tauc_min and particle_xy_stats against numpy: cells from
    searchsorted on z, sums by np.add.at, starting values of the maxima and
    minima as the reference sets them."""
    g = small_grid("stretched")
    c = Consts()
    o = ora_for(g, c)
    p = stats_population(g)
    ix, iy = np.floor(p["x"] / g.dx).astype(int), np.floor(p["y"] / g.dy).astype(int)
    kc = np.searchsorted(g.z, p["z"], side="right")
    ok = (ix < g.nnx) & (iy < g.nny) & (kc >= 1) & (kc <= g.nnz)
    cell = (ix + g.nnx * (iy + g.nny * (kc - 1)))[ok]
    nc = g.nnx * g.nny * g.nnz
    num = np.bincount(cell, minlength=nc)
    rad = np.bincount(cell, weights=p["radius"][ok], minlength=nc)
    mult = np.bincount(cell, weights=p.mult[ok].astype(float), minlength=nc)
    occ = num > 0
    dzw = g.dzw[1:g.nnz + 1][np.arange(nc) // (g.nnx * g.nny)]
    denom = c.pi2 * (c.nuf / c.Sc) * 2 * (rad[occ] / num[occ]) * (mult[occ] / (g.dx * g.dy * dzw[occ]))
    assert abs(o.calc_tauc_min(p.soa, p.mult) / (5 * np.min(1 / (denom + 1e-10))) - 1) < 1e-12
    s0, s1, lwp = o.xy_stats(p.soa, p.mult)
    r = p["radius"]
    volp = 4.0 / 3.0 * np.pi * r ** 3
    rep = 2 * r * np.sqrt((p["vx"] - p["ufx"]) ** 2 + (p["vy"] - p["ufy"]) ** 2 + (p["vz"] - p["ufz"]) ** 2) / c.nuf
    want0 = [p.n, np.sum(r > p["rc"]), np.sum(r <= p["rc"]), rep.sum(), (volp * c.rhow).sum(),
             (volp * c.rhow + p["m_s"]).sum(), volp.sum(), r.sum(), (r * r).sum(),
             max(r.max(), 0), max(p["Tp"].max(), 0), max(p["qinf"].max(), 0), max(p["vx"].max(), -1000),
             max(p["vy"].max(), -1000), max(p["vz"].max(), -1000), min(r.min(), 1000), min(p["Tp"].min(), 1000),
             min(p["qinf"].min(), 1000), min(p["vx"].min(), 1000), min(p["vy"].min(), 1000), min(p["vz"].min(), 1000)]
    assert np.allclose(s0, want0, rtol=1e-12, atol=0)
    m = p.mult.astype(float)
    qs = [np.ones(p.n), p["vx"], p["vy"], p["vz"], p["vx"] ** 2, p["vy"] ** 2, p["vz"] ** 2, p["vx"] * p["vz"],
          p["ufx"], p["ufy"], p["ufz"], p["ufx"] ** 2, p["ufy"] ** 2, p["ufz"] ** 2, p["Tp"], p["Tp"] ** 2, p["Tf"],
          r, r * r, p["qinf"], p["Tp"] * p["vz"], m, m * c.rhow * volp, p["qstar"]]
    for k, q in enumerate(qs):
        assert np.allclose(s1[k], np.bincount(kc, weights=q, minlength=g.nnz + 2), rtol=1e-12, atol=1e-300), k
    inxy = (ix < g.nnx) & (iy < g.nny)
    want_lwp = np.bincount((ix + g.nnx * iy)[inxy], weights=(m * c.rhow * volp)[inxy], minlength=g.nnx * g.nny)
    assert np.allclose(lwp.ravel(), want_lwp, rtol=1e-12, atol=0)


# --------------------------------------------------------------- parallel

def test_parallel_oracle_matches_sequential():
    """The OpenMP oracle the bench times as its all-core CPU baseline computes
    what the sequential oracle does: the same state bit for bit, the same
    counters, and the deposit to rounding (its atomic additions run in any
    order)."""
    grid, consts, flags, fields, p, ora = small_case(seed=3, n=5000)
    a, b = p.copy(), p.copy()
    assert ora.update_be(a.soa, 2) == ora.update_be_par(b.soa, 2)
    assert np.array_equal(a.soa, b.soa)
    sa, sb = ora.coupling(a.soa, a.mult), ora.coupling_par(b.soa, b.mult)
    assert np.max(np.abs(sa - sb)) <= 1e-14 * np.max(np.abs(sa))
