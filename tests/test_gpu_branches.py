"""GPU parity of the branches and outputs of the droplet step that the plain
parity cases do not reach, through the C ABI against the oracle:

- the Newton bookkeeping num100 / numimpos / actnum / denum (particles.f90:
  3127-3174), exact, with the Levenberg-Marquardt fallback and the guard that
  resets a failed solve to (r, Tf) forced;
- qinf <= 0 (the temperature-only branch, 3110, 3183-3189);
- the iexner = 1 advance (Tf * exner, rho_a from the base pressure,
  3064-3072);
- RK3 in mixed precision, the 6th-order gather through the fused step, droplets
  reflected at the top and bottom walls;
- the flux statistics (3208-3238);
- the stated error of every mixed-precision output against the FP64 oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from ntlp_amd import COL, MODE_BE, MODE_RK3, OPT_STAGED_GATHER, PREC_FP64, PREC_MIXED, DropletStep, Flags

from .helpers import rel_err, small_case
from .test_gpu_parity import TOL64, compare, make_step

pytestmark = pytest.mark.gpu

COUNTERS = ("num100", "numimpos", "actnum", "denum", "solves")
STATE = ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "qstar", "res", "actres", "numact", "ufx", "Tf", "qinf"]


def oracle_steps(ora, want, its):
    tot = {k: 0 for k in COUNTERS}
    for it in its:
        c = ora.update_be(want.soa, it)
        for k in COUNTERS:
            tot[k] += c[k]
    return tot


def device_counts(s):
    c = s.counters()
    return {k: c[k] for k in COUNTERS}


def counts_close(got, want, rel=0.01, slack=3):
    """Counters of a case full of ill-conditioned solves: the forward-difference
    Jacobian (step 1e-12) turns last-ulp differences of exp/log between the
    device and the host into 1e-4 differences of J, which flip a few of the
    Gauss-Newton failure decisions (a component below 0.01, 50 iterations) and
    Levenberg-Marquardt's stops: each counter within 1 % or 3 of the oracle's,
    the number of solves exact."""
    assert got["solves"] == want["solves"], (got, want)
    for k in ("num100", "numimpos", "actnum", "denum"):
        assert abs(got[k] - want[k]) <= max(slack, rel * want[k]), (k, got, want)


def borderline(want, before, tol=1e-7):
    """Droplets whose new radius sits within tol of their critical radius: the
    only ones whose activation count could flip with last-ulp differences."""
    return np.abs(want["radius"] / want["rc"] - 1.0) < tol


def ulp_spread(ora, before, idx, it=2, m=24):
    """(lo, hi) of the radius and Tp that the oracle's BE step gives for each
    droplet of idx when its radius or temperature is moved by 1..m ulps either
    way: how far last-ulp differences (like those of exp/log between the device
    and the host) can move an ill-conditioned solve.  Arrays (2, len(idx))."""
    k = np.arange(1, m + 1, dtype=float) * 2.0 ** -52
    scale = np.concatenate([1 + k, 1 - k])
    nv = 2 * scale.size
    idx = np.asarray(idx)
    q = before.copy()
    q.soa = np.repeat(before.soa[:, idx], nv, axis=1)
    q.mult = np.repeat(before.mult[idx], nv)
    sc = np.tile(scale, idx.size).reshape(idx.size, scale.size)
    rad = q["radius"].reshape(idx.size, nv)
    tp = q["Tp"].reshape(idx.size, nv)
    rad[:, : scale.size] *= sc
    tp[:, scale.size:] *= sc
    q["radius"] = rad.ravel()
    q["Tp"] = tp.ravel()
    ora.update_be(q.soa, it)
    r = q["radius"].reshape(idx.size, nv)
    t = q["Tp"].reshape(idx.size, nv)
    return np.stack([r.min(1), t.min(1)]), np.stack([r.max(1), t.max(1)])


def ulp_outcomes(ora, before, i, it=2, m=24):
    """The reset outcomes (radius kept and Tp = Tf) that the oracle gives for
    droplet i of the pre-step state when its radius or temperature is moved
    by 1..m ulps either way: the spread of outcomes of an ill-conditioned
    solve under last-ulp differences like those between the device's and the
    host's exp/log."""
    k = np.arange(1, m + 1, dtype=float) * 2.0 ** -52
    scale = np.concatenate([1 + k, 1 - k])
    nv = 2 * scale.size
    q = before.copy()
    q.soa = np.repeat(before.soa[:, i:i + 1], nv, axis=1)
    q.mult = np.repeat(before.mult[i:i + 1], nv)
    q["radius"][: scale.size] *= scale
    q["Tp"][scale.size:] *= scale
    r0 = q["radius"].copy()
    ora.update_be(q.soa, it)
    return set(((q["radius"] == r0) & (np.abs(q["Tp"] - q["Tf"]) < 1e-10)).tolist())


def test_newton_counters_lm_fallback_and_guard_match_oracle():
    """A large step (dt = 0.05 s) in RH 80-120 %: Gauss-Newton fails for a fifth
    of the droplets, which fall back to Levenberg-Marquardt (num100); a few fail
    there too and are reset to r, Tp = Tf (numimpos); droplets activate and
    deactivate.  Three more droplets of r > 1 cm take the guard
    r*rt > 1e-2 (particles.f90:3147).  The counters equal the oracle's, and
    the reset droplets are the same ones."""
    grid, consts, flags, fields, p, ora = small_case(seed=3, n=3000, dt=0.05, rh=(80, 120))
    p["radius"][:3] = [1.0001e-2, 1.2e-2, 1.05e-2]
    p["z"][:3] = 0.5 * grid.zl
    p["rc"][:3] = ora.crit_radius(p.soa, COL)[:3]
    want = p.copy()
    cnt_o = oracle_steps(ora, want, [2])
    assert cnt_o["num100"] > 300 and cnt_o["numimpos"] >= 3 and cnt_o["actnum"] > 0 and cnt_o["denum"] > 0
    assert not borderline(want, p).any()
    s = make_step(grid, consts, flags, fields, p)
    s.counters(reset=True)
    s.particle_update_BE(it=2, precision=PREC_FP64)
    got = s.get()
    cnt = device_counts(s)
    counts_close(cnt, cnt_o)
    # a reset keeps the radius (rt = 1) and sets Tp = (Tf/Tp)*Tp; the guard of
    # the three r > 1 cm droplets is deterministic
    reset_o = (want["radius"] == p["radius"]) & (np.abs(want["Tp"] - want["Tf"]) < 1e-10)
    reset_d = (got["radius"] == p["radius"]) & (np.abs(got["Tp"] - got["Tf"]) < 1e-10)
    assert reset_o[:3].all() and reset_d[:3].all() and reset_o.sum() == cnt_o["numimpos"]
    # a droplet reset on one side only is one whose Levenberg-Marquardt outcome
    # the oracle itself changes under a few ulps of input: the device's outcome
    # is among the oracle's
    flipped = np.flatnonzero(reset_d != reset_o)
    assert flipped.size <= max(4, 0.02 * cnt_o["num100"]), flipped
    for i in flipped:
        assert bool(reset_d[i]) in ulp_outcomes(ora, p, i), (i, bool(reset_d[i]), bool(reset_o[i]))
    # droplets whose Gauss-Newton solve converges agree to rounding: those that
    # differ are no more than the oracle's Levenberg-Marquardt fallbacks, and
    # each differs by no more than the oracle's own result moves under a few
    # ulps of input
    same = (np.abs(got["radius"] / want["radius"] - 1) < 1e-9) & (np.abs(got["Tp"] - want["Tp"]) < 3e-9)
    diff = np.flatnonzero(~same)
    assert diff.size <= cnt_o["num100"], (diff.size, cnt_o["num100"])
    lo, hi = ulp_spread(ora, p, diff)
    tol_r = 2 * (hi[0] - lo[0]) + 1e-9 * want["radius"][diff]
    tol_t = 2 * (hi[1] - lo[1]) + 3e-9
    dr, dt_ = np.abs(got["radius"][diff] - want["radius"][diff]), np.abs(got["Tp"][diff] - want["Tp"][diff])
    bad = diff[(dr > tol_r) | (dt_ > tol_t)]
    assert bad.size == 0, (bad[:10], dr.max(), dt_.max())
    s.close()


def test_qinf_not_positive_takes_the_temperature_only_branch():
    """Where the gathered humidity is zero or negative, particle_update_BE
    skips the radius solve and relaxes Tp implicitly (particles.f90:3183-3189):
    radius unchanged, no solve counted."""
    grid, consts, flags, fields, p, ora = small_case(seed=4, n=3000)
    u, v, w, T, q = fields
    q = q.copy()
    q[: grid.nnx // 2] = -1e-4  # x < xl/2 (and the columns it gathers from) dry
    q[grid.nnx // 2 - 1] = 0.0
    fields = (u, v, w, T, q)
    ora.set_fields(*fields)
    ora.interp(p.soa)
    want = p.copy()
    cnt_o = oracle_steps(ora, want, [2])
    dry = want["qinf"] <= 0
    assert dry.sum() > 500 and (~dry).sum() > 500
    assert np.array_equal(want["radius"][dry], p["radius"][dry]) and cnt_o["solves"] == (~dry).sum()
    s = make_step(grid, consts, flags, fields, p)
    s.counters(reset=True)
    s.particle_update_BE(it=2, precision=PREC_FP64)
    got = s.get()
    # the droplets beside the dry region see strongly subsaturated air and many
    # fall back to Levenberg-Marquardt (counters: see counts_close)
    counts_close(device_counts(s), cnt_o)
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "ufx", "Tf", "qinf", "res"])
    assert not bad, bad
    assert np.array_equal(got["radius"][dry], p["radius"][dry])
    # the temperature-only branch (no solve) matches to rounding
    assert rel_err(got["Tp"][dry], want["Tp"][dry], 300.0) <= 1e-11
    assert rel_err(got["qstar"][dry], want["qstar"][dry], 0.02) <= 1e-9
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_iexner_advance_matches_oracle(prec):
    """iexner = 1: Tf = Tf * exner(p_s, p_base(z)) and rho_a = p_base(z)/(Rd Tf)
    in the advance and the solve (particles.f90:3064-3072)."""
    grid, consts, flags, fields, p, ora = small_case(seed=5, n=3000, flags=Flags(iexner=1))
    want = p.copy()
    cnt_o = oracle_steps(ora, want, [2, 3])
    s = make_step(grid, consts, flags, fields, p)
    s.counters(reset=True)
    s.step(MODE_BE, prec, nsteps=2, it0=2)
    got = s.get()
    src = s.sources()
    if prec == PREC_FP64:
        assert device_counts(s) == cnt_o
        bad = compare(got, want, STATE)
        assert not bad, bad
        src_o = ora.coupling(want.soa, want.mult)
        for f in range(6):
            scale = np.max(np.abs(src_o[f])) + 1e-300
            assert np.max(np.abs(src[f] - src_o[f])) <= 1e-6 * scale, f
    else:
        assert rel_err(got["radius"], want["radius"], 1e-7) <= 1e-5
        assert rel_err(got["Tp"], want["Tp"], 300.0) <= 1e-6
    # Tf carries the exner factor (FP64 gather: to rounding; mixed: FP32 weights
    # on differences of T within a cell)
    assert rel_err(got["Tf"], want["Tf"], 300.0) <= (1e-12 if prec == PREC_FP64 else 1e-9)
    s.close()


def test_rk3_mixed_precision_within_stated_error():
    """Mixed-precision RK3 (FP32 drag, Reynolds/Nusselt/Sherwood numbers, q*
    and explicit right-hand sides; FP64 state) over two steps against the FP64
    oracle, with the bounds stated in the precision design: radius 1e-6
    relative, Tp 1e-7 relative, positions 1e-12 relative, velocities 1e-5
    relative to 0.1 m/s."""
    grid, consts, flags, fields, p, ora = small_case(seed=7, n=3000, dt=2e-5)
    want = p.copy()
    for it in (2, 3):
        for st in (1, 2, 3):
            ora.update_rk3(want.soa, st, it)
    s = make_step(grid, consts, flags, fields, p)
    s.step(MODE_RK3, PREC_MIXED, nsteps=2, it0=2)
    got = s.get()
    errs = {c: rel_err(got[c], want[c], f) for c, f in
            (("radius", 1e-6), ("Tp", 300.0), ("x", 1.0), ("z", 1.0), ("vx", 0.1), ("vz", 0.1))}
    print("mixed RK3 max relative errors:", errs)
    assert errs["radius"] <= 1e-6 and errs["Tp"] <= 1e-7, errs
    assert errs["x"] <= 1e-12 and errs["z"] <= 1e-12, errs
    assert errs["vx"] <= 1e-5 and errs["vz"] <= 1e-5, errs
    # the single-stage entry points in mixed precision give the same state
    s.put(p)
    for it in (2, 3):
        for st in (1, 2, 3):
            s.particle_update_rk3(st, it=it, precision=PREC_MIXED)
    again = s.get()
    for c in ("radius", "Tp", "x", "z", "vx", "vz"):
        assert rel_err(again[c], got[c], 1e-6 if c == "radius" else 0.1) <= 1e-13, c
    s.close()


@pytest.mark.parametrize("mode", [MODE_BE, MODE_RK3])
def test_sixth_order_gather_through_the_fused_step(mode):
    """ilin = 0 (uf_interp, particles.f90:554-765) inside ntlp_gpu_step."""
    dt = 5e-4 if mode == MODE_BE else 2e-5
    grid, consts, flags, fields, p, ora = small_case(seed=17, n=2000, dt=dt, flags=Flags(ilin=0), kind="stretched")
    want = p.copy()
    for it in (2, 3):
        if mode == MODE_BE:
            ora.update_be(want.soa, it)
        else:
            for st in (1, 2, 3):
                ora.update_rk3(want.soa, st, it)
    s = make_step(grid, consts, flags, fields, p)
    s.step(mode, PREC_FP64, nsteps=2, it0=2)
    got = s.get()
    tol = dict(TOL64, radius=(1e-8, 1e-6))
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "ufx", "ufy", "ufz", "Tf", "qinf"],
                  tol)
    assert not bad, bad
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_droplets_reflected_at_top_and_bottom_walls(prec):
    """Droplets placed just inside the walls and moving out of the domain are
    mirrored about z(nnz) - r and r with vz negated (particle_bcs_nonperiodic,
    2718-2723), in the single-step and the fused paths; FP64 equals the oracle
    to rounding, mixed precision within its stated error."""
    grid, consts, flags, fields, p, ora = small_case(seed=18, n=2000)
    k = 300
    r = p["radius"]
    p["z"][:k] = grid.z[grid.nnz] - r[:k] - 1e-5
    p["vz"][:k] = 0.2 + np.linspace(0, 0.2, k)
    p["z"][k:2 * k] = r[k:2 * k] + 1e-5
    p["vz"][k:2 * k] = -0.2 - np.linspace(0, 0.2, k)
    want = p.copy()
    ora.update_be(want.soa, 2)
    # every one of them crossed its wall (z + dt vz_old beyond it) and was mirrored
    top = grid.z[grid.nnz] - want["radius"][:k]
    assert np.all(p["z"][:k] + consts.dt * p["vz"][:k] > top) and np.all(want["z"][:k] < top)
    bot = want["radius"][k:2 * k]
    assert np.all(p["z"][k:2 * k] + consts.dt * p["vz"][k:2 * k] < bot) and np.all(want["z"][k:2 * k] > bot)
    s = make_step(grid, consts, flags, fields, p)
    for fused in (False, True):
        s.put(p)
        if fused:
            s.step(MODE_BE, prec, nsteps=1, it0=2)
        else:
            s.particle_update_BE(it=2, precision=prec)
        got = s.get()
        if prec == PREC_FP64:
            bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "radius", "Tp"])
            assert not bad, (fused, bad)
        else:
            assert rel_err(got["z"], want["z"], 1.0) <= 1e-9
            assert rel_err(got["vz"], want["vz"], 0.1) <= 1e-6
        assert np.all(got["z"][:k] < top) and np.all(got["z"][k:2 * k] > bot)
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_flux_statistics_match_oracle(prec):
    """pflux (exact: sums of integer multiplicities), pmassflux and penegflux
    (1e-12 of the largest level, summation order) of the last step of a call
    against the oracle's last step; fast droplets cross several levels and
    leave through the walls."""
    grid, consts, flags, fields, p, ora = small_case(seed=19, n=4000, kind="stretched")
    rng = np.random.default_rng(19)
    p["vz"] = rng.normal(0.0, 3.0, p.n)
    want = p.copy()
    ora.update_be(want.soa, 2)
    before_last = want.copy()
    cnt, flux_o = ora.update_be_flux(want.soa, want.mult, 3)
    s = make_step(grid, consts, flags, fields, p)
    s.step(MODE_BE, prec, nsteps=2, it0=2)
    flux = s.be_flux()
    assert np.abs(flux_o[0]).sum() > 100
    if prec == PREC_FP64:
        assert np.array_equal(flux[0], flux_o[0])
        # the masses carry the radius' agreement with the oracle (1e-9 relative,
        # tripled by r^3)
        for k in (1, 2):
            assert np.max(np.abs(flux[k] - flux_o[k])) <= 1e-8 * np.abs(flux_o[k]).max(), k
    else:
        # mixed positions differ from FP64 by ~1e-9 m: a droplet within that of a
        # level may cross it in one precision only
        assert np.abs(flux[0] - flux_o[0]).sum() <= 2 * before_last.mult.max()
    # a single-step call leaves the same statistics as the fused call's last step
    s.put(before_last)
    s.particle_update_BE(it=3, precision=prec)
    one = s.be_flux()
    if prec == PREC_FP64:
        assert np.array_equal(one[0], flux_o[0])
    s.close()


def test_mixed_precision_stated_errors():
    """Every floating-point output of the mixed step against the FP64 oracle
    after ten BE steps, with the bounds stated in the precision design:
    radius 1e-5, Tp 1e-6, positions 1e-9 (relative to 1 m), velocities 1e-5
    (relative to 0.1 m/s), qstar 1e-4, gathered u 1e-5 (FP32 fields), Tf 1e-9
    and qinf 1e-8 (FP32 weights on FP64 differences), and the coupling sources
    (rates of change: differences of states that carry the bounds above,
    divided by dt) 1e-2 of each field's largest value."""
    grid, consts, flags, fields, p, ora = small_case(seed=11, n=4000, kind="stretched")
    want = p.copy()
    for it in range(2, 12):
        ora.update_be(want.soa, it)
    s = make_step(grid, consts, flags, fields, p)
    s.step(MODE_BE, PREC_MIXED, nsteps=10, it0=2)
    got = s.get()
    src = s.sources()
    src_o = ora.coupling(want.soa, want.mult)
    bounds = {"radius": (1e-5, 1e-7), "Tp": (1e-6, 300.0), "x": (1e-9, 1.0), "y": (1e-9, 1.0), "z": (1e-9, 1.0),
              "vx": (1e-5, 0.1), "vy": (1e-5, 0.1), "vz": (1e-5, 0.1), "qstar": (1e-4, 0.02),
              "ufx": (1e-5, 0.1), "ufz": (1e-5, 0.1), "Tf": (1e-9, 300.0), "qinf": (1e-8, 0.02)}
    errs = {c: rel_err(got[c], want[c], f) for c, (b, f) in bounds.items()}
    src_errs = [float(np.max(np.abs(src[f] - src_o[f])) / (np.max(np.abs(src_o[f])) + 1e-300)) for f in range(6)]
    print("mixed BE max relative errors:", errs, "sources:", src_errs)
    bad = {c: e for c, e in errs.items() if e > bounds[c][0]}
    assert not bad, bad
    assert max(src_errs) <= 1e-2, src_errs
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_staged_gather_gives_the_same_bits(prec):
    """NTLP_OPT_STAGED_GATHER: the fused step's trilinear gathers from the
    workgroup's field box in LDS give the same bits as the gathers straight
    from HBM -- BE over 20 steps (a reordering inside, droplets crossing cells
    and walls, the coupling sources) and RK3 (three gathers per step, at
    positions that move between the stages).  The coupling sources are sums
    in LDS atomics, whose order varies from run to run: they agree to 1e-12
    of each field's largest value."""
    grid, consts, flags, fields, p, ora = small_case(seed=31, n=20000, kind="stretched")
    res = {}
    for staged in (0, 1):
        s = make_step(grid, consts, flags, fields, p)
        s.set_option(OPT_STAGED_GATHER, staged)
        s.step(MODE_BE, prec, nsteps=20, it0=2)
        be, src = s.get(), s.sources()
        s.put(p)
        s.step(MODE_RK3, prec, nsteps=3, it0=2)
        res[staged] = (be.soa, src, s.get().soa)
        s.close()
    assert np.array_equal(res[0][0], res[1][0], equal_nan=True)
    assert np.array_equal(res[0][2], res[1][2], equal_nan=True)
    for f in range(6):
        scale = np.abs(res[0][1][f]).max() + 1e-300
        assert np.abs(res[0][1][f] - res[1][1][f]).max() <= 1e-12 * scale, f


@pytest.mark.gpu
def test_mixed_step_fp32_amounts_match_fp64_amounts_of_the_same_state():
    """The coupling amounts a mixed-precision step forms in FP32 (after FP64
    differences over the step) against the reference-order FP64 amounts of the
    same droplet states: the sources of three mixed BE steps, then the same
    record put back and deposited again, which forms the amounts in FP64 from
    the record's old values (particles.f90:1187-1264).  Only the amount
    arithmetic and the summation order differ, so the sources agree to FP32
    rounding of each droplet's amount: within 1e-5 of each field's largest
    value.  The smallest term of the momentum amount, the vapour's momentum
    vap*v, is 2e-4 of the drag term on average over these droplets (up to
    3e-2), so dropping it would fail the check."""
    grid, consts, flags, fields, p, ora = small_case(seed=12, n=6000, kind="stretched")
    s = make_step(grid, consts, flags, fields, p)
    s.step(MODE_BE, PREC_MIXED, nsteps=3, it0=2)
    src32 = s.sources().copy()
    rec = s.get()
    s.put(rec)
    src64 = s.particle_coupling_update(deterministic=0)
    errs = [float(np.max(np.abs(src32[f] - src64[f])) / (np.max(np.abs(src64[f])) + 1e-300)) for f in range(6)]
    print("FP32 against FP64 amounts, max error per field / field max:", errs)
    assert all(np.max(np.abs(src64[f])) > 0 for f in range(6))
    assert max(errs) <= 1e-5, errs
    s.close()
