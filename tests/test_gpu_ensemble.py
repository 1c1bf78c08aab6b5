"""Droplet ensembles drawn on the device (ntlp_gpu_particles_generate,
ntlp_serve_submit_ensemble) against the host restatement of the draws
(tests/philox.py, pinned by Random123's known answers) and the oracle:
positions bit for bit, radii and salt masses to the last ulps of exp/log/cos,
the creation state (vp = uf, Tp = Tf, rc = crit_radius) as the oracle makes
it; the same bits however the ensemble is split; served ensembles equal the
plain path."""
from __future__ import annotations

import numpy as np
import pytest

from ntlp_amd import (COL, DIST_LOGNORMAL, DIST_LOGUNIFORM, MODE_BE, PREC_MIXED, SALT_CHOICE3, SALT_DRY_LOGNORMAL,
                      SALT_FIXED, SERVE_OUT, Ensemble, ServeLoop)

from .helpers import small_case
from .philox import droplet_draws
from .test_gpu_parity import TOL64, compare, make_step

pytestmark = pytest.mark.gpu

ENSEMBLES = {
    "lognormal_fixed": Ensemble(seed=11, radius_kind=DIST_LOGNORMAL, radius_p=(10e-6, 0.3), salt_kind=SALT_FIXED,
                                salt_p=(1e-18,)),
    "lognormal_dry": Ensemble(seed=12, radius_kind=DIST_LOGNORMAL, radius_p=(8e-6, 0.4),
                              salt_kind=SALT_DRY_LOGNORMAL, salt_p=(0.05e-6, 0.5), mult=17),
    "haze_mixture": Ensemble(seed=13, radius_kind=DIST_LOGUNIFORM, radius_p=(5e-6, 20e-6), radius2_kind=DIST_LOGUNIFORM,
                             radius2_p=(1e-6, 2e-6), frac2=0.6, salt_kind=SALT_FIXED, salt_p=(5e-18,)),
    "trimodal": Ensemble(seed=(1 << 40) + 14, radius_kind=DIST_LOGNORMAL, radius_p=(10e-6, 0.35),
                         salt_kind=SALT_CHOICE3, salt_p=(1e-19, 1e-17, 1e-15), first=123456),
}


def host_ensemble(e: Ensemble, grid, rhos: float, n: int):
    u = droplet_draws(e.seed, e.first, n)

    def radius(kind, p, ua, ub):
        if kind == DIST_LOGUNIFORM:
            return np.exp(np.log(p[0]) + ua * (np.log(p[1]) - np.log(p[0])))
        return p[0] * np.exp(p[1] * np.sqrt(-2.0 * np.log(ua)) * np.cos(6.283185307179586 * ub))

    second = (u[5] < e.frac2) if e.frac2 > 0 else np.zeros(n, dtype=bool)
    r = np.where(second, radius(e.radius2_kind, e.radius2_p, u[3], u[4]), radius(e.radius_kind, e.radius_p, u[3], u[4]))
    x, y = u[0] * grid.xl, u[1] * grid.yl
    z = r + u[2] * (grid.z[grid.nnz] - 2.0 * r)
    if e.salt_kind == SALT_FIXED:
        ms = np.full(n, e.salt_p[0])
    elif e.salt_kind == SALT_CHOICE3:
        ms = np.where(u[6] < 1 / 3, e.salt_p[0], np.where(u[6] < 2 / 3, e.salt_p[1], e.salt_p[2]))
    else:
        rd = e.salt_p[0] * np.exp(e.salt_p[1] * np.sqrt(-2.0 * np.log(u[6])) * np.cos(6.283185307179586 * u[7]))
        ms = rhos * (2.0 / 3.0 * (2.0 * np.pi)) * rd ** 3
    return x, y, z, r, ms, second


@pytest.mark.parametrize("name", list(ENSEMBLES))
def test_generated_ensemble_matches_host_draws_and_oracle(name):
    e = ENSEMBLES[name]
    grid, consts, flags, fields, p, ora = small_case(seed=40, n=10, kind="stretched")
    s = make_step(grid, consts, flags, fields, p)
    n = 40000
    s.generate(e, n)
    got = s.get()
    x, y, z, r, ms, second = host_ensemble(e, grid, consts.rhos, n)
    assert np.array_equal(got["x"], x) and np.array_equal(got["y"], y)
    assert np.max(np.abs(got["radius"] / r - 1)) < 1e-13
    assert np.max(np.abs(got["z"] - z)) < 1e-15
    assert np.max(np.abs(got["m_s"] / ms - 1)) < 1e-13
    assert np.all(got["kappa_s"] == e.kappa_s) and np.all(got.mult == e.mult)
    assert np.array_equal(got.pidx, np.arange(e.first + 1, e.first + n + 1))
    assert np.all(got["z"] >= got["radius"]) and np.all(got["z"] <= grid.z[grid.nnz] - got["radius"])
    # the creation state as the oracle makes it from the device's own positions
    want = got.copy()
    ora.interp(want.soa)
    for d in "xyz":
        want["v" + d] = want["uf" + d]
    want["Tp"] = want["Tf"]
    assert np.array_equal(got["rc"], ora.crit_radius(got.soa, COL))
    bad = compare(got, want, ["vx", "vy", "vz", "ufx", "ufy", "ufz", "Tf", "qinf", "Tp"])
    assert not bad, bad
    for c in ("res", "actres", "numact", "qstar", "radrhs", "vxold", "Tpold"):
        assert not got[c].any(), c
    if e.frac2 > 0:  # the mixture's components
        assert abs(second.mean() - e.frac2) < 0.01
        assert np.all((got["radius"][second] >= 1e-6) & (got["radius"][second] <= 2e-6))
    s.close()


def test_generated_ensemble_is_the_same_however_it_is_split():
    grid, consts, flags, fields, p, ora = small_case(seed=41, n=10)
    s = make_step(grid, consts, flags, fields, p)
    e = ENSEMBLES["lognormal_dry"]
    s.generate(e, 30000)
    whole = s.get()
    s.generate(e, 30000)
    assert np.array_equal(s.get().soa, whole.soa)
    import dataclasses
    s.generate(dataclasses.replace(e, first=e.first + 12345), 5000)
    part = s.get()
    assert np.array_equal(part.soa, whole.soa[:, 12345:17345])
    s.close()


def test_several_ensembles_in_one_call_are_their_concatenation():
    grid, consts, flags, fields, p, ora = small_case(seed=44, n=10)
    s = make_step(grid, consts, flags, fields, p)
    es = [ENSEMBLES["lognormal_fixed"], ENSEMBLES["haze_mixture"], ENSEMBLES["trimodal"]]
    ns = [1000, 1, 2500]
    parts = []
    for e, n in zip(es, ns):
        s.generate(e, n)
        parts.append(s.get())
    s.generate(es, ns)
    both = s.get()
    assert both.n == sum(ns)
    assert np.array_equal(both.soa, np.concatenate([q.soa for q in parts], axis=1))
    assert np.array_equal(both.mult, np.concatenate([q.mult for q in parts]))
    assert np.array_equal(both.pidx, np.concatenate([q.pidx for q in parts]))
    s.close()


def test_lognormal_ensemble_statistics():
    """One million droplets: median and sigma of ln r within five standard
    errors of the requested (10 um, 0.35); salt thirds within 1 %."""
    grid, consts, flags, fields, p, ora = small_case(seed=42, n=10)
    s = make_step(grid, consts, flags, fields, p)
    e = ENSEMBLES["trimodal"]
    n = 1_000_000
    s.generate(e, n)
    lr = np.log(s.column("radius"))
    se = 0.35 / np.sqrt(n)
    assert abs(np.median(lr) - np.log(10e-6)) < 5 * 1.2533 * se
    assert abs(lr.std() - 0.35) < 5 * 0.35 / np.sqrt(2 * n)
    ms = s.column("m_s")
    for m in (1e-19, 1e-17, 1e-15):
        assert abs(np.mean(ms == m) - 1 / 3) < 0.01
    s.close()


def test_served_ensembles_equal_plain_path_alone_and_in_a_mix():
    """Requests drawn on the device at admission: each result equals, bit for
    bit, the same ensemble generated and stepped through the plain entry
    points (generate + ntlp_gpu_step), alone and among uploaded requests."""
    from .test_serve import STEPS, make_request, run_served, serve_case

    grid, s = serve_case("stretched", seed=43)
    ens = [Ensemble(seed=500 + i, radius_kind=DIST_LOGNORMAL, radius_p=(rm, 0.4), salt_kind=SALT_DRY_LOGNORMAL,
                    salt_p=(am, 0.5), mult=3 + i)
           for i, (rm, am) in enumerate(((2e-6, 0.02e-6), (12e-6, 0.2e-6), (5e-6, 0.07e-6)))]
    sizes = [6000, 17, 25000]
    want = []
    for e, n in zip(ens, sizes):
        s.generate(e, n)
        s.step(MODE_BE, PREC_MIXED, nsteps=STEPS, it0=2)
        q = s.get()
        want.append((np.stack([q[c] for c in SERVE_OUT]), q.mult))
    loop = ServeLoop(s, hbm_budget_bytes=64 << 20, steps=STEPS, precision=PREC_MIXED)
    outs = [(np.zeros((14, n)), np.zeros(n, dtype=np.int64)) for n in sizes]
    t0 = loop.submit_ensemble(ens[0], sizes[0], *outs[0])
    loop.wait(t0)
    alone = (outs[0][0].copy(), outs[0][1].copy())
    mixed = [(np.zeros((14, n)), np.zeros(n, dtype=np.int64)) for n in sizes]
    uploaded = make_request(grid, 77, 3000, 7e-6, 0.05e-6)
    tickets = [loop.submit_ensemble(e, n, *o) for e, n, o in zip(ens, sizes, mixed)]
    up_out = (np.zeros((14, 3000)), np.zeros(3000, dtype=np.int64))
    tickets.append(loop.submit(uploaded[0], uploaded[1], *up_out))
    for t in tickets:
        loop.wait(t)
    tot = loop.close()
    assert np.array_equal(alone[0], want[0][0]) and np.array_equal(alone[1], want[0][1])
    for (o, om), (w, wm) in zip(mixed, want):
        assert np.array_equal(o, w), np.max(np.abs(o - w))
        assert np.array_equal(om, wm)
    assert tot["completed"] == 5
