"""The continuous-batching serving loop (ntlp_serve_*).

Every served result is compared bit for bit with the plain path on the same
droplets (put + interp(vp=uf, Tp=Tf) + crit_radius + ntlp_gpu_step), alone
and in a mix of concurrent requests of uneven size, with captured graphs and
with direct launches."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from ntlp_amd import (COL, MODE_BE, MODE_RK3, PREC_FP64, PREC_MIXED, SERVE_BYTES_PER_DROPLET, SERVE_IN, SERVE_OUT,
                      DropletStep, Flags, NtlpError, Particles, ServeLoop, synth)

from .helpers import small_grid

pytestmark = pytest.mark.gpu

STEPS = 10


def make_request(grid, seed: int, n: int, r_med: float, a_med: float):
    """A droplet ensemble: uniform positions, lognormal radius around r_med,
    dry aerosol lognormal around a_med (m_s from its volume), kappa 1.2."""
    rng = np.random.default_rng(seed)
    inp = np.empty((6, n))
    inp[0] = rng.uniform(0, grid.xl, n)
    inp[1] = rng.uniform(0, grid.yl, n)
    inp[2] = rng.uniform(0.05, 0.95, n) * grid.zl
    inp[3] = synth.lognormal(rng, n, r_med, 0.3)
    inp[4] = synth.dry_mass(synth.lognormal(rng, n, a_med, 0.5))
    inp[5] = 1.2
    mult = rng.integers(1, 1000, n).astype(np.int64)
    return np.ascontiguousarray(inp), mult


def plain_result(s: DropletStep, inp, mult, prec, mode=MODE_BE):
    """The same droplets through the plain entry points."""
    n = inp.shape[1]
    p = Particles.empty(n)
    for k, name in enumerate(SERVE_IN):
        p[name] = inp[k]
    p.mult[:] = mult
    s.put(p)
    s.interp(set_vp=True, set_tp=True)
    s.crit_radius()
    s.step(mode, prec, nsteps=STEPS, it0=2)
    q = s.get()
    return np.stack([q[c] for c in SERVE_OUT]), q.mult


def serve_case(kind="stretched", seed=3, dt=5e-4):
    grid = small_grid(kind, nnx=24, nny=20, nnz=24)
    from ntlp_amd import Consts
    consts = Consts(dt=dt, radius_init=5e-6)
    fields = synth.make_fields(grid, consts, np.random.default_rng(seed), T0=300.0, T_sd=0.1, rh_uniform=(99.7, 100.4))
    s = DropletStep(grid, consts, Flags())
    s.set_fields(*fields)
    return grid, s


def run_served(loop: ServeLoop, reqs, concurrent: bool = True):
    outs = []
    tickets = []
    for inp, mult in reqs:
        n = inp.shape[1]
        out, out_mult = np.zeros((14, n)), np.zeros(n, dtype=np.int64)
        outs.append((out, out_mult))
        tickets.append(loop.submit(inp, mult, out, out_mult))
        if not concurrent:
            loop.wait(tickets[-1])
    if concurrent:
        for t in tickets:
            lat, queued = loop.wait(t)
            assert lat >= queued >= 0.0
    return outs


@pytest.mark.parametrize("prec", [PREC_MIXED, PREC_FP64])
def test_served_results_equal_plain_steps_alone_and_in_a_mix(prec):
    grid, s = serve_case()
    sizes = [5000, 37, 1200, 20000, 1, 777]
    reqs = [make_request(grid, 100 + i, n, r, a)
            for i, (n, r, a) in enumerate(zip(sizes, (2e-6, 15e-6, 6e-6, 1.2e-6, 9e-6, 20e-6),
                                              (0.02e-6, 0.2e-6, 0.05e-6, 0.1e-6, 0.03e-6, 0.15e-6)))]
    want = [plain_result(s, inp, mult, prec) for inp, mult in reqs]
    loop = ServeLoop(s, hbm_budget_bytes=64 << 20, steps=STEPS, precision=prec)
    alone = run_served(loop, reqs[:1], concurrent=False)
    mixed = run_served(loop, reqs, concurrent=True)
    tot = loop.close()
    assert np.array_equal(alone[0][0], want[0][0]) and np.array_equal(alone[0][1], want[0][1])
    for (out, om), (w, wm), (inp, mult) in zip(mixed, want, reqs):
        assert np.array_equal(out, w), np.max(np.abs(out - w))
        assert np.array_equal(om, wm) and np.array_equal(om, mult)
    assert tot["completed"] == len(reqs) + 1 and tot["queued"] == 0
    assert tot["newton"]["solves"] == STEPS * (sizes[0] + sum(sizes))
    assert tot["graphs"] >= 1 and tot["graph_launches"] == tot["iterations"]


@pytest.mark.parametrize("prec", [PREC_MIXED, PREC_FP64])
def test_served_rk3_results_equal_plain_rk3_steps_alone_and_in_a_mix(prec):
    """Requests that advance RK3 steps (particle_update_rk3's three stages per
    step, particles.f90:2845-2998): each result equals the plain fused RK3
    path on the same droplets, bit for bit, alone and in a mix.  dt = 2e-6 s
    keeps the explicit stages stable for the smallest droplets of these
    ensembles (dt below the droplet response time down to about 0.3 um)."""
    grid, s = serve_case(dt=2e-6)
    sizes = [4000, 37, 1500, 9000, 1]
    reqs = [make_request(grid, 300 + i, n, r, a)
            for i, (n, r, a) in enumerate(zip(sizes, (2e-6, 15e-6, 6e-6, 1.2e-6, 9e-6),
                                              (0.02e-6, 0.2e-6, 0.05e-6, 0.1e-6, 0.03e-6)))]
    want = [plain_result(s, inp, mult, prec, mode=MODE_RK3) for inp, mult in reqs]
    with pytest.raises(NtlpError):
        ServeLoop(s, hbm_budget_bytes=64 << 20, steps=STEPS, precision=prec, mode=7)
    loop = ServeLoop(s, hbm_budget_bytes=64 << 20, steps=STEPS, precision=prec, mode=MODE_RK3)
    alone = run_served(loop, reqs[:1], concurrent=False)
    mixed = run_served(loop, reqs, concurrent=True)
    tot = loop.close()
    assert np.array_equal(alone[0][0], want[0][0]) and np.array_equal(alone[0][1], want[0][1])
    for (out, om), (w, wm), (inp, mult) in zip(mixed, want, reqs):
        assert np.isfinite(w).all()
        assert np.array_equal(out, w), np.max(np.abs(out - w))
        assert np.array_equal(om, wm) and np.array_equal(om, mult)
    assert tot["completed"] == len(reqs) + 1 and tot["queued"] == 0
    assert tot["newton"]["solves"] == 0  # the RK3 stages have no implicit solve
    s.close()


def test_direct_launches_match_graphs_and_requests_arriving_late_join():
    grid, s = serve_case("uniform", seed=5)
    reqs = [make_request(grid, 200 + i, n, 4e-6, 0.05e-6) for i, n in enumerate((3000, 4096, 10, 2500))]
    results = {}
    for use_graph in (True, False):
        loop = ServeLoop(s, hbm_budget_bytes=32 << 20, steps=STEPS, precision=PREC_MIXED, use_graph=use_graph)
        outs = [(np.zeros((14, inp.shape[1])), np.zeros(inp.shape[1], dtype=np.int64)) for inp, _ in reqs]
        tickets = [loop.submit(reqs[0][0], reqs[0][1], *outs[0])]
        # the others arrive while the first is in the pool
        done = threading.Event()

        def late():
            for (inp, mult), o in zip(reqs[1:], outs[1:]):
                tickets.append(loop.submit(inp, mult, *o))
            done.set()

        th = threading.Thread(target=late)
        th.start()
        done.wait(30)
        th.join()
        for t in tickets:
            loop.wait(t)
        tot = loop.close()
        assert tot["iterations"] >= STEPS
        results[use_graph] = outs
    for (a, am), (b, bm) in zip(results[True], results[False]):
        assert np.array_equal(a, b) and np.array_equal(am, bm)
    want = plain_result(s, *reqs[3], PREC_MIXED)
    assert np.array_equal(results[True][3][0], want[0])


def test_admission_respects_the_hbm_budget():
    grid, s = serve_case("uniform", seed=7)
    budget = 6000 * SERVE_BYTES_PER_DROPLET
    loop = ServeLoop(s, hbm_budget_bytes=budget, steps=STEPS, precision=PREC_MIXED)
    assert loop.capacity == 6000
    with pytest.raises(NtlpError):
        big = make_request(grid, 1, 6001, 5e-6, 0.05e-6)
        loop.submit(big[0], big[1], np.zeros((14, 6001)), np.zeros(6001, dtype=np.int64))
    with pytest.raises(NtlpError):  # the fields are in use
        s.set_fields(*[np.zeros((grid.nnx, grid.nny, grid.nnz + 2))] * 5)
    reqs = [make_request(grid, 300 + i, 2500, 5e-6, 0.05e-6) for i in range(5)]  # 12500 > 6000
    outs = run_served(loop, reqs)
    st = loop.stats()
    assert st["completed"] == 5 and st["queued"] == 0 and st["capacity"] == 6000
    assert st["iterations"] >= 3 * STEPS and st["droplet_steps"] == STEPS * 12500
    tot = loop.close()
    assert tot["max_pool"] <= 6000 and tot["completed"] == 5 and tot["iterations"] == st["iterations"]
    assert tot["iterations"] >= 3 * STEPS  # at most two requests fit at a time
    want = plain_result(s, *reqs[4], PREC_MIXED)
    assert np.array_equal(outs[4][0], want[0])


def test_strided_windows_of_wide_buffers_give_the_same_results():
    """Requests handed over as column windows of wider arrays (pinned here),
    results written into windows of a wide result buffer."""
    from ntlp_amd import PinnedArray
    grid, s = serve_case("uniform", seed=9)
    inp, mult = make_request(grid, 400, 9000, 7e-6, 0.08e-6)
    bank = PinnedArray((6, 12000))
    bank.array[:, 1500:10500] = inp
    res = PinnedArray((14, 30000))
    res.array[:] = 0.0
    res_mult = np.zeros(30000, dtype=np.int64)
    loop = ServeLoop(s, hbm_budget_bytes=16 << 20, steps=STEPS, precision=PREC_MIXED)
    t1 = loop.submit(bank.array[:, 1500:10500], mult, res.array[:, 20000:29000], res_mult[20000:29000])
    out, om = np.zeros((14, 9000)), np.zeros(9000, dtype=np.int64)
    t2 = loop.submit(inp, mult, out, om)
    loop.wait(t1)
    loop.wait(t2)
    loop.close()
    assert np.array_equal(res.array[:, 20000:29000], out) and np.array_equal(res_mult[20000:29000], om)
    assert not res.array[:, :20000].any() and not res.array[:, 29000:].any()
    bank.free()
    res.free()
