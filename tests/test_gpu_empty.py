"""A population of no droplets: an MPI rank whose particle list is empty,
where the reference's loops (particle_update_BE, particle_update_rk3,
particle_coupling_update, calc_tauc_min, particle_xy_stats) run over nothing.
Every update returns at once; the coupling sources, counters and flux
statistics are zero; tauc_min (its start value 1e10 times the fudge factor 5,
particles.f90:3327, 3363) and the minima and maxima of particle_xy_stats keep
the reference's start values (3407-3418).  Droplets put afterwards step
exactly as in a fresh context, and emptying the population again leaves no
stale sources."""
from __future__ import annotations

import numpy as np
import pytest

from ntlp_amd import MODE_BE, MODE_RK3, PREC_FP64, PREC_MIXED, Particles

from .helpers import small_case
from .test_gpu_parity import make_step

pytestmark = pytest.mark.gpu


def test_updates_of_an_empty_population_are_no_ops():
    grid, consts, flags, fields, p, ora = small_case(seed=5, n=1500)
    s = make_step(grid, consts, flags, fields, Particles.empty(0))
    s.counters(reset=True)
    s.interp(set_vp=True, set_tp=True)
    s.crit_radius()
    s.particle_update_BE(it=2, precision=PREC_MIXED)
    assert not s.be_flux().any()
    for st in (1, 2, 3):
        s.particle_update_rk3(st, it=2)
    for method in (0, 1, 2):
        assert not s.particle_coupling_update(deterministic=method).any()
    s.step(MODE_BE, PREC_MIXED, nsteps=5)
    s.step(MODE_RK3, PREC_FP64, nsteps=2)
    assert not s.sources().any()
    assert not s.be_flux().any()
    c = s.counters()
    assert c["solves"] == 0 and c["num100"] == 0 and c["actnum"] == 0 and c["gn_iters"] == 0
    assert c["steps"] == 8  # one particle_update_BE, five BE and two RK3 fused steps
    e = Particles.empty(0)
    assert s.calc_tauc_min() == 5.0e10 == ora.calc_tauc_min(e.soa, e.mult)  # 1e10 times the fudge factor 5
    # zero sums and counts; the minima and maxima keep the reference's start
    # values, as the oracle's do
    s0, s1, lwp = s.particle_xy_stats()
    w0, w1, wl = ora.xy_stats(e.soa, e.mult)
    assert np.array_equal(s0, w0) and np.array_equal(s1, w1) and np.array_equal(lwp, wl)
    assert not s1.any() and not lwp.any() and not s0[:9].any()
    cellidx, counts, offsets, perm = s.bin()
    assert cellidx.size == 0 and perm.size == 0 and not counts.any() and not offsets.any()
    assert not s.wcell_counts().any()
    assert s.get().n == 0
    s.close()


def test_droplets_after_an_empty_population_step_as_in_a_fresh_context():
    grid, consts, flags, fields, p, ora = small_case(seed=6, n=2500)
    s = make_step(grid, consts, flags, fields, Particles.empty(0))
    s.step(MODE_BE, PREC_FP64, nsteps=2)
    s.put(p)
    s.step(MODE_BE, PREC_FP64, nsteps=4)
    a, src_a = s.get(), s.sources()
    # and emptied again: the next step's sources are zero, not the last ones
    assert src_a.any()
    s.put(Particles.empty(0))
    s.step(MODE_BE, PREC_FP64, nsteps=1)
    assert not s.sources().any()
    s.close()
    t = make_step(grid, consts, flags, fields, p)
    t.step(MODE_BE, PREC_FP64, nsteps=4)
    b, src_b = t.get(), t.sources()
    t.close()
    assert np.array_equal(a.soa, b.soa, equal_nan=True)
    assert np.array_equal(a.mult, b.mult) and np.array_equal(a.pidx, b.pidx)
    # the sources are sums in LDS atomics, whose order varies from run to run
    for f in range(6):
        scale = np.abs(src_b[f]).max() + 1e-300
        assert np.abs(src_a[f] - src_b[f]).max() <= 1e-12 * scale, f
