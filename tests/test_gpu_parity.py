"""Parity of the HIP droplet step with the oracle, through the C ABI.

Bit-exact: cell indices, per-cell counts, the stable sort permutation, the
critical radius table lookup.  FP64 outputs: bounded relative error against
the oracle (the tolerances are per quantity in TOL64 below and absorb the
last-ulp differences of exp/log/pow between the device math library and the
host's).  Mixed precision: radius within 1e-5 relative of the FP64 oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from ntlp_amd import (COL, LAYOUT_CORE, LAYOUT_EXT, LAYOUT_XPENCIL, MODE_BE, MODE_RK3, PREC_FP64,
                      PREC_FP64_AJ, PREC_MIXED, DropletStep, Flags, NtlpError, synth)

from .helpers import edge_positions, rel_err, small_case

pytestmark = pytest.mark.gpu

# relative error bound and magnitude floor per column, FP64 device vs oracle
TOL64 = {
    "x": (1e-13, 1.0), "y": (1e-13, 1.0), "z": (1e-13, 1.0),
    "vx": (1e-10, 0.1), "vy": (1e-10, 0.1), "vz": (1e-10, 0.1),
    "ufx": (1e-13, 0.1), "ufy": (1e-13, 0.1), "ufz": (1e-13, 0.1),
    "Tf": (1e-14, 300.0), "qinf": (1e-13, 0.02),
    "Tp": (1e-11, 300.0), "radius": (1e-9, 1e-6), "qstar": (1e-9, 0.02),
    "res": (0.0, 1.0), "actres": (0.0, 1.0), "numact": (0.0, 1.0), "rc": (0.0, 1e-6),
}


def make_step(grid, consts, flags, fields, p, layout=LAYOUT_CORE):
    s = DropletStep(grid, consts, flags)
    s.set_fields(*fields, layout=layout)
    s.put(p)
    return s


def compare(got, want, cols, tol=TOL64):
    bad = {}
    for c in cols:
        r, floor = tol[c]
        e = rel_err(got[c], want[c], floor)
        if e > r:
            bad[c] = e
    return bad


def test_cell_indices_counts_and_sort_bit_exact():
    for kind in ("uniform", "stretched"):
        grid, consts, flags, fields, p, ora = small_case(seed=1, n=4000, kind=kind)
        edges = edge_positions(grid)
        p.soa[COL["x"], :len(edges)] = edges[:, 0]
        p.soa[COL["y"], :len(edges)] = edges[:, 1]
        p.soa[COL["z"], :len(edges)] = edges[:, 2]
        s = make_step(grid, consts, flags, fields, p)
        cellidx, counts, offsets, perm = s.bin()
        want = ora.cell_index(p["x"], p["y"], p["z"])
        assert np.array_equal(cellidx, want[:, :4]), kind
        keys, c_want, o_want, p_want = ora.bin(p["x"], p["y"], p["z"])
        assert np.array_equal(counts, c_want), kind
        assert np.array_equal(offsets, o_want), kind
        assert np.array_equal(perm, p_want), kind
        assert np.array_equal(s.wcell_counts(), ora.wcell_counts(p["x"], p["y"], p["z"])), kind
        s.close()


@pytest.mark.parametrize("ilin", [1, 0])
def test_interpolation(ilin):
    grid, consts, flags, fields, p, ora = small_case(seed=2, flags=Flags(ilin=ilin), kind="stretched")
    s = make_step(grid, consts, flags, fields, p)
    s.interp()
    got = s.get()
    want = p.copy()
    ora.interp(want.soa)
    bad = compare(got, want, ["ufx", "ufy", "ufz", "Tf", "qinf"])
    assert not bad, bad
    s.close()


def test_field_layouts_agree():
    grid, consts, flags, fields, p, ora = small_case(seed=3, n=500)
    outs = []
    ext = [ora.fill_ext(f) for f in fields]
    # flow-solver x-pencil layout u(1:nnx+2, 1:nny, 0:nnz+1): numpy (nnz+2, nny, nnx+2)
    xp = []
    for f in fields:
        a = np.zeros((grid.nnz + 2, grid.nny, grid.nnx + 2))
        a[:, :, :grid.nnx] = np.transpose(f, (2, 1, 0))
        xp.append(a)
    for layout, data in ((LAYOUT_CORE, fields), (LAYOUT_EXT, ext), (LAYOUT_XPENCIL, xp)):
        s = make_step(grid, consts, flags, data, p, layout=layout)
        s.interp()
        outs.append(s.get())
        s.close()
    for o in outs[1:]:
        assert np.array_equal(o.soa, outs[0].soa)


def test_crit_radius_bit_exact():
    grid, consts, flags, fields, p, ora = small_case(seed=4)
    q = p.copy()
    q["rc"] = 0.0
    s = make_step(grid, consts, flags, fields, q)
    s.crit_radius()
    assert np.array_equal(s.get()["rc"], p["rc"])
    s.close()


def test_crit_radius_bit_exact_across_the_table():
    """The device searches only the table entries around the analytic maximum
    of the Koehler curve: the same radius as the full 1024-entry scan for
    solute from none (m_s = 0) to far past either end of the table, at cold
    and warm Tp."""
    grid, consts, flags, fields, p, ora = small_case(seed=14, n=20000)
    rng = np.random.default_rng(14)
    q = p.copy()
    q["m_s"] = 10.0 ** rng.uniform(-32, -6, q.n)
    q["m_s"][::53] = 0.0
    q["kappa_s"] = rng.choice([0.0, 0.1, 0.6, 1.2, 1.28], q.n)
    q["Tp"] = rng.uniform(230.0, 320.0, q.n)
    q["rc"] = 0.0
    want = ora.crit_radius(q.soa, COL)
    s = make_step(grid, consts, flags, fields, q)
    s.crit_radius()
    got = s.get()["rc"]
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    # both ends of the table are reached
    assert want.min() == 1e-8 and want.max() > 9.8e-4
    s.close()


def test_put_get_and_records_roundtrip():
    grid, consts, flags, fields, p, ora = small_case(seed=5, n=777)
    s = make_step(grid, consts, flags, fields, p)
    got = s.get()
    assert np.array_equal(got.soa, p.soa) and np.array_equal(got.mult, p.mult)
    assert np.array_equal(got.pidx, p.pidx)
    rec = s.get_records()
    s.put_records(rec)
    again = s.get()
    assert np.array_equal(again.soa, p.soa) and np.array_equal(again.mult, p.mult)
    s.close()


def test_restart_file_roundtrip(tmp_path):
    """save_particles writes an int32 count and 352-byte records; read_part_res
    reads them back, keeping only 0 < x < xl, 0 < y < yl like the reference."""
    grid, consts, flags, fields, p, ora = small_case(seed=14, n=1500)
    p["x"][:3] = [0.0, grid.xl * 0.5, grid.xl * 0.25]
    p["y"][:3] = [grid.yl * 0.5, 0.0, grid.yl]
    p["vx"][0] = p["vy"][1] = p["vy"][2] = 0.0  # x += dt*vp_old keeps them on the bounds
    s = make_step(grid, consts, flags, fields, p)
    s.particle_update_BE(it=2)
    path = str(tmp_path / "part.res")
    s.save_particles(path)
    raw = np.fromfile(path, dtype=np.uint8)
    assert raw.size == 4 + 352 * p.n and int(raw[:4].view(np.int32)[0]) == p.n
    assert np.array_equal(raw[4:], s.get_records())
    saved = s.get()
    kept = s.read_part_res(path)
    inside = (saved["x"] > 0) & (saved["x"] < grid.xl) & (saved["y"] > 0) & (saved["y"] < grid.yl)
    assert kept == int(inside.sum()) and kept <= p.n - 3
    back = s.get()
    for c in ("x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "m_s", "kappa_s", "rc", "res", "numact", "qstar"):
        assert np.array_equal(back[c], saved[c][inside]), c
    assert np.array_equal(back.mult, saved.mult[inside]) and np.array_equal(back.pidx, saved.pidx[inside])
    s.close()


COUNTERS = ("num100", "numimpos", "actnum", "denum", "solves")


def add_counts(total, cnt):
    for k in COUNTERS:
        total[k] = total.get(k, 0) + cnt[k]
    return total


def _record_columns():
    """Particles column -> (member of `type particle`, component)."""
    m = {}
    for vec, cols in (("xp", ("x", "y", "z")), ("vp", ("vx", "vy", "vz")), ("uf", ("ufx", "ufy", "ufz")),
                      ("xrhs", ("xrhsx", "xrhsy", "xrhsz")), ("vrhs", ("vrhsx", "vrhsy", "vrhsz")),
                      ("vp_old", ("vxold", "vyold", "vzold"))):
        for k, c in enumerate(cols):
            m[c] = (vec, k)
    for c in ("Tp", "Tprhs_s", "Tprhs_L", "Tf", "radius", "radrhs", "qinf", "qstar", "res", "m_s", "kappa_s", "rc",
              "actres", "numact"):
        m[c] = (c, None)
    m["Tpold"], m["radiusold"] = ("Tp_old", None), ("radius_old", None)
    return m


def test_restart_records_follow_the_type_particle_layout(tmp_path):
    """A restart file whose 352-byte records come from a numpy dtype built from
    the declaration of `type particle` (particles.f90:76-85, the MPI datatype
    of 1929-1950: 4 int32, 39 real(8), 3 int64), with a distinct value in every
    member (tests/golden/part_res_7.bin), read through ntlp_gpu_read_part_res:
    every member lands in its column, the records outside 0 < x < xl,
    0 < y < yl are dropped, and save_particles writes the kept records back
    with every member at its declared offset."""
    import os
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_restart_fixture import PARTICLE

    path = os.path.join(os.path.dirname(__file__), "golden", "part_res_7.bin")
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    rec = raw[4:].view(PARTICLE)
    assert n == 7 and len(rec) == 7
    grid, consts, flags, fields, p, ora = small_case(seed=0, n=10)
    s = make_step(grid, consts, flags, fields, p)
    kept = s.read_part_res(path)
    inside = (rec["xp"][:, 0] > 0) & (rec["xp"][:, 0] < grid.xl) & (rec["xp"][:, 1] > 0) & (rec["xp"][:, 1] < grid.yl)
    assert kept == 5 == int(inside.sum())
    got = s.get()
    want = rec[inside]
    for c, (member, k) in _record_columns().items():
        col = want[member] if k is None else want[member][:, k]
        assert np.array_equal(got[c], col), c
    assert np.array_equal(got.mult, want["mult"]) and np.array_equal(got.pidx, want["pidx"])
    assert np.array_equal(got.procidx, want["procidx"])
    out = str(tmp_path / "again.res")
    s.save_particles(out)
    back = np.fromfile(out, dtype=np.uint8)
    assert int(back[:4].view(np.int32)[0]) == 5
    brec = back[4:].view(PARTICLE)
    for member in PARTICLE.names:
        if member in ("dist", "u_sub", "sigm_s", "prev", "next", "nbr_pidx", "nbr_procidx"):
            continue  # members of the stochastic subgrid model and the list links are not kept
        assert np.array_equal(brec[member], want[member]), member
    s.close()


@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_particle_update_be_fp64(kind):
    grid, consts, flags, fields, p, ora = small_case(seed=6, kind=kind)
    s = make_step(grid, consts, flags, fields, p)
    want = p.copy()
    s.counters(reset=True)
    cnt_o = {}
    for it in (1, 2):
        s.particle_update_BE(it=it)
        add_counts(cnt_o, ora.update_be(want.soa, it))
    got = s.get()
    cnt = s.counters()
    assert {k: cnt[k] for k in COUNTERS} == cnt_o
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "ufx", "ufy", "ufz", "Tf", "qinf", "Tp",
                              "radius", "qstar", "res", "actres", "numact"])
    assert not bad, bad
    for c in ("vxold", "vyold", "vzold", "Tpold", "radiusold"):
        assert rel_err(got[c], want[c], 1e-6) < 1e-9, c
    s.close()


def test_particle_update_rk3_fp64():
    grid, consts, flags, fields, p, ora = small_case(seed=7, n=2000, dt=2e-5)
    s = make_step(grid, consts, flags, fields, p)
    want = p.copy()
    for it in (1, 2):
        for st in (1, 2, 3):
            s.particle_update_rk3(st, it=it)
            ora.update_rk3(want.soa, st, it)
    got = s.get()
    tol = dict(TOL64)
    tol["Tp"] = (1e-12, 300.0)
    tol["radius"] = (1e-11, 1e-6)
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "qstar", "res"], tol)
    assert not bad, bad
    # right-hand sides carry differences such as Tp - Tf, so they are compared
    # on the scale of their column (1e-13 differences of Tp become 1e-8 of
    # Tp - Tf)
    for c in ("xrhsx", "vrhsz", "Tprhs_s", "Tprhs_L", "radrhs"):
        assert rel_err(got[c], want[c], np.max(np.abs(want[c])) + 1e-30) < 1e-9, c
    s.close()


@pytest.mark.parametrize("deterministic", [0, 1, 2])
@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_coupling_update(deterministic, kind):
    """particle_coupling_update on the oracle's own post-step state: only the
    deposit arithmetic and the summation order differ.  Methods: 0 LDS node
    boxes, 1 fixed order per node, 2 wavefront-segmented sums + atomics."""
    grid, consts, flags, fields, p, ora = small_case(seed=8, kind=kind)
    want = p.copy()
    ora.update_be(want.soa, 2)
    src_o = ora.coupling(want.soa, want.mult)
    s = make_step(grid, consts, flags, fields, want)
    src = s.particle_coupling_update(deterministic=deterministic)
    for f in range(6):
        scale = np.max(np.abs(src_o[f])) + 1e-300
        # field 5 is the reference's -A*Tp + A*Tf (particles.f90:1258-1260): two
        # terms ~Tp/(Tp-Tf) ~ 3000 times larger than their sum cancel
        tol = 1e-12 if f < 5 else 1e-9
        assert np.max(np.abs(src[f] - src_o[f])) <= tol * scale, f
    if deterministic == 1:
        again = s.particle_coupling_update(deterministic=1)
        assert np.array_equal(again, src)
    s.close()


def test_coupling_method_out_of_range_is_an_error():
    grid, consts, flags, fields, p, ora = small_case(seed=8, n=256)
    s = make_step(grid, consts, flags, fields, p)
    with pytest.raises(NtlpError):
        s.particle_coupling_update(deterministic=3)
    s.close()


def test_coupling_single_droplet_on_a_node_and_linearity():
    """A droplet exactly on node (i, j, k) deposits on that node only; the
    deposit is linear in the multiplicity."""
    grid, consts, flags, fields, p, ora = small_case(seed=9, n=64)
    want = p.copy()
    ora.update_be(want.soa, 2)
    one = want.copy()
    i, j, k = 5, 4, 7
    one.soa = one.soa[:, :1].copy()
    one.mult, one.pidx, one.procidx = one.mult[:1].copy(), one.pidx[:1].copy(), one.procidx[:1].copy()
    one["x"], one["y"], one["z"] = grid.dx * (i - 1), grid.dy * (j - 1), grid.zz[k]
    s = make_step(grid, consts, flags, fields, one)
    src = s.particle_coupling_update()
    big = np.abs(src).max()
    nz = np.argwhere(np.any(np.abs(src) > 1e-12 * big, axis=0))
    assert nz.tolist() == [[i - 1, j - 1, k]]
    one.mult[:] *= 3
    s.put(one)
    src3 = s.particle_coupling_update()
    assert np.allclose(src3, 3 * src, rtol=1e-14, atol=0)
    s.close()


@pytest.mark.parametrize("mode", [MODE_BE, MODE_RK3])
def test_fused_steps_match_single_steps(mode):
    dt = 5e-4 if mode == MODE_BE else 2e-5
    grid, consts, flags, fields, p, ora = small_case(seed=10, dt=dt)
    s = make_step(grid, consts, flags, fields, p)
    s.step(mode=mode, precision=PREC_FP64, nsteps=4, it0=2)
    got = s.get()
    src = s.sources()
    want = p.copy()
    for it in range(2, 6):
        if mode == MODE_BE:
            ora.update_be(want.soa, it)
        else:
            for st in (1, 2, 3):
                ora.update_rk3(want.soa, st, it)
    src_o = ora.coupling(want.soa, want.mult)
    cols = ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "res", "actres", "numact", "ufx", "Tf", "qinf",
            "qstar"]
    if mode == MODE_BE:
        assert want["numact"].sum() > 0  # an activation resets actres (particles.f90:3172)
    tol = dict(TOL64)
    tol["radius"] = (1e-8, 1e-6)
    bad = compare(got, want, cols, tol)
    assert not bad, bad
    assert np.array_equal(got.mult, p.mult) and np.array_equal(got.pidx, p.pidx)
    # after four steps Tp - Tpold and v - vold carry the 1e-13 state
    # differences at 1e-7 of their own size
    for f in range(6):
        scale = np.max(np.abs(src_o[f])) + 1e-300
        assert np.max(np.abs(src[f] - src_o[f])) <= 1e-5 * scale, f
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_residence_times_of_a_multi_step_call_equal_per_step_calls(prec):
    """A BE call of many steps adds the residence times at its end
    (particles.f90:3204-3205, actres restarting at activations, 3172); the
    bits equal those of one-step calls, across reorderings and activations."""
    wl = synth.workload(3, n=400_000, seed=3)
    s = DropletStep(wl.grid, wl.consts, wl.flags)
    s.set_fields(*wl.fields)
    s.put(wl.droplets)
    s.interp(set_vp=True, set_tp=True)
    s.crit_radius()
    state = s.get()
    s.counters(reset=True)
    s.step(MODE_BE, prec, nsteps=20, it0=2)
    many = s.get()
    activations = s.counters(reset=True)["actnum"]
    s.put(state)
    for k in range(20):
        s.step(MODE_BE, prec, nsteps=1, it0=2 + k)
    ones = s.get()
    s.close()
    assert activations > 0
    for c in ("res", "actres", "numact", "radius", "Tp"):
        assert np.array_equal(many[c], ones[c]), c
    assert np.all(many["actres"] <= many["res"])


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_fused_step_sources_match_oracle_deposit(prec):
    """The sources a fused step leaves (per-droplet amounts written by the step
    kernel, deposited by the box scatter through the fresh cell sort) equal the
    oracle's particle_coupling_update on the device's own post-step state (its
    positions, the gathered Tf and the pre-step velocity, Tp and radius).
    Droplets placed just short of the 8-cell box faces and of x = xl, moving
    across them, exercise the boxes' shared faces and the periodic wrap."""
    grid, consts, flags, fields, p, ora = small_case(seed=21, n=6000)
    q = p.copy()
    rng = np.random.default_rng(21)
    pick = rng.permutation(q.n)
    k = q.n // 8
    face_x, face_y = 8 * grid.dx, 8 * grid.dy
    q["x"][pick[:k]] = face_x - 2e-5
    q["vx"][pick[:k]] = 0.2
    q["x"][pick[k:2 * k]] = face_x + 2e-5
    q["vx"][pick[k:2 * k]] = -0.2
    q["y"][pick[2 * k:3 * k]] = face_y - 2e-5
    q["vy"][pick[2 * k:3 * k]] = 0.2
    q["x"][pick[3 * k:3 * k + 50]] = grid.xl - 1e-5
    q["vx"][pick[3 * k:3 * k + 50]] = 0.1
    s = make_step(grid, consts, flags, fields, q)
    s.step(mode=MODE_BE, precision=prec, nsteps=3, it0=2)
    src = s.sources()
    got = s.get()
    src_o = ora.coupling(got.soa, got.mult)
    for f in range(6):
        scale = np.max(np.abs(src_o[f])) + 1e-300
        # field 5 is -A*Tp + A*Tf: its two terms cancel to 1/3000; a mixed step
        # stores each droplet's amounts in FP32 (sums in FP64)
        tol = (1e-12 if f < 5 else 1e-9) if prec == PREC_FP64 else 1e-6
        assert np.max(np.abs(src[f] - src_o[f])) <= tol * scale, (f, np.max(np.abs(src[f] - src_o[f])) / scale)
    s.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_MIXED])
def test_set_dt_matches_a_context_made_with_that_dt(prec):
    """ntlp_gpu_set_dt (les.F's get_dt changes dt between steps) refreshes every
    constant derived from dt: two steps after set_dt equal, bit for bit, two
    steps of a context initialised with that dt (the sources to rounding: the
    scatter's LDS atomics add in whatever order the hardware runs them)."""
    import dataclasses

    grid, consts, flags, fields, p, ora = small_case(seed=23, n=3000)
    a = make_step(grid, consts, flags, fields, p)
    a.set_dt(2.5e-4)
    a.step(mode=MODE_BE, precision=prec, nsteps=2, it0=2)
    got = a.get()
    src = a.sources()
    a.close()
    b = make_step(grid, dataclasses.replace(consts, dt=2.5e-4), flags, fields, p)
    b.step(mode=MODE_BE, precision=prec, nsteps=2, it0=2)
    want = b.get()
    src_b = b.sources()
    b.close()
    assert np.array_equal(got.soa, want.soa)
    for f in range(6):
        assert np.max(np.abs(src[f] - src_b[f])) <= 1e-12 * (np.max(np.abs(src_b[f])) + 1e-300), f


def test_fused_steps_across_reorderings_match_oracle():
    """20 fused FP64 BE steps: in-place steps between the periodic
    reorderings of the slots into cell order, and the reorderings themselves,
    leave the same droplets as 20 oracle steps."""
    grid, consts, flags, fields, p, ora = small_case(seed=16, n=4000, kind="stretched")
    s = make_step(grid, consts, flags, fields, p)
    s.step(mode=MODE_BE, precision=PREC_FP64, nsteps=12, it0=2)
    s.step(mode=MODE_BE, precision=PREC_FP64, nsteps=8, it0=14)
    got = s.get()
    want = p.copy()
    for it in range(2, 22):
        ora.update_be(want.soa, it)
    assert want["numact"].sum() > 0  # activations restart actres inside the lazy 12- and 8-step calls
    tol = dict(TOL64)
    tol.update({"x": (1e-12, 1.0), "y": (1e-12, 1.0), "z": (1e-12, 1.0), "vx": (1e-9, 0.1), "vy": (1e-9, 0.1),
                "vz": (1e-9, 0.1), "Tp": (1e-10, 300.0), "radius": (1e-8, 1e-6)})
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "res", "actres", "numact"], tol)
    assert not bad, bad
    assert np.array_equal(got.mult, p.mult) and np.array_equal(got.pidx, p.pidx)
    s.close()


@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_fused_mixed_steps_match_single_mixed_steps(kind):
    """Dense droplets (about 7 per cell), mixed precision: three fused steps
    (cell-sorted, ping-ponged, wave-tiled) against three single-step calls in
    put order.  The gathered values and the state must agree to rounding."""
    grid, consts, flags, fields, p, ora = small_case(seed=15, n=30000, kind=kind)
    s = make_step(grid, consts, flags, fields, p)
    for it in (2, 3, 4):
        s.particle_update_BE(it=it, precision=PREC_MIXED)
    want = s.get()
    s.put(p)
    s.step(mode=MODE_BE, precision=PREC_MIXED, nsteps=3, it0=2)
    got = s.get()
    tol = {c: (1e-13, f) for c, f in (("x", 1.0), ("y", 1.0), ("z", 1.0), ("vx", 0.1), ("vy", 0.1), ("vz", 0.1),
                                     ("ufx", 0.1), ("ufy", 0.1), ("ufz", 0.1), ("Tf", 300.0), ("qinf", 0.02),
                                     ("Tp", 300.0), ("radius", 1e-6))}
    bad = compare(got, want, list(tol), tol)
    assert not bad, bad
    s.close()


@pytest.mark.parametrize("prec", [PREC_MIXED, PREC_FP64_AJ])
def test_mixed_precision_radius_within_1e5(prec):
    grid, consts, flags, fields, p, ora = small_case(seed=11, n=4000, kind="stretched")
    s = make_step(grid, consts, flags, fields, p)
    s.step(mode=MODE_BE, precision=prec, nsteps=10, it0=2)
    got = s.get()
    want = p.copy()
    for it in range(2, 12):
        ora.update_be(want.soa, it)
    assert rel_err(got["radius"], want["radius"], 1e-7) <= 1e-5
    assert rel_err(got["Tp"], want["Tp"], 300.0) <= 1e-6
    assert rel_err(got["z"], want["z"], 1.0) <= 1e-9
    s.close()


def test_counters_count_solves():
    grid, consts, flags, fields, p, ora = small_case(seed=12)
    s = make_step(grid, consts, flags, fields, p)
    s.counters(reset=True)
    s.step(mode=MODE_BE, precision=PREC_FP64, nsteps=3, it0=2)
    c = s.counters()
    assert c["steps"] == 3
    assert c["solves"] == 3 * p.n  # qinf > 0 and ievap = 1 for every droplet
    assert sum(c["gn_hist"]) == c["solves"]
    rows = s.step_counters()
    assert rows.shape[0] == 3 and rows[:, 6].sum() == c["solves"]
    s.close()


@pytest.mark.parametrize("iexner", [0, 1])
def test_coupling_exchange_and_apply_match_oracle(iexner):
    """particle_coupling_exchange + apply_particle_coupling after a BE step:
    the fold/transpose is bit-exact against the oracle on the device's own
    sources; the T/q update matches to the last ulps of exner's pow; the
    resident T and q the next step gathers are the updated ones."""
    grid, consts, flags, fields, p, ora = small_case(seed=12, flags=Flags(iexner=iexner))
    s = make_step(grid, consts, flags, fields, p)
    s.particle_update_BE(it=2)
    s.particle_coupling_update()
    src = np.ascontiguousarray(s.sources())
    partsrc, partT, partH, partTE = s.particle_coupling_exchange()
    xs_o = ora.coupling_exchange(src.copy())
    assert np.array_equal(partsrc, xs_o[:3]) and np.array_equal(partT, xs_o[3])
    assert np.array_equal(partH, xs_o[4]) and np.array_equal(partTE, xs_o[5])
    # T and q in the flow solver's x-pencil layout (1:nnx+2, 1:nny, 0:nnz+1)
    u, v, w, T, q = fields
    def xpencil(a):
        out = np.zeros((grid.nnz + 2, grid.nny, grid.nnx + 2))
        out[:, :, :grid.nnx] = np.transpose(a, (2, 1, 0))
        out[:, :, grid.nnx:] = -1.0
        return out
    Tx, qx = xpencil(T), xpencil(q)
    Tx_o, qx_o = Tx.copy(), qx.copy()
    s.apply_particle_coupling(Tx, qx, LAYOUT_XPENCIL)
    ora.apply_coupling(xs_o, Tx_o, qx_o)
    if iexner:
        assert np.max(np.abs(Tx - Tx_o)) <= 1e-13 * 300 and np.array_equal(qx, qx_o)
    else:
        assert np.array_equal(Tx, Tx_o) and np.array_equal(qx, qx_o)
    # the resident fields: gather at the droplets and compare with the oracle's
    T_new = np.ascontiguousarray(np.transpose(Tx_o[:, :, :grid.nnx], (2, 1, 0)))
    q_new = np.ascontiguousarray(np.transpose(qx_o[:, :, :grid.nnx], (2, 1, 0)))
    ora.set_fields(u, v, w, T_new, q_new)
    got = s.get()
    s.interp()
    mine = s.get()
    want = got.copy()
    ora.interp(want.soa)
    bad = compare(mine, want, ["Tf", "qinf"], dict(TOL64, Tf=(1e-13, 300.0)))
    assert not bad, bad
    s.close()


@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_statistics_match_oracle(kind):
    """calc_tauc_min and particle_xy_stats on the device's own state after
    two fused BE steps: counts, maxima and minima exact; sums within 1e-12
    (summation order)."""
    grid, consts, flags, fields, p, ora = small_case(seed=13, kind=kind)
    s = make_step(grid, consts, flags, fields, p)
    s.step(MODE_BE, PREC_FP64, nsteps=2, it0=2)
    st = s.get()
    assert abs(s.calc_tauc_min() / ora.calc_tauc_min(st.soa, st.mult) - 1) < 1e-12
    s0, s1, lwp = s.particle_xy_stats()
    w0, w1, wl = ora.xy_stats(st.soa, st.mult)
    assert np.array_equal(s0[:3], w0[:3]) and np.array_equal(s0[9:], w0[9:])
    assert np.allclose(s0[3:9], w0[3:9], rtol=1e-12, atol=0)
    assert np.array_equal(s1[0], w1[0])
    assert np.allclose(s1, w1, rtol=1e-12, atol=1e-300)
    assert np.allclose(lwp, wl, rtol=1e-12, atol=0)
    s.close()
