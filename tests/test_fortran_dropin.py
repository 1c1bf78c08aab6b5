"""The Fortran drop-in: fortran/ntlp_dropin_driver (built on the ntlp_gpu
ISO_C_BINDING module) runs particle_update_BE + particle_coupling_update for a
few steps the way les.F:395-431 calls them.  Its output must equal, bit for
bit, the same sequence through the ctypes binding, and match the oracle
within the FP64 parity tolerances."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from ntlp_amd import LAYOUT_CORE, NREAL, PREC_FP64, Particles
from ntlp_amd.build import FORTRAN_BUILD

from .helpers import small_case
from .test_gpu_parity import TOL64, compare, make_step

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(FORTRAN_BUILD, "ntlp_dropin_driver")


def write_input(path, grid, consts, flags, fields, p, nsteps, it0, prec):
    hdr = np.array([grid.nnx, grid.nny, grid.nnz, p.n, nsteps, it0, prec, LAYOUT_CORE], dtype=np.int32)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in fields]
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.array([grid.xl, grid.yl]).tobytes())
        for a in (grid.z, grid.zz, grid.dzw, grid.dzu):
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        fh.write(bytes(consts.to_c()))
        fh.write(bytes(flags.to_c()))
        fh.write(np.array([arrs[0].size], dtype=np.int32).tobytes())
        for a in arrs:
            fh.write(a.tobytes())
        # soa(n, 34) column-major is the (34, n) row-major block of Particles
        fh.write(np.ascontiguousarray(p.soa, dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(p.mult, dtype=np.int64).tobytes())
        fh.write(np.ascontiguousarray(p.pidx, dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(p.procidx, dtype=np.int32).tobytes())


def read_output(path, n, grid):
    raw = open(path, "rb").read()
    off = 0

    def take(count, dtype):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    soa = take(NREAL * n, np.float64).reshape(NREAL, n).copy()
    mult = take(n, np.int64).copy()
    pidx = take(n, np.int32).copy()
    procidx = take(n, np.int32).copy()
    m = (grid.nnz + 2) * (grid.nny + 1) * (grid.nnx + 1)
    # Fortran (0:nnz+1, 1:nny+1, 1:nnx+1) column-major == C (nnx+1, nny+1, nnz+2)
    src = take(6 * m, np.float64).reshape(6, grid.nnx + 1, grid.nny + 1, grid.nnz + 2).copy()
    cnt = take(8, np.int64)
    kept = int(take(1, np.int64)[0])
    maxdiff = float(take(1, np.float64)[0])
    assert off == len(raw)
    return Particles(soa, mult, pidx, procidx), src, cnt, (kept, maxdiff)


@pytest.mark.parametrize("kind", ["uniform", "stretched"])
def test_fortran_driver_matches_binding_and_oracle(kind, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (ntlp_amd.build.build_fortran)")
    grid, consts, flags, fields, p, ora = small_case(seed=11, n=2500, kind=kind)
    nsteps, it0 = 3, 2
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, grid, consts, flags, fields, p, nsteps, it0, PREC_FP64)
    proc = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    got, src, cnt, (kept, maxdiff) = read_output(fout, p.n, grid)
    # the restart round trip through the module's save_particles / read_part_res
    # wrappers keeps every droplet (all inside 0 < x < xl, 0 < y < yl) and
    # gives back the radii bit for bit
    assert kept == p.n and maxdiff == 0.0, (kept, maxdiff)
    assert os.path.getsize(fout + ".restart") == 4 + 352 * p.n

    # the same calls through the ctypes binding
    s = make_step(grid, consts, flags, fields, p)
    s.counters(reset=True)
    for k in range(nsteps):
        s.particle_update_BE(it=it0 + k)
        src_py = s.particle_coupling_update(deterministic=True)
    mine = s.get()
    c = s.counters()
    s.close()
    assert np.array_equal(got.soa, mine.soa)
    assert np.array_equal(got.mult, mine.mult) and np.array_equal(got.pidx, mine.pidx)
    assert np.array_equal(src, src_py)
    assert cnt[6] == c["solves"] == nsteps * p.n

    # and the oracle
    want = p.copy()
    for k in range(nsteps):
        ora.update_be(want.soa, it0 + k)
    src_o = ora.coupling(want.soa, want.mult)
    bad = compare(got, want, ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "qstar", "res", "numact"])
    assert not bad, bad
    # the sources are rates (new - old)/dt of the final step: a 1e-11 relative
    # difference of Tp (300 K) after three steps is 1e-6 of the heat source
    for f in range(6):
        assert np.max(np.abs(src[f] - src_o[f])) <= 1e-5 * (np.max(np.abs(src_o[f])) + 1e-300), f
