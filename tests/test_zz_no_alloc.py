"""ntlp_gpu_particle_xy_stats_noalloc against ntlp_gpu_particle_xy_stats.

The variant allocates nothing: its partial rows live in a static device
workspace and its three outputs are device buffers the caller owns (allocated
here with the HIP runtime the library uses).  Two calls in a row, the second with fewer droplets (fewer
partial rows, a workspace still holding the first call's rows), each compared
with the allocating entry point and with the oracle:
  - counts, maxima and minima (s0[0:3], s0[9:21]) and the per-level counts
    (s1[0]) bit-exact;
  - sums within 1e-12 relative (the per-level sums are LDS atomics and LWP
    global atomics, whose order varies from run to run).
"""
import ctypes

import numpy as np
import pytest

from ntlp_amd import MODE_BE, PREC_FP64, DropletStep, Particles
from tests.helpers import small_case


def _hip():
    """The HIP runtime the library itself uses (the copy mapped into this
    process), for the caller-side device buffers of the test."""
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 is not mapped: load the library first"
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(s: DropletStep):
    """Calls the no-allocation entry point with device buffers the caller
    allocated (and filled with NaN), and copies them back."""
    hip = _hip()
    g = s.grid
    shapes = [(21,), (24, g.nnz + 2), (g.nny, g.nnx)]
    bufs = []
    for shp in shapes:
        p = ctypes.c_void_p()
        nb = int(np.prod(shp)) * 8
        assert hip.hipMalloc(ctypes.byref(p), nb) == 0
        nan = np.full(shp, np.nan)
        assert hip.hipMemcpy(p, nan.ctypes.data, nb, 1) == 0  # host to device
        bufs.append((p, shp, nb))
    try:
        rc = s.lib.ntlp_gpu_particle_xy_stats_noalloc(*[b[0] for b in bufs])
        assert rc == 0, s.lib.ntlp_gpu_last_error()
        assert s.lib.ntlp_gpu_sync() == 0
        out = []
        for p, shp, nb in bufs:
            a = np.empty(shp)
            assert hip.hipMemcpy(a.ctypes.data, p, nb, 2) == 0  # device to host
            out.append(a)
    finally:
        for p, _, _ in bufs:
            hip.hipFree(p)
    return tuple(out)


def _same(got, want):
    g0, g1, gl = got
    w0, w1, wl = want
    assert np.isfinite(g0).all() and np.isfinite(g1).all() and np.isfinite(gl).all()
    assert np.array_equal(g0[:3], w0[:3]) and np.array_equal(g0[9:], w0[9:])
    assert np.allclose(g0[3:9], w0[3:9], rtol=1e-12, atol=0)
    assert np.array_equal(g1[0], w1[0])
    assert np.allclose(g1, w1, rtol=1e-12, atol=1e-300)
    assert np.allclose(gl, wl, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_xy_stats_noalloc_matches_allocating_entry_point():
    print("no-alloc statistics: 6000 then 1100 droplets, stretched 16x12x20", flush=True)
    grid, consts, flags, fields, p, ora = small_case(seed=23, n=6000, kind="stretched")
    s = DropletStep(grid, consts, flags)
    try:
        s.set_fields(*fields)
        for n in (6000, 1100):  # the second call has fewer rows than the workspace still holds
            s.put(Particles(p.soa[:, :n].copy(), p.mult[:n].copy(), p.pidx[:n].copy(), p.procidx[:n].copy()))
            s.step(MODE_BE, PREC_FP64, nsteps=2, it0=2)
            want = s.particle_xy_stats()
            got = _noalloc(s)
            _same(got, want)
            st = s.get()
            _same(got, ora.xy_stats(st.soa, st.mult))
        null = ctypes.c_void_p(0)
        assert s.lib.ntlp_gpu_particle_xy_stats_noalloc(null, null, null) == -1  # NTLP_E_INVALID
    finally:
        s.close()
