"""ntlp_amd -- MI355X-native Lagrangian droplet step of NTLP.

Python host binding of libntlp_gpu.so (C ABI in include/ntlp_gpu.h).  The
methods keep the names and argument meaning of the reference's particle
entry points (particles.f90): ``particle_update_rk3(istage)``,
``particle_update_BE()``, ``particle_coupling_update()``; failures raise
:class:`NtlpError` carrying the library's error text, where the reference
would stop.

There is no CPU fallback: importing works without a GPU, but every compute
call goes to the HIP library and raises if it cannot run.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import grid as _grid  # noqa: F401  (re-exported helpers)
from .grid import Grid

__all__ = ["Grid", "Consts", "Flags", "Particles", "Ensemble", "DropletStep", "ServeLoop", "PinnedArray", "NtlpError",
           "load_library", "DIST_LOGNORMAL", "DIST_LOGUNIFORM", "SALT_FIXED", "SALT_DRY_LOGNORMAL", "SALT_CHOICE3",
           "NREAL", "COL", "MODE_RK3", "MODE_BE", "PREC_FP64", "PREC_MIXED", "PREC_FP64_AJ", "OPT_STAGED_GATHER",
           "LAYOUT_EXT", "LAYOUT_CORE", "LAYOUT_XPENCIL", "LIB_PATH"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libntlp_gpu.so")

# droplet record columns (enum ntlp_column)
COLUMNS = ["x", "y", "z", "vx", "vy", "vz", "ufx", "ufy", "ufz", "xrhsx", "xrhsy", "xrhsz",
           "vrhsx", "vrhsy", "vrhsz", "Tp", "Tprhs_s", "Tprhs_L", "Tf", "radius", "radrhs", "qinf",
           "qstar", "res", "m_s", "kappa_s", "rc", "actres", "numact", "vxold", "vyold", "vzold",
           "Tpold", "radiusold"]
COL = {name: i for i, name in enumerate(COLUMNS)}
NREAL = len(COLUMNS)

MODE_RK3, MODE_BE = 1, 2
PREC_FP64, PREC_MIXED, PREC_FP64_AJ = 0, 1, 2
OPT_STAGED_GATHER = 1
LAYOUT_EXT, LAYOUT_CORE, LAYOUT_XPENCIL = 0, 1, 2
GN_HIST = 64
STEP_ROW = 72
DIST_LOGNORMAL, DIST_LOGUNIFORM = 0, 1
SALT_FIXED, SALT_DRY_LOGNORMAL, SALT_CHOICE3 = 0, 1, 2

NTLP_ERRORS = {-1: "NTLP_E_INVALID", -2: "NTLP_E_NOT_INIT", -3: "NTLP_E_HIP", -4: "NTLP_E_NOMEM",
               -5: "NTLP_E_RANGE", -6: "NTLP_E_UNSUPPORTED", -7: "NTLP_E_STATE"}


class NtlpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{NTLP_ERRORS.get(code, code)}: {msg}")
        self.code = code


class _CGrid(ctypes.Structure):
    _fields_ = [("nnx", ctypes.c_int32), ("nny", ctypes.c_int32), ("nnz", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("xl", ctypes.c_double), ("yl", ctypes.c_double),
                ("z", ctypes.c_void_p), ("zz", ctypes.c_void_p), ("dzw", ctypes.c_void_p),
                ("dzu", ctypes.c_void_p)]


_CONST_SCALARS = ["dt", "nuf", "Cpa", "Cpp", "CpaCpp", "Pra", "Sc", "Rd", "rhow", "rhos", "Mw", "Ru",
                  "Gam", "Lv", "Cpv", "surf_rho", "surf_p", "tsfcc1", "grav", "radius_init", "pi2"]


class _CConsts(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in _CONST_SCALARS] +
                [("part_grav", ctypes.c_double * 3), ("zetas", ctypes.c_double * 3),
                 ("gama", ctypes.c_double * 3)])


class _CFlags(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("ilin", "icouple", "iTcouple", "iHcouple", "ievap", "iexner", "icase", "reserved")]


class _CCounters(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int64) for n in
                 ("num100", "numimpos", "actnum", "denum", "gn_iters", "lm_iters", "solves", "steps")] +
                [("gn_hist", ctypes.c_int64 * GN_HIST)])


@dataclass
class Consts:
    """Namelist constants of the droplet path (defs.F:43-46, particles.f90:42,
    test_cases/pi_chamber/params.in:105-139) and the RK3 coefficients of setcon
    (les.F:856-861).  Defaults: the pi-chamber values with kappa_s = 1.2,
    part_grav = (0, 0, -9.81) and surf_rho = 1.2."""
    dt: float = 1.0e-3
    nuf: float = 1.57e-5
    Cpa: float = 1006.0
    Cpp: float = 4179.0
    Pra: float = 0.715
    Sc: float = 0.615
    Rd: float = 286.9
    rhow: float = 1000.0
    rhos: float = 2000.0
    Mw: float = 0.018015
    Ru: float = 8.3144
    Gam: float = 7.28e-2
    Lv: float = (25.0 - 0.02274 * 26.0) * 10.0 ** 5
    Cpv: float = 1952.0
    surf_rho: float = 1.2
    surf_p: float = 101325.0
    tsfcc1: float = 300.0
    grav: float = 9.81
    radius_init: float = 1.0e-5
    part_grav: tuple = (0.0, 0.0, -9.81)
    zetas: tuple = (0.0, -17.0 / 60.0, -5.0 / 12.0)
    gama: tuple = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)

    @property
    def CpaCpp(self) -> float:  # les.F:10759
        return self.Cpa / self.Cpp

    @property
    def pi2(self) -> float:  # les.F:5387
        return 2.0 * (4.0 * np.arctan(1.0))

    def to_c(self) -> _CConsts:
        c = _CConsts()
        for n in _CONST_SCALARS:
            setattr(c, n, float(getattr(self, n)))
        c.part_grav = (ctypes.c_double * 3)(*self.part_grav)
        c.zetas = (ctypes.c_double * 3)(*self.zetas)
        c.gama = (ctypes.c_double * 3)(*self.gama)
        return c


@dataclass
class Flags:
    """Hot-path flags of namelist &flags (defs.F:8-14); defaults are those of
    every shipped test case except icase (0: reflecting walls)."""
    ilin: int = 1
    icouple: int = 1
    iTcouple: int = 1
    iHcouple: int = 1
    ievap: int = 1
    iexner: int = 0
    icase: int = 0

    def to_c(self) -> _CFlags:
        return _CFlags(self.ilin, self.icouple, self.iTcouple, self.iHcouple, self.ievap, self.iexner,
                       self.icase, 0)


@dataclass
class Particles:
    """Droplets as a structure of arrays: ``soa[COL[name], i]`` (FP64),
    multiplicity, and the (pidx, procidx) identity of the reference record."""
    soa: np.ndarray
    mult: np.ndarray
    pidx: np.ndarray
    procidx: np.ndarray

    @property
    def n(self) -> int:
        return int(self.soa.shape[1])

    def __getitem__(self, name: str) -> np.ndarray:
        return self.soa[COL[name]]

    def __setitem__(self, name: str, value) -> None:
        self.soa[COL[name]] = value

    def copy(self) -> "Particles":
        return Particles(self.soa.copy(), self.mult.copy(), self.pidx.copy(), self.procidx.copy())

    @staticmethod
    def empty(n: int) -> "Particles":
        return Particles(np.zeros((NREAL, n)), np.ones(n, dtype=np.int64),
                         np.arange(1, n + 1, dtype=np.int32), np.zeros(n, dtype=np.int32))


class _CEnsemble(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("radius_kind", ctypes.c_int32), ("radius2_kind", ctypes.c_int32),
                ("radius_p", ctypes.c_double * 2), ("radius2_p", ctypes.c_double * 2), ("frac2", ctypes.c_double),
                ("salt_kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("salt_p", ctypes.c_double * 3),
                ("kappa_s", ctypes.c_double), ("mult", ctypes.c_int64), ("first", ctypes.c_int64)]


@dataclass
class Ensemble:
    """A droplet ensemble drawn on the device (ntlp_ensemble): radius
    distribution (DIST_LOGNORMAL: (median, sigma_ln); DIST_LOGUNIFORM: (lo,
    hi)), optionally a second one taken with probability frac2, salt mass
    (SALT_FIXED: m_s; SALT_DRY_LOGNORMAL: (median dry radius, sigma_ln);
    SALT_CHOICE3: three masses), kappa_s and multiplicity; droplet i uses the
    Philox counter first + i under the key seed."""
    seed: int = 0
    radius_kind: int = DIST_LOGNORMAL
    radius_p: tuple = (10e-6, 0.3)
    radius2_kind: int = DIST_LOGNORMAL
    radius2_p: tuple = (1e-6, 0.0)
    frac2: float = 0.0
    salt_kind: int = SALT_FIXED
    salt_p: tuple = (1e-18, 0.0, 0.0)
    kappa_s: float = 1.2
    mult: int = 1
    first: int = 0

    def to_c(self) -> _CEnsemble:
        e = _CEnsemble()
        e.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        e.radius_kind, e.radius2_kind = int(self.radius_kind), int(self.radius2_kind)
        e.radius_p = (ctypes.c_double * 2)(*self.radius_p)
        e.radius2_p = (ctypes.c_double * 2)(*self.radius2_p)
        e.frac2 = float(self.frac2)
        e.salt_kind = int(self.salt_kind)
        sp = tuple(self.salt_p) + (0.0,) * (3 - len(self.salt_p))
        e.salt_p = (ctypes.c_double * 3)(*sp)
        e.kappa_s, e.mult, e.first = float(self.kappa_s), int(self.mult), int(self.first)
        return e


class _CServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("steps", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("mode", ctypes.c_int32)]


class _CServeTotals(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "queued", "iterations", "droplet_steps",
                                              "max_pool", "graph_launches", "graphs", "capacity")] + \
               [("newton", _CCounters)] + \
               [(n, ctypes.c_double) for n in ("ms_engine", "ms_copy_calls", "ms_wait_download", "ms_wait_upload",
                                               "ms_wait_control")]


_lib_handle = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libntlp_gpu.so (built by ``python -m ntlp_amd.build``).  Raises if
    it is missing: there is no other implementation to fall back to."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing; build it with `python -m ntlp_amd.build` "
                           "(hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    i32, i64, vp, dp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
    sig = {
        "ntlp_gpu_init": (ctypes.c_int, [ctypes.POINTER(_CGrid), ctypes.POINTER(_CConsts), ctypes.POINTER(_CFlags)]),
        "ntlp_gpu_finalize": (ctypes.c_int, []),
        "ntlp_gpu_last_error": (ctypes.c_char_p, []),
        "ntlp_gpu_set_dt": (ctypes.c_int, [dp]),
        "ntlp_gpu_set_fields": (ctypes.c_int, [vp, vp, vp, vp, vp, i32]),
        "ntlp_gpu_particles_put": (ctypes.c_int, [vp, vp, vp, vp, i64]),
        "ntlp_gpu_particles_get": (ctypes.c_int, [vp, vp, vp, vp, i64]),
        "ntlp_gpu_particles_put_records": (ctypes.c_int, [vp, i64]),
        "ntlp_gpu_particles_get_records": (ctypes.c_int, [vp, i64]),
        "ntlp_gpu_particles_count": (i64, []),
        "ntlp_gpu_particles_get_column": (ctypes.c_int, [i32, vp, i64]),
        "ntlp_gpu_particles_generate": (ctypes.c_int, [vp, vp, i32]),
        "ntlp_serve_submit_ensemble": (ctypes.c_int, [ctypes.POINTER(_CEnsemble), i64, vp, i64, vp,
                                                       ctypes.POINTER(i64)]),
        "ntlp_gpu_save_particles": (ctypes.c_int, [ctypes.c_char_p]),
        "ntlp_gpu_read_part_res": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(i64)]),
        "ntlp_gpu_interp": (ctypes.c_int, [i32, i32]),
        "ntlp_gpu_crit_radius": (ctypes.c_int, []),
        "ntlp_gpu_particle_update_rk3": (ctypes.c_int, [i32, i32]),
        "ntlp_gpu_particle_update_be": (ctypes.c_int, [i32]),
        "ntlp_gpu_particle_update_rk3_prec": (ctypes.c_int, [i32, i32, i32]),
        "ntlp_gpu_particle_update_be_prec": (ctypes.c_int, [i32, i32]),
        "ntlp_gpu_particle_coupling_update": (ctypes.c_int, [vp, i32]),
        "ntlp_gpu_get_sources": (ctypes.c_int, [vp]),
        "ntlp_gpu_particle_coupling_exchange": (ctypes.c_int, [vp, vp, vp, vp]),
        "ntlp_gpu_apply_particle_coupling": (ctypes.c_int, [vp, vp, i32]),
        "ntlp_gpu_bin": (ctypes.c_int, [vp, vp, vp, vp]),
        "ntlp_gpu_wcell_counts": (ctypes.c_int, [vp]),
        "ntlp_gpu_num_ucells": (i64, []),
        "ntlp_gpu_calc_tauc_min": (ctypes.c_int, [ctypes.POINTER(dp)]),
        "ntlp_gpu_particle_xy_stats": (ctypes.c_int, [vp, vp, vp]),
        "ntlp_gpu_particle_xy_stats_noalloc": (ctypes.c_int, [vp, vp, vp]),
        "ntlp_gpu_step": (ctypes.c_int, [i32, i32, i32, i32]),
        "ntlp_gpu_be_flux": (ctypes.c_int, [vp, vp, vp]),
        "ntlp_gpu_counters": (ctypes.c_int, [ctypes.POINTER(_CCounters), i32]),
        "ntlp_gpu_step_counters": (ctypes.c_int, [vp, i32]),
        "ntlp_gpu_timing": (ctypes.c_int, [i32]),
        "ntlp_gpu_kernel_times": (ctypes.c_int, [vp, vp, vp, i32]),
        "ntlp_gpu_copy_bandwidth": (ctypes.c_int, [i64, i32, vp]),
        "ntlp_gpu_sync": (ctypes.c_int, []),
        "ntlp_gpu_set_option": (ctypes.c_int, [i32, i64]),
        "ntlp_serve_open": (ctypes.c_int, [ctypes.POINTER(_CServeConfig), ctypes.POINTER(i64)]),
        "ntlp_serve_submit": (ctypes.c_int, [vp, i64, vp, i64, vp, i64, vp, ctypes.POINTER(i64)]),
        "ntlp_host_alloc": (ctypes.c_int, [i64, ctypes.POINTER(vp)]),
        "ntlp_host_free": (ctypes.c_int, [vp]),
        "ntlp_serve_poll": (ctypes.c_int, [i64]),
        "ntlp_serve_wait": (ctypes.c_int, [i64, vp]),
        "ntlp_serve_stats": (ctypes.c_int, [ctypes.POINTER(_CServeTotals)]),
        "ntlp_serve_close": (ctypes.c_int, [ctypes.POINTER(_CServeTotals)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib_handle = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DropletStep:
    """The droplet-step context on the current GPU (the module state of
    particles.f90).  One per process, like the reference module."""

    def __init__(self, grid: Grid, consts: Consts | None = None, flags: Flags | None = None):
        self.lib = load_library()
        self.grid = grid
        self.consts = consts or Consts()
        self.flags = flags or Flags()
        self._cgrid = grid.to_c(_CGrid)
        self._check(self.lib.ntlp_gpu_init(ctypes.byref(self._cgrid), ctypes.byref(self.consts.to_c()),
                                           ctypes.byref(self.flags.to_c())))
        self.n = 0

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int) -> int:
        if rc < 0:
            raise NtlpError(rc, self.lib.ntlp_gpu_last_error().decode())
        return rc

    def close(self) -> None:
        self._check(self.lib.ntlp_gpu_finalize())

    def set_dt(self, dt: float) -> None:
        self.consts.dt = dt
        self._check(self.lib.ntlp_gpu_set_dt(dt))

    # --------------------------------------------------------------- fields
    def set_fields(self, u, v, w, T, q, layout: int = LAYOUT_CORE) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, T, q)]
        self._check(self.lib.ntlp_gpu_set_fields(*[_ptr(a) for a in arrs], layout))

    # ------------------------------------------------------------- droplets
    def put(self, p: Particles) -> None:
        soa = np.ascontiguousarray(p.soa, dtype=np.float64)
        mult = np.ascontiguousarray(p.mult, dtype=np.int64)
        pidx = np.ascontiguousarray(p.pidx, dtype=np.int32)
        procidx = np.ascontiguousarray(p.procidx, dtype=np.int32)
        self._check(self.lib.ntlp_gpu_particles_put(_ptr(soa), _ptr(mult), _ptr(pidx), _ptr(procidx), p.n))
        self.n = p.n

    def get(self) -> Particles:
        n = self.n
        p = Particles(np.empty((NREAL, n)), np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int32),
                      np.empty(n, dtype=np.int32))
        self._check(self.lib.ntlp_gpu_particles_get(_ptr(p.soa), _ptr(p.mult), _ptr(p.pidx), _ptr(p.procidx), n))
        return p

    def generate(self, ens, n) -> None:
        """Droplets drawn on the device (ntlp_gpu_particles_generate), replacing
        those there: n droplets of the Ensemble ens, or, with lists, n[k]
        droplets of ens[k] one ensemble after the other; vp = uf, Tp = Tf and
        rc set as create_particle does."""
        if isinstance(ens, Ensemble):
            ens, n = [ens], [n]
        arr = (_CEnsemble * len(ens))(*[e.to_c() for e in ens])
        counts = np.ascontiguousarray(n, dtype=np.int64)
        self._check(self.lib.ntlp_gpu_particles_generate(ctypes.cast(arr, ctypes.c_void_p), _ptr(counts), len(ens)))
        self.n = int(counts.sum())

    def column(self, name: str) -> np.ndarray:
        """One record column of every droplet, in put order."""
        out = np.empty(self.n)
        self._check(self.lib.ntlp_gpu_particles_get_column(COL[name], _ptr(out), self.n))
        return out

    def save_particles(self, path: str) -> None:
        """save_particles (particles.f90:1988-2046): int32 count + 352-byte records."""
        self._check(self.lib.ntlp_gpu_save_particles(os.fsencode(path)))

    def read_part_res(self, path: str) -> int:
        """read_part_res (particles.f90:2048-2146): returns the droplets kept."""
        k = ctypes.c_int64()
        self._check(self.lib.ntlp_gpu_read_part_res(os.fsencode(path), ctypes.byref(k)))
        self.n = int(k.value)
        return self.n

    def put_records(self, records: np.ndarray) -> None:
        rec = np.ascontiguousarray(records)
        n = rec.nbytes // 352
        self._check(self.lib.ntlp_gpu_particles_put_records(_ptr(rec), n))
        self.n = n

    def get_records(self) -> np.ndarray:
        rec = np.empty(self.n * 352, dtype=np.uint8)
        self._check(self.lib.ntlp_gpu_particles_get_records(_ptr(rec), self.n))
        return rec

    # ----------------------------------------------- reference entry points
    def interp(self, set_vp: bool = False, set_tp: bool = False) -> None:
        """uf_interp_lin / uf_interp for every droplet (particles.f90:554-847)."""
        self._check(self.lib.ntlp_gpu_interp(int(set_vp), int(set_tp)))

    def crit_radius(self) -> None:
        """rc = crit_radius(m_s, kappa_s, Tp) for every droplet (particles.f90:2264-2266)."""
        self._check(self.lib.ntlp_gpu_crit_radius())

    def particle_update_rk3(self, istage: int, it: int = 2, precision: int = PREC_FP64) -> None:
        self._check(self.lib.ntlp_gpu_particle_update_rk3_prec(istage, it, precision))

    def particle_update_BE(self, it: int = 2, precision: int = PREC_FP64) -> None:
        self._check(self.lib.ntlp_gpu_particle_update_be_prec(it, precision))

    def particle_coupling_update(self, deterministic: int = 0, out: bool = True):
        """Returns partsrc_t(:,:,:,1:3), partTsrc_t, partHsrc_t, partTEsrc_t as
        an array (6, nnx+1, nny+1, nnz+2) (C order; z fastest), or None with
        out=False (the sources stay on the device).  deterministic picks the
        summation: 0 (or False) LDS node boxes, 1 (or True) a fixed order per
        node, 2 wavefront-segmented run sums with global atomics."""
        g = self.grid
        arr = np.empty((6, g.nnx + 1, g.nny + 1, g.nnz + 2)) if out else None
        self._check(self.lib.ntlp_gpu_particle_coupling_update(_ptr(arr), int(deterministic)))
        return arr

    def sources(self) -> np.ndarray:
        """The resident coupling sources, (6, nnx+1, nny+1, nnz+2)."""
        g = self.grid
        out = np.empty((6, g.nnx + 1, g.nny + 1, g.nnz + 2))
        self._check(self.lib.ntlp_gpu_get_sources(_ptr(out)))
        return out

    def particle_coupling_exchange(self):
        """particle_coupling_exchange (particles.f90:1066-1149): the folded
        sources in the flow solver's x-pencil layout, as C arrays
        partsrc (3, nnz+2, nny, nnx) and partTsrc, partHsrc, partTEsrc
        (nnz+2, nny, nnx)."""
        g = self.grid
        partsrc = np.empty((3, g.nnz + 2, g.nny, g.nnx))
        rest = [np.empty((g.nnz + 2, g.nny, g.nnx)) for _ in range(3)]
        self._check(self.lib.ntlp_gpu_particle_coupling_exchange(_ptr(partsrc), *[_ptr(a) for a in rest]))
        return (partsrc, *rest)

    def apply_particle_coupling(self, T: np.ndarray, q: np.ndarray, layout: int = LAYOUT_XPENCIL) -> None:
        """apply_particle_coupling (particles.f90:1278-1313) on T and q in place
        (x-pencil C arrays (nnz+2, nny, nnx+2), or core (nnx, nny, nnz+2)); the
        resident T and q become the results."""
        g = self.grid
        shape = (g.nnz + 2, g.nny, g.nnx + 2) if layout == LAYOUT_XPENCIL else (g.nnx, g.nny, g.nnz + 2)
        for a in (T, q):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape != shape:
                raise ValueError(f"T and q must be contiguous float64 arrays of shape {shape}")
        self._check(self.lib.ntlp_gpu_apply_particle_coupling(_ptr(T), _ptr(q), layout))

    # -------------------------------------------------------------- binning
    def bin(self):
        n, nc = self.n, int(self.lib.ntlp_gpu_num_ucells())
        cellidx = np.empty((n, 4), dtype=np.int32)
        counts = np.empty(nc, dtype=np.int32)
        offsets = np.empty(nc + 1, dtype=np.int32)
        perm = np.empty(n, dtype=np.int32)
        self._check(self.lib.ntlp_gpu_bin(_ptr(cellidx), _ptr(counts), _ptr(offsets), _ptr(perm)))
        return cellidx, counts, offsets, perm

    def wcell_counts(self) -> np.ndarray:
        g = self.grid
        out = np.empty((g.nnz, g.nny, g.nnx), dtype=np.int32)
        self._check(self.lib.ntlp_gpu_wcell_counts(_ptr(out)))
        return out

    # ----------------------------------------------------------- statistics
    def calc_tauc_min(self) -> float:
        """calc_tauc_min (particles.f90:3310-3368)."""
        v = ctypes.c_double()
        self._check(self.lib.ntlp_gpu_calc_tauc_min(ctypes.byref(v)))
        return float(v.value)

    def particle_xy_stats(self):
        """particle_xy_stats (particles.f90:3370-3770) before its divisions:
        (s0[21], s1 (24, nnz+2), lwp (nny, nnx))."""
        g = self.grid
        s0, s1, lwp = np.empty(21), np.empty((24, g.nnz + 2)), np.empty((g.nny, g.nnx))
        self._check(self.lib.ntlp_gpu_particle_xy_stats(_ptr(s0), _ptr(s1), _ptr(lwp)))
        return s0, s1, lwp

    # ------------------------------------------------------------ stepping
    def step(self, mode: int = MODE_BE, precision: int = PREC_FP64, nsteps: int = 1, it0: int = 2) -> None:
        self._check(self.lib.ntlp_gpu_step(mode, precision, nsteps, it0))

    def be_flux(self) -> np.ndarray:
        """Flux statistics of the last particle_update_BE step
        (particles.f90:3208-3238): (3, nnz+1) = pflux, pmassflux, penegflux
        over w levels 0..nnz."""
        out = np.empty((3, self.grid.nnz + 1))
        self._check(self.lib.ntlp_gpu_be_flux(_ptr(out[0]), _ptr(out[1]), _ptr(out[2])))
        return out

    def counters(self, reset: bool = False) -> dict:
        c = _CCounters()
        self._check(self.lib.ntlp_gpu_counters(ctypes.byref(c), int(reset)))
        d = {n: int(getattr(c, n)) for n in
             ("num100", "numimpos", "actnum", "denum", "gn_iters", "lm_iters", "solves", "steps")}
        d["gn_hist"] = [int(v) for v in c.gn_hist]
        return d

    def step_counters(self, max_rows: int = 4095) -> np.ndarray:
        rows = np.zeros((max_rows, STEP_ROW), dtype=np.int64)
        nr = self._check(self.lib.ntlp_gpu_step_counters(_ptr(rows), max_rows))
        return rows[:nr]

    def timing(self, enable: bool) -> None:
        self._check(self.lib.ntlp_gpu_timing(int(enable)))

    def copy_bandwidth(self, gib: float = 2.0, reps: int = 10) -> float:
        """GB/s (read + write) of a plain device copy of `gib` GiB: the
        achievable HBM bandwidth the kernels are compared with."""
        out = np.zeros(1)
        self._check(self.lib.ntlp_gpu_copy_bandwidth(int(gib * (1 << 30)), int(reps), _ptr(out)))
        return float(out[0])

    def kernel_times(self) -> dict:
        cap = 64
        names = ctypes.create_string_buffer(32 * cap)
        launches = np.zeros(cap, dtype=np.int64)
        total = np.zeros(cap, dtype=np.float64)
        k = self._check(self.lib.ntlp_gpu_kernel_times(ctypes.cast(names, ctypes.c_void_p), _ptr(launches),
                                                       _ptr(total), cap))
        out = {}
        for i in range(k):
            nm = names.raw[32 * i:32 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (int(launches[i]), float(total[i]))
        return out

    def set_option(self, option: int, value: int) -> None:
        """ntlp_gpu_set_option: OPT_STAGED_GATHER 1 stages each workgroup's
        field nodes in LDS before the trilinear gathers (0, the default:
        gathers through the caches)."""
        self._check(self.lib.ntlp_gpu_set_option(int(option), int(value)))

    def sync(self) -> None:
        self._check(self.lib.ntlp_gpu_sync())


# serving request columns (in) and result columns (out), include/ntlp_gpu.h
SERVE_IN = ["x", "y", "z", "radius", "m_s", "kappa_s"]
SERVE_OUT = ["x", "y", "z", "vx", "vy", "vz", "Tp", "radius", "m_s", "kappa_s", "rc", "res", "actres", "numact"]
SERVE_BYTES_PER_DROPLET = 436


class ServeLoop:
    """The continuous-batching serving loop (ntlp_serve_*) over the fields
    resident in an open :class:`DropletStep`.

    Every request advances ``steps`` steps of ``mode`` (MODE_BE or MODE_RK3).
    ``submit(inp, mult, out, out_mult)`` queues one droplet ensemble --
    ``inp`` is (6, n) FP64 in ``SERVE_IN`` order, ``out`` (14, n) FP64 receives
    ``SERVE_OUT`` -- and returns a ticket at once; the arrays must stay alive
    until ``wait(ticket)`` returns (latency ms, queueing ms).  Pinned host
    arrays make the copies asynchronous DMA."""

    def __init__(self, step: DropletStep, hbm_budget_bytes: int, steps: int = 10, precision: int = PREC_MIXED,
                 use_graph: bool = True, mode: int = MODE_BE):
        self.step = step
        self.lib = step.lib
        cfg = _CServeConfig(int(hbm_budget_bytes), int(steps), int(precision), int(use_graph), int(mode))
        cap = ctypes.c_int64()
        self._check(self.lib.ntlp_serve_open(ctypes.byref(cfg), ctypes.byref(cap)))
        self.capacity = int(cap.value)
        self._held = {}

    def _check(self, rc: int) -> int:
        return self.step._check(rc)

    def submit(self, inp: np.ndarray, mult: np.ndarray, out: np.ndarray, out_mult: np.ndarray) -> int:
        """inp (6, n) and out (14, n) may be column slices of wider arrays (rows
        contiguous, any row stride); mult and out_mult are contiguous."""
        n = int(inp.shape[1])
        for a, shape in ((inp, (6, n)), (out, (14, n))):
            if a.shape != shape or a.dtype != np.float64 or a.strides[1] != 8 or a.strides[0] % 8:
                raise ValueError(f"serving arrays must be float64 of shape {shape} with contiguous rows")
        for a in (mult, out_mult):
            if a.shape != (n,) or a.dtype != np.int64 or not a.flags.c_contiguous:
                raise ValueError(f"multiplicities must be contiguous int64 of shape ({n},)")
        t = ctypes.c_int64()
        self._check(self.lib.ntlp_serve_submit(_ptr(inp), inp.strides[0] // 8, _ptr(mult), n, _ptr(out),
                                               out.strides[0] // 8, _ptr(out_mult), ctypes.byref(t)))
        self._held[t.value] = (inp, mult, out, out_mult)
        return int(t.value)

    def submit_ensemble(self, ens: Ensemble, n: int, out: np.ndarray, out_mult: np.ndarray) -> int:
        """Queues n droplets of `ens`, drawn on the device at admission (no
        upload); out (14, n) and out_mult as for submit."""
        if out.shape != (14, n) or out.dtype != np.float64 or out.strides[1] != 8 or out.strides[0] % 8:
            raise ValueError(f"serving output must be float64 of shape (14, {n}) with contiguous rows")
        if out_mult.shape != (n,) or out_mult.dtype != np.int64 or not out_mult.flags.c_contiguous:
            raise ValueError(f"multiplicities must be contiguous int64 of shape ({n},)")
        t = ctypes.c_int64()
        self._check(self.lib.ntlp_serve_submit_ensemble(ctypes.byref(ens.to_c()), int(n), _ptr(out),
                                                        out.strides[0] // 8, _ptr(out_mult), ctypes.byref(t)))
        self._held[t.value] = (out, out_mult)
        return int(t.value)

    def poll(self, ticket: int) -> bool:
        return self._check(self.lib.ntlp_serve_poll(int(ticket))) == 1

    def wait(self, ticket: int) -> tuple[float, float]:
        times = np.zeros(2)
        self._check(self.lib.ntlp_serve_wait(int(ticket), _ptr(times)))
        self._held.pop(ticket, None)
        return float(times[0]), float(times[1])

    @staticmethod
    def _totals(t: _CServeTotals, newton: bool) -> dict:
        d = {n: int(getattr(t, n)) for n in ("submitted", "completed", "queued", "iterations", "droplet_steps",
                                             "max_pool", "graph_launches", "graphs", "capacity")}
        if newton:
            c = t.newton
            d["newton"] = {n: int(getattr(c, n)) for n in
                           ("num100", "numimpos", "actnum", "denum", "gn_iters", "lm_iters", "solves", "steps")}
        d["engine_ms"] = {n[3:]: float(getattr(t, n)) for n in
                          ("ms_engine", "ms_copy_calls", "ms_wait_download", "ms_wait_upload", "ms_wait_control")}
        return d

    def stats(self) -> dict:
        """Totals so far (the engine's counters as of its last iteration;
        no Newton counts until close)."""
        t = _CServeTotals()
        self._check(self.lib.ntlp_serve_stats(ctypes.byref(t)))
        return self._totals(t, newton=False)

    def close(self) -> dict:
        t = _CServeTotals()
        self._check(self.lib.ntlp_serve_close(ctypes.byref(t)))
        self._held.clear()
        return self._totals(t, newton=True)


class PinnedArray:
    """Page-locked host memory from the library (hipHostMalloc), viewed as a
    numpy array; freed by :meth:`free` or when garbage collected."""

    def __init__(self, shape, dtype=np.float64):
        self.lib = load_library()
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        rc = self.lib.ntlp_host_alloc(nbytes, ctypes.byref(p))
        if rc != 0:
            raise NtlpError(rc, self.lib.ntlp_gpu_last_error().decode())
        self._p = p
        buf = (ctypes.c_char * max(nbytes, 1)).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self) -> None:
        if self._p is not None:
            self.array = None
            self.lib.ntlp_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
