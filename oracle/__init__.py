"""Python access to the oracle (oracle/ntlp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

NREAL = 34


class _G(ctypes.Structure):
    _fields_ = [("nnx", ctypes.c_int), ("nny", ctypes.c_int), ("nnz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("xl", ctypes.c_double), ("yl", ctypes.c_double),
                ("z", ctypes.c_void_p), ("zz", ctypes.c_void_p), ("dzw", ctypes.c_void_p), ("dzu", ctypes.c_void_p)]


_CS = ["dt", "nuf", "Cpa", "Cpp", "CpaCpp", "Pra", "Sc", "Rd", "rhow", "rhos", "Mw", "Ru", "Gam", "Lv", "Cpv",
       "surf_rho", "surf_p", "tsfcc1", "grav", "radius_init", "pi2"]


class _C(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in _CS] + [("part_grav", ctypes.c_double * 3),
                                                       ("zetas", ctypes.c_double * 3), ("gama", ctypes.c_double * 3)]


class _F(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("ilin", "icouple", "iTcouple", "iHcouple", "ievap", "iexner", "icase")]


class _Fl(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("u", "v", "w", "T", "q")]


class _Cnt(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("num100", "numimpos", "actnum", "denum", "gn_iters", "lm_iters", "solves")]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build_oracle
            build_oracle()
        _lib = ctypes.CDLL(LIB)
        vp, i64, i32, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        for name, res, args in [
            ("or_mod_magnus", dp, [dp]),
            ("or_exner", dp, [dp, dp, vp]),
            ("or_func_p_base", dp, [dp, dp, dp, vp]),
            ("or_cell_index_many", None, [vp, vp, vp, vp, i64, vp]),
            ("or_num_ucells", i64, [vp]),
            ("or_bin", None, [vp, vp, vp, vp, i64, vp, vp, vp, vp]),
            ("or_wcell_counts", None, [vp, vp, vp, vp, i64, vp]),
            ("or_fill_ext", None, [vp, vp, vp]),
            ("or_interp", None, [vp, vp, vp, vp, vp, i64]),
            ("or_crit_radius", dp, [vp, dp, dp, dp]),
            ("or_crit_radius_many", None, [vp, vp, vp, vp, i64, vp]),
            ("or_update_rk3", None, [vp, vp, vp, vp, vp, i64, i32, i32]),
            ("or_update_be", None, [vp, vp, vp, vp, vp, i64, i32, vp]),
            ("or_update_be_flux", None, [vp, vp, vp, vp, vp, vp, i64, i32, vp, vp]),
            ("or_coupling", None, [vp, vp, vp, vp, vp, i64, vp]),
            ("or_update_be_par", None, [vp, vp, vp, vp, vp, i64, i32, vp]),
            ("or_coupling_par", None, [vp, vp, vp, vp, vp, i64, vp]),
            ("or_max_threads", i32, []),
            ("or_ie_vrt_nd", None, [vp, vp, dp, vp, dp, dp, dp, vp]),
            ("or_rad_solver2", dp, [vp, vp, dp, vp]),
            ("or_gauss_newton_2d", None, [vp, vp, vp, dp, dp, vp, vp, vp, vp]),
            ("or_lv_solver", None, [vp, vp, vp, dp, dp, vp, vp, vp, vp]),
            ("or_coupling_exchange", None, [vp, vp, vp]),
            ("or_calc_tauc_min", dp, [vp, vp, vp, vp, i64]),
            ("or_xy_stats", None, [vp, vp, vp, vp, i64, vp, vp, vp]),
            ("or_apply_coupling", None, [vp, vp, vp, vp, vp, vp]),
        ]:
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


class Oracle:
    """FP64 CPU restatement of the reference droplet routines on one domain."""

    def __init__(self, grid, consts, flags):
        self.grid = grid
        self.L = lib()
        self._arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (grid.z, grid.zz, grid.dzw, grid.dzu)]
        g = _G()
        g.nnx, g.nny, g.nnz = grid.nnx, grid.nny, grid.nnz
        g.dx, g.dy, g.xl, g.yl = grid.dx, grid.dy, grid.xl, grid.yl
        g.z, g.zz, g.dzw, g.dzu = [_p(a) for a in self._arr]
        self.g = g
        self.set_consts(consts)
        f = _F()
        for n in ("ilin", "icouple", "iTcouple", "iHcouple", "ievap", "iexner", "icase"):
            setattr(f, n, int(getattr(flags, n)))
        self.f = f
        self.ext = None

    def set_consts(self, consts):
        c = _C()
        for n in _CS:
            setattr(c, n, float(getattr(consts, n)))
        c.part_grav = (ctypes.c_double * 3)(*consts.part_grav)
        c.zetas = (ctypes.c_double * 3)(*consts.zetas)
        c.gama = (ctypes.c_double * 3)(*consts.gama)
        self.c = c

    def exner(self, p0: float, p: float) -> float:
        return self.L.or_exner(p0, p, ctypes.byref(self.c))

    def p_base(self, p_surf: float, T_surf: float, z: float) -> float:
        return self.L.or_func_p_base(p_surf, T_surf, z, ctypes.byref(self.c))

    # fields --------------------------------------------------------------
    def fill_ext(self, core: np.ndarray) -> np.ndarray:
        g = self.grid
        core = np.ascontiguousarray(core, dtype=np.float64)
        ext = np.empty((g.nnx + 5, g.nny + 5, g.nnz + 2))
        self.L.or_fill_ext(ctypes.byref(self.g), _p(core), _p(ext))
        return ext

    def set_fields(self, u, v, w, T, q, ext: bool = False):
        self.ext = [np.ascontiguousarray(a, dtype=np.float64) if ext else self.fill_ext(a) for a in (u, v, w, T, q)]
        fl = _Fl()
        fl.u, fl.v, fl.w, fl.T, fl.q = [_p(a) for a in self.ext]
        self.fl = fl

    # droplets ------------------------------------------------------------
    def interp(self, soa: np.ndarray):
        self.L.or_interp(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), ctypes.byref(self.fl),
                         _p(soa), soa.shape[1])

    def crit_radius(self, soa: np.ndarray, col):
        n = soa.shape[1]
        out = np.empty(n)
        ms, ka, tp = (np.ascontiguousarray(soa[col[k]]) for k in ("m_s", "kappa_s", "Tp"))
        self.L.or_crit_radius_many(ctypes.byref(self.c), _p(ms), _p(ka), _p(tp), n, _p(out))
        return out

    def update_rk3(self, soa: np.ndarray, istage: int, it: int):
        self.L.or_update_rk3(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), ctypes.byref(self.fl),
                             _p(soa), soa.shape[1], istage, it)

    def update_be(self, soa: np.ndarray, it: int) -> dict:
        cnt = _Cnt()
        self.L.or_update_be(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), ctypes.byref(self.fl),
                            _p(soa), soa.shape[1], it, ctypes.byref(cnt))
        return {n: int(getattr(cnt, n)) for n, _ in _Cnt._fields_}

    def update_be_flux(self, soa: np.ndarray, mult: np.ndarray, it: int):
        """update_be plus the step's flux statistics: (counters, flux (3, nnz+1)
        = pflux, pmassflux, penegflux over w levels 0..nnz)."""
        cnt = _Cnt()
        mult = np.ascontiguousarray(mult, dtype=np.int64)
        flux = np.zeros((3, self.grid.nnz + 1))
        self.L.or_update_be_flux(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f),
                                 ctypes.byref(self.fl), _p(soa), _p(mult), soa.shape[1], it, ctypes.byref(cnt),
                                 _p(flux))
        return {n: int(getattr(cnt, n)) for n, _ in _Cnt._fields_}, flux

    def update_be_par(self, soa: np.ndarray, it: int) -> dict:
        """update_be over the host's cores (OpenMP): CPU timing baseline."""
        cnt = _Cnt()
        self.L.or_update_be_par(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f),
                                ctypes.byref(self.fl), _p(soa), soa.shape[1], it, ctypes.byref(cnt))
        return {n: int(getattr(cnt, n)) for n, _ in _Cnt._fields_}

    def coupling_par(self, soa: np.ndarray, mult: np.ndarray) -> np.ndarray:
        """coupling over the host's cores (OpenMP, atomic deposits): CPU timing baseline."""
        g = self.grid
        out = np.empty((6, g.nnx + 1, g.nny + 1, g.nnz + 2))
        mult = np.ascontiguousarray(mult, dtype=np.int64)
        self.L.or_coupling_par(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), _p(soa), _p(mult),
                               soa.shape[1], _p(out))
        return out

    def max_threads(self) -> int:
        return int(self.L.or_max_threads())

    def coupling(self, soa: np.ndarray, mult: np.ndarray) -> np.ndarray:
        g = self.grid
        out = np.empty((6, g.nnx + 1, g.nny + 1, g.nnz + 2))
        mult = np.ascontiguousarray(mult, dtype=np.int64)
        self.L.or_coupling(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), _p(soa), _p(mult),
                           soa.shape[1], _p(out))
        return out

    def coupling_exchange(self, src: np.ndarray) -> np.ndarray:
        """Folds src (6, nnx+1, nny+1, nnz+2) in place and returns the x-pencil
        arrays (6, nnz+2, nny, nnx)."""
        g = self.grid
        assert src.shape == (6, g.nnx + 1, g.nny + 1, g.nnz + 2) and src.flags.c_contiguous
        out = np.empty((6, g.nnz + 2, g.nny, g.nnx))
        self.L.or_coupling_exchange(ctypes.byref(self.g), _p(src), _p(out))
        return out

    def apply_coupling(self, xsrc: np.ndarray, T: np.ndarray, q: np.ndarray) -> None:
        """T, q: x-pencil (nnz+2, nny, nnx+2) arrays, updated in place."""
        g = self.grid
        for a in (T, q):
            assert a.shape == (g.nnz + 2, g.nny, g.nnx + 2) and a.flags.c_contiguous and a.dtype == np.float64
        self.L.or_apply_coupling(ctypes.byref(self.g), ctypes.byref(self.c), ctypes.byref(self.f), _p(xsrc),
                                 _p(T), _p(q))

    def calc_tauc_min(self, soa: np.ndarray, mult: np.ndarray) -> float:
        mult = np.ascontiguousarray(mult, dtype=np.int64)
        return self.L.or_calc_tauc_min(ctypes.byref(self.g), ctypes.byref(self.c), _p(soa), _p(mult), soa.shape[1])

    def xy_stats(self, soa: np.ndarray, mult: np.ndarray):
        """(s0[21], s1 (24, nnz+2), lwp (nny, nnx)) of particle_xy_stats."""
        g = self.grid
        mult = np.ascontiguousarray(mult, dtype=np.int64)
        s0, s1, lwp = np.empty(21), np.empty((24, g.nnz + 2)), np.empty((g.nny, g.nnx))
        self.L.or_xy_stats(ctypes.byref(self.g), ctypes.byref(self.c), _p(soa), _p(mult), soa.shape[1], _p(s0),
                           _p(s1), _p(lwp))
        return s0, s1, lwp

    # binning -------------------------------------------------------------
    def cell_index(self, x, y, z) -> np.ndarray:
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        out = np.empty((len(x), 5), dtype=np.int32)
        self.L.or_cell_index_many(ctypes.byref(self.g), _p(x), _p(y), _p(z), len(x), _p(out))
        return out

    def bin(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        n = len(x)
        nc = int(self.L.or_num_ucells(ctypes.byref(self.g)))
        keys = np.empty(n, dtype=np.int32)
        counts = np.empty(nc, dtype=np.int32)
        offsets = np.empty(nc + 1, dtype=np.int32)
        perm = np.empty(n, dtype=np.int32)
        self.L.or_bin(ctypes.byref(self.g), _p(x), _p(y), _p(z), n, _p(keys), _p(counts), _p(offsets), _p(perm))
        return keys, counts, offsets, perm

    def wcell_counts(self, x, y, z) -> np.ndarray:
        g = self.grid
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        out = np.empty((g.nnz, g.nny, g.nnx), dtype=np.int32)
        self.L.or_wcell_counts(ctypes.byref(self.g), _p(x), _p(y), _p(z), len(x), _p(out))
        return out

    # single-droplet pieces -----------------------------------------------
    def ie_vrt_nd(self, rec, rhoa, vnext, tr, tt, h):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        vn = np.ascontiguousarray(vnext, dtype=np.float64)
        F = np.empty(2)
        self.L.or_ie_vrt_nd(ctypes.byref(self.c), _p(rec), rhoa, _p(vn), tr, tt, h, _p(F))
        return F

    def rad_solver2(self, rec, rhoa):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        mflag = ctypes.c_int()
        g = self.L.or_rad_solver2(ctypes.byref(self.c), _p(rec), rhoa, ctypes.byref(mflag))
        return g, mflag.value

    def _solver(self, fn, rec, vnext, h, rhoa, start):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        vn = np.ascontiguousarray(vnext, dtype=np.float64)
        v1 = np.ascontiguousarray(start, dtype=np.float64)
        v2 = np.zeros(2)
        flag, iters = ctypes.c_int(), ctypes.c_int()
        fn(ctypes.byref(self.c), _p(rec), _p(vn), h, rhoa, _p(v1), _p(v2), ctypes.byref(flag), ctypes.byref(iters))
        return v2, flag.value, iters.value

    def gauss_newton(self, rec, vnext, h, rhoa, start):
        return self._solver(self.L.or_gauss_newton_2d, rec, vnext, h, rhoa, start)

    def lv_solver(self, rec, vnext, h, rhoa, start):
        return self._solver(self.L.or_lv_solver, rec, vnext, h, rhoa, start)
