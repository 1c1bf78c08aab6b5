"""Grid metadata of the droplet path.

Builds the vertical levels the way the reference's setup does: w levels
z(0:nnz+1) (les.F:5465-5482), then get_dz (les.F:5836-5874) derives
dzw(0:nnz+2), u levels zz(0:nnz+1) with zz(0) = -zz(1), and dzu(0:nnz+2);
dx = xl/nnx, dy = yl/nny (les.F:5488-5489).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np


@dataclass
class Grid:
    nnx: int
    nny: int
    nnz: int
    xl: float
    yl: float
    z: np.ndarray    # (nnz+2,)  w levels 0..nnz+1
    zz: np.ndarray   # (nnz+2,)  u levels 0..nnz+1
    dzw: np.ndarray  # (nnz+3,)
    dzu: np.ndarray  # (nnz+3,)

    @property
    def dx(self) -> float:
        return self.xl / self.nnx

    @property
    def dy(self) -> float:
        return self.yl / self.nny

    @property
    def zl(self) -> float:
        return float(self.z[self.nnz])

    @property
    def ncell(self) -> int:
        return self.nnx * self.nny * self.nnz

    def to_c(self, struct):
        c = struct()
        c.nnx, c.nny, c.nnz, c.reserved = self.nnx, self.nny, self.nnz, 0
        c.dx, c.dy, c.xl, c.yl = self.dx, self.dy, self.xl, self.yl
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (self.z, self.zz, self.dzw, self.dzu)]
        c.z, c.zz, c.dzw, c.dzu = [a.ctypes.data_as(ctypes.c_void_p) for a in self._keep]
        return c

    # ------------------------------------------------------------ builders
    @staticmethod
    def from_w_levels(nnx: int, nny: int, xl: float, yl: float, z: np.ndarray) -> "Grid":
        """get_dz (les.F:5836-5874) on w levels z(0:nnz+1)."""
        z = np.asarray(z, dtype=np.float64)
        nnz = len(z) - 2
        dzw = np.empty(nnz + 3)
        for iz in range(1, nnz + 2):
            dzw[iz] = z[iz] - z[iz - 1]
        dzw[0] = dzw[1]
        dzw[nnz + 2] = dzw[nnz + 1]
        zz = np.empty(nnz + 2)
        for iz in range(1, nnz + 2):
            zz[iz] = 0.5 * (z[iz] + z[iz - 1])
        zz[0] = -zz[1]
        dzu = np.empty(nnz + 3)
        for iz in range(1, nnz + 2):
            dzu[iz] = zz[iz] - zz[iz - 1]
        dzu[0] = dzu[1]
        dzu[nnz + 2] = dzu[nnz + 1]
        return Grid(nnx, nny, nnz, float(xl), float(yl), z, zz, dzw, dzu)

    @staticmethod
    def uniform(nnx: int, nny: int, nnz: int, dx: float, dy: float, dz: float) -> "Grid":
        """iz_space = 0: z(iz) = dz*float(iz) (les.F:5474-5477)."""
        z = np.array([dz * float(iz) for iz in range(nnz + 2)])
        return Grid.from_w_levels(nnx, nny, nnx * dx, nny * dy, z)

    @staticmethod
    def tanh_stretched(nnx: int, nny: int, nnz: int, dx: float, dy: float, dz_min: float, dz_max: float,
                       gamma: float = 4.0) -> "Grid":
        """Vertical spacing rising from dz_min at the bottom to dz_max at the
        top along a normalised tanh: dz_k = dz_min + (dz_max - dz_min) s(xi),
        s(xi) = (tanh(gamma (xi - 1/2)) + tanh(gamma/2)) / (2 tanh(gamma/2)),
        xi = (k - 1/2)/nnz; the ghost level nnz+1 repeats the top spacing."""
        xi = (np.arange(1, nnz + 1) - 0.5) / nnz
        s = (np.tanh(gamma * (xi - 0.5)) + np.tanh(gamma / 2)) / (2 * np.tanh(gamma / 2))
        dz = dz_min + (dz_max - dz_min) * s
        z = np.zeros(nnz + 2)
        for k in range(1, nnz + 1):
            z[k] = z[k - 1] + dz[k - 1]
        z[nnz + 1] = z[nnz] + dz[-1]
        return Grid.from_w_levels(nnx, nny, nnx * dx, nny * dy, z)
