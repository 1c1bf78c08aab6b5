"""Writes part_res_7.bin: a restart file in the layout save_particles writes
(particles.f90:1988-2046) and read_part_res reads (2048-2146) -- an int32
droplet count, then one record per droplet with the MPI datatype of
particle_setup (particles.f90:1929-1950): 4 integer(4), 39 real(8) and 3
integer(8), i.e. the members of `type particle` (particles.f90:76-85) in
declaration order:

  integer :: pidx, procidx, nbr_pidx, nbr_procidx
  real :: vp(3), xp(3), uf(3), xrhs(3), vrhs(3), Tp, Tprhs_s
  real :: Tprhs_L, Tf, radius, radrhs, qinf, qstar, dist
  real :: res, m_s, kappa_s, rc, actres, numact
  real :: u_sub(3), sigm_s
  real :: vp_old(3), Tp_old, radius_old
  integer*8 :: mult
  type(particle), pointer :: prev, next

The dtype below is built from that declaration alone.  Every real member of
every record gets a distinct value (record i, member slot k -> (i+1)*1000 +
k + 0.25), so a reader that takes a member from the wrong offset fails.  Two
records lie outside the domain 0 < x < xl = 0.32, 0 < y < yl = 0.3 of the
16x12 test grid (one at x = xl exactly, one at y < 0) and must be dropped.

Run: python tests/golden/make_restart_fixture.py
"""
import os

import numpy as np

PARTICLE = np.dtype([
    ("pidx", "<i4"), ("procidx", "<i4"), ("nbr_pidx", "<i4"), ("nbr_procidx", "<i4"),
    ("vp", "<f8", 3), ("xp", "<f8", 3), ("uf", "<f8", 3), ("xrhs", "<f8", 3), ("vrhs", "<f8", 3),
    ("Tp", "<f8"), ("Tprhs_s", "<f8"), ("Tprhs_L", "<f8"), ("Tf", "<f8"), ("radius", "<f8"), ("radrhs", "<f8"),
    ("qinf", "<f8"), ("qstar", "<f8"), ("dist", "<f8"),
    ("res", "<f8"), ("m_s", "<f8"), ("kappa_s", "<f8"), ("rc", "<f8"), ("actres", "<f8"), ("numact", "<f8"),
    ("u_sub", "<f8", 3), ("sigm_s", "<f8"),
    ("vp_old", "<f8", 3), ("Tp_old", "<f8"), ("radius_old", "<f8"),
    ("mult", "<i8"), ("prev", "<i8"), ("next", "<i8"),
])
assert PARTICLE.itemsize == 352

XL, YL = 16 * 0.02, 12 * 0.025


def records(n=7):
    rec = np.zeros(n, dtype=PARTICLE)
    for i in range(n):
        k = 0
        for name in PARTICLE.names:
            shape = PARTICLE.fields[name][0].shape
            if PARTICLE.fields[name][0].base.kind != "f":
                continue
            size = int(np.prod(shape)) if shape else 1
            vals = (i + 1) * 1000.0 + k + np.arange(size) + 0.25
            rec[name][i] = vals if shape else vals[0]
            k += size
        rec["pidx"][i], rec["procidx"][i] = 100 + i, 7
        rec["nbr_pidx"][i], rec["nbr_procidx"][i] = -1, -1
        rec["mult"][i] = 10 ** (i + 3)
        rec["prev"][i], rec["next"][i] = 0x1111 * (i + 1), 0x2222 * (i + 1)
        # positions inside the domain (z anywhere: read_part_res tests only x, y)
        rec["xp"][i] = (0.01 + 0.04 * i, 0.02 + 0.035 * i, 0.05 + 0.01 * i)
    rec["xp"][5, 0] = XL          # x = xl: outside (strict <)
    rec["xp"][6, 1] = -1e-3       # y < 0: outside
    return rec


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "part_res_7.bin")
    rec = records()
    with open(path, "wb") as fh:
        fh.write(np.int32(len(rec)).tobytes())
        fh.write(rec.tobytes())
    print(f"wrote {path}: {4 + rec.nbytes} bytes")


if __name__ == "__main__":
    main()
