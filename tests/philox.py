"""Host restatement of the droplet-ensemble draws (test infrastructure): the
Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC 2011)
vectorised in numpy, and the uniforms built from it the way
ntlp_amd/csrc/ensemble.hpp builds them, so tests can recompute a device
ensemble.  Pinned by the published known-answer vectors of Random123
(tests/test_ensemble_draws.py)."""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Ten rounds on arrays of 32-bit counter words under the key (k0, k1)."""
    c = [np.asarray(v, dtype=np.uint32) for v in (c0, c1, c2, c3)]
    k0, k1 = np.uint32(k0), np.uint32(k1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c[0].astype(np.uint64)
            p1 = M1 * c[2].astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            k0 = np.uint32(k0 + W0)
            k1 = np.uint32(k1 + W1)
    return c


def u01(a, b):
    m = ((a.astype(np.uint64) >> np.uint64(5)) << np.uint64(26)) | (b.astype(np.uint64) >> np.uint64(6))
    return (m.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def droplet_draws(seed: int, first: int, n: int) -> np.ndarray:
    """u[8, n]: the eight uniforms of droplets first .. first+n-1."""
    i = np.arange(first, first + n, dtype=np.uint64)
    lo, hi = (i & MASK).astype(np.uint32), (i >> np.uint64(32)).astype(np.uint32)
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    u = np.empty((8, n))
    zero = np.zeros(n, dtype=np.uint32)
    for b in range(4):
        w = philox4x32_10(lo, hi, zero, np.full(n, b, dtype=np.uint32), k0, k1)
        u[2 * b] = u01(w[0], w[1])
        u[2 * b + 1] = u01(w[2], w[3])
    return u
