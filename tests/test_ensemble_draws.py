"""The host restatement of the ensemble generator's Philox4x32-10 against the
published known-answer vectors of Random123 (kat_vectors, philox4x32 with 10
rounds), and the uniforms it yields."""
import numpy as np

from .philox import droplet_draws, philox4x32_10

KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        got = philox4x32_10(*[np.array([v], dtype=np.uint32) for v in ctr], *key)
        assert tuple(int(w[0]) for w in got) == want


def test_uniforms_are_open_unit_and_uniform():
    u = droplet_draws(seed=12345, first=0, n=200_000)
    assert u.min() > 0.0 and u.max() < 1.0
    # mean and variance of U(0,1) within 5 sigma of the sampling error
    assert np.all(np.abs(u.mean(axis=1) - 0.5) < 5 * np.sqrt(1 / 12 / u.shape[1]))
    # different seeds and counters give different streams; the same give the same
    assert not np.array_equal(u, droplet_draws(seed=12346, first=0, n=200_000))
    assert np.array_equal(droplet_draws(12345, 1000, 10), u[:, 1000:1010])
