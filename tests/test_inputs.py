"""Counter-based inputs: Philox4x32-10 known-answer vectors (Random123,
Salmon et al. SC'11) and the properties of the ensemble generators."""
import numpy as np

from simplediffeq_hip import inputs


def test_philox4x32_10_known_answers():
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        got = inputs.philox4x32(*[np.array([c], dtype=np.uint64) for c in ctr], key[0], key[1])
        assert tuple(int(g[0]) for g in got) == want


def test_generators_ranges_and_subsets():
    u0, p = inputs.lorenz_wide(4096, seed=1)
    assert (p[0] >= 8).all() and (p[0] < 12).all() and (p[1] >= 0).all() and (p[1] < 21).all()
    assert (p[2] >= 2).all() and (p[2] < 3).all() and (np.abs(u0) <= 1).all()
    # any slice regenerates identically
    u0b, pb = inputs.lorenz_wide(100, seed=1, start=1000)
    assert np.array_equal(u0b, u0[:, 1000:1100]) and np.array_equal(pb, p[:, 1000:1100])
    u0, p, ts = inputs.lorenz_uneven(20000, seed=2)
    assert (ts[1] >= 0.5).all() and (ts[1] <= 20).all()
    assert abs(np.mean(np.log(ts[1])) - 0.5 * (np.log(0.5) + np.log(20))) < 0.05
    assert abs(np.mean(u0)) < 0.03 and abs(np.std(u0) - 1) < 0.03
    u0, r, A = inputs.glv16(64, seed=3)
    assert np.all(np.diag(A.reshape(16, 16)) == -1.0)
    off = A.reshape(16, 16)[~np.eye(16, dtype=bool)]
    assert abs(np.std(off) - 0.1) < 0.03
