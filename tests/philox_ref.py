"""Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
as easy as 1, 2, 3", SC'11) in numpy: the CPU check of the counter-based
generator the device emission uses (pinned by the published known-answer
vectors in test_philox.py)."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(v, np.uint64) & MASK for v in (c0, c1, c2, c3)]
    k0 = np.asarray(k0, np.uint64) & MASK
    k1 = np.asarray(k1, np.uint64) & MASK
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [(hi1 ^ c[1] ^ k0) & MASK, lo1, (hi0 ^ c[3] ^ k1) & MASK, lo0]
        k0 = (k0 + np.uint64(W0)) & MASK
        k1 = (k1 + np.uint64(W1)) & MASK
    return [v.astype(np.uint32) for v in c]


def uniform2(seed, pid, draw):
    """The two doubles in [0,1) the device forms for (seed, particle id, draw)."""
    pid = np.asarray(pid, np.uint64)
    r = philox4x32_10(pid & MASK, pid >> np.uint64(32), np.full(pid.shape, draw, np.uint64), np.zeros(pid.shape, np.uint64),
                      np.uint64(seed) & MASK, np.uint64(seed) >> np.uint64(32))
    r = [v.astype(np.uint64) for v in r]
    s = 1.0 / 9007199254740992.0
    u0 = (((r[0] << np.uint64(32)) | r[1]) >> np.uint64(11)).astype(np.float64) * s
    u1 = (((r[2] << np.uint64(32)) | r[3]) >> np.uint64(11)).astype(np.float64) * s
    return u0, u1
