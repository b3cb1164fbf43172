"""CPU restatement of the device initial-state generator (test
infrastructure only): Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC 2011)
in NumPy uint32/uint64 arithmetic with the device's counter layout
(particle index, draw block) and key (seed), the 53-bit uniform, and the
position / momentum transforms.  The reference draws its initial states from
the compiler's random_number (sources/random.F90:104-220) and has no bitwise
counterpart; the restatement pins the generator by its published
known-answer vectors (tests/test_init.py) and the distributions by their
moments.  Only tests/ import this module."""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (n, 4) uint32, key: (n, 2) uint32 -> (n, 4) uint32."""
    c = np.array(ctr, dtype=np.uint32).copy()
    k = np.array(key, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k[:, 0] += W0
                k[:, 1] += W1
            p0 = M0 * c[:, 0].astype(np.uint64)
            p1 = M1 * c[:, 2].astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
            c = np.stack([hi1 ^ c[:, 1] ^ k[:, 0], lo1, hi0 ^ c[:, 3] ^ k[:, 1], lo0], axis=1)
    return c


def draws(n: int, seed: int) -> np.ndarray:
    """(n, 8) uniforms in [0, 1) of particles 0..n-1."""
    i = np.arange(n, dtype=np.uint64)
    key = np.stack([np.full(n, seed & 0xFFFFFFFF), np.full(n, (seed >> 32) & 0xFFFFFFFF)],
                   1).astype(np.uint32)
    out = np.zeros((n, 8))
    for blk in range(4):
        ctr = np.stack([(i & MASK).astype(np.uint32), (i >> np.uint64(32)).astype(np.uint32),
                        np.full(n, blk, np.uint32), np.zeros(n, np.uint32)], 1)
        w = philox4x32_10(ctr, key).astype(np.uint64)
        for h in range(2):
            a, b = w[:, 2 * h], w[:, 2 * h + 1]
            out[:, 2 * blk + h] = (((a >> np.uint64(5)) << np.uint64(26)) |
                                   (b >> np.uint64(6))).astype(np.float64) / 9007199254740992.0
    return out


def initial_state(n, seed, kind, lo, hi, a, b=1.0):
    """(6, n) positions and momenta as the device generates them."""
    d = draws(n, seed)
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    s = np.zeros((6, n))
    for c in range(3):
        s[c] = lo[c] + (hi[c] - lo[c]) * d[:, c]
    two_pi = 2.0 * np.pi
    if kind == "maxwell":
        r0 = np.sqrt(-2.0 * np.log(1.0 - d[:, 3]))
        r1 = np.sqrt(-2.0 * np.log(1.0 - d[:, 5]))
        s[3] = a * r0 * np.cos(two_pi * d[:, 4])
        s[4] = a * r0 * np.sin(two_pi * d[:, 4])
        s[5] = a * r1 * np.cos(two_pi * d[:, 6])
    else:
        g = np.exp(np.log(a) + (np.log(b) - np.log(a)) * d[:, 3]) if kind == "gamma_log" else np.full(n, a)
        um = np.sqrt(np.maximum(g * g - 1.0, 0.0))
        ct = 2.0 * d[:, 4] - 1.0
        st = np.sqrt(np.maximum(1.0 - ct * ct, 0.0))
        ph = two_pi * d[:, 5]
        s[3], s[4], s[5] = um * st * np.cos(ph), um * st * np.sin(ph), um * ct
    return s
