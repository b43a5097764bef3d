"""Philox4x32-10 counter-based generator and the spectral-draw mapping of the
device path, restated in numpy (uint64 arithmetic).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Philox4x32-10 is the published generator of Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3" (SC'11): per round
  (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,
  c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),
with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the key bumped by
(0x9E3779B9, 0xBB67AE85) between rounds.  Pinned by the Random123 known-answer
vectors in tests/test_oracle.py.

The stream layout (word 0 = radial jitter, 1 + k = angle k, 1 + P + 2k and
2 + P + 2k = Box-Muller pair of amplitude k, counter = (rid_lo, rid_hi,
screen, word // 4), key = seed) is this project's own; the distributions are
those of grids.py:98-107 and phase_screens.py:98-103.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32-valued, key: (2,) -> (..., 4) uint64 holding uint32 words."""
    c = np.asarray(ctr, dtype=np.uint64) & MASK
    k0 = np.uint64(key[0]) & MASK
    k1 = np.uint64(key[1]) & MASK
    c0, c1, c2, c3 = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    for i in range(10):
        if i:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        c0, c1, c2, c3 = ((p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & MASK, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & MASK)
    return np.stack([c0, c1, c2, c3], axis=-1)


def stream_words(rid: int, screen: int, words: np.ndarray, seed: int) -> np.ndarray:
    words = np.asarray(words, dtype=np.uint64)
    ctr = np.stack([np.full_like(words, rid & 0xFFFFFFFF), np.full_like(words, (rid >> 32) & 0xFFFFFFFF),
                    np.full_like(words, screen), words >> np.uint64(2)], axis=-1)
    out = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    return np.take_along_axis(out, (words & np.uint64(3)).astype(np.int64)[..., None], axis=-1)[..., 0]


def u01(w):
    return (np.asarray(w, dtype=np.float64) + 0.5) * 2.0 ** -32


def spectral_draw(rid: int, screen: int, seed: int, base: np.ndarray, sqrt_psd: np.ndarray):
    """(rho f32, theta f32, value c64) of one screen of realization `rid`."""
    P = base.shape[0]
    k = np.arange(P, dtype=np.uint64)
    U = np.float32(u01(stream_words(rid, screen, np.array([0]), seed))[0])
    f = base.astype(np.float32)
    fp = np.concatenate([[np.float32(0)], f[:-1]]).astype(np.float32)
    f2 = f.astype(np.float64) * f.astype(np.float64)
    fp2 = fp.astype(np.float64) * fp.astype(np.float64)
    rho = np.sqrt(fp2 + np.float64(U) * (f2 - fp2)).astype(np.float32)
    theta = (2 * np.pi * u01(stream_words(rid, screen, 1 + k, seed))).astype(np.float32)
    u1 = u01(stream_words(rid, screen, 1 + P + 2 * k, seed))
    u2 = u01(stream_words(rid, screen, 2 + P + 2 * k, seed))
    R = np.sqrt(-2.0 * np.log(u1))
    nr = (R * np.cos(2 * np.pi * u2)).astype(np.float32)
    ni = (R * np.sin(2 * np.pi * u2)).astype(np.float32)
    value = (nr * sqrt_psd + 1j * (ni * sqrt_psd)).astype(np.complex64)
    return rho, theta, value
