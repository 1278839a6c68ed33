"""Shared input generators for the tests (seeded numpy generators)."""
import numpy as np


def band(rng, kl, ku, n, dtype=np.float64, dist="u", nan_corners=False, m=None):
    """(kl+ku+1) x n band array; with nan_corners the slots outside the m x n
    matrix hold NaN, so a kernel that reads them poisons its result."""
    m = n if m is None else m
    h = kl + ku + 1
    d = rng.random((h, n)) if dist == "u" else rng.standard_normal((h, n))
    d = np.asfortranarray(d.astype(dtype))
    if nan_corners:
        for j in range(n):
            for t in range(h):
                i = j - ku + t
                if i < 0 or i >= m:
                    d[t, j] = np.nan
    return d


def in_band_mask(m, n, l, u, hrows=None, shift=0):
    """Boolean mask over a band array marking slots inside the m x n matrix.
    shift = number of extra rows above (u already includes them)."""
    h = hrows if hrows is not None else l + u + 1
    i = np.arange(n)[None, :] - u + np.arange(h)[:, None]  # row of slot (t, j)
    return (i >= 0) & (i < m)


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
