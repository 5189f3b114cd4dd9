"""Oracle (test infrastructure only): numpy restatement of the learned tuner's
40 statistics of one sample block (block_features_kernel in
faz_amd/csrc/faz_tune.hip; listed in include/faz_gpu.h).  The statistics are
this build's own (the reference's tuner has no learned part), so this file
restates the kernel's definitions, and tests/test_oracle_features.py pins it
by closed forms (constant, linear, quadratic, checkerboard and AR(1)
blocks).  Differences are formed in float32 from u = (v - vmin) / rng in
float32, as the kernel does; sums run in float64.
"""
import numpy as np

STRIDES = (1, 2, 4)


def _shift(u, a, s):
    """(u[x + s e_a], u[x]) over the points with x_a + s < d_a."""
    n = u.shape[a]
    hi = [slice(None)] * u.ndim
    lo = [slice(None)] * u.ndim
    hi[a] = slice(s, n)
    lo[a] = slice(0, n - s)
    return u[tuple(hi)], u[tuple(lo)]


def _laplacian(u, s):
    """Sum over axes of u[x + s e_a] - 2 u[x] + u[x - s e_a] at the points
    interior along every axis (float32, added axis by axis)."""
    inner = tuple(slice(s, n - s) for n in u.shape)
    c = u[inner]
    L = np.zeros(c.shape, np.float32)
    for a in range(u.ndim):
        p = [slice(s, n - s) for n in u.shape]
        m = [slice(s, n - s) for n in u.shape]
        p[a] = slice(2 * s, u.shape[a])
        m[a] = slice(0, u.shape[a] - 2 * s)
        L = (L + ((u[tuple(p)] - np.float32(2.0) * c) + u[tuple(m)])).astype(np.float32)
    return L


def block_features(block, vmin, rng, eb_norm):
    """The 40 statistics (float32) of one block of shape (d0, d1[, d2])."""
    u = ((np.asarray(block, np.float32) - np.float32(vmin)) / np.float32(rng)).astype(np.float32)
    N = u.ndim
    n = float(u.size)
    o = np.zeros(40, np.float64)
    mean = u.astype(np.float64).sum() / n
    var = max((u.astype(np.float64) ** 2).sum() / n - mean * mean, 0.0)
    o[0], o[1], o[2], o[3] = mean, np.sqrt(var), u.min(), u.max()
    for si, s in enumerate(STRIDES):
        for a in range(3):
            if a < N and u.shape[a] > s:
                hi, lo = _shift(u, a, s)
                d = (hi - lo).astype(np.float32).astype(np.float64)
                o[4 + 9 * si + a] = np.abs(d).sum() / d.size
                o[7 + 9 * si + a] = np.sqrt((d * d).sum() / d.size)
        if all(m > 2 * s for m in u.shape):
            L = _laplacian(u, s).astype(np.float64)
            o[10 + 9 * si] = np.abs(L).sum() / L.size
            o[11 + 9 * si] = np.sqrt((L * L).sum() / L.size)
            o[12 + 9 * si] = np.abs(L).max()
    for a in range(3):
        for lag, base in ((1, 31), (2, 34)):
            if a < N and u.shape[a] > lag and var > 0:
                hi, lo = _shift(u, a, lag)
                cross = (hi.astype(np.float64) * lo.astype(np.float64)).sum() / hi.size
                o[base + a] = (cross - mean * mean) / var
    if all(m > 2 for m in u.shape):
        L1 = _laplacian(u, 1).astype(np.float64)
        m2 = (L1 * L1).sum() / L1.size
        o[37] = ((L1 ** 4).sum() / L1.size) / (m2 * m2) if m2 > 0 else 0.0
        o[38] = np.count_nonzero(np.abs(L1) > 2.0 * np.sqrt(m2)) / L1.size
    fr, na = 0.0, 0
    for a in range(N):
        if u.shape[a] > 1:
            hi, lo = _shift(u, a, 1)
            d = (hi - lo).astype(np.float32)
            fr += np.count_nonzero(np.abs(d) < np.float32(eb_norm)) / d.size
            na += 1
    o[39] = fr / na if na else 0.0
    return o.astype(np.float32)
