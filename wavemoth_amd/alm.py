"""Seeded synthetic a_lm in the reference's m-major layout.

Re, Im ~ N(0, C_l/2) for m > 0; a_l0 real ~ N(0, C_l); C_l = (l+1)^-2
(power law) or 1 (flat).  numpy's PCG64 stream makes the values identical
on every machine running the same numpy.
"""
import numpy as np


def nalm(lmax, mmax):
    return (mmax + 1) * (lmax + 1) - mmax * (mmax + 1) // 2


def lm_to_idx_mmajor(l, m, lmax):
    """wavemoth/psht.pyx:48-49, src/wavemoth.c:1016-1017."""
    return m * (2 * lmax - m + 3) // 2 + (l - m)


def l_of_index(lmax, mmax):
    """(l, m) of every row of the m-major triangle."""
    ls, ms = [], []
    for m in range(mmax + 1):
        ls.append(np.arange(m, lmax + 1))
        ms.append(np.full(lmax - m + 1, m))
    return np.concatenate(ls), np.concatenate(ms)


def power_law_alm(lmax, mmax, nmaps, seed, flat=False):
    ls, ms = l_of_index(lmax, mmax)
    cl = np.ones(ls.shape) if flat else (ls + 1.0) ** -2
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((ls.shape[0], nmaps, 2))
    sd = np.sqrt(cl / 2.0)[:, None]
    alm = (g[..., 0] * sd) + 1j * (g[..., 1] * sd)
    m0 = ms == 0
    alm[m0] = (g[m0, :, 0] * np.sqrt(cl[m0])[:, None]) + 0j
    return np.ascontiguousarray(alm, dtype=np.complex128)
