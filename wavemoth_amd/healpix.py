"""HEALPix ring geometry as the reference uses it (wavemoth/healpix.py:5-41,
src/wavemoth.c:1199-1232)."""
import numpy as np
from numpy import pi


def get_ring_thetas(Nside, positive_only=False):
    """Ring colatitudes; with positive_only the 2N northern rings ordered
    from the equator towards the pole (wavemoth/healpix.py:5-18)."""
    nrings = 4 * Nside - 1 if not positive_only else 2 * Nside
    i = np.arange(1, nrings + 1)
    z = np.empty(nrings, np.double)
    z[:Nside] = 1 - i[:Nside] ** 2 / (3 * Nside ** 2)
    z[Nside:2 * Nside] = (4 / 3) - (2 * i[Nside:2 * Nside]) / (3 * Nside)
    if positive_only:
        return np.arccos(z[::-1])
    z[2 * Nside:] = -z[2 * Nside - 2::-1]
    return np.arccos(z)


def get_ring_pixel_counts(Nside):
    """Pixels per ring, north to south (wavemoth/healpix.py:30-41)."""
    cap = np.arange(4, 4 * Nside, 4)
    belt = 4 * Nside * np.ones(2 * Nside + 1, dtype=np.int64)
    return np.hstack([cap, belt, cap[::-1]]).astype(np.int64)


def get_ring_offsets(Nside):
    """First pixel of every ring plus npix (src/wavemoth.c:1199-1232)."""
    c = get_ring_pixel_counts(Nside)
    return np.concatenate([[0], np.cumsum(c)]).astype(np.int64)


def get_grid_phi0s(Nside):
    """phi0 of every ring exactly as src/wavemoth.c:1212-1224 sets it (the
    belt uses iring % 2, i.e. HEALPix's convention for even Nside)."""
    nrings = 4 * Nside - 1
    it = np.arange(nrings)
    phi = (pi / (4.0 * Nside)) * (it % 2)
    cap = it <= Nside - 1
    phi[cap] = pi / (4.0 * (it[cap] + 1))
    sc = it > 3 * Nside - 1
    phi[sc] = pi / (4.0 * (nrings - it[sc]))
    return phi
