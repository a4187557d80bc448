"""CPU oracle for a_lm -> HEALPix synthesis (ctypes wrapper + numpy c2r).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only there as the reference that results
are compared with.  The product package (wavemoth_amd) never imports this
module.

The heavy lifting is the plain-C restatement in sht_oracle.c (see its
header for the reference file:line each function follows).  The one step
done here in numpy is the per-ring unnormalised complex-to-real FFT, which
the reference delegates to FFTW (src/wavemoth.c:821-833, 1181, 1185):
``numpy.fft.irfft(X, n) * n`` has FFTW's c2r semantics (Hermitian half
spectrum of n//2+1 bins, imaginary parts of the DC and Nyquist bins
ignored, no normalisation).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int)


def build(ref=True):
    """Compile liboracle.so (and _ref/ when the reference tree exists)."""
    subprocess.run(["make", "-C", _HERE, "-B", "all"], check=True,
                   stdout=subprocess.DEVNULL)
    refdir = os.environ.get("WAVEMOTH_REF", "/root/reference")
    if ref and os.path.isdir(os.path.join(refdir, "libpshtlight")):
        subprocess.run(["make", "-C", _HERE, "-B", "ref", "REF=" + refdir], check=True,
                       stdout=subprocess.DEVNULL)


def _ptr(a, t=_d):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        L.or_grid.argtypes = [ctypes.c_int, _d, _i64]
        L.or_ring_z.argtypes = [ctypes.c_int, _d]
        L.or_lambda_column.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, _d]
        L.or_lambda_column.restype = ctypes.c_int
        L.or_auxdata.argtypes = [ctypes.c_int64] * 3 + [_d]
        L.or_legendre_transform.argtypes = [ctypes.c_int64] * 3 + [_d] * 6
        L.or_plan_new.argtypes = [ctypes.c_int] * 7
        L.or_plan_new.restype = ctypes.c_void_p
        L.or_plan_free.argtypes = [ctypes.c_void_p]
        L.or_legendre_stage.argtypes = [ctypes.c_void_p, _d, _d, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_legendre_stage_fast.argtypes = [ctypes.c_void_p, _d, _d] + [ctypes.c_int] * 5
        L.or_legendre_stage_fast.restype = ctypes.c_int
        L.or_fast_width.restype = ctypes.c_int
        L.or_ring_fold_sample.argtypes = [ctypes.c_void_p, _d, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _i64, _d, ctypes.c_int]
        L.or_ring_fold_sample.restype = ctypes.c_int
        L.or_ring_spectra.argtypes = [ctypes.c_void_p, _d, ctypes.c_int, _i64, _d]
        L.or_plan_trimmed_triples.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int]
        L.or_plan_trimmed_triples.restype = ctypes.c_int64
        L.or_num_threads.restype = ctypes.c_int
        L.or_ylm_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.or_ylm_new.restype = ctypes.c_void_p
        L.or_ylm_free.argtypes = [ctypes.c_void_p]
        L.or_ylm_column.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _d]
        L.or_ylm_column.restype = ctypes.c_int
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own Ylmgen (oracle/_ref), or None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libylmgen_ref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_lambda_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _d, ctypes.c_int,
                                        ctypes.c_double, _d, _i32]
        L.ref_lambda_matrix.restype = ctypes.c_int
        _REF = L
    return _REF


# ------------------------------------------------------------------ grid

def grid(N):
    """(phi0s[4N-1], ring_offsets[4N]) as src/wavemoth.c:1199-1232."""
    phi0s = np.zeros(4 * N - 1)
    offs = np.zeros(4 * N, np.int64)
    lib().or_grid(N, _ptr(phi0s), _ptr(offs, _i64))
    return phi0s, offs


def ring_z(N):
    z = np.zeros(2 * N)
    lib().or_ring_z(N, _ptr(z))
    return z


def ring_thetas_positive(N):
    return np.arccos(ring_z(N))


# ------------------------------------------------------------ Lambda

def lambda_column(m, lmax, theta, eps=1e-300):
    out = np.zeros(lmax + 1)
    fl = lib().or_lambda_column(m, lmax, float(theta), eps, _ptr(out))
    return out, fl


def lambda_matrix(m, thetas, lmax, eps=1e-300):
    """Lambda_l^m(theta) for l = m..lmax as (len(thetas), lmax-m+1), one
    fresh generator per column like the reference's
    compute_normalized_associated_legendre (legendre.pyx:24-57) *without*
    its cross-column m_crit memory."""
    rows = [lambda_column(m, lmax, t, eps)[0][m:] for t in np.atleast_1d(thetas)]
    return np.array(rows)


def lambda_matrix_shared(m, thetas, lmax, eps=1e-300):
    """Like lambda_matrix but with one generator over all columns in order,
    as legendre.pyx:24-57 does (so the m_crit skip carries across columns)."""
    L = lib()
    g = L.or_ylm_new(lmax, lmax, eps)
    out = np.zeros(lmax + 1)
    rows = []
    try:
        for t in np.atleast_1d(thetas):
            L.or_ylm_column(g, m, float(t), _ptr(out))
            rows.append(out[m:].copy())
    finally:
        L.or_ylm_free(g)
    return np.array(rows)


def ref_lambda_matrix(m, thetas, lmax, eps=1e-300):
    """Same matrix from the compiled reference Ylmgen (None if unbuilt)."""
    L = ref_lib()
    if L is None:
        return None
    th = np.ascontiguousarray(np.atleast_1d(thetas), dtype=np.float64)
    out = np.zeros((th.shape[0], lmax + 1))
    fl = np.zeros(th.shape[0], np.intc)
    rc = L.ref_lambda_matrix(m, lmax, _ptr(th), th.shape[0], eps, _ptr(out), _ptr(fl, _i32))
    assert rc == 0
    return out[:, m:]


def auxdata(m, lmin, nk):
    if nk < 3:
        return np.zeros(0)
    out = np.zeros(3 * (nk - 2))
    lib().or_auxdata(m, lmin, nk, _ptr(out))
    return out


def legendre_transform(m, lmin, a, x_squared, P, Pp1, aux=None):
    """The reference's scalar per-strip transform: y = sum_k P_k a[k, :]."""
    a = np.ascontiguousarray(a, np.float64)
    nk, nvecs = a.shape
    abig = np.zeros((2 * nk, nvecs))
    abig[0::2] = a
    if aux is None:
        aux = auxdata(m, lmin, nk)
    aux = np.ascontiguousarray(aux if aux.size else np.zeros(1))
    x2 = np.ascontiguousarray(x_squared, np.float64)
    y = np.zeros((x2.shape[0], nvecs))
    lib().or_legendre_transform(x2.shape[0], nk, nvecs, _ptr(abig), _ptr(y), _ptr(x2),
                                _ptr(aux), _ptr(np.ascontiguousarray(P, np.float64)),
                                _ptr(np.ascontiguousarray(Pp1, np.float64)))
    return y


# ------------------------------------------------------------ full path

def nalm(lmax, mmax):
    return sum(lmax - m + 1 for m in range(mmax + 1))


def lm_to_idx_mmajor(l, m, lmax):
    return m * (2 * lmax - m + 3) // 2 + (l - m)


class OraclePlan:
    """Restated plan: precompute (the resource file) in the constructor,
    then legendre_stage / ring stage / synthesize like wavemoth_execute."""

    def __init__(self, Nside, lmax, mmax, nmaps, m_start=0, m_stop=None, m_step=1):
        self.Nside, self.lmax, self.mmax, self.nmaps = Nside, lmax, mmax, nmaps
        self.R = 2 * Nside
        self.m_range = (m_start, mmax + 1 if m_stop is None else m_stop, m_step)
        self._p = lib().or_plan_new(Nside, lmax, mmax, nmaps, *self.m_range)
        self.phi0s, self.ring_offsets = grid(Nside)
        counts = np.diff(self.ring_offsets)
        self.counts = counts
        nb = counts // 2 + 1
        self.spec_off = np.zeros(4 * Nside - 1, np.int64)
        self.spec_off[1:] = np.cumsum(nb)[:-1]
        self.nspec = int(nb.sum())

    def __del__(self):
        if getattr(self, "_p", None):
            lib().or_plan_free(self._p)
            self._p = None

    def trimmed_triples(self):
        """T_eps counted per (m, odd, ring) column, both parities."""
        return lib().or_plan_trimmed_triples(self._p, *self.m_range)

    def legendre_stage(self, alm, compact=False):
        """work[m, odd, ring, map]; with compact=True only the plan's sampled
        m values, in order (for timing samples of large transforms)."""
        alm = np.ascontiguousarray(alm, np.complex128)
        assert alm.shape == (nalm(self.lmax, self.mmax), self.nmaps)
        nm = len(range(*self.m_range)) if compact else self.mmax + 1
        work = np.zeros((nm, 2, self.R, self.nmaps), np.complex128)
        lib().or_legendre_stage(self._p, _ptr(alm.view(np.float64)),
                                _ptr(work.view(np.float64)), *self.m_range, int(bool(compact)))
        return work

    def legendre_stage_fast(self, alm, compact=False, width=0, work=None):
        """legendre_stage through the vectorised kernel of sht_fast.c (the CPU
        baseline's timed leg) at vector width 4 (AVX2) or 8 (AVX-512), 0 for
        the widest the CPU has.  Returns (work, width run)."""
        alm = np.ascontiguousarray(alm, np.complex128)
        assert alm.shape == (nalm(self.lmax, self.mmax), self.nmaps)
        nm = len(range(*self.m_range)) if compact else self.mmax + 1
        if work is None:
            work = np.zeros((nm, 2, self.R, self.nmaps), np.complex128)
        assert work.shape == (nm, 2, self.R, self.nmaps) and work.flags.c_contiguous
        w = lib().or_legendre_stage_fast(self._p, _ptr(alm.view(np.float64)),
                                         _ptr(work.view(np.float64)), *self.m_range,
                                         int(bool(compact)), int(width))
        if w == 0:
            raise RuntimeError("vector width %d not supported by this CPU" % width)
        return work, w

    def ring_spectra(self, work, phase_shifts=True):
        work = np.ascontiguousarray(work, np.complex128)
        spec = np.zeros((self.nspec, self.nmaps), np.complex128)
        lib().or_ring_spectra(self._p, _ptr(work.view(np.float64)), int(bool(phase_shifts)),
                              _ptr(self.spec_off, _i64), _ptr(spec.view(np.float64)))
        return spec

    def c2r(self, spec):
        """Per-ring unnormalised c2r (FFTW semantics) into a (npix, nmaps) map."""
        npix = int(self.ring_offsets[-1])
        out = np.zeros((npix, self.nmaps))
        nrings = 4 * self.Nside - 1
        by_len = {}
        for i in range(nrings):
            by_len.setdefault(int(self.counts[i]), []).append(i)
        for n, rings in by_len.items():
            nb = n // 2 + 1
            X = np.stack([spec[self.spec_off[i]:self.spec_off[i] + nb] for i in rings])
            f = np.fft.irfft(X, n, axis=1) * n
            for j, i in enumerate(rings):
                o = int(self.ring_offsets[i])
                out[o:o + n] = f[j]
        return out

    def ring_stage_sample(self, work_s, r_step, workers=1, phase_shifts=True, width=0):
        """Ring stage of the CPU baseline on the ring pairs r = 0, r_step, ...
        of the northern half and their mirrors, all maps at once: the fold of
        sht_fast.c, then per pair one pocketfft c2r over its top and bottom
        rings and every map (the reference's plan_many c2r with howmany =
        nmaps, src/wavemoth.c:828-831; the fold writes each ring map-major so
        that the c2r runs over contiguous bins), pairs spread over `workers`
        threads.
        work_s: work[m, odd, s, map] of the sampled rings.  Returns
        (rings, pixels): the physical ring indices, top then bottom per pair,
        and their pixels (sum of their lengths, nmaps), in that order."""
        import scipy.fft as sf
        from concurrent.futures import ThreadPoolExecutor
        work_s = np.ascontiguousarray(work_s, np.complex128)
        K = self.nmaps
        nsamp = work_s.shape[2]
        assert work_s.shape == (self.mmax + 1, 2, nsamp, K)
        assert nsamp == len(range(0, self.R, r_step))
        mid = 2 * self.Nside - 1
        soff, poff, rings, tot, npx = [], [], [], 0, 0
        for s in range(nsamp):
            r = s * r_step
            n = int(self.counts[mid + r])
            cnt = 2 if r > 0 else 1
            soff.append(tot)
            poff.append(npx)
            rings += [mid - r, mid + r] if r > 0 else [mid]
            tot += cnt * (n // 2 + 1)
            npx += cnt * n
        soff = np.array(soff, np.int64)
        spec = np.empty((tot, K), np.complex128)
        w = lib().or_ring_fold_sample(self._p, _ptr(work_s.view(np.float64)),
                                      int(bool(phase_shifts)), int(r_step), nsamp,
                                      _ptr(soff, _i64), _ptr(spec.view(np.float64)), int(width))
        if w == 0:
            raise RuntimeError("vector width %d not supported by this CPU" % width)
        out = np.empty((npx, K))

        def one(s):
            r = s * r_step
            n = int(self.counts[mid + r])
            nb = n // 2 + 1
            cnt = 2 if r > 0 else 1
            X = spec[soff[s]:soff[s] + cnt * nb].reshape(cnt, K, nb)
            out[poff[s]:poff[s] + cnt * n].reshape(cnt, n, K)[...] = sf.irfft(
                X, n, axis=2, norm="forward").transpose(0, 2, 1)

        if workers > 1:
            with ThreadPoolExecutor(workers) as ex:
                list(ex.map(one, range(nsamp)))
        else:
            for s in range(nsamp):
                one(s)
        return rings, out

    def synthesize(self, alm, phase_shifts=True):
        work = self.legendre_stage(alm)
        return self.c2r(self.ring_spectra(work, phase_shifts))


def fast_width():
    """Widest vector width (doubles) sht_fast.c can run on this CPU, 0 if none."""
    return lib().or_fast_width()


def num_threads():
    return lib().or_num_threads()
