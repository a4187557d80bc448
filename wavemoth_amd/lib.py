"""Python drop-in for the reference's ShtPlan (wavemoth/lib.pyx:113-195).

Same constructor signature, exceptions and methods; the work runs on the
current HIP device through libwavemoth_hip.so (include/wavemoth.h).
"""
import ctypes
import os

import numpy as np

from . import _native
from ._native import PLAN_INFO_KEYS, lib, last_error, plan_info  # noqa: F401
from .alm import nalm

_configured = False


class ShtPlan:
    """a_lm (complex128 (nalm, nmaps), m-major) -> HEALPix map
    (float64 (12 Nside^2, nmaps)), mirroring wavemoth/lib.pyx:113-195.

    Extensions: ``flags`` (e.g. WAVEMOTH_SEEDS_ON_THE_FLY), ``info()``,
    ``legendre_flops(m, odd)``.
    """

    def __init__(self, Nside, lmax, mmax, input, output, ordering, phase_shifts=True,
                 matrix_data_filename=None, nthreads=1, flags=0):
        global _configured
        if ordering != 'mmajor':
            raise ValueError("Invalid ordering: %s" % ordering)
        if not (isinstance(input, np.ndarray) and input.dtype == np.complex128
                and input.ndim == 2 and input.flags.c_contiguous):
            raise ValueError("input must be a C-contiguous complex128 array (nalm, nmaps)")
        if not (isinstance(output, np.ndarray) and output.dtype == np.float64
                and output.ndim == 2 and output.flags.c_contiguous):
            raise ValueError("output must be a C-contiguous float64 array (npix, nmaps)")
        self.input = input
        self.output = output
        if not (input.shape[1] == output.shape[1]):
            raise ValueError("Nonconforming arrays")
        if output.shape[0] != 12 * Nside * Nside:
            raise ValueError("Output must have shape (Npix, nmaps), has %r" % (output.shape,))
        if input.shape[0] != nalm(lmax, mmax):
            raise ValueError("Input must have shape (nalm, nmaps) with nalm=%d, has %r"
                             % (nalm(lmax, mmax), input.shape))
        L = lib()
        if not _configured and matrix_data_filename is None:
            L.wavemoth_configure(os.environ.get('SHTRESOURCES', '').encode())
            _configured = True
        fname = None
        if matrix_data_filename is not None:
            fname = matrix_data_filename if isinstance(matrix_data_filename, bytes) \
                else str(matrix_data_filename).encode()
        self._plan = L.wavemoth_plan_to_healpix(
            Nside, lmax, mmax, input.shape[1], nthreads,
            input.ctypes.data, output.ctypes.data, _native.WAVEMOTH_MMAJOR,
            _native.WAVEMOTH_ESTIMATE | int(flags), fname)
        if not self._plan:
            raise Exception("Plan creation failed: %s" % last_error())
        self.Nside = Nside
        self.lmax = lmax
        self.mmax = mmax
        self.nmaps = input.shape[1]
        if not phase_shifts:
            L.wavemoth_disable_phase_shifting(self._plan)
            self._check()

    def _check(self):
        e = last_error()
        if e:
            raise RuntimeError(e)

    def __del__(self):
        p = getattr(self, "_plan", None)
        if p:
            lib().wavemoth_destroy_plan(p)
            self._plan = None

    def get_work(self):
        """Legendre-stage output as complex (mmax+1, 2, 2 Nside, nmaps):
        work[m, odd, ring, map], rings from the equator to the pole."""
        work = np.empty((self.mmax + 1, 2, 2 * self.Nside, self.nmaps), np.complex128)
        if lib().wavemoth_get_work(self._plan, work.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return work

    def execute(self, repeat=1):
        for _ in range(repeat):
            lib().wavemoth_execute(self._plan)
            self._check()
        return self.output

    def perform_backward_ffts(self):
        lib().wavemoth_perform_backward_ffts(self._plan)
        self._check()

    def perform_legendre_transform(self, repeat=1):
        for _ in range(repeat):
            lib().wavemoth_perform_legendre_transforms(self._plan)
            self._check()

    # ---- extensions
    def execute_adjoint(self):
        """input <- Y^T output (the transpose of execute, in place on the
        plan's arrays); returns input."""
        lib().wavemoth_execute_adjoint(self._plan)
        self._check()
        return self.input

    def execute_device(self, d_input, d_output, stream=None):
        """Asynchronous execute on device pointers (ints)."""
        rc = lib().wavemoth_execute_device(self._plan, ctypes.c_void_p(int(d_input)),
                                           ctypes.c_void_p(int(d_output)),
                                           ctypes.c_void_p(int(stream or 0)))
        if rc != 0:
            raise RuntimeError(last_error())

    def legendre_flops(self, m, odd):
        return lib().wavemoth_get_legendre_flops(self._plan, m, odd)

    def info(self):
        return plan_info(self._plan)


def _c_array(x, name):
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("%s must be a C-contiguous float64 array" % name)
    return x


def legendre_transform_auxdata(m, lmin, nk):
    """Recurrence coefficients of one parity strip (wavemoth/lib.pyx:282-290):
    3 (nk - 2) values, empty for nk < 3."""
    if nk < 3:
        return np.zeros(0)
    out = np.empty(3 * (nk - 2))
    lib().wavemoth_legendre_transform_auxdata(int(m), int(lmin), int(nk), out.ctypes.data)
    return out


def legendre_transform(m, lmin, a, y, x_squared, P, Pp1, repeat=1, use_sse=False,
                       auxdata=None):
    """Per-strip transform y[i, j] = sum_k P_k(x_i) a[k, j] for rows
    l = lmin, lmin + 2, ... (wavemoth/lib.pyx:220-280), run on the GPU.
    Same argument checks and ValueErrors as the reference.  ``use_sse``
    selects between the reference's two CPU kernels, which compute the same
    sums; both settings run the one GPU kernel here."""
    a = _c_array(a, "a")
    y = _c_array(y, "y")
    for nm, v in (("x_squared", x_squared), ("P", P), ("Pp1", Pp1)):
        _c_array(v, nm)
    if a.ndim != 2 or y.ndim != 2 or x_squared.ndim != 1 or P.ndim != 1 or Pp1.ndim != 1:
        raise ValueError("nonconforming arrays")
    nx = x_squared.shape[0]
    if not nx == P.shape[0] == Pp1.shape[0]:
        raise ValueError("nonconforming arrays")
    nk, nvecs = a.shape
    if not nvecs == y.shape[1] or y.shape[0] != nx:
        raise ValueError("nonconforming arrays")
    abig = np.zeros((2 * nk, nvecs))
    abig[0::2, :] = a
    if auxdata is None:
        auxdata = legendre_transform_auxdata(m, lmin, nk)
    else:
        _c_array(auxdata, "auxdata")
        if auxdata.shape[0] != 3 * (nk - 2):
            raise ValueError("auxdata.shape[0] != 3 * (nk - 2)")
    if nvecs % 2 != 0:
        raise ValueError("nvecs not divisible by 2")
    aux = auxdata if auxdata.size else np.zeros(1)
    L = lib()
    for _ in range(repeat):
        L.wavemoth_legendre_transform(nx, nk, nvecs, abig.ctypes.data, y.ctypes.data,
                                      x_squared.ctypes.data, aux.ctypes.data, P.ctypes.data,
                                      Pp1.ctypes.data)
        e = last_error()
        if e:
            raise RuntimeError(e)


def query_resourcefile(filename):
    """(Nside, lmax) from a resource-file header, or None
    (src/wavemoth.c:144-155)."""
    n, l = ctypes.c_int(), ctypes.c_int()
    f = filename if isinstance(filename, bytes) else str(filename).encode()
    if lib().wavemoth_query_resourcefile(f, ctypes.byref(n), ctypes.byref(l)) != 0:
        return None
    return n.value, l.value
