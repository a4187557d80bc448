"""Device-tensor entry points: synthesis of torch tensors that already live
on the GPU, on torch's current stream, through wavemoth_execute_device."""
import ctypes

import torch

from . import _native
from ._native import lib, last_error
from .alm import nalm

_PLANS = {}


class DevicePlan:
    """A plan with no host buffers; executes on device pointers."""

    def __init__(self, Nside, lmax, mmax, nmaps, flags=0):
        self.Nside, self.lmax, self.mmax, self.nmaps = Nside, lmax, mmax, nmaps
        self.npix = 12 * Nside * Nside
        self.nalm = nalm(lmax, mmax)
        self._plan = lib().wavemoth_plan_to_healpix(Nside, lmax, mmax, nmaps, 1, None, None,
                                                    _native.WAVEMOTH_MMAJOR, int(flags), None)
        if not self._plan:
            raise RuntimeError("Plan creation failed: %s" % last_error())

    def __del__(self):
        p = getattr(self, "_plan", None)
        if p:
            lib().wavemoth_destroy_plan(p)
            self._plan = None

    def synthesize_into(self, alm, out, stream=None):
        if stream is None:
            stream = torch.cuda.current_stream(alm.device).cuda_stream
        rc = lib().wavemoth_execute_device(self._plan, ctypes.c_void_p(alm.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()),
                                           ctypes.c_void_p(int(stream)))
        if rc != 0:
            raise RuntimeError(last_error())
        return out

    def set_profiling(self, on):
        if lib().wavemoth_set_profiling(self._plan, int(bool(on))) != 0:
            raise RuntimeError(last_error())

    def stage_times(self):
        """Average ms of (prescale, legendre sweep, ring stage) over the
        executes recorded since the last call, and their count."""
        avg = (ctypes.c_double * 3)()
        n = ctypes.c_int()
        if lib().wavemoth_stage_times(self._plan, avg, ctypes.byref(n)) != 0:
            raise RuntimeError(last_error())
        return list(avg), n.value

    def info(self):
        keys = ["Nside", "lmax", "mmax", "nmaps", "npix", "triples_dense", "triples_eps",
                "flags", "legendre_launches", "device_bytes", "triples_executed"]
        return {k: int(lib().wavemoth_plan_info(self._plan, i)) for i, k in enumerate(keys)}


def get_plan(Nside, lmax, mmax, nmaps, device, flags=0):
    key = (int(device.index if device.index is not None else torch.cuda.current_device()),
           Nside, lmax, mmax, nmaps, flags)
    p = _PLANS.get(key)
    if p is None:
        with torch.cuda.device(key[0]):
            p = DevicePlan(Nside, lmax, mmax, nmaps, flags)
        _PLANS[key] = p
    return p


def synthesis(alm, Nside, lmax, mmax=None, flags=0):
    """alm: complex128 cuda tensor (nalm, nmaps) in the m-major layout.
    Returns the float64 map (12 Nside^2, nmaps) on the same device."""
    if mmax is None:
        mmax = lmax
    if not (alm.is_cuda and alm.dtype == torch.complex128 and alm.dim() == 2):
        raise ValueError("alm must be a 2-D complex128 CUDA tensor")
    if alm.shape[0] != nalm(lmax, mmax):
        raise ValueError("alm has %d rows, expected %d" % (alm.shape[0], nalm(lmax, mmax)))
    alm = alm.contiguous()
    p = get_plan(Nside, lmax, mmax, alm.shape[1], alm.device, flags)
    out = torch.empty((p.npix, alm.shape[1]), dtype=torch.float64, device=alm.device)
    with torch.cuda.device(alm.device):
        p.synthesize_into(alm, out)
    return out
