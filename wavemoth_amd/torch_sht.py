"""Device-tensor entry points: synthesis of torch tensors that already live
on the GPU, on torch's current stream, through wavemoth_execute_device."""
import ctypes

import torch

from . import _native
from ._native import lib, last_error, plan_info
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

    def adjoint_into(self, gmap, out, stream=None):
        """out (nalm, nmaps) complex128 <- Y^T gmap (npix, nmaps) float64."""
        if stream is None:
            stream = torch.cuda.current_stream(gmap.device).cuda_stream
        rc = lib().wavemoth_execute_adjoint_device(self._plan, ctypes.c_void_p(gmap.data_ptr()),
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
        return plan_info(self._plan)


def random_alm(lmax, mmax, nmaps, seed, device, flat=False):
    """Seeded a_lm generated on the device: Re, Im ~ N(0, C_l/2) for m > 0,
    a_l0 real ~ N(0, C_l), C_l = (l+1)^-2 or 1 (flat).  Same law as
    alm.power_law_alm but drawn by torch's device generator (fast at large
    sizes; the values differ from the numpy stream)."""
    from .alm import l_of_index
    ls, ms = l_of_index(lmax, mmax)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    re = torch.randn((len(ls), nmaps), generator=g, device=device, dtype=torch.float64)
    im = torch.randn((len(ls), nmaps), generator=g, device=device, dtype=torch.float64)
    l_t = torch.from_numpy(ls).to(device=device, dtype=torch.float64)
    m0 = torch.from_numpy(ms == 0).to(device)[:, None]
    cl = torch.ones_like(l_t) if flat else (l_t + 1.0) ** -2
    sd = torch.sqrt(cl / 2.0)[:, None]
    sd = torch.where(m0, sd * (2.0 ** 0.5), sd)
    im = torch.where(m0, torch.zeros_like(im), im)
    return torch.complex(re * sd, im * sd).contiguous()


def get_plan(Nside, lmax, mmax, nmaps, device, flags=0):
    key = (int(device.index if device.index is not None else torch.cuda.current_device()),
           Nside, lmax, mmax, nmaps, flags)
    p = _PLANS.get(key)
    if p is None:
        with torch.cuda.device(key[0]):
            p = DevicePlan(Nside, lmax, mmax, nmaps, flags)
        _PLANS[key] = p
    return p


class _Synthesis(torch.autograd.Function):
    """Synthesis with the hand-written adjoint as its backward: for a real
    loss, grad_alm = dL/dRe(a) + i dL/dIm(a) = Y^T (dL/dmap)."""

    @staticmethod
    def forward(ctx, alm, plan):
        out = torch.empty((plan.npix, plan.nmaps), dtype=torch.float64, device=alm.device)
        with torch.cuda.device(alm.device):
            plan.synthesize_into(alm, out)
        ctx.plan = plan
        return out

    @staticmethod
    def backward(ctx, gmap):
        plan = ctx.plan
        gmap = gmap.contiguous().to(torch.float64)
        out = torch.empty((plan.nalm, plan.nmaps), dtype=torch.complex128, device=gmap.device)
        with torch.cuda.device(gmap.device):
            plan.adjoint_into(gmap, out)
        return out, None


def _check_alm(alm, lmax, mmax):
    if not (alm.is_cuda and alm.dtype == torch.complex128 and alm.dim() == 2):
        raise ValueError("alm must be a 2-D complex128 CUDA tensor")
    if alm.shape[0] != nalm(lmax, mmax):
        raise ValueError("alm has %d rows, expected %d" % (alm.shape[0], nalm(lmax, mmax)))


def synthesis(alm, Nside, lmax, mmax=None, flags=0):
    """alm: complex128 cuda tensor (nalm, nmaps) in the m-major layout.
    Returns the float64 map (12 Nside^2, nmaps) on the same device;
    differentiable with respect to alm (backward = the adjoint kernels)."""
    if mmax is None:
        mmax = lmax
    _check_alm(alm, lmax, mmax)
    p = get_plan(Nside, lmax, mmax, alm.shape[1], alm.device, flags)
    return _Synthesis.apply(alm.contiguous(), p)


def adjoint(gmap, Nside, lmax, mmax=None, flags=0):
    """Y^T gmap: gmap float64 cuda (12 Nside^2, nmaps) -> complex128 (nalm, nmaps)."""
    if mmax is None:
        mmax = lmax
    if not (gmap.is_cuda and gmap.dtype == torch.float64 and gmap.dim() == 2):
        raise ValueError("gmap must be a 2-D float64 CUDA tensor")
    p = get_plan(Nside, lmax, mmax, gmap.shape[1], gmap.device, flags)
    out = torch.empty((p.nalm, gmap.shape[1]), dtype=torch.complex128, device=gmap.device)
    with torch.cuda.device(gmap.device):
        p.adjoint_into(gmap.contiguous(), out)
    return out
