"""Batched synthesis of many uneven requests (wavemoth_plan_batch).

The reference has no batch entry point; a caller of it makes one ShtPlan
per request (wavemoth/lib.pyx:113-195).  ``BatchPlan`` takes the list of
requests at once and runs them together: persistent work-queue kernels over
all requests, replayed from a HIP graph.  Each request's map equals the one
a single plan of the same (Nside, lmax, mmax, nmaps) produces.
"""
import ctypes

import numpy as np

from . import _native
from ._native import last_error, lib
from .alm import nalm


def random_requests(nreq, seed, nsides=(128, 256, 512, 1024, 2048), nmaps=(1, 3, 8)):
    """Uneven request list: Nside uniform over `nsides`, lmax = mmax =
    2 Nside, nmaps uniform over `nmaps` (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreq):
        N = int(rng.choice(nsides))
        out.append((N, 2 * N, 2 * N, int(rng.choice(nmaps))))
    return out


class BatchPlan:
    """requests: sequence of (Nside, lmax, mmax, nmaps).  flags may hold
    WAVEMOTH_MEASURE (autotune the sweep shapes) and
    WAVEMOTH_BATCH_NO_GRAPH (launch directly instead of replaying a graph)."""

    def __init__(self, requests, flags=0):
        self.requests = [tuple(int(v) for v in r) for r in requests]
        n = len(self.requests)
        arr = [(ctypes.c_int * n)(*[r[i] for r in self.requests]) for i in range(4)]
        self._b = lib().wavemoth_plan_batch(n, arr[0], arr[1], arr[2], arr[3], int(flags))
        if not self._b:
            raise RuntimeError("Batch plan creation failed: %s" % last_error())
        self.nreq = n

    def __del__(self):
        b = getattr(self, "_b", None)
        if b:
            lib().wavemoth_destroy_batch(b)
            self._b = None

    def shapes(self, r):
        N, L, M, K = self.requests[r]
        return (nalm(L, M), K), (12 * N * N, K)

    def _check(self, alms, outs, device):
        if len(alms) != self.nreq or len(outs) != self.nreq:
            raise ValueError("need one input and one output per request")
        for r, (a, o) in enumerate(zip(alms, outs)):
            (na, K), (npix, _) = self.shapes(r)
            if device:
                import torch
                ok = (a.is_cuda and o.is_cuda and a.dtype == torch.complex128 and
                      o.dtype == torch.float64 and a.is_contiguous() and o.is_contiguous())
            else:
                ok = (a.dtype == np.complex128 and o.dtype == np.float64 and
                      a.flags.c_contiguous and o.flags.c_contiguous)
            if not ok or tuple(a.shape) != (na, K) or tuple(o.shape) != (npix, K):
                raise ValueError("request %d: need input (%d, %d) complex128 and output "
                                 "(%d, %d) float64, contiguous" % (r, na, K, npix, K))

    def synthesize_into(self, alms, outs, stream=None):
        """Device tensors, on torch's current stream unless `stream` is given."""
        import torch
        self._check(alms, outs, True)
        if stream is None:
            stream = torch.cuda.current_stream(alms[0].device).cuda_stream
        pin = (ctypes.c_void_p * self.nreq)(*[a.data_ptr() for a in alms])
        pout = (ctypes.c_void_p * self.nreq)(*[o.data_ptr() for o in outs])
        rc = lib().wavemoth_execute_batch_device(self._b, pin, pout,
                                                 ctypes.c_void_p(int(stream)))
        if rc != 0:
            raise RuntimeError(last_error())
        return outs

    def execute(self, inputs, outputs):
        """Host numpy arrays, synchronous."""
        self._check(inputs, outputs, False)
        pin = (ctypes.c_void_p * self.nreq)(*[a.ctypes.data for a in inputs])
        pout = (ctypes.c_void_p * self.nreq)(*[o.ctypes.data for o in outputs])
        if lib().wavemoth_execute_batch(self._b, pin, pout) != 0:
            raise RuntimeError(last_error())
        return outputs

    def set_profiling(self, on):
        if lib().wavemoth_batch_set_profiling(self._b, int(bool(on))) != 0:
            raise RuntimeError(last_error())

    def stage_times(self):
        avg = (ctypes.c_double * 3)()
        n = ctypes.c_int()
        if lib().wavemoth_batch_stage_times(self._b, avg, ctypes.byref(n)) != 0:
            raise RuntimeError(last_error())
        return list(avg), n.value

    def info(self):
        keys = ["requests", "maps", "triples_eps", "triples_dense", "flops", "device_bytes",
                "sweep_items", "ring_items", "graph", "slots_executed", "sweep_workgroups"]
        return {k: int(lib().wavemoth_batch_info(self._b, -1, i)) for i, k in enumerate(keys)}

    def request_info(self, r):
        keys = ["Nside", "lmax", "mmax", "nmaps", "rings_per_lane", "maps_per_pass",
                "triples_eps", "triples_dense", "npix", "slots_executed", "measured",
                "wrap_register"]
        return {k: int(lib().wavemoth_batch_info(self._b, r, i)) for i, k in enumerate(keys)}


WAVEMOTH_BATCH_NO_GRAPH = _native.WAVEMOTH_BATCH_NO_GRAPH
