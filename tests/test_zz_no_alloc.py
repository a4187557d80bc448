"""The allocation-free adjoint (wavemoth_execute_adjoint_device_noalloc)
against wavemoth_execute_adjoint_device, both through the C ABI.

Two calls in a row on one GPU, the second with a smaller input, so that the
second runs on the contents the first left in the library's fixed scratch
array: bitwise equal to the allocating entry where both use the same ring
groups (plan_info keys 12 and 13), within 1e-13 relative rms where the
scratch holds fewer groups (only the summation order differs).  A plan whose
single ring group does not fit is refused with an error and leaves the
output untouched.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan_and_inputs(N, L, K, seed):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    dev = torch.device("cuda:0")
    p = DevicePlan(N, L, L, K)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    gmap = torch.randn((12 * N * N, K), generator=gen, device=dev, dtype=torch.float64)
    return p, gmap, p.nalm


def _noalloc(p, gmap, out):
    import torch
    from wavemoth_amd._native import last_error, lib
    rc = lib().wavemoth_execute_adjoint_device_noalloc(
        p._plan, ctypes.c_void_p(gmap.data_ptr()), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream)))
    return rc, last_error()


def _compare(N, L, K, seed):
    import torch
    p, gmap, nalm = _plan_and_inputs(N, L, K, seed)
    info = p.info()
    g_alloc, g_free = info["adjoint_groups"], info["adjoint_groups_noalloc"]
    print("case noalloc adjoint Nside=%d lmax=%d nmaps=%d groups %d / %d"
          % (N, L, K, g_alloc, g_free), flush=True)
    assert g_free >= 1
    ref = torch.empty((nalm, K), dtype=torch.complex128, device="cuda")
    p.adjoint_into(gmap, ref)
    got = torch.full_like(ref, complex(float("nan"), float("nan")))
    rc, err = _noalloc(p, gmap, got)
    assert rc == 0, err
    torch.cuda.synchronize()
    assert bool(torch.isfinite(torch.view_as_real(got)).all())
    if g_free == g_alloc:
        assert torch.equal(got, ref)
    else:
        rel = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
        print("rel rms %.3e" % rel, flush=True)
        assert rel <= 1e-13, rel


def test_noalloc_adjoint_matches_allocating_entry():
    # the differentiable-synthesis size (ring groups equal), then a smaller
    # plan on the same scratch
    _compare(512, 1024, 3, 21)
    _compare(64, 128, 1, 22)


def test_noalloc_adjoint_fewer_groups_then_smaller():
    # Nside 1024 / lmax 2048 / 2 maps: the plan's 4 groups need 268 MB, the
    # scratch holds 2 of them; then a smaller plan
    _compare(1024, 2048, 2, 23)
    _compare(32, 64, 3, 24)


def test_noalloc_adjoint_refuses_plan_too_large():
    import torch
    p, gmap, nalm = _plan_and_inputs(16, 4096, 2, 25)
    assert p.info()["adjoint_groups_noalloc"] == 0
    out = torch.zeros((nalm, 2), dtype=torch.complex128, device="cuda")
    rc, err = _noalloc(p, gmap, out)
    assert rc == -1 and "scratch" in err, err
    torch.cuda.synchronize()
    assert not bool(out.abs().any())
    # the same plan still runs through the allocating entry
    p.adjoint_into(gmap, out)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(torch.view_as_real(out)).all())
