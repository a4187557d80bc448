"""ctypes binding of libwavemoth_hip.so (the C ABI of include/wavemoth.h).

The library is built in-tree by wavemoth_amd.build; loading fails loudly
when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwavemoth_hip.so")
_LIB = None

# Every symbol include/wavemoth.h declares, with its ctypes signature.
_c_int, _c_u, _c_i64 = ctypes.c_int, ctypes.c_uint, ctypes.c_int64
_vp, _dp, _cp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
_sz = ctypes.c_size_t
SIGNATURES = {
    "wavemoth_configure": (None, [_cp]),
    "wavemoth_plan_to_healpix": (_vp, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                       _c_int, _c_u, _cp]),
    "wavemoth_destroy_plan": (None, [_vp]),
    "wavemoth_execute": (None, [_vp]),
    "wavemoth_get_legendre_flops": (_c_i64, [_vp, _c_int, _c_int]),
    "wavemoth_query_resourcefile": (_c_int, [_cp, ctypes.POINTER(_c_int),
                                             ctypes.POINTER(_c_int)]),
    "wavemoth_resource_info": (_c_int, [_cp, ctypes.POINTER(_c_i64), _c_int]),
    "wavemoth_perform_legendre_transforms": (None, [_vp]),
    "wavemoth_perform_backward_ffts": (None, [_vp]),
    "wavemoth_disable_phase_shifting": (None, [_vp]),
    "wavemoth_last_error": (_cp, []),
    "wavemoth_get_work": (_c_int, [_vp, _vp]),
    "wavemoth_execute_device": (_c_int, [_vp, _vp, _vp, _vp]),
    "wavemoth_execute_adjoint_device": (_c_int, [_vp, _vp, _vp, _vp]),
    "wavemoth_execute_adjoint": (None, [_vp]),
    "wavemoth_execute_adjoint_device_noalloc": (_c_int, [_vp, _vp, _vp, _vp]),
    "wavemoth_set_profiling": (_c_int, [_vp, _c_int]),
    "wavemoth_stage_times": (_c_int, [_vp, _dp, ctypes.POINTER(_c_int)]),
    "wavemoth_plan_info": (_c_i64, [_vp, _c_int]),
    "wavemoth_plan_batch": (_vp, [_c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                  ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), _c_u]),
    "wavemoth_destroy_batch": (None, [_vp]),
    "wavemoth_execute_batch_device": (_c_int, [_vp, _vp, _vp, _vp]),
    "wavemoth_execute_batch": (_c_int, [_vp, _vp, _vp]),
    "wavemoth_batch_set_profiling": (_c_int, [_vp, _c_int]),
    "wavemoth_batch_stage_times": (_c_int, [_vp, _dp, ctypes.POINTER(_c_int)]),
    "wavemoth_batch_info": (_c_i64, [_vp, _c_int, _c_int]),
    "wavemoth_legendre_transform_auxdata": (None, [_sz, _sz, _sz, _vp]),
    "wavemoth_legendre_transform": (None, [_sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wavemoth_legendre_transform_device": (_c_int, [_sz, _sz, _sz, _vp, _vp, _vp, _vp, _vp,
                                                    _vp, _vp]),
}

WAVEMOTH_MMAJOR = 0x0
WAVEMOTH_ESTIMATE = 0x0
WAVEMOTH_MEASURE = 0x1
WAVEMOTH_NO_RESOURCE_COPY = 0x10
WAVEMOTH_SEEDS_ON_THE_FLY = 0x100
WAVEMOTH_RESOURCE_APPLY = 0x400
WAVEMOTH_BATCH_NO_GRAPH = 0x200
WAVEMOTH_MIXED_F32 = 0x1000
WAVEMOTH_MIXED_BF16X3 = 0x2000
WAVEMOTH_MIXED_BF16X6 = 0x4000
WAVEMOTH_MIXED_F16X3 = 0x8000


def _bind_torch_runtime():
    """PyTorch-ROCm ships its own libamdhip64 / libhsa-runtime64 (same
    sonames as /opt/rocm's).  Two HIP runtimes cannot share a process, so
    torch is loaded first when it is installed: the dynamic loader then
    resolves this library's runtime dependencies to torch's copies."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        _bind_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libwavemoth_hip.so is not built (run `python -m wavemoth_amd.build`); "
                "wavemoth_amd has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def last_error():
    e = lib().wavemoth_last_error()
    return e.decode() if e else ""


# wavemoth_plan_info keys, in key order
PLAN_INFO_KEYS = ["Nside", "lmax", "mmax", "nmaps", "npix", "triples_dense", "triples_eps",
                  "flags", "legendre_launches", "device_bytes", "triples_executed", "scheme",
                  "adjoint_groups", "adjoint_groups_noalloc", "rings_general", "rings_belt",
                  "rings_cap", "rings_wrap", "max_register_fft", "rings_long"]


def plan_info(plan):
    """Every wavemoth_plan_info fact of a plan handle, by name."""
    return {k: int(lib().wavemoth_plan_info(plan, i)) for i, k in enumerate(PLAN_INFO_KEYS)}
