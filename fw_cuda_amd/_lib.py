"""ctypes binding of libfwx.so, the C ABI declared in include/fwx.h.

The library is built in-tree (fw_cuda_amd/lib/libfwx.so by
`make -C fw_cuda_amd`, or __graft_entry__.build()).  Loading fails loudly if
it is missing: there is no fallback implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfwx.so")

INF = 1_000_000_000      # include/constants.h:3
MAX_N = 32769            # include/constants.h:5
MAX_EDGE_WEIGHT = 25_000  # include/constants.h:7

OK = 0
ERR_INVALID_ARG = 1
ERR_UNKNOWN_ALGO = 2
ERR_HIP = 3
ERR_OOM = 4
ERR_UNSUPPORTED = 5
ERR_TIMEOUT = 6


class FwxError(RuntimeError):
    """A libfwx call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libfwx status {code}: {msg}")
        self.code = code


_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "fwx_last_error": (ctypes.c_char_p, []),
    "fwx_version": (ctypes.c_int, []),
    "fwx_gpu_algorithm_count": (ctypes.c_int, []),
    "fwx_gpu_algorithm_name": (ctypes.c_char_p, [ctypes.c_int]),
    "fwx_is_gpu_algorithm": (ctypes.c_int, [ctypes.c_char_p]),
    "fwx_run": (ctypes.c_int, [ctypes.c_char_p, _vp, _i32, _i32, _i32, _vp]),
    "fwx_baseline_gpu": (ctypes.c_int, [_vp, _i32, _vp]),
    "fwx_tiling_gpu": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "fwx_tiling_gpu_refshape": (ctypes.c_int, [_vp, _i32, _vp]),
    "fwx_multi_layer_tiling_gpu": (ctypes.c_int, [_vp, _i32, _i32, _i32, _vp]),
    "fwx_tuned_config": (ctypes.c_int, [_i32, _i32p, _i32p]),
    "fwx_relaxation_form": (ctypes.c_int, [_vp, _i32, _i32p, _vp]),
    "fwx_last_solve_form": (ctypes.c_int, [_i32p, _i32p]),
    "fwx_last_solve_packed": (ctypes.c_int, [_i32p]),
    "fwx_event_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "fwx_event_record": (ctypes.c_int, [_vp, _vp]),
    "fwx_event_synchronize": (ctypes.c_int, [_vp]),
    "fwx_event_elapsed_ms": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), _vp, _vp]),
    "fwx_event_destroy": (ctypes.c_int, [_vp]),
    "fwx_device_alloc": (ctypes.c_int, [ctypes.POINTER(_vp), _i32]),
    "fwx_copy_to_device": (ctypes.c_int, [_vp, _vp, _i32]),
    "fwx_copy_to_host": (ctypes.c_int, [_vp, _vp, _i32]),
    "fwx_device_free": (ctypes.c_int, [_vp]),
    "fwx_device_synchronize": (ctypes.c_int, []),
    "fwx_device_memory": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "fwx_warmup": (ctypes.c_int, []),
    "fwx_generate_graph": (ctypes.c_int, [_vp, _i32, ctypes.c_double, ctypes.c_uint32, _i32, _i32]),
    "fwx_has_negative_cycle": (ctypes.c_int, [_vp, _i32]),
    "fwx_has_negative_cycle_device": (ctypes.c_int, [_vp, _i32, _i32p, _vp]),
    "fwx_generate_graph_device": (ctypes.c_int, [_vp, _i32, ctypes.c_double, ctypes.c_uint64, _i32, _i32, _vp]),
    "fwx_host_alloc_pinned": (ctypes.c_int, [ctypes.POINTER(_vp), _i32]),
    "fwx_host_free_pinned": (ctypes.c_int, [_vp]),
    "fwx_copy_to_device_async": (ctypes.c_int, [_vp, _vp, _i32, _vp]),
    "fwx_copy_to_host_async": (ctypes.c_int, [_vp, _vp, _i32, _vp]),
    "fwx_stream_synchronize": (ctypes.c_int, [_vp]),
    "fwx_predecessors": (ctypes.c_int, [_vp, _vp, _vp, _i32, _vp]),
    "fwx_certify_rows": (ctypes.c_int, [_vp, _i32, _vp, _i32, _vp, _vp, _i32p, _vp]),
    "fwx_batch_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _vp, _i32, _vp, _vp]),
    "fwx_batch_plan_run": (ctypes.c_int, [_vp, _i32]),
    "fwx_batch_plan_stats": (ctypes.c_int, [_vp, _i32p, _i32p, ctypes.POINTER(_i64)]),
    "fwx_batch_plan_last_form": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "fwx_batch_plan_persistent_info": (ctypes.c_int, [_vp, _i32p, _i32p, ctypes.POINTER(_i64)]),
    "fwx_batch_plan_destroy": (ctypes.c_int, [_vp]),
    "fwx_batch_loop": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "fwx_diff_forward": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp]),
    "fwx_diff_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "fwx_diff_backward_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "fwx_profile_begin": (ctypes.c_int, []),
    "fwx_profile_end": (ctypes.c_int, [_vp, _vp, _vp]),
}

PROF_REST, PROF_LEAD, PROF_BASELINE, PROF_BAND, PROF_NCLASS = 0, 1, 2, 3, 4
FORM_MIN3, FORM_PK64, FORM_PK16 = 0, 1, 2
PK16_CAP = 15359  # include/fwx.h FWX_PK16_CAP

_lib = None


def _share_torch_runtime():
    """A process must hold one HIP runtime.  PyTorch-ROCm wheels carry their
    own libamdhip64; loading it before libfwx makes libfwx bind to that same
    runtime instead of starting a second one from /opt/rocm (two runtimes in
    one process do not share the device)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libfwx.so once (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        _share_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built: run `make -C fw_cuda_amd` "
                          "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise FwxError(rc, lib().fwx_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Address of a numpy array or torch tensor's data as a ctypes void*."""
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return ctypes.c_void_p(a.ctypes.data)
