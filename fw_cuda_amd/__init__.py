"""fw_cuda_amd -- MI355X-native Floyd-Warshall all-pairs shortest paths.

Python side of the engine.  All computation happens in libfwx.so (HIP
kernels for gfx950 behind the C ABI in include/fwx.h); this module only moves
buffers and names the entry points.  The names mirror the reference's API:
the GPU algorithm registry (src/cuda_utils.cu:7-11), GPUMemory
(src/gpu_memory.cu), generateGraph1D (src/graph_generator.cpp:48) and
hasNegativeCycle (src/utils.cpp:5).
"""
import ctypes

import numpy as np

from ._lib import (INF, MAX_EDGE_WEIGHT, MAX_N, FwxError, check, lib, ptr)  # noqa: F401

# reference GPU_ALGORITHMS, src/cuda_utils.cu:7-11
GPU_ALGORITHMS = ("baseline_gpu", "tiling_gpu", "multi_layer_tiling_gpu")


def get_gpu_algorithm_names():
    L = lib()
    return tuple(L.fwx_gpu_algorithm_name(i).decode() for i in range(L.fwx_gpu_algorithm_count()))


def is_gpu_algorithm(name):
    """reference isGPUAlgorithm (src/cuda_utils.cu:21-23)."""
    return bool(lib().fwx_is_gpu_algorithm(name.encode()))


def tuned_config(n):
    t, k = ctypes.c_int32(), ctypes.c_int32()
    check(lib().fwx_tuned_config(int(n), ctypes.byref(t), ctypes.byref(k)))
    return t.value, k.value


def generate_graph(n, density, seed=0, wmin=0, wmax=MAX_EDGE_WEIGHT):
    """Seeded random directed graph (reference generateGraph1D,
    src/graph_generator.cpp:48-80; identical output for wmin=0, wmax=25000)."""
    out = np.empty((int(n), int(n)), dtype=np.int32)
    check(lib().fwx_generate_graph(ptr(out), int(n), float(density), int(seed) & 0xFFFFFFFF,
                                   int(wmin), int(wmax)))
    return out


def has_negative_cycle(D):
    """reference hasNegativeCycle (src/utils.cpp:5-12)."""
    D = np.ascontiguousarray(D, dtype=np.int32)
    return bool(lib().fwx_has_negative_cycle(ptr(D), int(D.shape[0])))


class DeviceMatrix:
    """n x n int32 device buffer owned by the caller (reference GPUMemory,
    src/gpu_memory.cu:6-27): allocate, copy in, run in place, copy out, free."""

    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        check(lib().fwx_device_alloc(ctypes.byref(p), self.n))
        self.ptr = p

    def upload(self, h):
        h = np.ascontiguousarray(h, dtype=np.int32)
        assert h.shape == (self.n, self.n)
        check(lib().fwx_copy_to_device(self.ptr, ptr(h), self.n))

    def download(self):
        out = np.empty((self.n, self.n), dtype=np.int32)
        check(lib().fwx_copy_to_host(ptr(out), self.ptr, self.n))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            check(lib().fwx_device_free(self.ptr))
        self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def run_device(algo, d_ptr, n, tile=0, kappa=0, stream=None):
    """In-place FW on a device buffer through the registry (fwx_run)."""
    s = ctypes.c_void_p(stream) if stream else None
    if algo == "tiling_gpu_refshape":
        return check(lib().fwx_tiling_gpu_refshape(d_ptr, int(n), s))
    return check(lib().fwx_run(algo.encode(), d_ptr, int(n), int(tile), int(kappa), s))


def floyd_warshall(D, algo="multi_layer_tiling_gpu", tile=0, kappa=0):
    """Host matrix in, host matrix out, computed on the GPU."""
    D = np.ascontiguousarray(D, dtype=np.int32)
    n = D.shape[0]
    with DeviceMatrix(n) as dm:
        dm.upload(D)
        run_device(algo, dm.ptr, n, tile, kappa)
        return dm.download()


def floyd_warshall_(t, algo="multi_layer_tiling_gpu", tile=0, kappa=0):
    """In place on a square int32 torch CUDA tensor, on torch's current stream."""
    import torch

    if not (t.is_cuda and t.dtype == torch.int32 and t.dim() == 2 and t.shape[0] == t.shape[1]
            and t.is_contiguous()):
        raise ValueError("expected a contiguous square int32 CUDA tensor")
    stream = torch.cuda.current_stream(t.device).cuda_stream
    run_device(algo, ptr(t), t.shape[0], tile, kappa, stream)
    return t
