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


def generate_graph_device(d_ptr, n, density, seed, wmin=1, wmax=1000, stream=None):
    """fwx_generate_graph_device: Philox-based graph generated in place on the
    GPU (same distributions as generate_graph, a different stream)."""
    s = ctypes.c_void_p(stream) if stream else None
    check(lib().fwx_generate_graph_device(d_ptr, int(n), float(density), int(seed) & (2**64 - 1),
                                          int(wmin), int(wmax), s))


def floyd_warshall_with_pred(W, algo="multi_layer_tiling_gpu", tile=0, kappa=0):
    """Distances and the predecessor matrix (fwx_predecessors) of host matrix
    W: P[i][j] is the vertex before j on a shortest i->j path (-1 on the
    diagonal)."""
    W = np.ascontiguousarray(W, dtype=np.int32)
    n = W.shape[0]
    with DeviceMatrix(n) as dw, DeviceMatrix(n) as dd, DeviceMatrix(n) as dp:
        dw.upload(W)
        dd.upload(W)
        run_device(algo, dd.ptr, n, tile, kappa)
        check(lib().fwx_predecessors(dw.ptr, dd.ptr, dp.ptr, n, None))
        return dd.download(), dp.download()


def has_negative_cycle_device(d_ptr, n, stream=None):
    """fwx_has_negative_cycle_device: the diagonal check on a device matrix."""
    r = ctypes.c_int32()
    s = ctypes.c_void_p(stream) if stream else None
    check(lib().fwx_has_negative_cycle_device(d_ptr, int(n), ctypes.byref(r), s))
    return bool(r.value)


def certify_rows_device(d_W, n, d_src, nsrc, d_rows, d_ok, stream=None):
    """fwx_certify_rows on device buffers (raw pointers): d_ok[q] = 1 iff claimed
    row q is the exact shortest-distance row from d_src[q]; returns the number
    of failing rows."""
    failed = ctypes.c_int32()
    s = ctypes.c_void_p(stream) if stream else None
    check(lib().fwx_certify_rows(d_W, int(n), d_src, int(nsrc), d_rows, d_ok, ctypes.byref(failed), s))
    return int(failed.value)


def certify_rows(W, src, rows):
    """Host convenience wrapper of fwx_certify_rows: W (n x n), src (nsrc,),
    rows (nsrc x n) claimed distance rows -> per-row bool array."""
    import torch
    W = np.ascontiguousarray(W, dtype=np.int32)
    src = np.ascontiguousarray(src, dtype=np.int32)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    n = W.shape[0]
    assert W.shape == (n, n) and rows.shape == (src.size, n)
    dev = torch.device("cuda")
    tw, ts, tr = (torch.from_numpy(a).to(dev) for a in (W, src, rows))
    ok = torch.zeros(src.size, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    certify_rows_device(ctypes.c_void_p(tw.data_ptr()), n, ctypes.c_void_p(ts.data_ptr()), src.size,
                        ctypes.c_void_p(tr.data_ptr()), ctypes.c_void_p(ok.data_ptr()), stream)
    return ok.cpu().numpy().astype(bool)


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


def last_solve_form():
    """(form, exact_finish) of this thread's last plain blocked solve
    (fwx_last_solve_form): form is FORM_MIN3, FORM_PK64, FORM_PK16 or -1 for
    the fused small-n rounds; exact_finish is 1 when the pk16 run left cells
    at or above the cap and an exact pk64 run finished the solve."""
    f, e = ctypes.c_int32(), ctypes.c_int32()
    check(lib().fwx_last_solve_form(ctypes.byref(f), ctypes.byref(e)))
    return f.value, e.value


def last_solve_packed():
    """True when this thread's last plain blocked solve ran its pk16 rest
    passes on packed stage panels (fwx_last_solve_packed)."""
    p = ctypes.c_int32()
    check(lib().fwx_last_solve_packed(ctypes.byref(p)))
    return bool(p.value)


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


BATCH_MODES = {"direct": 0, "graph": 1, "persistent": 2, "persistent_graph": 3}


class BatchPlan:
    """Many graphs of uneven size solved together (fwx_batch_plan_*).

    `sizes[g]` and element `offsets[g]` locate graph g inside the int32 device
    buffer at address `d_data`.  run(mode) with mode one of BATCH_MODES:
    "graph" replays the captured hipGraph of the step launches, "direct"
    enqueues them, "persistent" runs the whole batch as one persistent
    cooperative kernel (grid-wide barriers between steps) and
    "persistent_graph" replays that one launch from a hipGraph.  A bool is
    accepted for the first two (True = "graph")."""

    def __init__(self, sizes, offsets, d_data, stream=None):
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        p = ctypes.c_void_p()
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().fwx_batch_plan_create(ctypes.byref(p), ptr(self.sizes), ptr(self.offsets),
                                          len(self.sizes), d_data, s))
        self.plan = p

    def run(self, mode="graph"):
        if isinstance(mode, bool):
            mode = "graph" if mode else "direct"
        check(lib().fwx_batch_plan_run(self.plan, BATCH_MODES[mode]))

    def persistent_info(self):
        """The persistent mode's cooperative grid (blocks; 0 without
        cooperative launches), grid-wide steps and tasks."""
        g, st, t = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(lib().fwx_batch_plan_persistent_info(self.plan, ctypes.byref(g), ctypes.byref(st), ctypes.byref(t)))
        return {"grid": g.value, "steps": st.value, "tasks": t.value}

    def stats(self):
        st, la, it = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(lib().fwx_batch_plan_stats(self.plan, ctypes.byref(st), ctypes.byref(la), ctypes.byref(it)))
        return {"steps": st.value, "launches": la.value, "items": it.value}

    def last_form(self):
        """(form, exact_graphs) of the last run: FORM_MIN3 / FORM_PK64 /
        FORM_PK16, and the graphs its pk16 check finished in pk64."""
        f, e = ctypes.c_int32(), ctypes.c_int32()
        check(lib().fwx_batch_plan_last_form(self.plan, ctypes.byref(f), ctypes.byref(e)))
        return f.value, e.value

    def destroy(self):
        if self.plan is not None and self.plan.value:
            check(lib().fwx_batch_plan_destroy(self.plan))
        self.plan = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def batch_loop(sizes, offsets, d_data, stream=None):
    """Comparator: one plain multi_layer_tiling_gpu call per graph."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    s = ctypes.c_void_p(stream) if stream else None
    check(lib().fwx_batch_loop(ptr(sizes), ptr(offsets), len(sizes), d_data, s))


def batch_layout(sizes, align=4):
    """Element offsets packing the graphs back to back, each start rounded up
    to `align` elements (16 bytes for align=4)."""
    offs, pos = [], 0
    for n in sizes:
        offs.append(pos)
        pos += int(n) * int(n)
        pos = (pos + align - 1) // align * align
    return np.array(offs, dtype=np.int64), pos


def mixed_batch(seed=11, count=512, lo=64, hi=2048):
    """The batched workload: `count` graphs with n log-uniform in [lo, hi]
    (numpy PCG64 seeded with `seed`), density per graph U[0.05, 0.9], int32
    weights U[1, 1000]; graph g uses generator seed seed*100000 + g + 1."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    ns = np.clip(ns, lo, hi).astype(np.int32)
    dens = rng.uniform(0.05, 0.9, count)
    seeds = [seed * 100000 + g + 1 for g in range(count)]
    return ns, dens, seeds
