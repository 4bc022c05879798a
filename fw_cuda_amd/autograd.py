"""torch.autograd binding of the differentiable fp32 Floyd-Warshall.

    D = floyd_warshall_diff(W)          # W: (n, n) or (B, n, n) float32 CUDA
    loss = (G * D).sum(); loss.backward()  # W.grad from the hand-written kernel

Forward and backward run in libfwx.so (fwx_diff_forward / fwx_diff_backward,
include/fwx.h) on torch's current stream; this module only allocates the
tensors.  The forward keeps the int32 predecessor matrix (4 n^2 B per graph)
for the backward when it fits: by default (keep_pred=None) when it takes at
most PRED_FREE_FRACTION of the device memory free at the forward (the
device's free memory plus what torch's caching allocator holds unused).
Otherwise, or with keep_pred=False, nothing is stored between the passes and
the backward recomputes predecessors 256 sources at a time.  keep_pred=True
always stores it.  Absent edges are +inf in W and receive zero gradient;
weights must be > 0 (a predecessor row that is not a tree makes the backward
raise).
"""
import ctypes

import torch

from . import _lib

# The predecessor matrix is kept when it needs at most this fraction of the
# free device memory.
PRED_FREE_FRACTION = 0.25

# Whether the last forward kept its predecessor matrix (for callers and tests
# that need to know which backward ran).
last_kept_pred = None


def _shape(W):
    if W.dim() == 2:
        return 1, W.shape[0]
    if W.dim() == 3:
        return W.shape[0], W.shape[1]
    raise ValueError("W must be (n, n) or (batch, n, n)")


def free_device_memory(device):
    """Bytes a new allocation on `device` can use: the device's free memory
    (fwx_device_memory) plus torch's cached, unallocated blocks."""
    free = ctypes.c_uint64()
    total = ctypes.c_uint64()
    with torch.cuda.device(device):
        _lib.check(_lib.lib().fwx_device_memory(ctypes.byref(free), ctypes.byref(total)))
    cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return int(free.value) + max(0, int(cached))


_total_memory = {}


def choose_keep_pred(n, batch, device):
    """True when the n x n x batch int32 predecessor matrix fits in
    PRED_FREE_FRACTION of the free device memory.  A matrix below 1/256 of
    the device's total memory is kept without asking for the free memory
    (the query costs a few hundred microseconds per forward) unless
    PRED_FREE_FRACTION is set below 1/64."""
    need = 4 * n * n * batch
    idx = torch.device(device).index or 0
    if idx not in _total_memory:
        free = ctypes.c_uint64()
        total = ctypes.c_uint64()
        with torch.cuda.device(device):
            _lib.check(_lib.lib().fwx_device_memory(ctypes.byref(free), ctypes.byref(total)))
        _total_memory[idx] = int(total.value)
    if need * 256 <= _total_memory[idx] and PRED_FREE_FRACTION >= 1.0 / 64:
        return True
    return need <= PRED_FREE_FRACTION * free_device_memory(device)


class FloydWarshallFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, keep_pred=None):
        global last_kept_pred
        if not (W.is_cuda and W.dtype == torch.float32 and W.shape[-1] == W.shape[-2]):
            raise ValueError("expected a float32 CUDA tensor of square matrices")
        Wc = W.contiguous()
        batch, n = _shape(Wc)
        if keep_pred is None:
            keep_pred = choose_keep_pred(n, batch, Wc.device)
        last_kept_pred = bool(keep_pred)
        D = torch.empty_like(Wc)
        P = torch.empty(Wc.shape, dtype=torch.int32, device=Wc.device) if keep_pred else None
        stream = torch.cuda.current_stream(Wc.device).cuda_stream
        _lib.check(_lib.lib().fwx_diff_forward(_lib.ptr(Wc), _lib.ptr(D),
                                               _lib.ptr(P) if P is not None else None,
                                               n, batch, stream))
        ctx.save_for_backward(Wc, D, P)
        return D

    @staticmethod
    def backward(ctx, G):
        W, D, P = ctx.saved_tensors
        batch, n = _shape(W)
        G = G.contiguous().to(torch.float32)
        dW = torch.empty_like(W)
        stream = torch.cuda.current_stream(W.device).cuda_stream
        _lib.check(_lib.lib().fwx_diff_backward(_lib.ptr(W), _lib.ptr(D),
                                                _lib.ptr(P) if P is not None else None,
                                                _lib.ptr(G), _lib.ptr(dW), n, batch, stream))
        return dW, None


def floyd_warshall_diff(W, keep_pred=None):
    """Differentiable all-pairs shortest path lengths of W (see module doc)."""
    return FloydWarshallFn.apply(W, keep_pred)


def forward_with_pred(W):
    """(D, P) without autograd bookkeeping: P[i][j] is the vertex before j on
    the chosen i->j path, -1 if none."""
    Wc = W.contiguous()
    batch, n = _shape(Wc)
    D = torch.empty_like(Wc)
    P = torch.empty(Wc.shape, dtype=torch.int32, device=Wc.device)
    stream = torch.cuda.current_stream(Wc.device).cuda_stream
    _lib.check(_lib.lib().fwx_diff_forward(_lib.ptr(Wc), _lib.ptr(D), _lib.ptr(P), n, batch, stream))
    return D, P
