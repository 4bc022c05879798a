"""torch.autograd binding of the differentiable fp32 Floyd-Warshall.

    D = floyd_warshall_diff(W)          # W: (n, n) or (B, n, n) float32 CUDA
    loss = (G * D).sum(); loss.backward()  # W.grad from the hand-written kernel

Forward and backward run in libfwx.so (fwx_diff_forward / fwx_diff_backward,
include/fwx.h) on torch's current stream; this module only allocates the
tensors.  keep_pred=False stores no predecessor matrix between the passes;
the backward then recomputes predecessors 256 sources at a time.
Absent edges are +inf in W and receive zero gradient; weights must be > 0.
"""
import torch

from . import _lib


def _shape(W):
    if W.dim() == 2:
        return 1, W.shape[0]
    if W.dim() == 3:
        return W.shape[0], W.shape[1]
    raise ValueError("W must be (n, n) or (batch, n, n)")


class FloydWarshallFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, W, keep_pred=True):
        if not (W.is_cuda and W.dtype == torch.float32 and W.shape[-1] == W.shape[-2]):
            raise ValueError("expected a float32 CUDA tensor of square matrices")
        Wc = W.contiguous()
        batch, n = _shape(Wc)
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


def floyd_warshall_diff(W, keep_pred=True):
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
