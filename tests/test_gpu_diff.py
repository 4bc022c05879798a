"""Differentiable fp32 FW: forward against the oracle's fp32 FW, backward against
the CPU path-enumeration gradient and float64 central finite differences.

Finite differences: L = sum(G * D) is piecewise linear in W, so central
differences equal the derivative exactly wherever L is linear on
[w - eps, w + eps].  At eps = 1e-3 a minority of edges sit within eps of a
path switch (a kink); those are detected by their one-sided slopes
disagreeing and are reported, not compared.  A second pass at eps = 1e-6
compares every sampled edge.  Tolerance: relative L2 error < 1e-4.
"""
import ctypes
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def graphs(batch, n, p, seed):
    rng = np.random.default_rng(seed)
    W = np.where(rng.random((batch, n, n)) < p, rng.uniform(0.5, 10.0, (batch, n, n)), np.inf)
    W = W.astype(np.float32)
    for b in range(batch):
        np.fill_diagonal(W[b], 0.0)
    G = rng.standard_normal((batch, n, n)).astype(np.float32)
    return W, G


def gpu_fwd_bwd(W, G, keep_pred=True):
    import torch
    from fw_cuda_amd.autograd import floyd_warshall_diff
    Wt = torch.from_numpy(W).cuda().requires_grad_(True)
    D = floyd_warshall_diff(Wt, keep_pred)
    (D * torch.from_numpy(G).cuda()).sum().backward()
    torch.cuda.synchronize()
    return D.detach().cpu().numpy(), Wt.grad.cpu().numpy()


def test_forward_and_backward_small():
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    W, G = graphs(5, 70, 0.3, 1)
    Dg, Pg = forward_with_pred(torch.from_numpy(W).cuda())
    Dg, Pg = Dg.cpu().numpy(), Pg.cpu().numpy()
    D, dW = gpu_fwd_bwd(W, G)
    assert (D == Dg).all()
    for b in range(W.shape[0]):
        Do, Po = po.fw_f32_pred(W[b])
        fin = np.isfinite(Do)
        assert (np.isfinite(Dg[b]) == fin).all()
        assert np.allclose(Dg[b][fin], Do[fin], rtol=2e-6, atol=0)
        same = Pg[b] == Po
        assert same.mean() > 0.999, same.mean()
        ref = po.grad_paths(Po, G[b])
        mine = po.grad_paths(Pg[b], G[b])  # same enumeration over the GPU's tree
        assert np.allclose(dW[b], mine, rtol=1e-5, atol=1e-4)
        if same.all():
            assert np.allclose(dW[b], ref, rtol=1e-5, atol=1e-4)
    # absent edges and the diagonal carry no gradient
    assert (dW[~np.isfinite(W)] == 0).all()


@pytest.mark.parametrize("batch,n", [(3, 129), (2, 300), (1, 640), (4, 200), (5, 70), (64, 512)])
def test_forward_schedules_agree(monkeypatch, batch, n):
    """The single-stream, look-ahead and side-band (two-stream) forwards give
    the same bits for D and P."""
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    W, _ = graphs(batch, n, 0.3, n + batch)
    Wt = torch.from_numpy(W).cuda()
    out = {}
    for la in ("0", "1", "2"):
        monkeypatch.setenv("FWX_LOOKAHEAD", la)
        D, P = forward_with_pred(Wt)
        out[la] = (D.cpu().numpy(), P.cpu().numpy())
    for la in ("1", "2"):
        assert np.array_equal(out["0"][0], out[la][0], equal_nan=True), la
        assert (out["0"][1] == out[la][1]).all(), la
    Do, _ = po.fw_f32_pred(W[0])
    fin = np.isfinite(Do)
    assert np.allclose(out["1"][0][0][fin], Do[fin], rtol=2e-6, atol=0)


@pytest.mark.parametrize("batch,n", [(3, 129), (2, 300), (1, 640), (4, 200), (2, 258), (1, 1023)])
def test_forward_rest_kernels_agree(monkeypatch, batch, n):
    """The vectorised rest product and two-step diagonal closure (default)
    and the first versions (FWX_DIFF_REST=0, FWX_DIFF_DIAG=0) give the same
    bits for D and P, with P stored and without, on 16-byte (n % 4 == 0) and
    scalar (ragged n) paths."""
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    from fw_cuda_amd import _lib
    W, _ = graphs(batch, n, 0.3, 7 * n + batch)
    Wt = torch.from_numpy(W).cuda()
    out = {}
    for v in ("0", "1"):
        monkeypatch.setenv("FWX_DIFF_REST", v)
        monkeypatch.setenv("FWX_DIFF_DIAG", v)
        D, P = forward_with_pred(Wt)
        Dn = torch.empty_like(Wt)
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().fwx_diff_forward(_lib.ptr(Wt), _lib.ptr(Dn), None, n, batch, ctypes.c_void_p(s)))
        torch.cuda.synchronize()
        out[v] = (D.cpu().numpy(), P.cpu().numpy(), Dn.cpu().numpy())
    assert np.array_equal(out["0"][0], out["1"][0], equal_nan=True)
    assert (out["0"][1] == out["1"][1]).all()
    assert np.array_equal(out["0"][2], out["1"][2], equal_nan=True)
    assert np.array_equal(out["1"][0], out["1"][2], equal_nan=True)


@pytest.mark.parametrize("kappa,batch,n", [(2, 2, 1000), (3, 1, 1300), (4, 3, 700), (2, 1, 2100), (3, 2, 1003), (2, 1, 1537)])
def test_forward_multilayer(monkeypatch, kappa, batch, n):
    """Multi-layer stages (kappa k-tiles per stage): the single-stream and
    side-band schedules give the same bits, and D matches the oracle's fp32
    FW (predecessors on >= 99.9 % of cells: a different decomposition may
    round a near-tie differently)."""
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    monkeypatch.setenv("FWX_DIFF_KAPPA", str(kappa))
    W, _ = graphs(batch, n, 0.3, 11 * n + kappa)
    Wt = torch.from_numpy(W).cuda()
    out = {}
    for la in ("0", "2"):
        monkeypatch.setenv("FWX_LOOKAHEAD", la)
        D, P = forward_with_pred(Wt)
        out[la] = (D.cpu().numpy(), P.cpu().numpy())
    assert np.array_equal(out["0"][0], out["2"][0], equal_nan=True)
    assert (out["0"][1] == out["2"][1]).all()
    Do, Po = po.fw_f32_pred(W[0])
    fin = np.isfinite(Do)
    assert (np.isfinite(out["2"][0][0]) == fin).all()
    assert np.allclose(out["2"][0][0][fin], Do[fin], rtol=2e-6, atol=0)
    assert (out["2"][1][0] == Po).mean() > 0.999


def test_backward_deep_trees():
    """Shortest-path trees hundreds of levels deep (a ring with a few light
    chords, and a bare chain): the level-by-level accumulation must give the
    path-enumeration gradient of the GPU's own tree."""
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    rng = np.random.default_rng(12)
    n = 600
    W = np.full((2, n, n), np.inf, dtype=np.float32)
    for b in range(2):
        np.fill_diagonal(W[b], 0.0)
        for v in range(n - 1):
            W[b, v, v + 1] = rng.uniform(0.5, 1.5)
        if b == 0:  # ring + sparse chords
            W[b, n - 1, 0] = 1.0
            for _ in range(20):
                u, v = rng.integers(0, n, 2)
                if u != v:
                    W[b, u, v] = rng.uniform(5.0, 10.0)
    G = rng.standard_normal((2, n, n)).astype(np.float32)
    _, Pg = forward_with_pred(torch.from_numpy(W).cuda())
    Pg = Pg.cpu().numpy()
    _, dW = gpu_fwd_bwd(W, G)
    for b in range(2):
        ref = po.grad_paths(Pg[b], G[b])
        # fp32 sums over subtrees of up to n vertices in different orders:
        # the error scales with the partial sums, so compare globally
        err = float(np.linalg.norm(dW[b] - ref) / np.linalg.norm(ref))
        mx = float(np.abs(dW[b] - ref).max() / np.abs(ref).max())
        say("deep tree", b, "rel L2", err, "max rel", mx)
        assert err < 1e-5 and mx < 1e-4


def test_backward_recomputed_predecessors_match_stored():
    W, G = graphs(3, 300, 0.3, 2)
    _, g1 = gpu_fwd_bwd(W, G, keep_pred=True)
    _, g2 = gpu_fwd_bwd(W, G, keep_pred=False)
    assert np.allclose(g1, g2, rtol=1e-5, atol=1e-4)


def test_keep_pred_follows_free_memory(monkeypatch):
    """With keep_pred left to the binding, the predecessor matrix is stored
    when it fits in the set fraction of free device memory and recomputed in
    the backward when it does not (the threshold forced both ways); the two
    gradients agree, and against path enumeration over the stored P."""
    import fw_cuda_amd.autograd as ag
    W, G = graphs(4, 257, 0.3, 17)
    grads = {}
    for frac, expect in ((1.0, True), (0.0, False)):
        monkeypatch.setattr(ag, "PRED_FREE_FRACTION", frac)
        _, grads[expect] = gpu_fwd_bwd(W, G, keep_pred=None)
        assert ag.last_kept_pred is expect
    assert np.allclose(grads[True], grads[False], rtol=1e-5, atol=1e-4)
    import torch
    _, Pg = ag.forward_with_pred(torch.from_numpy(W).cuda())
    Pg = Pg.cpu().numpy()
    for b in range(W.shape[0]):
        assert np.allclose(grads[False][b], po.grad_paths(Pg[b], G[b]), rtol=1e-4, atol=1e-3)
    # the default threshold keeps P for small problems
    monkeypatch.setattr(ag, "PRED_FREE_FRACTION", 0.25)
    gpu_fwd_bwd(W, G, keep_pred=None)
    assert ag.last_kept_pred is True


def test_backward_rejects_predecessor_cycles():
    """Zero-weight edges 1 <-> 2 reached from source 0 through 3 and 4: the
    recomputed predecessors (smallest tight u) point 1 -> 2 and 2 -> 1, a
    cycle that is not a shortest-path tree.  The backward reports it as an
    error instead of returning a gradient with that cycle's terms dropped."""
    from fw_cuda_amd import FwxError
    n = 64
    W = np.full((1, n, n), np.inf, dtype=np.float32)
    np.fill_diagonal(W[0], 0.0)
    for u, v, w in ((0, 3, 1.0), (0, 4, 1.0), (3, 1, 1.0), (4, 2, 1.0), (1, 2, 0.0), (2, 1, 0.0)):
        W[0, u, v] = w
    G = np.ones((1, n, n), dtype=np.float32)
    # premise, on the oracle: D[0][1] == D[0][2], so both are tight for each other
    D64 = po.fw_f64(W[0].astype(np.float64))
    assert D64[0, 1] == D64[0, 2] == 2.0
    with pytest.raises(FwxError, match="not a tree"):
        gpu_fwd_bwd(W, G, keep_pred=False)
    # the library is usable afterwards: a valid problem still runs
    W2, G2 = graphs(1, 64, 0.3, 9)
    gpu_fwd_bwd(W2, G2, keep_pred=False)


@pytest.mark.parametrize("batch,n,keep", [(3, 129, True), (2, 700, True), (1100, 200, True), (1100, 200, False),
                                          (3, 300, False), (2, 700, False)])
def test_backward_column_form_matches_atomics(monkeypatch, batch, n, keep):
    """dW from each source's g row summed per target column in LDS (default)
    against per-vertex global atomics (FWX_DIFF_BACK=0), and shallow trees
    accumulated by ancestor walks (default) against the level-by-level pass
    (FWX_DIFF_WALK=0): equal up to fp32 summation order.  Stored predecessors put the g rows in dW itself;
    recomputed ones come in chunks of 256 sources (n = 300 and 700: a first
    chunk that writes dW, then chunks that add to it)."""
    W, G = graphs(batch, n, 0.3, batch + n)
    monkeypatch.setenv("FWX_DIFF_BACK", "0")
    monkeypatch.setenv("FWX_DIFF_WALK", "0")
    _, g0 = gpu_fwd_bwd(W, G, keep_pred=keep)
    monkeypatch.delenv("FWX_DIFF_WALK")
    _, gw = gpu_fwd_bwd(W, G, keep_pred=keep)
    monkeypatch.delenv("FWX_DIFF_BACK")
    _, g1 = gpu_fwd_bwd(W, G, keep_pred=keep)
    say("column form", batch, n, keep, float(np.abs(g1 - g0).max()), float(np.abs(gw - g0).max()))
    for g in (g1, gw):
        assert (g[~np.isfinite(W)] == 0).all()
        assert np.allclose(g, g0, rtol=1e-5, atol=1e-4)
        err = float(np.linalg.norm(g - g0) / np.linalg.norm(g0))
        assert err < 1e-6, err


def _fd(W64, G64, u, v, eps):
    Wp, Wm = W64.copy(), W64.copy()
    Wp[u, v] += eps
    Wm[u, v] -= eps
    Dp, Dm = po.fw_f64(Wp), po.fw_f64(Wm)
    fin = np.isfinite(Dp)
    return float((G64[fin] * Dp[fin]).sum()), float((G64[fin] * Dm[fin]).sum())


def test_fd_64_graphs_n512():
    """64 graphs at n=512, density 0.3, weights U[0.5, 10], seed 3: dL/dW on
    256 random present edges (4 per graph) against float64 central FD."""
    W, G = graphs(64, 512, 0.3, 3)
    D, dW = gpu_fwd_bwd(W, G)
    rng = np.random.default_rng(3)
    picks = []
    for b in range(64):
        pres = np.argwhere(np.isfinite(W[b]) & ~np.eye(512, dtype=bool))
        for e in pres[rng.choice(len(pres), 4, replace=False)]:
            picks.append((b, int(e[0]), int(e[1])))
    W64, G64 = W.astype(np.float64), G.astype(np.float64)
    L0 = {}
    for b in range(64):
        D0 = po.fw_f64(W64[b])
        fin = np.isfinite(D0)
        L0[b] = float((G64[b][fin] * D0[fin]).sum())

    def run(eps):
        def one(p):
            b, u, v = p
            return _fd(W64[b], G64[b], u, v, eps)
        with ThreadPoolExecutor(16) as ex:
            return list(ex.map(one, picks))

    ana = np.array([dW[b, u, v] for b, u, v in picks], dtype=np.float64)
    r3 = run(1e-3)
    fd3 = np.array([(lp - lm) / 2e-3 for lp, lm in r3])
    right = np.array([(lp - L0[b]) / 1e-3 for (lp, _), (b, _, _) in zip(r3, picks)])
    left = np.array([(L0[b] - lm) / 1e-3 for (_, lm), (b, _, _) in zip(r3, picks)])
    kink = np.abs(right - left) > 1e-6 * (1.0 + np.abs(fd3))
    ok = ~kink
    rel3 = np.linalg.norm(ana[ok] - fd3[ok]) / max(np.linalg.norm(fd3[ok]), 1e-30)
    say(f"eps=1e-3: {kink.sum()} of {len(picks)} edges within eps of a path switch; rel L2 on the rest {rel3:.2e}")
    assert rel3 < 1e-4
    r6 = run(1e-6)
    fd6 = np.array([(lp - lm) / 2e-6 for lp, lm in r6])
    rel6 = np.linalg.norm(ana - fd6) / max(np.linalg.norm(fd6), 1e-30)
    say(f"eps=1e-6: rel L2 over all {len(picks)} edges {rel6:.2e}")
    assert rel6 < 1e-4
    assert (ana != 0).sum() >= 10  # the sample does exercise used edges


@pytest.mark.slow
def test_fd_n4096():
    """One graph at n=4096 (density 0.3, weights U[0.5, 10], seed 3): D rows of
    16 sampled sources against the oracle's float64 Dijkstra within 2e-6
    relative, every predecessor of those rows tight against the float64
    distances (a path-consistent but non-shortest forward fails here: a
    planted one is checked to fail), path consistency of walks through the
    GPU's predecessors, and dW against path enumeration over them."""
    import torch
    from fw_cuda_amd.autograd import forward_with_pred
    W, G = graphs(1, 4096, 0.3, 3)
    Dg, Pg = forward_with_pred(torch.from_numpy(W).cuda())
    Dg, Pg = Dg.cpu().numpy()[0], Pg.cpu().numpy()[0]
    _, dW = gpu_fwd_bwd(W, G)
    dW = dW[0]
    rows = np.random.default_rng(5).choice(4096, 16, replace=False).astype(np.int32)
    ref = po.dijkstra_f64_rows(W[0], rows)
    bad = po.f32_row_failures(Dg[rows], Pg[rows], W[0], rows, ref)
    say(f"n=4096: {len(rows)} rows vs float64 Dijkstra, {len(bad)} failures {bad[:5]}")
    assert bad == []
    # planted: one cell set to a longer direct edge, its predecessor to the source
    s = int(rows[0])
    t = int(np.flatnonzero(np.isfinite(W[0, s]) & (W[0, s] > 1.5 * Dg[s]) & (np.arange(4096) != s))[0])
    D2, P2 = Dg[rows].copy(), Pg[rows].copy()
    D2[0, t], P2[0, t] = W[0, s, t], s
    assert po.f32_row_failures(D2, P2, W[0], rows, ref) != []
    # walking P from sampled sources reproduces D within fp32 rounding
    rng = np.random.default_rng(4)
    for s in rng.choice(4096, 16, replace=False):
        for t in rng.choice(4096, 64, replace=False):
            if t == s:
                continue
            v, tot, hops = t, 0.0, 0
            while v != s:
                u = Pg[s, v]
                assert u >= 0
                tot += float(W[0, u, v])
                v = u
                hops += 1
                assert hops < 4096
            assert abs(tot - Dg[s, t]) <= 1e-5 * max(1.0, tot)
    ref = po.grad_paths(Pg, G[0])
    assert np.allclose(dW, ref, rtol=1e-4, atol=1e-3)
