"""CPU check of the batched plan's side branches, on a Python model of the
plan's per-graph schedule.

The model restates `schedule()` and `hoist_closures()` of
fw_cuda_amd/csrc/fwx_batched.hip step by step: per stage of kappa k-tiles,
the lead rounds (closure of the diagonal tile; panels of its row and column;
cross updates of the stage's other rows and columns), then the rest pass over
the blocks outside the stage; then each closure step that follows a product
step moves, with the one product item whose C block holds its diagonal tile,
onto a side branch of that step.  For every step with a side branch the test
checks, cell by cell, that the two branches touch disjoint cells (no cell
written by one is read or written by the other), and it runs the plan on
random graphs -- side branch before the main items and after them -- against
the oracle's serial Floyd-Warshall.  On the GPU the product path itself is
checked against the oracle in tests/test_gpu_batched.py, with and without
side branches.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

INF = 1_000_000_000


def schedule(n, T, kappa, maxR=128):
    """Steps of one graph: lists of product items (block, (cr, cc, ar, ac,
    br, bc, depth)) and of closures (T, c0), as fwx_batched.hip builds them."""
    N = (n + T - 1) // T
    span = kappa * T
    R = min(maxR, 128 if span % 128 == 0 else 64 if span % 64 == 0 else 32)
    ps, ds = [], []

    def step():
        ps.append([])
        ds.append([])

    def prod(RR, cr, cc, ar, ac, br, bc, depth):
        ps[-1].append((RR, (cr, cc, ar, ac, br, bc, depth)))

    for kb in range(0, N, kappa):
        B = min(kappa, N - kb)
        for l in range(B):
            c = kb + l
            step()
            ds[-1].append((T, c * T))
            if N > 1:
                step()
                for j in range(N):
                    if j != c:
                        prod(T, c * T, j * T, c * T, c * T, c * T, j * T, T)
                for i in range(N):
                    if i != c:
                        prod(T, i * T, c * T, i * T, c * T, c * T, c * T, T)
            if B > 1:
                first = True
                for i in range(N):
                    for j in range(N):
                        cross = (kb <= i < kb + B) or (kb <= j < kb + B)
                        if not cross or i == c or j == c:
                            continue
                        if first:
                            step()
                            first = False
                        prod(T, i * T, j * T, i * T, c * T, c * T, j * T, T)
        s0, S = kb * T, B * T
        nr = (n + R - 1) // R
        sr0 = s0 // R
        srn = min((S + R - 1) // R, nr - sr0)
        if nr - srn > 0:
            step()
            for I in range(nr):
                if sr0 <= I < sr0 + srn:
                    continue
                for J in range(nr):
                    if sr0 <= J < sr0 + srn:
                        continue
                    prod(R, I * R, J * R, I * R, s0, s0, J * R, S)
    return ps, ds


def hoist(ps, ds, T):
    sps = [[] for _ in ps]
    sds = [[] for _ in ps]
    j = 1
    while j < len(ps):
        moved = False
        if len(ds[j]) == 1 and not ps[j] and not sds[j - 1] and ps[j - 1]:
            c0 = ds[j][0][1]
            for i, (RR, it) in enumerate(ps[j - 1]):
                if it[0] <= c0 and c0 + T <= it[0] + RR and it[1] <= c0 and c0 + T <= it[1] + RR:
                    sps[j - 1].append(ps[j - 1].pop(i))
                    sds[j - 1].append(ds[j][0])
                    del ps[j], ds[j], sps[j], sds[j]
                    moved = True
                    break
        if not moved:
            j += 1
    return ps, ds, sps, sds


def prod_cells(n, RR, it):
    cr, cc, ar, ac, br, bc, depth = it
    rd = np.zeros((n, n), bool)
    wr = np.zeros((n, n), bool)
    wr[cr:cr + RR, cc:cc + RR] = True
    rd[cr:cr + RR, cc:cc + RR] = True
    rd[ar:ar + RR, ac:ac + depth] = True
    rd[br:br + depth, bc:bc + RR] = True
    return rd, wr


def diag_cells(n, T, c0):
    w = np.zeros((n, n), bool)
    w[c0:c0 + T, c0:c0 + T] = True
    return w, w


def run_prod(D, RR, it):
    cr, cc, ar, ac, br, bc, depth = it
    n = D.shape[0]
    if cr >= n or cc >= n or ac >= n:
        return
    A = D[cr:cr + RR, ac:ac + depth].astype(np.int64)
    B = D[ac:ac + depth, cc:cc + RR].astype(np.int64)
    C = D[cr:cr + RR, cc:cc + RR].astype(np.int64)
    for k in range(A.shape[1]):
        C = np.minimum(C, A[:, k:k + 1] + B[k:k + 1, :])
    D[cr:cr + RR, cc:cc + RR] = C


def run_diag(D, T, c0):
    t = D[c0:c0 + T, c0:c0 + T].astype(np.int64)
    for k in range(t.shape[0]):
        t = np.minimum(t, t[:, k:k + 1] + t[k:k + 1, :])
    D[c0:c0 + T, c0:c0 + T] = t


@pytest.mark.parametrize("n", [31, 100, 129, 256, 300])
@pytest.mark.parametrize("T,kappa", [(32, 1), (32, 4), (64, 2), (64, 3), (128, 1), (128, 3)])
def test_side_branches_disjoint_and_exact(n, T, kappa):
    ps, ds = schedule(n, T, kappa)
    closures = sum(len(d) for d in ds)
    ps, ds, sps, sds = hoist(ps, ds, T)
    # every closure but the graph's first moves onto a side branch
    assert sum(len(d) for d in sds) == closures - 1
    for j in range(len(ps)):
        if not sps[j]:
            continue
        mr, mw, sr, sw = (np.zeros((n, n), bool) for _ in range(4))
        for RR, it in ps[j]:
            r, w = prod_cells(n, RR, it)
            mr |= r
            mw |= w
        for RR, it in sps[j]:
            r, w = prod_cells(n, RR, it)
            sr |= r
            sw |= w
        for Td, c0 in sds[j]:
            r, w = diag_cells(n, Td, c0)
            sr |= r
            sw |= w
        assert not (mw & (sr | sw)).any(), ("main writes a side cell", n, T, kappa, j)
        assert not (sw & mr).any(), ("side writes a main read", n, T, kappa, j)
    rng = np.random.default_rng(n * 1000 + T * 10 + kappa)
    W = np.where(rng.random((n, n)) < 0.4, rng.integers(1, 1001, (n, n)), INF).astype(np.int32)
    np.fill_diagonal(W, 0)
    ref = po.fw_serial(W)
    for side_first in (True, False):
        D = W.copy()
        for j in range(len(ps)):
            def main():
                for Td, c0 in ds[j]:
                    run_diag(D, Td, c0)
                for RR, it in ps[j]:
                    run_prod(D, RR, it)

            def side():
                for RR, it in sps[j]:
                    run_prod(D, RR, it)
                for Td, c0 in sds[j]:
                    run_diag(D, Td, c0)
            if side_first:
                side()
                main()
            else:
                main()
                side()
        assert np.array_equal(D, ref), (n, T, kappa, side_first)
