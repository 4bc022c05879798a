"""CPU model of the block lead phase of the pk16 multi-layer schedule
(fwx_blocked.hip, `lead_block`), checked against the oracle's fwBaselineCPU
restatement (src/fw_baseline_cpu.cpp:3-16).

Per stage of kappa k-tiles K (width L = kappa T), instead of kappa lead
rounds of (closure, panels, cross):
  1. the L x L stage block closes by min-plus squaring in the pk16 form,
     block <- min(block, block (x) block), one launch per squaring with every
     tile of the block updated in place by its own workgroup (so a tile reads
     a mix of other tiles' old and new values), and a launch after one that
     changed no cell returns at once; log2(L) launches at most;
  2. every row-panel tile (k, j) and column-panel tile (i, k) outside the
     block takes one product of depth L with the closed block, in place
     (again a mix of old and new panel values);
  3. the rest pass as before.
The model runs the tile updates of each launch in a random order, each
reading the matrix as it stands (the in-place mix), with the pk16 arithmetic
of tests/test_pk16_model.py.  Every cell below the cap must be exact and the
exact finish must restore the rest.  Planted mistakes -- a single squaring,
stopping the squarings when one tile is unchanged, panels of depth T --
fail.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from test_pk16_model import CAP, exact_finish, graph, product_pk16


def lead_block_fw(W, T, kappa, seed=0, squarings=None, stop_on_any_unchanged=False, panel_depth=None):
    n = W.shape[0]
    D = W.astype(np.int64)
    N = (n + T - 1) // T
    rng = np.random.default_rng(seed)
    span = lambda t: (t * T, min(n, (t + 1) * T))  # noqa: E731
    for kb in range(0, N, kappa):
        B = min(kappa, N - kb)
        K = list(range(kb, kb + B))
        s0, s1 = kb * T, min(n, (kb + B) * T)
        # 1. closure of the stage block by squaring, tile by tile in place
        nsq = squarings if squarings is not None else max(1, math.ceil(math.log2(B * T)))
        changed_prev = True
        for it in range(nsq):
            if not changed_prev:
                break
            changed = False
            unchanged_tiles = 0
            for t in rng.permutation(B * B):
                a, b = K[t // B], K[t % B]
                a0, a1 = span(a)
                b0, b1 = span(b)
                C = D[a0:a1, b0:b1]
                before = C.copy()
                product_pk16(C, D[a0:a1, s0:s1].copy(), D[s0:s1, b0:b1].copy())
                if (C != before).any():
                    changed = True
                else:
                    unchanged_tiles += 1
            changed_prev = changed and not (stop_on_any_unchanged and unchanged_tiles > 0)
        # 2. row and column panels: one product of depth L each, in place
        p1 = s1 if panel_depth is None else min(s1, s0 + panel_depth)
        tiles = [("r", a, j) for a in K for j in range(N) if j not in K] + \
                [("c", i, b) for b in K for i in range(N) if i not in K]
        for t in rng.permutation(len(tiles)):
            kind, x, y = tiles[t]
            x0, x1 = span(x)
            y0, y1 = span(y)
            C = D[x0:x1, y0:y1]
            product_pk16(C, D[x0:x1, s0:p1].copy(), D[s0:p1, y0:y1].copy())
        # 3. rest pass
        for i in range(N):
            for j in range(N):
                if i in K or j in K:
                    continue
                i0, i1 = span(i)
                j0, j1 = span(j)
                product_pk16(D[i0:i1, j0:j1], D[i0:i1, s0:s1].copy(), D[s0:s1, j0:j1].copy())
    return D


def chain(n, back=3000):
    """i -> i+1 weighs 1, every other edge `back`: shortest paths run through
    every vertex between their ends, so closing a block of width L needs walks
    of L - 1 edges (every squaring)."""
    W = np.full((n, n), back, dtype=np.int32)
    W[np.arange(n - 1), np.arange(1, n)] = 1
    np.fill_diagonal(W, 0)
    return W


CASES = [
    ("dense, small weights", graph(96, 0.9, 1, 1000, 11), False),
    ("sparse: unreachable pairs", graph(90, 0.02, 1, 1000, 12), True),
    ("weights around the cap", graph(80, 0.5, CAP - 3, CAP + 3, 13), True),
    ("distances straddling the cap", graph(100, 0.08, 1000, 9000, 14), None),
    ("chain", chain(96), False),
]


@pytest.mark.parametrize("name,W,saturates", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("T,kappa", [(8, 2), (8, 4), (16, 3)])
def test_block_lead_is_exact_below_the_cap(name, W, saturates, T, kappa):
    want = po.fw_serial(W).astype(np.int64)
    for seed in (0, 1):
        got = lead_block_fw(W, T, kappa, seed=seed)
        below = got < CAP
        assert (got[below] == want[below]).all(), f"{name}: a cell below the cap is not exact"
        assert (got[~below] >= want[~below]).all() and (want[~below] >= CAP).all(), name
        if saturates is not None:
            assert bool((~below).any()) == saturates, name
        assert (exact_finish(got) == want).all(), f"{name}: the exact finish does not restore the distances"


def test_planted_single_squaring_fails():
    W = chain(96)
    want = po.fw_serial(W).astype(np.int64)
    assert not (lead_block_fw(W, 8, 4, squarings=1) == want).all()


def test_planted_stop_when_one_tile_is_unchanged_fails():
    W = chain(96)
    want = po.fw_serial(W).astype(np.int64)
    assert not (lead_block_fw(W, 8, 4, stop_on_any_unchanged=True) == want).all()


def test_planted_shallow_panels_fail():
    W = graph(96, 0.3, 1, 1000, 15)
    want = po.fw_serial(W).astype(np.int64)
    assert not (lead_block_fw(W, 8, 4, panel_depth=8) == want).all()
