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
fail.  The first stage's squarings may split each tile's depth over several
workgroups that merge by atomic min (k_block_square16s); the model
interleaves their reads and writes at random, and plain stores in place of
the atomic min (lost updates) fail.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from test_pk16_model import CAP, exact_finish, graph, product_pk16


def square_split(D, K, span, s0, s1, parts, rng, plain_store=False, reads_first=False):
    """One squaring launch of k_block_square16s: each tile of the stage
    block by `parts` workgroups over disjoint depth ranges.  Every part's
    read (its C tile and operands as the matrix stands) and write are two
    events of a random interleaving, the read first; the write lowers the
    cells whose result differs from the value read and lies below the cap by
    an atomic min (`plain_store` stores them instead: a planted lost
    update).  `reads_first` puts every read before every write (one
    synchronous squaring per launch).  Returns whether any part changed a
    cell."""
    B = len(K)
    depth = (s1 - s0) // parts
    items = [(t, q) for t in range(B * B) for q in range(parts)]
    if reads_first:
        events = np.concatenate([rng.permutation(len(items)), rng.permutation(len(items))])
    else:
        events = rng.permutation(np.repeat(np.arange(len(items)), 2))
    pending = {}
    changed = False
    for e in events:
        t, q = items[e]
        a0, a1 = span(K[t // B])
        b0, b1 = span(K[t % B])
        k0, k1 = s0 + q * depth, s0 + (q + 1) * depth
        if e not in pending:
            was = D[a0:a1, b0:b1].copy()
            res = was.copy()
            product_pk16(res, D[a0:a1, k0:k1].copy(), D[k0:k1, b0:b1].copy())
            pending[e] = (was, res)
            continue
        was, res = pending[e]
        m = (res != was) & (res < CAP)
        changed = changed or bool((res != was).any())
        C = D[a0:a1, b0:b1]
        C[m] = res[m] if plain_store else np.minimum(C[m], res[m])
    return changed


def lead_block_fw(W, T, kappa, seed=0, squarings=None, stop_on_any_unchanged=False, panel_depth=None,
                  split=1, plain_store=False, reads_first=False):
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
            if split > 1 and kb == 0:
                # the first stage's launches split each tile's depth
                changed_prev = square_split(D, K, span, s0, s1, split, rng, plain_store, reads_first)
                continue
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


@pytest.mark.parametrize("name,W,saturates", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("T,kappa,split", [(8, 4, 4), (8, 4, 2), (16, 2, 4)])
def test_split_squarings_are_exact_below_the_cap(name, W, saturates, T, kappa, split):
    """The first stage's squarings with each tile's depth split over `split`
    workgroups, merged by atomic min, in random read/write interleavings."""
    want = po.fw_serial(W).astype(np.int64)
    for seed, reads_first in ((0, False), (1, False), (2, True)):
        got = lead_block_fw(W, T, kappa, seed=seed, split=split, reads_first=reads_first)
        below = got < CAP
        assert (got[below] == want[below]).all(), f"{name}: a cell below the cap is not exact"
        assert (exact_finish(got) == want).all(), name


def test_planted_split_with_plain_stores_fails():
    """Parts that store their result instead of lowering by atomic min lose
    other parts' lower values.  A chain with decoy shortcuts i -> i+2 of
    weight 3 gives a cell several improving parts of different values; with
    every read ahead of every write (no in-place mix to make up for it) the
    fixed launch count then leaves the stage block unclosed, while the
    atomic min closes it."""
    W = chain(64)
    W[np.arange(62), np.arange(2, 64)] = 3
    want = po.fw_serial(W).astype(np.int64)
    assert (lead_block_fw(W, 8, 8, split=8, reads_first=True) == want).all()
    bad = [not (lead_block_fw(W, 8, 8, seed=s, split=8, plain_store=True, reads_first=True) == want).all()
           for s in range(4)]
    assert all(bad)


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
