"""CPU model of the pk16 rest passes (fw_cuda_amd/csrc/fwx_tile16.hpp) inside
the multi-layer blocked schedule, checked against the oracle's fwBaselineCPU
restatement (src/fw_baseline_cpu.cpp:3-16).

The model runs the schedule of fwx_blocked.hip at small sizes: per stage of
kappa k-tiles, exact lead rounds (closure, panels, cross) and one rest product
over the other tiles of depth kappa*T.  The rest product takes the pk16
arithmetic: operands and C are clamped at CAP (A and C carry +BIAS), the
minimum runs on the encoded values, and a cell whose result is CAP keeps its
old value.  After the run every cell below CAP must hold the exact distance;
cells at or above CAP hold walk lengths, and one exact FW over the matrix
(the library's exact finish) must give the exact distances.  The 16-bit lane
bounds (no carry between lanes, every compared pattern in [BIAS, 0x7BFF])
are asserted on every product.  Planted mistakes -- storing a saturated
result, not clamping operands, a cap above the lane bound -- fail.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

INF = po.INF
BIAS = 0x0400
CAP = 0x3BFF
F16_FINITE_MAX = 0x7BFF


def product_exact(C, A, B):
    """C <- min(C, A (x) B) in int64 (no overflow for these inputs)."""
    for k in range(A.shape[1]):
        np.minimum(C, A[:, k:k + 1] + B[k:k + 1, :], out=C)


def product_pk16(C, A, B, cap=CAP, clamp=True, keep_saturated=True):
    """The pk16 rest product on the int64 block C (in place)."""
    if clamp:
        ea = np.minimum(A, cap) + BIAS
        eb = np.minimum(B, cap)
    else:
        ea = (A + BIAS) & 0xFFFF
        eb = B & 0xFFFF
    ec = np.minimum(C, cap) + BIAS
    r = ec.copy()
    for k in range(A.shape[1]):
        s = ea[:, k:k + 1] + eb[k:k + 1, :]
        if clamp:
            assert s.max() <= F16_FINITE_MAX and s.min() >= BIAS, "lane sum outside the finite normal range"
        np.minimum(r, s & 0xFFFF, out=r)
    dec = r - BIAS
    if keep_saturated:
        store = dec < cap
        C[store] = dec[store]
    else:
        C[...] = dec


def close(D, lo, hi):
    """FW restricted to intermediates and cells in [lo, hi) (the diagonal tile)."""
    for k in range(lo, hi):
        np.minimum(D[lo:hi, lo:hi], D[lo:hi, k:k + 1] + D[k:k + 1, lo:hi], out=D[lo:hi, lo:hi])


def blocked_fw(W, T, kappa, rest=product_pk16, **kw):
    """fwx_blocked.hip's plain multi-layer schedule, rest passes via `rest`."""
    n = W.shape[0]
    D = W.astype(np.int64)
    N = (n + T - 1) // T
    span = lambda t: (t * T, min(n, (t + 1) * T))  # noqa: E731
    for kb in range(0, N, kappa):
        B = min(kappa, N - kb)
        K = list(range(kb, kb + B))
        for c in K:  # lead rounds: exact
            c0, c1 = span(c)
            close(D, c0, c1)
            for j in range(N):
                if j != c:
                    j0, j1 = span(j)
                    product_exact(D[c0:c1, j0:j1], D[c0:c1, c0:c1].copy(), D[c0:c1, j0:j1].copy())
                    product_exact(D[j0:j1, c0:c1], D[j0:j1, c0:c1].copy(), D[c0:c1, c0:c1].copy())
            for i in range(N):
                for j in range(N):
                    if i == c or j == c or not (i in K or j in K):
                        continue
                    i0, i1 = span(i)
                    j0, j1 = span(j)
                    product_exact(D[i0:i1, j0:j1], D[i0:i1, c0:c1].copy(), D[c0:c1, j0:j1].copy())
        s0, s1 = kb * T, min(n, (kb + B) * T)
        for i in range(N):
            for j in range(N):
                if i in K or j in K:
                    continue
                i0, i1 = span(i)
                j0, j1 = span(j)
                rest(D[i0:i1, j0:j1], D[i0:i1, s0:s1].copy(), D[s0:s1, j0:j1].copy(), **kw)
    return D


def exact_finish(D):
    """The library's exact pk64 finish: FW over the matrix as it stands."""
    D = D.copy()
    for k in range(D.shape[0]):
        np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :], out=D)
    return D


def graph(n, p, lo, hi, seed):
    rng = np.random.default_rng(seed)
    W = np.where(rng.random((n, n)) < p, rng.integers(lo, hi + 1, (n, n)), INF).astype(np.int32)
    np.fill_diagonal(W, 0)
    return W


CASES = [
    ("dense, small weights", graph(96, 0.9, 1, 1000, 1), False),
    ("sparse: unreachable pairs", graph(90, 0.02, 1, 1000, 2), True),
    ("weights around the cap", graph(80, 0.5, CAP - 3, CAP + 3, 3), True),
    ("distances straddling the cap", graph(100, 0.08, 1000, 9000, 4), None),
]


@pytest.mark.parametrize("name,W,saturates", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("T,kappa", [(8, 2), (16, 1), (16, 3)])
def test_pk16_rest_passes_are_exact_below_the_cap(name, W, saturates, T, kappa):
    want = po.fw_serial(W).astype(np.int64)
    got = blocked_fw(W, T, kappa)
    below = got < CAP
    assert (got[below] == want[below]).all(), f"{name}: a cell below the cap is not exact"
    assert (want[~below] >= CAP).all() and (got[~below] >= want[~below]).all(), \
        f"{name}: a saturated cell is not an upper bound at or above the cap"
    if saturates is not None:
        assert bool((~below).any()) == saturates, name
    assert (exact_finish(got) == want).all(), f"{name}: the exact finish does not restore the distances"


def test_pk16_agrees_with_int32_products_when_nothing_saturates():
    W = graph(96, 0.9, 1, 1000, 5)
    assert (blocked_fw(W, 16, 2) == blocked_fw(W, 16, 2, rest=product_exact)).all()


def test_planted_storing_saturated_results_fails():
    """Storing CAP for a saturated cell loses its walk length: the exact
    finish then keeps values that are too small."""
    W = graph(90, 0.02, 1, 1000, 2)
    want = po.fw_serial(W).astype(np.int64)
    got = blocked_fw(W, 16, 2, keep_saturated=False)
    assert not (exact_finish(got) == want).all()


def test_planted_unclamped_operands_fail():
    """Without the clamp, operands of 2^16 and more wrap in their 16-bit
    lanes and make small, wrong sums."""
    W = graph(80, 0.5, 64000, 66000, 7)
    want = po.fw_serial(W).astype(np.int64)
    got = blocked_fw(W, 16, 2, clamp=False)
    below = got < CAP
    assert not (got[below] == want[below]).all()


def test_planted_cap_above_the_lane_bound_fails():
    """A cap whose sums leave the finite normal binary16 range is caught by
    the lane-bound assertion (the f16 minimum would see infinities or NaNs)."""
    W = graph(64, 0.5, 20000, 30000, 6)
    with pytest.raises(AssertionError, match="lane sum"):
        blocked_fw(W, 16, 2, cap=0x4000)
