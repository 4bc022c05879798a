"""The input pipeline around the solve: the Philox device generator (bit-exact
against the oracle's restatement of the same definition) and pinned host
matrices with asynchronous copies, end to end through a solve."""
import ctypes
import sys

import numpy as np
import pytest

import fw_cuda_amd as fx
from fw_cuda_amd import _lib
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

INF = fx.INF


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,density,seed,wmin,wmax", [
    (1, 0.5, 1, 1, 1000), (2, 1.0, 2, 1, 1000), (3, 0.0, 3, 1, 1000), (37, 0.3, 4, 1, 1000),
    (64, 0.9, 5, 1, 25000), (257, 0.1, 2**40 + 17, 0, 50), (1000, 0.9, 42, 1, 1000),
    (1024, 0.5, 2**63 + 1, -5, 5), (2047, 0.05, 99, 7, 7)])
def test_device_generator_matches_oracle(n, density, seed, wmin, wmax):
    with fx.DeviceMatrix(n) as dm:
        say("generate", n, density, seed)
        fx.generate_graph_device(dm.ptr, n, density, seed, wmin, wmax)
        got = dm.download()
    assert (got == po.generate_philox(n, density, seed, wmin, wmax)).all()


def test_generated_graph_solves_exactly():
    """Generate on the device, solve in place, compare with the oracle solve
    of the oracle-generated matrix (no host generation on the product side)."""
    n = 1500
    with fx.DeviceMatrix(n) as dm:
        fx.generate_graph_device(dm.ptr, n, 0.4, 2024)
        fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
        got = dm.download()
    assert (got == po.fw_rowpar(po.generate_philox(n, 0.4, 2024))).all()


def test_pinned_async_round_trip():
    """Pinned host matrix -> async H2D -> solve -> async D2H -> sync."""
    L = _lib.lib()
    n = 777
    g = fx.generate_graph(n, 0.5, 31, 1, 1000)
    expect = po.fw_rowpar(g)
    h = ctypes.c_void_p()
    _lib.check(L.fwx_host_alloc_pinned(ctypes.byref(h), n))
    try:
        buf = np.ctypeslib.as_array(ctypes.cast(h, ctypes.POINTER(ctypes.c_int32)), shape=(n, n))
        buf[:] = g
        with fx.DeviceMatrix(n) as dm:
            _lib.check(L.fwx_copy_to_device_async(dm.ptr, h, n, None))
            fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
            buf[:] = 0
            _lib.check(L.fwx_copy_to_host_async(h, dm.ptr, n, None))
            _lib.check(L.fwx_stream_synchronize(None))
        assert (buf == expect).all()
    finally:
        _lib.check(L.fwx_host_free_pinned(h))


def positive_cycle_graph(rng, n, p):
    """Negative edges, every cycle positive: w(u,v) + h(u) - h(v), w >= 1."""
    W = fx.generate_graph(n, p, int(rng.integers(1, 2**31)), 1, 50).astype(np.int64)
    h = rng.integers(0, 40, n)
    W2 = np.where(W != fx.INF, W + h[:, None] - h[None, :], fx.INF).astype(np.int32)
    np.fill_diagonal(W2, 0)
    return W2


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 200, 511, 1000])
def test_predecessors_match_oracle(n):
    """fwx_predecessors bit-exact against the oracle's restatement, on
    generated graphs and on graphs with negative edges (ragged sizes)."""
    rng = np.random.default_rng(n)
    for W in (fx.generate_graph(n, float(rng.uniform(0.05, 0.9)), int(rng.integers(1, 2**31)), 1, 1000),
              positive_cycle_graph(rng, n, 0.3)):
        say("predecessors", n)
        D, P = fx.floyd_warshall_with_pred(W)
        assert (D == po.fw_rowpar(W)).all()
        assert (P == po.predecessors(W, D)).all()


def test_predecessor_keys_saturate():
    """The key (sat16(t) << 16) | u must clamp t, not truncate it: a
    candidate u=1 that misses by exactly 2^16 (and 2^17 for another cell)
    would otherwise look tight and beat the true predecessor u=2.  Also
    large weights (U[1, 10^6]), where t often exceeds 2^16, against the
    oracle."""
    W = np.full((5, 5), INF, dtype=np.int32)
    np.fill_diagonal(W, 0)
    W[0, 1], W[0, 2], W[2, 3], W[1, 3] = 5, 3, 7, 65541     # t(u=1 -> 3) = 65536
    W[2, 4], W[1, 4] = 9, 5 + 131072 - 5 + 7                 # t(u=1 -> 4) = 131072 + ...
    D, P = fx.floyd_warshall_with_pred(W)
    assert (D == po.fw_rowpar(W)).all()
    assert (P == po.predecessors(W, D)).all()
    assert P[0, 3] == 2 and P[0, 4] == 2
    rng = np.random.default_rng(17)
    for n in (200, 513):
        W = fx.generate_graph(n, 0.3, int(rng.integers(1, 2**31)), 1, 1_000_000)
        D, P = fx.floyd_warshall_with_pred(W)
        assert (P == po.predecessors(W, D)).all(), n


def test_predecessors_walks_at_4096():
    """n=4096: walking P from j along 64 sampled rows reaches i with the
    weights summing to D[i][j] (every cell of those rows)."""
    n = 4096
    W = fx.generate_graph(n, 0.3, 77, 1, 1000)
    D, P = fx.floyd_warshall_with_pred(W)
    rows = np.random.default_rng(1).choice(n, 64, replace=False)
    for i in rows:
        for j in range(n):
            if j == i:
                assert P[i, j] == -1
                continue
            total, v, steps = 0, j, 0
            while v != i:
                u = int(P[i, v])
                assert 0 <= u < n and steps <= n, (i, j)
                total += int(W[u, v])
                v, steps = u, steps + 1
            assert total == int(D[i, j]), (i, j)
