"""Pins the CPU oracle (oracle/fw_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle fails it:
the reference's own known-answer tests and its own fwBaselineCPU output
(golden vectors and, where built, oracle/_ref), brute-force shortest walks by
min-plus squaring (no FW ordering at all), Dijkstra (another algorithm), and
exactness of the fp32 / gradient restatements on integer-valued inputs.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

INF = po.INF


def random_graph(rng, n, p, lo=1, hi=1000):
    W = np.where(rng.random((n, n)) < p, rng.integers(lo, hi + 1, (n, n)), INF).astype(np.int32)
    np.fill_diagonal(W, 0)
    return W


def reweighted(rng, n, p):
    """Graph with negative edges but no negative cycle: w(u,v) + h(u) - h(v)
    with w >= 0 (Johnson reweighting in reverse); INF cells stay INF."""
    W = random_graph(rng, n, p, 0, 50)
    h = rng.integers(0, 40, n)
    adj = W != INF
    W2 = W.astype(np.int64) + h[:, None] - h[None, :]
    W2 = np.where(adj, W2, INF).astype(np.int32)
    np.fill_diagonal(W2, 0)
    return W2


def test_kats(kats):
    for k in kats:
        n = k["n"]
        D = np.array(k["input"], dtype=np.int32).reshape(n, n)
        out = po.fw_serial(D)
        if k.get("check") == "negative_cycle":
            assert po.has_negative_cycle(out), k["name"]
        else:
            assert (out.ravel() == np.array(k["expected"])).all(), k["name"]


def test_reference_golden_outputs(gen_ref):
    # fwBaselineCPU outputs captured from the reference's own sources
    for c in gen_ref:
        n = c["n"]
        g = np.array(c["graph"], dtype=np.int32).reshape(n, n)
        assert (po.fw_serial(g).ravel() == np.array(c["fw"])).all(), (n, c["seed"])


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_against_reference_build():
    rng = np.random.default_rng(5)
    for n, p in [(1, 0.5), (2, 1.0), (3, 0.5), (17, 0.3), (64, 0.9), (100, 0.1), (257, 0.5)]:
        W = random_graph(rng, n, p)
        assert (po.fw_serial(W) == po.ref_fw_baseline_cpu(W)).all(), n
        W = reweighted(rng, n, p)
        assert (po.fw_serial(W) == po.ref_fw_baseline_cpu(W)).all(), n
    # negative cycle: both orders must agree cell for cell (same k-i-j loop)
    W = np.array([[0, 2, INF], [INF, 0, 1], [INF, -3, 0]], dtype=np.int32)
    assert (po.fw_serial(W) == po.ref_fw_baseline_cpu(W)).all()


def test_brute_force_walks():
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 9, 16, 31):
        for p in (0.1, 0.5, 1.0):
            W = random_graph(rng, n, p)
            assert (po.fw_serial(W) == po.brute_walks(W)).all(), (n, p)
            W = reweighted(rng, n, p)
            assert (po.fw_serial(W) == po.brute_walks(W)).all(), (n, p, "neg")


def test_dijkstra_independent_algorithm():
    rng = np.random.default_rng(11)
    for n, p in [(50, 0.05), (200, 0.3), (333, 0.9)]:
        W = random_graph(rng, n, p)
        rows = np.arange(n, dtype=np.int32)
        assert (po.dijkstra_rows(W, rows) == po.fw_serial(W)).all(), n


def test_rowpar_equals_serial():
    rng = np.random.default_rng(13)
    for n, p in [(1, 1.0), (7, 0.5), (128, 0.9), (1000, 0.1)]:
        W = random_graph(rng, n, p)
        assert (po.fw_rowpar(W, 4) == po.fw_serial(W)).all(), n
        W = reweighted(rng, n, p)
        assert (po.fw_rowpar(W, 4) == po.fw_serial(W)).all(), n


def test_planted_mistake_is_caught():
    """A classic wrong loop order (k innermost) must fail the same checks."""
    rng = np.random.default_rng(17)
    W = random_graph(rng, 24, 0.2)
    bad = W.astype(np.int64).copy()
    n = W.shape[0]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                bad[i, j] = min(bad[i, j], bad[i, k] + bad[k, j])
    assert not (bad == po.brute_walks(W)).all()
    assert (po.fw_serial(W) == po.brute_walks(W)).all()


def test_certificate_accepts_exact_rows_only():
    """oracle_certify_rows accepts exactly the shortest-distance rows: it
    passes FW/Dijkstra rows (ragged n, sparse to dense, INF edges), and fails
    every planted mistake -- one cell too low or too high, a row of another
    source, a source cell that is not 0, the k-innermost wrong loop order,
    an unreachable (INF) pair claimed reachable."""
    rng = np.random.default_rng(19)
    for n, p in [(2, 1.0), (3, 0.0), (33, 0.1), (130, 0.5), (301, 0.9), (2500, 0.05)]:
        W = random_graph(rng, n, p)
        rows = np.unique(rng.integers(0, n, min(n, 70))).astype(np.int32)
        D = po.fw_rowpar(W)
        assert po.certify_rows(W, rows, D[rows]).all(), n
        assert po.certify_rows(W, rows, po.dijkstra_rows(W, rows)).all(), n
        if n < 3:
            continue
        bad = D[rows].copy()
        q, v = 0, (int(rows[0]) + 1) % n
        lo = bad.copy(); lo[q, v] -= 1
        hi = bad.copy(); hi[q, v] += 1
        other = bad.copy(); other[q] = D[(int(rows[0]) + 1) % n]
        src = bad.copy(); src[q, rows[0]] = 1
        for planted in (lo, hi, other, src):
            ok = po.certify_rows(W, rows, planted)
            assert not ok[0] and ok[1:].all(), n
    # the k-innermost loop order leaves some rows unfinished; all of them fail
    W = random_graph(rng, 40, 0.08)
    bad = W.astype(np.int64).copy()
    for i in range(40):
        for j in range(40):
            for k in range(40):
                bad[i, j] = min(bad[i, j], bad[i, k] + bad[k, j])
    bad = bad.astype(np.int32)
    D = po.fw_serial(W)
    rows = np.arange(40, dtype=np.int32)
    wrong = np.where((bad != D).any(axis=1))[0]
    assert wrong.size > 0
    ok = po.certify_rows(W, rows, bad)
    assert (~ok[wrong]).all() and ok[np.setdiff1d(rows, wrong)].all()
    # an unreachable pair (INF) claimed one lower is caught too
    W = random_graph(rng, 50, 0.02)
    D = po.fw_serial(W)
    rows = np.arange(50, dtype=np.int32)
    i, j = np.argwhere(D == INF)[0]
    bad = D.copy(); bad[i, j] = INF - 1
    ok = po.certify_rows(W, rows, bad)
    assert not ok[i] and ok[np.setdiff1d(rows, [i])].all()


def walk_ok(W, D, P, i, j):
    """Walking P from j reaches i; the weights along the walk sum to D[i][j]."""
    n = W.shape[0]
    total, v, steps = 0, j, 0
    while v != i:
        u = int(P[i, v])
        if u < 0 or steps > n:
            return False
        total += int(W[u, v])
        v, steps = u, steps + 1
    return total == int(D[i, j])


def test_predecessors_hand_worked():
    """A 5-vertex graph worked by hand: two equal-length routes 0->1->3 and
    0->2->3 (the tie goes to the smaller vertex, 1), an unreachable vertex 4
    (reached only by the INF edge itself: predecessor 0), a direct edge
    beaten by a detour (0->3 of weight 9 vs 2+2)."""
    W = np.array([[0, 2, 2, 9, INF],
                  [INF, 0, INF, 2, INF],
                  [INF, INF, 0, 2, INF],
                  [INF, INF, INF, 0, INF],
                  [INF, INF, INF, INF, 0]], dtype=np.int32)
    D = po.fw_serial(W)
    P = po.predecessors(W, D)
    assert list(P[0]) == [-1, 0, 0, 1, 0]
    assert P[1, 3] == 1 and P[2, 3] == 2
    assert (np.diag(P) == -1).all()


def test_predecessors_walk_to_shortest_paths():
    """On random graphs with positive weights, and with negative edges but
    positive cycles, every predecessor walk is a shortest path."""
    rng = np.random.default_rng(23)

    def positive_cycles_reweighted(n, p):
        # w(u,v) + h(u) - h(v) with w >= 1: every cycle keeps its positive weight
        W = random_graph(rng, n, p, 1, 50).astype(np.int64)
        h = rng.integers(0, 40, n)
        W2 = np.where(W != INF, W + h[:, None] - h[None, :], INF).astype(np.int32)
        np.fill_diagonal(W2, 0)
        return W2

    for n, p in [(2, 1.0), (17, 0.3), (60, 0.1), (90, 0.6)]:
        for W in (random_graph(rng, n, p), positive_cycles_reweighted(n, p)):
            D = po.fw_serial(W)
            P = po.predecessors(W, D)
            for i in range(n):
                for j in range(n):
                    if i != j:
                        assert walk_ok(W, D, P, i, j), (n, i, j)
    # a planted mistake -- predecessors of the input graph instead of the
    # closure -- breaks the walk
    W = random_graph(rng, 30, 0.2)
    P_bad = po.predecessors(W, W)
    D = po.fw_serial(W)
    assert not all(walk_ok(W, D, P_bad, i, j) for i in range(30) for j in range(30) if i != j)


def test_philox_known_answers():
    """Philox4x32-10 against known-answer vectors of the Random123
    distribution (kat_vectors: zero, all-ones and the digits of pi)."""
    kat = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
           ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
           ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, expect in kat:
        assert list(po.philox4x32_10(ctr, key)) == expect


def test_philox_generator_definition():
    """The generator's cell rule on top of Philox: diagonal 0, INF where the
    first word is above the density threshold, weight from the second word;
    density 0 / 1 give no / every edge; the edge fraction and the weight
    range match; the cell depends on (seed, i*n+j) only."""
    g = po.generate_philox(300, 0.3, 123456789, 5, 9)
    assert (np.diag(g) == 0).all()
    off = g[~np.eye(300, dtype=bool)]
    present = off != INF
    assert abs(present.mean() - 0.3) < 0.01
    assert off[present].min() == 5 and off[present].max() == 9
    for w in range(5, 10):  # uniform over the 5 weights
        assert abs((off[present] == w).mean() - 0.2) < 0.01
    assert (po.generate_philox(50, 0.0, 7)[~np.eye(50, dtype=bool)] == INF).all()
    assert (po.generate_philox(50, 1.0, 7)[~np.eye(50, dtype=bool)] != INF).all()
    # one cell by hand: (i, j) = (3, 17) of n = 40, seed 2^40 + 5
    n, seed, i, j = 40, (1 << 40) + 5, 3, 17
    r = po.philox4x32_10([i * n + j, 0, 0, 0], [5, 1 << 8])
    expect = (1 + ((int(r[1]) * 1000) >> 32)) if int(r[0]) < int(0.5 * 2**32) else INF
    assert po.generate_philox(n, 0.5, seed)[i, j] == expect
    # a different seed gives a different matrix; the same seed the same one
    assert (po.generate_philox(64, 0.5, 1) != po.generate_philox(64, 0.5, 2)).any()
    assert (po.generate_philox(64, 0.5, 1) == po.generate_philox(64, 0.5, 1)).all()


def test_f32_pred_matches_integer_oracle():
    rng = np.random.default_rng(19)
    for n, p in [(1, 0.5), (40, 0.2), (150, 0.6)]:
        Wi = random_graph(rng, n, p)
        Wf = np.where(Wi == INF, np.inf, Wi).astype(np.float32)
        D, P = po.fw_f32_pred(Wf)
        Di = po.fw_serial(Wi)
        # with INF edges the integer oracle may route through them; compare
        # only cells whose integer distance is below INF (truly reachable)
        reach = Di < INF
        assert (D[reach] == Di[reach].astype(np.float32)).all()
        assert np.isinf(D[~reach]).all()
        # walking P reproduces D exactly (integer-valued weights)
        for s in range(n):
            for t in range(n):
                if s == t or not reach[s, t]:
                    continue
                v, tot = t, 0.0
                while v != s:
                    u = P[s, v]
                    tot += Wf[u, v]
                    v = u
                assert tot == D[s, t]


def test_dijkstra_f64_matches_fw_f64():
    """The float64 single-source rows equal float64 Floyd-Warshall bit for
    bit: fp32 weights in [0.5, 10] are multiples of 2^-24, so every path sum
    below 2^29 is exact in float64 and the two orders cannot round
    differently.  Sparse graphs leave vertices unreachable (+inf)."""
    rng = np.random.default_rng(41)
    for n, p in [(1, 0.5), (2, 0.0), (60, 0.02), (200, 0.05), (257, 0.3)]:
        W = np.where(rng.random((n, n)) < p, rng.uniform(0.5, 10.0, (n, n)), np.inf).astype(np.float32)
        np.fill_diagonal(W, 0)
        full = po.fw_f64(W.astype(np.float64))
        rows = np.arange(n, dtype=np.int32)
        got = po.dijkstra_f64_rows(W, rows)
        assert np.array_equal(got, full), n
    # hand-worked: 0->1 (1), 1->2 (2), 0->2 (5), 2->0 (1), vertex 3 unreachable
    W = np.full((4, 4), np.inf, dtype=np.float32)
    np.fill_diagonal(W, 0)
    W[0, 1], W[1, 2], W[0, 2], W[2, 0] = 1, 2, 5, 1
    got = po.dijkstra_f64_rows(W, np.array([0, 2, 3], dtype=np.int32))
    inf = np.inf
    assert np.array_equal(got, [[0, 1, 3, inf], [1, 2, 0, inf], [inf, inf, inf, 0]])
    with pytest.raises(ValueError):
        po.dijkstra_f64_rows(-W, np.array([0], dtype=np.int32))


def test_f32_row_check_catches_planted_mistakes():
    """f32_row_failures accepts the fp32 oracle's own rows and rejects a row
    with a path-consistent but non-shortest distance (a direct edge kept
    where a shorter detour exists, its predecessor pointing at the source),
    a distance off by 1e-4 relative, and a non-tight predecessor."""
    rng = np.random.default_rng(43)
    n = 300
    W = np.where(rng.random((n, n)) < 0.3, rng.uniform(0.5, 10.0, (n, n)), np.inf).astype(np.float32)
    np.fill_diagonal(W, 0)
    D, P = po.fw_f32_pred(W)
    rows = rng.choice(n, 16, replace=False).astype(np.int32)
    ref = po.dijkstra_f64_rows(W, rows)
    assert po.f32_row_failures(D[rows], P[rows], W, rows, ref) == []
    s = int(rows[0])
    # a present direct edge much longer than the shortest distance
    cand = np.flatnonzero(np.isfinite(W[s]) & (W[s] > 1.5 * D[s]) & (np.arange(n) != s))
    t = int(cand[0])
    D2, P2 = D[rows].copy(), P[rows].copy()
    D2[0, t], P2[0, t] = W[s, t], s
    kinds = {k for (_, tt, k) in po.f32_row_failures(D2, P2, W, rows, ref) if tt == t}
    assert kinds == {"distance", "predecessor not tight"}
    D3 = D[rows].copy()
    D3[0, t] *= 1.0001
    assert [k for (_, tt, k) in po.f32_row_failures(D3, P[rows], W, rows, ref)] == ["distance"]
    # a predecessor u that is not even joined to t by an edge
    P4 = P[rows].copy()
    u = int(np.flatnonzero(~np.isfinite(W[:, t]))[0])
    P4[0, t] = u
    assert [k for (_, tt, k) in po.f32_row_failures(D[rows], P4, W, rows, ref)] == ["predecessor not tight"]


def _fd_grad(W, G, edges, eps):
    g = []
    W = W.astype(np.float64)
    for (u, v) in edges:
        Wp = W.copy()
        Wm = W.copy()
        Wp[u, v] += eps
        Wm[u, v] -= eps
        Dp = po.fw_f64(Wp)
        Dm = po.fw_f64(Wm)
        fin = np.isfinite(Dp)
        g.append(((G[fin] * Dp[fin]).sum() - (G[fin] * Dm[fin]).sum()) / (2 * eps))
    return np.array(g)


def test_grad_paths_matches_finite_differences():
    rng = np.random.default_rng(23)
    n = 24
    W = np.where(rng.random((n, n)) < 0.3, rng.uniform(0.5, 10.0, (n, n)), np.inf).astype(np.float32)
    np.fill_diagonal(W, 0)
    G = rng.standard_normal((n, n)).astype(np.float32)
    D, P = po.fw_f32_pred(W)
    dW = po.grad_paths(P, G)
    edges = [(u, v) for u in range(n) for v in range(n) if u != v and np.isfinite(W[u, v])]
    fd = _fd_grad(W, G.astype(np.float64), edges, 1e-3)
    ana = np.array([dW[u, v] for (u, v) in edges], dtype=np.float64)
    assert np.allclose(ana, fd, rtol=1e-5, atol=1e-4)
    # a planted mistake (dropping the last edge of every path) is caught
    bad = dW.copy()
    for s in range(n):
        for t in range(n):
            if s != t and P[s, t] >= 0:
                bad[P[s, t], t] -= G[s, t]
    assert not np.allclose([bad[u, v] for (u, v) in edges], fd, rtol=1e-5, atol=1e-4)
