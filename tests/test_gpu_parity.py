"""Bit-exact parity of the HIP kernels (through the C ABI of libfwx.so) with the
CPU oracle, on the reference's own test cases (src/tests_gpu.cu:156-185) and
on the gaps it leaves: ragged n, every tile/kappa, negative weights, the
workload sizes of the benchmark configurations.
"""
import sys

import numpy as np
import pytest

import fw_cuda_amd as fx
from oracle import pyoracle as po

INF = fx.INF
pytestmark = pytest.mark.gpu

# (algorithm, tile, kappa) combinations exercised on every case
CONFIGS = [
    ("baseline_gpu", 0, 0),
    ("tiling_gpu_refshape", 0, 0),
    ("tiling_gpu", 32, 0),
    ("tiling_gpu", 64, 0),
    ("tiling_gpu", 128, 0),
    ("multi_layer_tiling_gpu", 0, 0),
    ("multi_layer_tiling_gpu", 32, 4),
    ("multi_layer_tiling_gpu", 64, 1),
    ("multi_layer_tiling_gpu", 64, 2),
    ("multi_layer_tiling_gpu", 64, 3),
    ("multi_layer_tiling_gpu", 64, 4),
    ("multi_layer_tiling_gpu", 64, 8),
    ("multi_layer_tiling_gpu", 128, 2),
    ("multi_layer_tiling_gpu", 32, 16),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu(D, algo, tile, kappa):
    return fx.floyd_warshall(D, algo, tile, kappa)


def reweighted(rng, n, p):
    W = fx.generate_graph(n, p, int(rng.integers(1, 2**31)), 0, 50).astype(np.int64)
    h = rng.integers(0, 40, n)
    adj = W != INF
    W2 = np.where(adj, W + h[:, None] - h[None, :], INF).astype(np.int32)
    np.fill_diagonal(W2, 0)
    return W2


@pytest.mark.parametrize("algo,tile,kappa", CONFIGS)
def test_kats(kats, algo, tile, kappa):
    for k in kats:
        n = k["n"]
        D = np.array(k["input"], dtype=np.int32).reshape(n, n)
        say("KAT", k["name"], algo, tile, kappa)
        out = gpu(D, algo, tile, kappa)
        if k.get("check") == "negative_cycle":
            assert fx.has_negative_cycle(out), k["name"]
        else:
            assert (out.ravel() == np.array(k["expected"])).all(), k["name"]


# src/tests_gpu.cu:172-180
REF_CASES = [(128, 0.9, 42), (256, 0.9, 123), (512, 0.9, 456), (1024, 0.9, 789),
             (128, 0.1, 101), (256, 0.1, 202), (512, 0.1, 303), (1024, 0.1, 404)]


@pytest.fixture(scope="module")
def ref_cases():
    out = []
    for n, p, seed in REF_CASES:
        g = fx.generate_graph(n, p, seed)
        out.append((n, p, seed, g, po.fw_rowpar(g)))
    return out


@pytest.mark.parametrize("algo,tile,kappa", CONFIGS)
def test_reference_seeded_graphs(ref_cases, algo, tile, kappa):
    for n, p, seed, g, expect in ref_cases:
        if algo == "baseline_gpu" and n > 512:
            continue  # n launches; covered at 1024 by test_n1024_every_algorithm
        say("seeded", n, p, seed, algo, tile, kappa)
        assert (gpu(g, algo, tile, kappa) == expect).all(), (n, p, seed)


RAGGED = [1, 2, 3, 5, 31, 33, 63, 65, 97, 127, 129, 130, 191, 257, 300, 383, 513, 777]


@pytest.mark.parametrize("algo,tile,kappa", CONFIGS)
def test_ragged_sizes(algo, tile, kappa):
    rng = np.random.default_rng(1234)
    for n in RAGGED:
        p = float(rng.uniform(0.05, 0.9))
        g = fx.generate_graph(n, p, int(rng.integers(1, 2**31)), 1, 1000)
        say("ragged", n, algo, tile, kappa)
        assert (gpu(g, algo, tile, kappa) == po.fw_rowpar(g)).all(), n


@pytest.mark.parametrize("algo,tile,kappa", CONFIGS)
def test_negative_weights(algo, tile, kappa):
    """Negative edges without negative cycles: exercises the min3 (signed)
    arithmetic path; the result must still be bit-exact."""
    rng = np.random.default_rng(99)
    for n in (7, 64, 100, 256, 401):
        g = reweighted(rng, n, 0.3)
        assert g.min() < 0
        say("negative", n, algo, tile, kappa)
        assert (gpu(g, algo, tile, kappa) == po.fw_rowpar(g)).all(), n


def test_negative_cycle_check_on_device(kats):
    """fwx_has_negative_cycle_device agrees with the host check (reference
    hasNegativeCycle) on the KAT with a negative cycle after a GPU solve,
    on solved graphs without one, and on planted negative diagonals at
    the first, last and a middle position of ragged sizes."""
    k = [k for k in kats if k.get("check") == "negative_cycle"][0]
    n = k["n"]
    D = np.array(k["input"], dtype=np.int32).reshape(n, n)
    with fx.DeviceMatrix(n) as dm:
        dm.upload(D)
        fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
        assert fx.has_negative_cycle_device(dm.ptr, n)
        assert fx.has_negative_cycle(dm.download())
    rng = np.random.default_rng(41)
    for n in (1, 2, 255, 256, 257, 1000):
        g = fx.generate_graph(n, 0.5, int(rng.integers(1, 2**31)), 1, 1000)
        with fx.DeviceMatrix(n) as dm:
            dm.upload(g)
            fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
            assert not fx.has_negative_cycle_device(dm.ptr, n)
            for pos in {0, n // 2, n - 1}:
                h = g.copy()
                h[pos, pos] = -1
                dm.upload(h)
                say("device neg-cycle", n, pos)
                assert fx.has_negative_cycle_device(dm.ptr, n), (n, pos)


def test_large_values_take_exact_path():
    """Operands >= 2^30 (above INF) must take the exact min3 form.  They are
    placed in one column only, so no relaxation adds two of them and no sum
    leaves int32: the shortest-path matrix stays unique."""
    rng = np.random.default_rng(5)
    n = 300
    g = fx.generate_graph(n, 0.5, 77, 1, 1000)
    col = 123
    big = (rng.random(n) < 0.5) & (g[:, col] == INF)
    g[big, col] = 1_100_000_000
    np.fill_diagonal(g, 0)
    assert (g >= 2**30).any()
    expect = po.fw_rowpar(g)
    for algo, tile, kappa in CONFIGS:
        assert (gpu(g, algo, tile, kappa) == expect).all(), (algo, tile, kappa)


@pytest.mark.parametrize("lookahead,prio", [("0", "1"), ("1", "1"), ("1", "0"), ("2", "1"), ("2", "0"), ("3", "1"),
                                            ("3", "0")])
def test_schedules_agree(monkeypatch, lookahead, prio):
    """The single-stream, look-ahead, side-band and side-band-ahead
    (two-stream) schedules, with and without the lead kernels' raised wave
    priority, graph-replayed and launched directly."""
    monkeypatch.setenv("FWX_LOOKAHEAD", lookahead)
    monkeypatch.setenv("FWX_LEAD_PRIO", prio)
    rng = np.random.default_rng(8)
    for n, t, k in [(1000, 64, 2), (777, 64, 3), (1537, 128, 2), (900, 32, 4), (2048, 64, 4), (1100, 64, 1)]:
        g = fx.generate_graph(n, float(rng.uniform(0.1, 0.9)), int(rng.integers(1, 2**31)), 1, 1000)
        expect = po.fw_rowpar(g)
        for graph in ("1", "0"):
            monkeypatch.setenv("FWX_GRAPH", graph)
            say("schedule", lookahead, prio, graph, n, t, k)
            assert (gpu(g, "multi_layer_tiling_gpu", t, k) == expect).all(), (n, t, k, graph)


@pytest.mark.parametrize("var,val", [("FWX_BAND", "1"), ("FWX_DPANEL", "1"), ("FWX_EVENT_SYSFENCE", "1")])
def test_switches_agree(monkeypatch, var, val):
    """The remaining schedule switches (128-wide band pass, fused closure +
    panels, system-scope event fences), graph-replayed and direct."""
    monkeypatch.setenv(var, val)
    rng = np.random.default_rng(44)
    for n, t, k in [(1537, 128, 2), (777, 64, 3), (900, 32, 4), (2048, 64, 4), (1100, 128, 1)]:
        g = fx.generate_graph(n, float(rng.uniform(0.1, 0.9)), int(rng.integers(1, 2**31)), 1, 1000)
        expect = po.fw_rowpar(g)
        for graph in ("1", "0"):
            monkeypatch.setenv("FWX_GRAPH", graph)
            say("switch", var, val, graph, n, t, k)
            assert (gpu(g, "multi_layer_tiling_gpu", t, k) == expect).all(), (n, t, k, graph)


@pytest.mark.parametrize("variant", ["0", "1", "2"])
def test_rest_kernel_variants(monkeypatch, variant):
    """Every shape of the 128-wide rest kernel (FWX_REST_VARIANT), ragged and
    whole-tile sizes, both relaxation forms."""
    monkeypatch.setenv("FWX_REST_VARIANT", variant)
    rng = np.random.default_rng(31)
    for n in (128, 256, 384, 640, 1024, 300, 1000):
        g = fx.generate_graph(n, float(rng.uniform(0.1, 0.9)), int(rng.integers(1, 2**31)), 1, 1000)
        gneg = reweighted(rng, n, 0.4)
        expect, expect_neg = po.fw_rowpar(g), po.fw_rowpar(gneg)
        for t, k in ((128, 1), (128, 2), (64, 2), (64, 4), (32, 4), (128, 3)):
            say("rest variant", variant, n, t, k)
            assert (gpu(g, "multi_layer_tiling_gpu", t, k) == expect).all(), (n, t, k)
            assert (gpu(gneg, "multi_layer_tiling_gpu", t, k) == expect_neg).all(), (n, t, k)


@pytest.mark.parametrize("nth", ["256", "512", "1024"])
def test_diag_block_sizes(monkeypatch, nth):
    """The diagonal closure with 256, 512 and 1024 threads per block (tile
    64 and 128), on ragged sizes and negative weights."""
    monkeypatch.setenv("FWX_DIAG_NTH", nth)
    rng = np.random.default_rng(21)
    for n in (1, 37, 64, 127, 128, 200, 385):
        g = fx.generate_graph(n, float(rng.uniform(0.1, 0.9)), int(rng.integers(1, 2**31)), 1, 1000)
        gneg = reweighted(rng, n, 0.4)
        for t, k in ((64, 1), (64, 3), (128, 1), (128, 2)):
            say("diag threads", nth, n, t, k)
            assert (gpu(g, "multi_layer_tiling_gpu", t, k) == po.fw_rowpar(g)).all(), (n, t, k)
            assert (gpu(gneg, "multi_layer_tiling_gpu", t, k) == po.fw_rowpar(gneg)).all(), (n, t, k)


def test_n1024_every_algorithm():
    """n=1024, density 0.9, weights U[1,1000], seed 42: every algorithm."""
    g = fx.generate_graph(1024, 0.9, 42, 1, 1000)
    expect = po.fw_serial(g)
    for algo, tile, kappa in CONFIGS:
        say("n=1024", algo, tile, kappa)
        assert (gpu(g, algo, tile, kappa) == expect).all(), (algo, tile, kappa)


def test_idempotent_on_solved_matrix():
    """The reference times its second call on the already-solved matrix
    (src/timer.cu:43-55); running again must leave it unchanged."""
    g = fx.generate_graph(700, 0.4, 3, 1, 1000)
    once = gpu(g, "multi_layer_tiling_gpu", 0, 0)
    twice = gpu(once, "multi_layer_tiling_gpu", 0, 0)
    assert (once == twice).all()


@pytest.mark.slow
def test_n8192_all_kappas():
    """n=8192, density 0.9, U[1,1000], seed 42: multi_layer_tiling_gpu with
    the autotuned tile and kappa in {1,2,4,6,8}, tile 64 with the tuned
    kappa's neighbour, and tiling_gpu, bit-exact against the row-parallel
    oracle (the restatement of fwBaselineCPU); the solves whose rest passes
    are 128 wide run them on packed stage panels."""
    g = fx.generate_graph(8192, 0.9, 42, 1, 1000)
    say("n=8192: oracle start")
    expect = po.fw_rowpar(g)
    say("n=8192: oracle done")
    tile, _ = fx.tuned_config(8192)
    runs = [("multi_layer_tiling_gpu", tile, k) for k in (1, 2, 4, 6, 8)]
    runs += [("multi_layer_tiling_gpu", 64, 3), ("tiling_gpu", 0, 0)]
    with fx.DeviceMatrix(8192) as dm:
        for algo, t, k in runs:
            dm.upload(g)
            say("n=8192", algo, t, k)
            fx.run_device(algo, dm.ptr, 8192, t, k)
            assert (dm.download() == expect).all(), (algo, t, k)
            # 128-wide rest passes (stage width a multiple of 128) run on
            # packed stage panels (tiling_gpu's tuned tile is 128 here)
            span = 128 if algo == "tiling_gpu" else t * k
            assert fx.last_solve_packed() == (span % 128 == 0), (algo, t, k)


@pytest.mark.slow
def test_n32768_spot_rows():
    """n=32768 (4 GiB int32), density 0.1, weights U[1,1000], seed 7:
    multi_layer_tiling_gpu (tuned) certified exact on every row by the device
    certificate (fwx_certify_rows), on 4096 random rows by the oracle's
    certificate, 64 of them also equal to host Dijkstra over the INF-completed
    graph; tiling_gpu bit-identical to it on every cell."""
    n = 32768
    g = fx.generate_graph(n, 0.1, 7, 1, 1000)
    rows = np.sort(np.random.default_rng(7).choice(n, 4096, replace=False)).astype(np.int32)
    say("n=32768: dijkstra on 64 rows")
    dij = po.dijkstra_rows(g, rows[::64])
    import ctypes
    import torch
    with fx.DeviceMatrix(n) as dm, fx.DeviceMatrix(n) as dw:
        dm.upload(g)
        dw.upload(g)
        say("n=32768 multi_layer_tiling_gpu")
        fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
        ml = dm.download()
        say("n=32768: device certificate of all", n, "rows")
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        ok = torch.zeros(n, dtype=torch.int32, device="cuda")
        failed = fx.certify_rows_device(dw.ptr, n, ctypes.c_void_p(src.data_ptr()), n, dm.ptr,
                                        ctypes.c_void_p(ok.data_ptr()))
        all_rows_ok = failed == 0 and bool((ok == 1).all())
        dm.upload(g)
        say("n=32768 tiling_gpu")
        fx.run_device("tiling_gpu", dm.ptr, n, 0, 0)
        same = bool((dm.download() == ml).all())
    assert all_rows_ok, failed
    assert same, "tiling_gpu and multi_layer_tiling_gpu differ"
    assert (ml[rows[::64]] == dij).all()
    say("n=32768: certificate on", rows.size, "rows")
    ok = po.certify_rows(g, rows, ml[rows])
    assert ok.all(), rows[~ok][:10]


def test_concurrent_solves_on_two_streams():
    """Two host threads solving at once on their own streams: one captures and
    replays graphs (n = 1000, a new buffer each time), the other launches
    directly (n = 4500) through the shared side stream.  Every result equals
    the oracle."""
    import threading
    import torch
    from fw_cuda_amd import _lib
    ga = fx.generate_graph(1000, 0.4, 91, 1, 1000)
    gb = fx.generate_graph(4500, 0.2, 92, 1, 1000)
    ea, eb = po.fw_rowpar(ga), po.fw_rowpar(gb)
    errors, results = [], {}

    def worker(name, g, reps):
        try:
            s = torch.cuda.Stream()
            out = []
            for _ in range(reps):
                t = torch.from_numpy(g).cuda()
                torch.cuda.synchronize()
                fx.run_device("multi_layer_tiling_gpu", _lib.ptr(t), g.shape[0], 0, 0, s.cuda_stream)
                out.append(t.cpu().numpy())
            results[name] = out
        except Exception as e:  # reported below
            errors.append((name, repr(e)))

    th = [threading.Thread(target=worker, args=("graphs", ga, 6)),
          threading.Thread(target=worker, args=("direct", gb, 3))]
    say("concurrent solves")
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    assert all((r == ea).all() for r in results["graphs"])
    assert all((r == eb).all() for r in results["direct"])


def test_max_n_32769():
    """n = MAX_N = 32769 (include/constants.h:5; 4.3 GB int32, ragged: no
    16-byte path, edge tiles everywhere): multi_layer_tiling_gpu certified
    exact on every row by the device certificate, 32 rows equal to host
    Dijkstra over the INF-completed graph."""
    import ctypes
    import torch
    n = fx.MAX_N
    assert n == 32769
    g = fx.generate_graph(n, 0.05, 11, 1, 1000)
    rows = np.linspace(0, n - 1, 32).astype(np.int32)
    say("max n: dijkstra on", rows.size, "rows")
    dij = po.dijkstra_rows(g, rows)
    with fx.DeviceMatrix(n) as dm, fx.DeviceMatrix(n) as dw:
        dm.upload(g)
        dw.upload(g)
        del g
        say("max n: multi_layer_tiling_gpu")
        fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 0, 0)
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        ok = torch.zeros(n, dtype=torch.int32, device="cuda")
        say("max n: device certificate of every row")
        failed = fx.certify_rows_device(dw.ptr, n, ctypes.c_void_p(src.data_ptr()), n, dm.ptr,
                                        ctypes.c_void_p(ok.data_ptr()))
        got = dm.download()[rows]
    assert failed == 0
    assert (got == dij).all()


@pytest.mark.parametrize("lookahead", ["0", "1", "2", "3"])
def test_graph_replay_cache(monkeypatch, lookahead):
    """Repeated solves on one buffer replay a captured hipGraph; every replay
    must equal the oracle, whichever relaxation form (pk64 / min3) the input
    selects, interleaved with direct (FWX_GRAPH=0) and profiled solves --
    profiled solves launch directly and their totals add up to n^3 each."""
    import ctypes
    from fw_cuda_amd import _lib
    monkeypatch.setenv("FWX_LOOKAHEAD", lookahead)
    n = 1000
    g = fx.generate_graph(n, 0.5, 11, 1, 1000)
    expect = po.fw_rowpar(g)
    gneg = reweighted(np.random.default_rng(4), n, 0.5)
    expect_neg = po.fw_rowpar(gneg)
    with fx.DeviceMatrix(n) as dm:
        def solve(src, exp, tag):
            dm.upload(src)
            fx.run_device("multi_layer_tiling_gpu", dm.ptr, n, 64, 2)
            assert (dm.download() == exp).all(), tag

        for graph in ("0", "1", "1", "1", "0", "1"):
            monkeypatch.setenv("FWX_GRAPH", graph)
            say("graph replay", graph)
            solve(g, expect, "pk64 " + graph)
            solve(gneg, expect_neg, "min3 " + graph)
        L = _lib.lib()
        _lib.check(L.fwx_profile_begin())
        for _ in range(3):
            solve(g, expect, "profiled")
        ms = (ctypes.c_double * _lib.PROF_NCLASS)()
        wk = (ctypes.c_double * _lib.PROF_NCLASS)()
        grp = (ctypes.c_int64 * _lib.PROF_NCLASS)()
        _lib.check(L.fwx_profile_end(ms, wk, grp))
        # three solves' worth of rest + band + lead work = 3 n^3 relaxations
        total = wk[_lib.PROF_REST] + wk[_lib.PROF_LEAD] + wk[_lib.PROF_BAND]
        assert abs(total - 3 * float(n) ** 3) < 1e-6 * 3e9, total
        assert ms[_lib.PROF_REST] > 0 and grp[_lib.PROF_REST] > 0
        solve(g, expect, "replay after profiling")


@pytest.mark.parametrize("region,threads", [("64", "512"), ("128", "512"), ("64", "256"), ("128", "256")])
def test_small_rounds_every_form_and_mode(monkeypatch, region, threads):
    """Tile 32 with kappa 1 runs the fused small-n rounds (one launch per
    k-tile round, each block picking the relaxation form from its operands),
    with 64- and 128-wide regions on 256 or 512 threads per block, and the
    next round's diagonal tile closed ahead by one more block per launch.  On one buffer per size, inputs
    that select pk64 and min3 (negative weights; operands >= 2^30)
    alternate, replayed from a graph and launched directly, plus the
    three-launch rounds (FWX_SMALL=0); every result equals the oracle.
    Sizes: one tile, ragged edges, up to 65 k-tiles."""
    monkeypatch.setenv("FWX_SMALL_R", region)
    monkeypatch.setenv("FWX_SMALL_NT", threads)
    rng = np.random.default_rng(321)
    for n in (1, 2, 31, 32, 33, 64, 100, 255, 513, 1024, 1500, 2049):
        pos = fx.generate_graph(n, float(rng.uniform(0.05, 0.9)), int(rng.integers(1, 2**31)), 1, 1000)
        neg = reweighted(rng, n, 0.4) if n > 2 else pos
        big = pos.copy()
        if n > 2:
            col = n // 2
            sel = (rng.random(n) < 0.5) & (big[:, col] == INF)
            big[sel, col] = 1_100_000_000
            np.fill_diagonal(big, 0)
        cases = [("pk64", pos, po.fw_rowpar(pos)), ("min3", neg, po.fw_rowpar(neg)),
                 ("big", big, po.fw_rowpar(big))]
        with fx.DeviceMatrix(n) as dm:
            for small, graph in (("1", "1"), ("1", "1"), ("1", "0"), ("0", "1"), ("0", "0"), ("1", "1")):
                monkeypatch.setenv("FWX_SMALL", small)
                monkeypatch.setenv("FWX_GRAPH", graph)
                monkeypatch.setenv("FWX_SMALL_GRAPH", graph)
                for tag, src, exp in cases:
                    say("small rounds", region, threads, n, small, graph, tag)
                    dm.upload(src)
                    fx.run_device("tiling_gpu", dm.ptr, n, 32, 1)
                    assert (dm.download() == exp).all(), (region, threads, n, small, graph, tag)


def test_small_rounds_profiled_and_graph_eviction(monkeypatch):
    """Profiled small solves launch directly and account n^3 relaxations each;
    with small-n graph replay on (FWX_SMALL_GRAPH=1), more distinct buffers
    than the graph cache holds evict small-n graphs, and solves after the
    evictions stay exact."""
    import ctypes
    from fw_cuda_amd import _lib
    monkeypatch.setenv("FWX_SMALL_GRAPH", "1")
    n = 300
    g = fx.generate_graph(n, 0.5, 17, 1, 1000)
    gneg = reweighted(np.random.default_rng(6), n, 0.5)
    expect, expect_neg = po.fw_rowpar(g), po.fw_rowpar(gneg)
    L = _lib.lib()
    with fx.DeviceMatrix(n) as dm:
        _lib.check(L.fwx_profile_begin())
        for src, exp in ((g, expect), (gneg, expect_neg)):
            dm.upload(src)
            fx.run_device("tiling_gpu", dm.ptr, n, 32, 1)
            assert (dm.download() == exp).all()
        ms = (ctypes.c_double * _lib.PROF_NCLASS)()
        wk = (ctypes.c_double * _lib.PROF_NCLASS)()
        grp = (ctypes.c_int64 * _lib.PROF_NCLASS)()
        _lib.check(L.fwx_profile_end(ms, wk, grp))
        assert abs(wk[_lib.PROF_LEAD] - 2 * float(n) ** 3) < 1e-3, wk[_lib.PROF_LEAD]
        assert ms[_lib.PROF_LEAD] > 0 and grp[_lib.PROF_LEAD] == 2
    mats = [fx.DeviceMatrix(n) for _ in range(12)]
    try:
        for rep in range(3):
            for k, dm in enumerate(mats):
                src, exp = ((g, expect), (gneg, expect_neg))[(k + rep) % 2]
                say("eviction", rep, k)
                dm.upload(src)
                fx.run_device("tiling_gpu", dm.ptr, n, 32, 1)
                assert (dm.download() == exp).all(), (rep, k)
    finally:
        for dm in mats:
            dm.free()


def chain_graph(n, back=4000):
    """Every edge i -> j present: weight 1 when j = i + 1, else `back` (all
    below 4096, so every small-n block runs pk16).  The shortest path from i
    to j > i is the chain through every vertex between them, so closing a
    diagonal tile needs paths of 31 edges: five squarings."""
    W = np.full((n, n), back, dtype=np.int32)
    W[np.arange(n - 1), np.arange(1, n)] = 1
    np.fill_diagonal(W, 0)
    return W


@pytest.mark.parametrize("small16,squaring", [("1", "1"), ("1", "0"), ("0", "1")])
def test_small_rounds_pk16_blocks(monkeypatch, small16, squaring):
    """The small-n rounds' pk16 blocks (every loaded operand below 4096, so no
    sum reaches the 16-bit cap): dense small weights (almost every block
    pk16 after round 0), weights straddling the threshold (blocks of both
    forms in one round), distances growing past it during the solve, and a
    chain whose diagonal tiles need the full five squarings of the
    look-ahead closure, at ragged and aligned sizes, graph-replayed and
    direct; FWX_SMALL16=0 keeps pk64/min3 and FWX_SMALL_SQ=0 the 32-step
    closure, with the same results."""
    monkeypatch.setenv("FWX_SMALL16", small16)
    monkeypatch.setenv("FWX_SMALL_SQ", squaring)
    rng = np.random.default_rng(99)
    for n in (33, 64, 200, 512, 1024, 1537):
        for tag, lo, hi, p in (("small", 1, 1000, 0.9), ("straddle", 1000, 5000, 0.5),
                               ("growing", 500, 2000, 0.02), ("chain", 0, 0, 0)):
            if tag == "chain":
                W = chain_graph(n)
            else:
                W = fx.generate_graph(n, p, int(rng.integers(1, 2**31)), lo, hi)
            exp = po.fw_rowpar(W)
            with fx.DeviceMatrix(n) as dm:
                for graph in ("1", "0", "1"):
                    monkeypatch.setenv("FWX_GRAPH", graph)
                    monkeypatch.setenv("FWX_SMALL_GRAPH", graph)
                    say("small pk16", small16, squaring, n, tag, graph)
                    dm.upload(W)
                    fx.run_device("tiling_gpu", dm.ptr, n, 32, 1)
                    assert (dm.download() == exp).all(), (small16, squaring, n, tag, graph)
