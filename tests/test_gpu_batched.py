"""Batched FW: every graph of a batch bit-exact against the oracle, through the
hipGraph replay, the direct launch sequence and the per-graph loop."""
import sys

import numpy as np
import pytest

import fw_cuda_amd as fx
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
INF = fx.INF


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_batch(ns, dens, seeds, neg=()):
    graphs = []
    for g, (n, p, s) in enumerate(zip(ns, dens, seeds)):
        G = fx.generate_graph(int(n), float(p), int(s), 1, 1000)
        if g in neg:  # negative edges, no negative cycle (potential reweighting)
            h = np.random.default_rng(s).integers(0, 500, int(n))
            adj = G != INF
            G = np.where(adj, G.astype(np.int64) + h[:, None] - h[None, :], INF).astype(np.int32)
            np.fill_diagonal(G, 0)
        graphs.append(G)
    return graphs


def upload(graphs):
    import torch
    sizes = np.array([g.shape[0] for g in graphs], dtype=np.int32)
    offs, total = fx.batch_layout(sizes)
    flat = np.full(total, 7, dtype=np.int32)  # gaps hold junk that must survive
    for g, o in zip(graphs, offs):
        flat[o:o + g.size] = g.ravel()
    return sizes, offs, torch.from_numpy(flat).to("cuda:0"), flat


def unpack(t, sizes, offs):
    flat = t.cpu().numpy()
    return [flat[o:o + int(n) * int(n)].reshape(int(n), int(n)) for n, o in zip(sizes, offs)], flat


def check_all(got, expect, flat_after, flat_before, sizes, offs):
    for g, (a, b) in enumerate(zip(got, expect)):
        assert (a == b).all(), f"graph {g} (n={a.shape[0]})"
    # bytes outside the graphs untouched
    mask = np.ones(flat_before.size, dtype=bool)
    for n, o in zip(sizes, offs):
        mask[o:o + int(n) * int(n)] = False
    assert (flat_after[mask] == flat_before[mask]).all()


@pytest.mark.parametrize("deep", ["default", "default/nohoist", "off", "off/nohoist", "64,64,4", "100,128,3",
                                  "off/diag256"])
def test_batch_small_mixed(monkeypatch, deep):
    """Mixed sizes (negative weights in some graphs) under the default batch
    configuration rule, the plain path's tuned entries ("off", also with
    256-thread 64-wide closures), and two other per-size overrides, with the
    closures on side branches (default) and in steps of their own
    ("nohoist"): every graph equals the oracle."""
    deep, *flags = deep.split("/")
    if "diag256" in flags:
        monkeypatch.setenv("FWX_BATCH_DIAG", "256")
    if "nohoist" in flags:
        monkeypatch.setenv("FWX_BATCH_HOIST", "0")
    if deep != "default":
        monkeypatch.setenv("FWX_BATCH_DEEP", deep)
    rng = np.random.default_rng(3)
    ns = [1, 2, 3, 5, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 300, 399, 513, 640]
    ns = ns + list(rng.integers(1, 400, 18))
    dens = rng.uniform(0.05, 0.9, len(ns))
    seeds = list(rng.integers(1, 2**31, len(ns)))
    graphs = make_batch(ns, dens, seeds, neg={3, 9, 14, 20, 25})
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    from fw_cuda_amd import _lib
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    say("plan", plan.stats())
    for mode in ("graph", "graph", "direct"):
        t.copy_(__import__("torch").from_numpy(flat0).to("cuda:0"))
        plan.run(mode)
        got, flat = unpack(t, sizes, offs)
        say("checking", mode)
        check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()
    t.copy_(__import__("torch").from_numpy(flat0).to("cuda:0"))
    fx.batch_loop(sizes, offs, _lib.ptr(t))
    got, flat = unpack(t, sizes, offs)
    check_all(got, expect, flat, flat0, sizes, offs)


PERSISTENT_OFF = ("exercised the per-XCD dependency-counter persistent kernel, which was removed after its "
                  "first two GPU runs did not complete (cause not established); the persistent modes now run "
                  "the cooperative kernel that test_batch_cooperative_* check")


@pytest.mark.skip(reason=PERSISTENT_OFF)
def test_batch_persistent_modes():
    """The persistent kernel (direct and from its hipGraph, twice each) equals
    the oracle on the mixed batch of test_batch_small_mixed."""
    import torch
    from fw_cuda_amd import _lib
    rng = np.random.default_rng(3)
    ns = [1, 2, 3, 5, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 300, 399, 513, 640]
    ns = ns + list(rng.integers(1, 400, 18))
    dens = rng.uniform(0.05, 0.9, len(ns))
    seeds = list(rng.integers(1, 2**31, len(ns)))
    graphs = make_batch(ns, dens, seeds, neg={3, 9, 14, 20, 25})
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    for mode in ("persistent", "persistent", "persistent_graph", "persistent_graph"):
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        say("checking", mode)
        plan.run(mode)
        got, flat = unpack(t, sizes, offs)
        check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()


def _mixed_for_cooperative():
    rng = np.random.default_rng(5)
    ns = [1, 2, 3, 5, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257, 300, 399, 513, 640,
          700, 1024]
    ns = ns + list(rng.integers(1, 600, 15))
    dens = rng.uniform(0.05, 0.9, len(ns))
    seeds = list(rng.integers(1, 2**31, len(ns)))
    return make_batch(ns, dens, seeds, neg={3, 9, 14, 20, 26})


def test_batch_cooperative_direct():
    """One cooperative launch solves a mixed batch (ragged sizes from 1 to
    1024, graphs below n = 128 as one closure, negative weights in some
    graphs: the min3 form) bit-exactly against the oracle, twice on restored
    input, and again after a run in the pk64 form; bytes between graphs stay
    untouched."""
    import torch
    from fw_cuda_amd import _lib
    graphs = _mixed_for_cooperative()
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    info = plan.persistent_info()
    say("cooperative plan", info, plan.stats())
    assert info["grid"] > 0 and info["steps"] > 0 and info["tasks"] > 0
    for rep in range(2):
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        say("cooperative run", rep, "(min3 form: negative weights present)")
        plan.run("persistent")
        got, flat = unpack(t, sizes, offs)
        check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()
    # non-negative weights only: the pk64 form
    pos = [np.where(g < 0, 1, g).astype(np.int32) for g in graphs]
    for g in pos:
        np.fill_diagonal(g, 0)
    expect = [po.fw_rowpar(g) for g in pos]
    sizes, offs, t, flat0 = upload(pos)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    say("cooperative run (pk64 form)")
    plan.run("persistent")
    got, flat = unpack(t, sizes, offs)
    check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()


def test_batch_cooperative_graph():
    """The cooperative launch captured into a hipGraph and replayed twice on
    restored input: bit-exact against the oracle.  A runtime that cannot
    capture a cooperative launch makes the mode return FWX_ERR_UNSUPPORTED
    (checked, then the test is skipped with that message)."""
    import torch
    from fw_cuda_amd import _lib
    graphs = _mixed_for_cooperative()
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    try:
        for rep in range(2):
            t.copy_(torch.from_numpy(flat0).to("cuda:0"))
            say("cooperative graph replay", rep)
            try:
                plan.run("persistent_graph")
            except _lib.FwxError as e:
                assert e.code == _lib.ERR_UNSUPPORTED, e
                pytest.skip(f"cooperative launch not capturable here: {e}")
            got, flat = unpack(t, sizes, offs)
            check_all(got, expect, flat, flat0, sizes, offs)
        # the step-launch replay of the same plan afterwards is unaffected
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        plan.run("graph")
        got, flat = unpack(t, sizes, offs)
        check_all(got, expect, flat, flat0, sizes, offs)
    finally:
        plan.destroy()


@pytest.mark.skip(reason=PERSISTENT_OFF)
def test_batch_persistent_wait_limit():
    """The persistent kernel's bounded wait: with a zero limit every task that
    finds its dependencies unfinished stops the run, which reports
    FWX_ERR_TIMEOUT (no fault, no hang); with the default limit the same plan
    then solves the batch exactly."""
    import torch
    from fw_cuda_amd import _lib
    rng = np.random.default_rng(8)
    ns = [700, 640, 512, 300, 129, 64]
    graphs = make_batch(ns, rng.uniform(0.2, 0.9, len(ns)), list(rng.integers(1, 2**31, len(ns))))
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    plan.set_wait_limit(0.0)
    say("persistent with a zero wait limit")
    with pytest.raises(_lib.FwxError) as e:
        plan.run("persistent")
    assert e.value.code == _lib.ERR_TIMEOUT
    plan.set_wait_limit(10.0)
    t.copy_(torch.from_numpy(flat0).to("cuda:0"))
    plan.run("persistent")
    got, flat = unpack(t, sizes, offs)
    check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()


@pytest.mark.slow
def test_batch_512_mixed_sizes():
    """512 graphs, n log-uniform in [64, 2048] (seed 11), density U[0.05, 0.9],
    weights U[1, 1000]: one hipGraph replay of the step launches, every graph
    bit-exact against the oracle; a second replay and the persistent
    cooperative launch (from its hipGraph) give the same matrices."""
    import torch
    from fw_cuda_amd import _lib
    ns, dens, seeds = fx.mixed_batch()
    graphs = make_batch(ns, dens, seeds)
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    plan.run(True)
    got, flat = unpack(t, sizes, offs)
    say("512 graphs: gpu done; oracle", plan.stats())
    for g in range(len(graphs)):
        assert (got[g] == po.fw_rowpar(graphs[g])).all(), g
    # a second replay on restored input, and the one persistent cooperative
    # launch from its hipGraph, give the same matrices
    for mode in ("graph", "persistent_graph"):
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        say("512 graphs", mode)
        plan.run(mode)
        got2, _ = unpack(t, sizes, offs)
        for g in range(len(graphs)):
            assert (got2[g] == got[g]).all(), (mode, g)
    plan.destroy()


def _graph(n, p, lo, hi, seed):
    rng = np.random.default_rng(seed)
    W = np.where(rng.random((n, n)) < p, rng.integers(lo, hi + 1, (n, n)), INF).astype(np.int32)
    np.fill_diagonal(W, 0)
    return W


@pytest.mark.parametrize("pk16,table", [("1", "64:32:2:512,127:128:1:512,32769:128:3:512"),
                                         ("1", "64:32:2:512,127:128:1:512,32769:64:4:512"),
                                         ("0", "64:32:2:512,127:128:1:512,32769:128:3:512")])
def test_batch_pk16_products_and_exact_finish(monkeypatch, pk16, table):
    """pk16 128-wide products in a batch whose graphs do and do not reach the
    16-bit cap (dense small weights; unreachable pairs; weights around the
    cap; a graph too small for 128-wide products): every graph equals the
    oracle in both step-launch modes, the run reports the pk16 form, and the
    graphs with 128-wide products that hold cells at or above the cap are the
    ones finished exactly; with 128-wide and with 64-wide products for the
    graphs above n = 128.  FWX_PK16=0 runs pk64 with the same results."""
    import torch
    from fw_cuda_amd import _lib
    cap = _lib.PK16_CAP
    monkeypatch.setenv("FWX_PK16", pk16)
    monkeypatch.setenv("FWX_BATCH_TABLE", table)
    graphs = [_graph(300, 0.9, 1, 1000, 1), _graph(300, 0.003, 1, 1000, 2), _graph(520, 0.5, 1, 1000, 3),
              _graph(400, 0.5, cap - 3, cap + 3, 4), _graph(100, 0.01, 1, 1000, 5), _graph(777, 0.02, 1, 1000, 6),
              _graph(260, 0.05, 2000, 9000, 7)]
    expect = [po.fw_rowpar(g) for g in graphs]
    # graphs with 64- or 128-wide products under the forced tables: n > 128
    # (a single 128-wide tile has no products; every graph above 128 has
    # 64-wide ones under the second table)
    wide = [g.shape[0] > 128 for g in graphs]
    want_finished = sum(1 for w, e in zip(wide, expect) if w and (e >= cap).any())
    assert want_finished >= 2  # the batch exercises the exact finish
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    for mode in ("graph", "graph", "direct"):
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        say("pk16 batch", pk16, mode)
        plan.run(mode)
        form, finished = plan.last_form()
        if pk16 == "1":
            assert (form, finished) == (_lib.FORM_PK16, want_finished), mode
        else:
            assert (form, finished) == (_lib.FORM_PK64, 0), mode
        got, flat = unpack(t, sizes, offs)
        check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()


def _chain(n, back=4000):
    """i -> i+1 weighs 1, every other edge `back` (< 4096): closing a diagonal
    tile of width T needs walks of T - 1 edges (every squaring)."""
    W = np.full((n, n), back, dtype=np.int32)
    W[np.arange(n - 1), np.arange(1, n)] = 1
    np.fill_diagonal(W, 0)
    return W


@pytest.mark.parametrize("sq", ["1", "0"])
@pytest.mark.parametrize("table", ["64:32:2:512,127:128:1:512,32769:128:3:512",
                                   "64:32:2:512,127:128:1:512,32769:64:4:512"])
def test_batch_diag_closure_by_squaring(monkeypatch, sq, table):
    """Diagonal closures of 32-, 64- and 128-wide tiles by pk16 squaring
    (k_batch_diag_sq) in a batch of dense small weights (tiles square after
    the first stage), chains (every squaring), sparse graphs with INF cells
    and ragged last tiles (the T-step closure), and weights around the 4096
    eligibility bound; FWX_DIAG_SQ=0 closes every tile by the T steps; each
    graph equals the oracle in both step-launch modes."""
    import torch
    from fw_cuda_amd import _lib
    monkeypatch.setenv("FWX_DIAG_SQ", sq)
    monkeypatch.setenv("FWX_BATCH_TABLE", table)
    graphs = [_graph(300, 0.9, 1, 1000, 11), _chain(384), _chain(50), _graph(64, 0.9, 1, 1000, 12),
              _graph(517, 0.01, 1, 1000, 13), _graph(260, 0.3, 3000, 5000, 14), _chain(700)]
    expect = [po.fw_rowpar(g) for g in graphs]
    sizes, offs, t, flat0 = upload(graphs)
    plan = fx.BatchPlan(sizes, offs, _lib.ptr(t))
    for mode in ("graph", "direct"):
        t.copy_(torch.from_numpy(flat0).to("cuda:0"))
        say("batch diag squaring", sq, table, mode)
        plan.run(mode)
        got, flat = unpack(t, sizes, offs)
        check_all(got, expect, flat, flat0, sizes, offs)
    plan.destroy()
