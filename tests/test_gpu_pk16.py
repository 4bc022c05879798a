"""The pk16 rest passes on the GPU (fwx_tile16.hpp, through the C ABI):
bit-exact parity with the oracle in every regime of the 16-bit cap --
nothing saturates (the benchmark graphs), unreachable pairs, weights around
the cap, distances straddling it -- on graph-replayed and direct sizes,
ragged n, every (tile, kappa) whose rest pass is 128 wide; the form and the
exact finish reported by fwx_last_solve_form; FWX_PK16=0 identical.
"""
import sys

import numpy as np
import pytest

import fw_cuda_amd as fx
from fw_cuda_amd import _lib
from oracle import pyoracle as po

INF = fx.INF
CAP = _lib.PK16_CAP
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def graph(n, p, lo, hi, seed):
    rng = np.random.default_rng(seed)
    W = np.where(rng.random((n, n)) < p, rng.integers(lo, hi + 1, (n, n)), INF).astype(np.int32)
    np.fill_diagonal(W, 0)
    return W


# (name, n, density, wmin, wmax, seed, exact finish expected)
CASES = [
    ("dense small weights", 1024, 0.9, 1, 1000, 1, 0),
    ("ragged dense", 1001, 0.5, 1, 1000, 2, 0),
    ("sparse, unreachable pairs", 1000, 0.002, 1, 1000, 3, 1),
    ("weights around the cap", 768, 0.5, CAP - 3, CAP + 3, 4, 1),
    ("distances straddling the cap", 900, 0.02, 1000, 9000, 5, None),
]


@pytest.fixture(scope="module")
def cases():
    out = {}
    for c in CASES:
        W = graph(*c[1:6])
        out[c[0]] = (c, po.fw_rowpar(W), W)
    return out


@pytest.mark.parametrize("tile,kappa", [(64, 2), (64, 4), (128, 1), (128, 2), (32, 4)])
@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_pk16_bit_exact(cases, name, tile, kappa):
    (_, n, _, _, _, _, finish), want, W = cases[name]
    say("pk16", name, n, tile, kappa)
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
    form, exact = fx.last_solve_form()
    assert form == _lib.FORM_PK16
    if finish is not None:
        assert exact == finish, name
    assert (got == want).all(), name


@pytest.mark.parametrize("name", ["dense small weights", "sparse, unreachable pairs"])
def test_pk16_off_is_identical(monkeypatch, cases, name):
    (_, n, _, _, _, _, _), want, W = cases[name]
    monkeypatch.setenv("FWX_PK16", "0")
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 64, 2)
    assert fx.last_solve_form() == (_lib.FORM_PK64, 0)
    assert (got == want).all()


@pytest.mark.parametrize("kc", ["16", "32"])
def test_pk16_chunk_depths(monkeypatch, cases, kc):
    (_, n, _, _, _, _, _), want, W = cases["ragged dense"]
    monkeypatch.setenv("FWX_P16_KC", kc)
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 128, 2)
    assert (got == want).all()


def test_pk16_forms_follow_the_range_check():
    """Negative weights keep min3, operands >= 2^30 keep min3, a 64-wide rest
    pass keeps pk64; each exact."""
    rng = np.random.default_rng(9)
    W = graph(640, 0.3, 1, 1000, 9)
    W[rng.integers(0, 640, 50), rng.integers(0, 640, 50)] = 1 << 30
    np.fill_diagonal(W, 0)
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 64, 2)
    assert fx.last_solve_form()[0] == _lib.FORM_MIN3
    assert (got == po.fw_rowpar(W)).all()
    W = graph(640, 0.3, 1, 1000, 10)
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 64, 1)
    assert fx.last_solve_form()[0] == _lib.FORM_PK64
    assert (got == po.fw_rowpar(W)).all()


def test_pk16_direct_launches_large_ragged():
    """n above the graph-replay limit: direct launches, 16-byte-unaligned rows."""
    W = graph(4102, 0.02, 1, 1000, 11)
    say("pk16 direct n=4102")
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 64, 2)
    assert fx.last_solve_form()[0] == _lib.FORM_PK16
    assert (got == po.fw_rowpar(W)).all()


def test_pk16_repeated_replays_alternate_finish(cases):
    """One buffer solved repeatedly with inputs that do and do not saturate
    (graph replay of both forms, the exact finish in between)."""
    (_, n, _, _, _, _, _), want_a, Wa = cases["dense small weights"]
    Wb = graph(1024, 0.001, 1, 1000, 12)
    want_b = po.fw_rowpar(Wb)
    with fx.DeviceMatrix(1024) as dm:
        for i in range(4):
            W, want = (Wa, want_a) if i % 2 == 0 else (Wb, want_b)
            dm.upload(W)
            fx.run_device("multi_layer_tiling_gpu", dm.ptr, 1024, 64, 2)
            assert fx.last_solve_form() == (_lib.FORM_PK16, i % 2)
            assert (dm.download() == want).all(), i


def chain(n, back=4000):
    """i -> i+1 weighs 1, every other edge `back` (< 4096): the shortest
    path from i to j > i runs through every vertex between them, so closing
    a diagonal tile of width T needs walks of T - 1 edges (log2(T)
    squarings)."""
    W = np.full((n, n), back, dtype=np.int32)
    W[np.arange(n - 1), np.arange(1, n)] = 1
    np.fill_diagonal(W, 0)
    return W


@pytest.mark.parametrize("sq", ["1", "0"])
@pytest.mark.parametrize("tile,kappa", [(64, 4), (64, 1), (32, 4)])
def test_diag_closure_by_squaring(monkeypatch, sq, tile, kappa):
    """The lead rounds' diagonal closure by pk16 squaring (k_diag_sq) on
    dense small weights (every tile after the first stage squares), a chain
    (tiles needing every squaring), sparse ragged graphs (INF cells and a
    padded last tile: the T-step closure) and weights around the 4096
    eligibility bound; FWX_DIAG_SQ=0 closes by the T steps; all exact."""
    monkeypatch.setenv("FWX_DIAG_SQ", sq)
    cases = [("dense", graph(1024, 0.9, 1, 1000, 21)), ("chain", chain(1000)),
             ("sparse ragged", graph(1001, 0.01, 1, 1000, 22)), ("around 4096", graph(777, 0.3, 3000, 5000, 23))]
    for tag, W in cases:
        say("diag squaring", sq, tile, kappa, tag)
        got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
        assert (got == po.fw_rowpar(W)).all(), (tag, sq, tile, kappa)


PACK_GRAPHS = [
    # (tag, W, packed expected, exact finish expected)
    ("dense n=1024", lambda: graph(1024, 0.9, 1, 1000, 31), True, 0),
    ("unreachable pairs n=1280", lambda: graph(1280, 0.002, 1, 1000, 32), True, 1),
    ("partial last stage n=1152", lambda: graph(1152, 0.05, 1, 1000, 33), True, None),
    ("ragged n=1000", lambda: graph(1000, 0.5, 1, 1000, 34), False, 0),
    ("chain n=1024 (every block squaring)", lambda: chain(1024, 3000), True, 0),
]


def negative_matching(n, seed):
    """Weights in [2100, 3000] plus 32 edges of weight -2000 from the first
    half of the vertices to the second: no vertex has a negative edge both in
    and out, so every cycle pairs each negative edge with a following edge
    of weight >= 2100 and no cycle is negative.  -2000 lies below the pk16
    form's bias, so a pk16 kernel that ran on it would produce wrong bits."""
    rng = np.random.default_rng(seed)
    W = graph(n, 0.3, 2100, 3000, seed)
    src = rng.choice(n // 2, 32, replace=False)
    dst = n // 2 + rng.choice(n // 2, 32, replace=False)
    W[src, dst] = -2000
    return W


def unreachable_vertex(n, v, seed):
    """A dense graph in which no edge enters vertex v: column v stays INF
    (at the pk16 cap) off the diagonal, every other cell is reachable."""
    W = graph(n, 0.5, 1, 1000, seed)
    W[:, v] = INF
    W[v, v] = 0
    return W


@pytest.mark.parametrize("la", ["0", "1", "2", "3"])
@pytest.mark.parametrize("kappa", [2, 4])
def test_deferred_range_check(monkeypatch, la, kappa):
    """Packed solves on 64-wide tiles with kappa > 1 enqueue the range check
    and every kernel without waiting for it; the kernels return at entry
    when the check fails, and the host, reading the check after the solve,
    solves the untouched input in the min3 form.  In range: pk16 on packed
    panels, and the cells left at the cap are found by the last stage's pack
    (its rows and columns: vertex n-1 unreachable) and its rest pass (every
    other cell: vertex 0 unreachable), or by the pass over the matrix
    (FWX_CAP_FUSED=0); either way the exact finish follows.  Out of range
    (negative weights, operands >= 2^30): min3, not packed.
    FWX_RANGE_DEFER=0 waits for the check first.  Each bit-exact against the
    oracle, in every schedule."""
    monkeypatch.setenv("FWX_GRAPH", "0")
    monkeypatch.setenv("FWX_LOOKAHEAD", la)
    n = 1024
    rng = np.random.default_rng(45)
    big = graph(n, 0.3, 1, 1000, 44)
    big[rng.integers(0, n, 40), rng.integers(0, n, 40)] = 1 << 30
    np.fill_diagonal(big, 0)
    cases = [
        # (tag, W, form, packed, exact finish)
        ("in range", graph(n, 0.5, 1, 1000, 43), _lib.FORM_PK16, True, 0),
        ("last vertex unreachable", unreachable_vertex(n, n - 1, 47), _lib.FORM_PK16, True, 1),
        ("vertex 0 unreachable", unreachable_vertex(n, 0, 48), _lib.FORM_PK16, True, 1),
        ("negative weights", negative_matching(n, 46), _lib.FORM_MIN3, False, 0),
        ("operands >= 2^30", big, _lib.FORM_MIN3, False, 0),
    ]
    for defer, fused in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("FWX_RANGE_DEFER", defer)
        monkeypatch.setenv("FWX_CAP_FUSED", fused)
        for tag, W, form, packed, finish in cases:
            say("deferred range check", la, kappa, defer, fused, tag)
            got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", 64, kappa)
            assert fx.last_solve_form() == (form, finish), (tag, defer, fused)
            assert fx.last_solve_packed() == packed, (tag, defer, fused)
            assert (got == po.fw_rowpar(W)).all(), (tag, la, kappa, defer, fused)


@pytest.fixture(scope="module")
def pack_graphs():
    out = []
    for tag, make, packed, finish in PACK_GRAPHS:
        W = make()
        out.append((tag, W, po.fw_rowpar(W), packed, finish))
    return out


@pytest.mark.parametrize("la", ["0", "1", "2", "3"])
@pytest.mark.parametrize("tile,kappa", [(64, 2), (64, 4), (128, 1), (128, 2), (32, 4)])
def test_pk16_packed_panels(monkeypatch, pack_graphs, la, tile, kappa):
    """Direct launches (FWX_GRAPH=0) of a pk16 solve with n % 128 == 0 run
    the rest passes on packed stage panels (k_pack16 after each stage's lead
    rounds, k_rest16p reading buffer stage % 3) in every schedule: bit-exact
    against the oracle with and without the exact finish and with a partial
    last stage; ragged n takes the int32 panels; with 64-wide tiles and
    kappa > 1 the lead phase runs as block squarings plus depth-L panels (a
    chain needs every squaring; the first stage's squarings split each tile's
    depth over 4 workgroups, or 8 or none with FWX_SQ_SPLIT); the kappa lead
    rounds (FWX_LEAD_BLOCK=0), four waves per SIMD (FWX_REST16P_WAVES=4) and
    FWX_PACK16=0 give the same bits."""
    monkeypatch.setenv("FWX_GRAPH", "0")
    monkeypatch.setenv("FWX_LOOKAHEAD", la)
    for tag, W, want, packed, finish in pack_graphs:
        say("packed panels", la, tile, kappa, tag)
        got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
        form, exact = fx.last_solve_form()
        assert form == _lib.FORM_PK16, tag
        assert fx.last_solve_packed() == packed, tag
        if finish is not None:
            assert exact == finish, tag
        assert (got == want).all(), (tag, la, tile, kappa)
    for split in ("8", "0"):
        # the first stage's block squarings with each tile's depth in 8 parts
        # (atomic-min merge) and in one workgroup per tile
        monkeypatch.setenv("FWX_SQ_SPLIT", split)
        for tag, W, want, packed, _ in (pack_graphs[0], pack_graphs[4]):
            got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
            assert fx.last_solve_packed() == packed
            assert (got == want).all(), (tag, "squaring split", split)
    monkeypatch.delenv("FWX_SQ_SPLIT")
    tag, W, want, _, _ = pack_graphs[1]
    monkeypatch.setenv("FWX_LEAD_BLOCK", "0")
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
    assert fx.last_solve_packed()
    assert (got == want).all(), (tag, "packed, lead rounds")
    monkeypatch.delenv("FWX_LEAD_BLOCK")
    monkeypatch.setenv("FWX_REST16P_WAVES", "4")
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
    assert fx.last_solve_packed()
    assert (got == want).all(), (tag, "packed, four waves per SIMD")
    monkeypatch.setenv("FWX_PACK16", "0")
    got = fx.floyd_warshall(W, "multi_layer_tiling_gpu", tile, kappa)
    assert not fx.last_solve_packed()
    assert (got == want).all(), (tag, "unpacked")
