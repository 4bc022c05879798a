"""fwx_certify_rows: the on-device optimality certificate of claimed distance
rows, checked row by row against the oracle's certificate (oracle_certify_rows)
on the same inputs -- exact rows, every kind of corrupted row, ragged sizes,
unreachable vertices -- and then used at full size: every row of an n=8192
solve and of the n=32768 configuration certified on the device."""
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


def corrupt(rng, D, src):
    """Claimed rows D[src] with about half of them corrupted in one of six
    ways; returns (rows, mask of the rows left exact)."""
    n = D.shape[0]
    rows = D[src].copy()
    exact = np.ones(src.size, dtype=bool)
    for q in range(src.size):
        kind = int(rng.integers(0, 12))
        s = int(src[q])
        others = [v for v in range(n) if v != s]
        if kind >= 6 or not others:
            continue
        v = int(rng.choice(others))
        if kind == 0:
            rows[q, v] += 1                      # one distance too long
        elif kind == 1:
            if rows[q, v] == 0:
                continue
            rows[q, v] -= 1                      # one distance too short
        elif kind == 2:
            rows[q, s] = 1                       # source not at 0
        elif kind == 3:
            rows[q, v] = INF + 1                 # above INF
        elif kind == 4:
            rows[q, v] = -1                      # negative
        else:
            other = int(rng.integers(0, n))
            if other == s or (D[other] == D[s]).all():
                continue
            rows[q] = D[other]                   # another source's row
        exact[q] = False
    return rows, exact


@pytest.mark.parametrize("n,density", [(1, 0.5), (2, 0.9), (3, 0.0), (37, 0.3), (64, 0.9), (100, 0.05),
                                       (257, 0.5), (1000, 0.9), (1030, 0.02)])
def test_certificate_matches_oracle(n, density):
    rng = np.random.default_rng(n)
    W = fx.generate_graph(n, density, int(rng.integers(1, 2**31)), 1, 1000)
    D = po.fw_rowpar(W)
    nsrc = min(n, 200) + 3
    src = rng.integers(0, n, nsrc).astype(np.int32)
    src[0], src[-1] = 0, n - 1
    say("certificate exact rows", n, density, nsrc)
    assert fx.certify_rows(W, src, D[src]).all()
    assert po.certify_rows(W, src, D[src]).all()
    rows, exact = corrupt(rng, D, src)
    say("certificate corrupted rows", n, int((~exact).sum()))
    got = fx.certify_rows(W, src, rows)
    want = po.certify_rows(W, src, rows)
    assert (want == exact).all()
    assert (got == want).all(), np.nonzero(got != want)[0][:10]


def test_certificate_counts_failures():
    n = 300
    rng = np.random.default_rng(3)
    W = fx.generate_graph(n, 0.2, 5, 1, 1000)
    D = po.fw_rowpar(W)
    src = np.arange(n, dtype=np.int32)
    rows, exact = corrupt(rng, D, src)
    import torch
    dev = torch.device("cuda")
    tw, ts, tr = (torch.from_numpy(a).to(dev) for a in (W, src, rows))
    ok = torch.full((n,), 7, dtype=torch.int32, device=dev)
    failed = fx.certify_rows_device(ctypes.c_void_p(tw.data_ptr()), n, ctypes.c_void_p(ts.data_ptr()), n,
                                    ctypes.c_void_p(tr.data_ptr()), ctypes.c_void_p(ok.data_ptr()))
    assert failed == int((~exact).sum())
    assert (ok.cpu().numpy() == exact.astype(np.int32)).all()


@pytest.mark.parametrize("what", ["zero weight", "diagonal", "above INF", "negative"])
def test_certificate_rejects_weights_outside_domain(what):
    n = 65
    W = fx.generate_graph(n, 0.5, 9, 1, 1000)
    if what == "zero weight":
        W[3, 7] = 0
    elif what == "diagonal":
        W[5, 5] = 1
    elif what == "above INF":
        W[1, 2] = INF + 1
    else:
        W[60, 0] = -3
    src = np.arange(4, dtype=np.int32)
    with pytest.raises(_lib.FwxError):
        fx.certify_rows(W, src, np.zeros((4, n), dtype=np.int32))


def test_certificate_of_every_row_at_8192():
    """All 8192 rows of a device solve at the headline size, certified on the
    device against the input (8192^3 relaxations); one corrupted cell in a
    copy of the result is found in exactly its row."""
    import torch
    n = 8192
    dev = torch.device("cuda")
    with fx.DeviceMatrix(n) as dw, fx.DeviceMatrix(n) as dd:
        fx.generate_graph_device(dw.ptr, n, 0.9, 42, 1, 1000)
        dd.upload(dw.download())
        say("cert 8192: solve")
        fx.run_device("multi_layer_tiling_gpu", dd.ptr, n, 0, 0)
        src = torch.arange(n, dtype=torch.int32, device=dev)
        ok = torch.zeros(n, dtype=torch.int32, device=dev)
        say("cert 8192: certificate of every row")
        failed = fx.certify_rows_device(dw.ptr, n, ctypes.c_void_p(src.data_ptr()), n, dd.ptr,
                                        ctypes.c_void_p(ok.data_ptr()))
        assert failed == 0
        assert bool((ok == 1).all())
        d = dd.download()
        d[4321, 777] += 1
        dd.upload(d)
        say("cert 8192: one corrupted cell")
        failed = fx.certify_rows_device(dw.ptr, n, ctypes.c_void_p(src.data_ptr()), n, dd.ptr,
                                        ctypes.c_void_p(ok.data_ptr()))
        bad = torch.nonzero(ok == 0).flatten().cpu().numpy()
    assert failed == 1 and bad.tolist() == [4321]
