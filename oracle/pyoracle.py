"""ctypes bindings to the C oracle (liboracle.so) and, when built, to the
reference's own CPU sources (_ref/libfwref.so).  TEST INFRASTRUCTURE ONLY.

Each wrapper names the C function it calls; the C file cites the reference
file:line each one restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 1_000_000_000  # include/constants.h:3

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return ctypes.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(os.path.join(HERE, "liboracle.so"))
        _lib.oracle_fw_serial.argtypes = [_i32p, ctypes.c_int32]
        _lib.oracle_fw_rowpar.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32]
        _lib.oracle_has_negative_cycle.argtypes = [_i32p, ctypes.c_int32]
        _lib.oracle_has_negative_cycle.restype = ctypes.c_int
        _lib.oracle_dijkstra_rows.argtypes = [_i32p, ctypes.c_int32, _i32p, ctypes.c_int32,
                                              _i32p, ctypes.c_int32]
        _lib.oracle_dijkstra_rows.restype = ctypes.c_int
        _lib.oracle_certify_rows.argtypes = [_i32p, ctypes.c_int32, _i32p, ctypes.c_int32, _i32p, _i32p,
                                             ctypes.c_int32]
        _lib.oracle_certify_rows.restype = ctypes.c_int
        _lib.oracle_brute_walks.argtypes = [_i32p, ctypes.c_int32, _i32p]
        _lib.oracle_predecessors.argtypes = [_i32p, _i32p, _i32p, ctypes.c_int32]
        _u32p = ctypes.POINTER(ctypes.c_uint32)
        _lib.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        _lib.oracle_generate_philox.argtypes = [_i32p, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64,
                                                ctypes.c_int32, ctypes.c_int32]
        _lib.oracle_fw_f32_pred.argtypes = [_f32p, _f32p, _i32p, ctypes.c_int32]
        _lib.oracle_grad_paths.argtypes = [_i32p, _f32p, _f32p, ctypes.c_int32]
        _lib.oracle_grad_paths.restype = ctypes.c_int
        _lib.oracle_fw_f64.argtypes = [_f64p, _f64p, ctypes.c_int32]
        _lib.oracle_dijkstra_f64_rows.argtypes = [_f32p, ctypes.c_int32, _i32p, ctypes.c_int32, _f64p,
                                                  ctypes.c_int32]
        _lib.oracle_dijkstra_f64_rows.restype = ctypes.c_int
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libfwref.so"))


def ref():
    """The reference's own fwBaselineCPU / generateGraph1D / hasNegativeCycle."""
    global _ref
    if _ref is None:
        _ref = _load(os.path.join(HERE, "_ref", "libfwref.so"))
        _ref.ref_fw_baseline_cpu.argtypes = [_i32p, ctypes.c_int]
        _ref.ref_generate_graph.argtypes = [_i32p, ctypes.c_int, ctypes.c_double, ctypes.c_uint]
        _ref.ref_generate_graph.restype = ctypes.c_int
        _ref.ref_has_negative_cycle.argtypes = [_i32p, ctypes.c_int]
        _ref.ref_has_negative_cycle.restype = ctypes.c_int
    return _ref


def _sq(D):
    D = np.ascontiguousarray(D, dtype=np.int32)
    n = int(round(D.size ** 0.5))
    assert n * n == D.size
    return D.reshape(n, n), n


def fw_serial(D):
    """oracle_fw_serial: returns the closed copy (input untouched)."""
    out, n = _sq(np.array(D, dtype=np.int32, copy=True))
    lib().oracle_fw_serial(_ptr(out, _i32p), n)
    return out


def fw_rowpar(D, threads=0):
    """oracle_fw_rowpar (OpenMP row-parallel, bit-identical without negative cycles)."""
    out, n = _sq(np.array(D, dtype=np.int32, copy=True))
    lib().oracle_fw_rowpar(_ptr(out, _i32p), n, int(threads))
    return out


def has_negative_cycle(D):
    D, n = _sq(D)
    return bool(lib().oracle_has_negative_cycle(_ptr(D, _i32p), n))


def dijkstra_rows(W, rows, threads=0):
    """oracle_dijkstra_rows: exact rows of the APSP over the INF-completed graph."""
    W, n = _sq(W)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.empty((rows.size, n), dtype=np.int32)
    rc = lib().oracle_dijkstra_rows(_ptr(W, _i32p), n, _ptr(rows, _i32p), rows.size,
                                    _ptr(out, _i32p), int(threads))
    if rc != 0:
        raise ValueError(f"oracle_dijkstra_rows failed ({rc}): negative weights?")
    return out


def certify_rows(W, rows, D_rows, threads=0):
    """oracle_certify_rows: per-row True iff D_rows[q] is exactly the shortest
    distance row from source rows[q] (weights must be in [1, INF], diagonal 0)."""
    W, n = _sq(W)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    D_rows = np.ascontiguousarray(D_rows, dtype=np.int32)
    assert D_rows.shape == (rows.size, n)
    ok = np.zeros(rows.size, dtype=np.int32)
    rc = lib().oracle_certify_rows(_ptr(W, _i32p), n, _ptr(rows, _i32p), rows.size, _ptr(D_rows, _i32p),
                                   _ptr(ok, _i32p), int(threads))
    if rc < 0:
        raise ValueError(f"oracle_certify_rows failed ({rc}): weights outside [1, INF] or no memory")
    return ok.astype(bool)


def predecessors(W, D):
    """oracle_predecessors: smallest tight predecessor u != j of every (i, j)."""
    W, n = _sq(W)
    D, _ = _sq(D)
    P = np.empty((n, n), dtype=np.int32)
    lib().oracle_predecessors(_ptr(W, _i32p), _ptr(D, _i32p), _ptr(P, _i32p), n)
    return P


def philox4x32_10(ctr, key):
    """oracle_philox4x32_10: one Philox4x32-10 block."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    u = ctypes.POINTER(ctypes.c_uint32)
    lib().oracle_philox4x32_10(c.ctypes.data_as(u), k.ctypes.data_as(u), out.ctypes.data_as(u))
    return out


def generate_philox(n, density, seed, wmin=1, wmax=1000):
    """oracle_generate_philox: the device generator's matrix, on the host."""
    out = np.empty((n, n), dtype=np.int32)
    lib().oracle_generate_philox(_ptr(out, _i32p), int(n), float(density), int(seed) & (2**64 - 1),
                                 int(wmin), int(wmax))
    return out


def brute_walks(W):
    W, n = _sq(W)
    out = np.empty_like(W)
    lib().oracle_brute_walks(_ptr(W, _i32p), n, _ptr(out, _i32p))
    return out


def fw_f32_pred(W):
    """oracle_fw_f32_pred: fp32 FW over W (inf = absent, diagonal taken as 0)."""
    W = np.ascontiguousarray(W, dtype=np.float32)
    n = W.shape[0]
    D = np.empty_like(W)
    P = np.empty(W.shape, dtype=np.int32)
    lib().oracle_fw_f32_pred(_ptr(W, _f32p), _ptr(D, _f32p), _ptr(P, _i32p), n)
    return D, P


def grad_paths(P, G):
    """oracle_grad_paths: dL/dW for L = sum(G*D) by walking every path."""
    P = np.ascontiguousarray(P, dtype=np.int32)
    G = np.ascontiguousarray(G, dtype=np.float32)
    n = P.shape[0]
    dW = np.zeros((n, n), dtype=np.float32)
    rc = lib().oracle_grad_paths(_ptr(P, _i32p), _ptr(G, _f32p), _ptr(dW, _f32p), n)
    if rc != 0:
        raise ValueError("corrupt predecessor matrix")
    return dW


def fw_f64(W):
    W = np.ascontiguousarray(W, dtype=np.float64)
    n = W.shape[0]
    D = np.empty(W.shape, dtype=np.float64)
    lib().oracle_fw_f64(_ptr(W, _f64p), _ptr(D, _f64p), n)
    return D


def dijkstra_f64_rows(W, rows, threads=0):
    """oracle_dijkstra_f64_rows: float64 shortest path lengths from each source
    in `rows` over the fp32 graph W (+inf = no edge, diagonal taken as 0)."""
    W = np.ascontiguousarray(W, dtype=np.float32)
    n = W.shape[0]
    assert W.shape == (n, n)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.empty((rows.size, n), dtype=np.float64)
    rc = lib().oracle_dijkstra_f64_rows(_ptr(W, _f32p), n, _ptr(rows, _i32p), rows.size,
                                        _ptr(out, _f64p), int(threads))
    if rc != 0:
        raise ValueError(f"oracle_dijkstra_f64_rows failed ({rc}): negative weights?")
    return out


def f32_row_failures(D_rows, P_rows, W, rows, ref, rtol=2e-6):
    """Checks claimed fp32 distance rows D_rows and predecessor rows P_rows of
    the sources `rows` against float64 reference rows `ref`
    (dijkstra_f64_rows): reachability agrees, every finite D is within rtol
    of the reference, and every predecessor is tight, i.e.
    ref[P[t]] + W[P[t]][t] equals ref[t] within rtol (P[s] = -1, P of an
    unreachable t = -1).  Returns a list of (source, target, reason); empty
    when every row passes."""
    W = np.asarray(W, dtype=np.float32)
    bad = []
    for q, s in enumerate(np.asarray(rows)):
        d, p, r = np.asarray(D_rows[q]), np.asarray(P_rows[q]), ref[q]
        fin = np.isfinite(r)
        for t in np.flatnonzero(np.isfinite(d) != fin):
            bad.append((int(s), int(t), "reachability"))
        close = np.abs(d[fin].astype(np.float64) - r[fin]) <= rtol * np.abs(r[fin])
        for t in np.flatnonzero(fin)[~close]:
            bad.append((int(s), int(t), "distance"))
        if p[s] != -1:
            bad.append((int(s), int(s), "source predecessor"))
        tgt = np.flatnonzero(fin)
        tgt = tgt[tgt != s]
        u = p[tgt]
        valid = (u >= 0) & (u < len(d))
        for t in tgt[~valid]:
            bad.append((int(s), int(t), "missing predecessor"))
        tgt, u = tgt[valid], u[valid]
        via = r[u] + W[u, tgt].astype(np.float64)
        tight = np.abs(via - r[tgt]) <= rtol * np.abs(r[tgt])
        for t in tgt[~tight]:
            bad.append((int(s), int(t), "predecessor not tight"))
        for t in np.flatnonzero(~fin):
            if p[t] != -1:
                bad.append((int(s), int(t), "predecessor of an unreachable vertex"))
    return bad


def ref_fw_baseline_cpu(D):
    out, n = _sq(np.array(D, dtype=np.int32, copy=True))
    ref().ref_fw_baseline_cpu(_ptr(out, _i32p), n)
    return out


def ref_generate_graph(n, density, seed):
    out = np.empty((n, n), dtype=np.int32)
    rc = ref().ref_generate_graph(_ptr(out, _i32p), n, float(density), int(seed))
    if rc != 0:
        raise ValueError("reference generator rejected the arguments")
    return out
