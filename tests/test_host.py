"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point include/fwx.h declares, argument errors come back as status
codes, the generator reproduces the reference's matrices bit for bit, and the
drop-in CLIs keep the reference's argv rules and output lines.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import fw_cuda_amd as fx
from fw_cuda_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fw_cuda_amd", "bin")


def declared_symbols():
    syms = set()
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(fwx_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 18
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert syms <= set(_lib.SIGNATURES), syms - set(_lib.SIGNATURES)


def test_registry_matches_reference_names():
    # src/cuda_utils.cu:7-11
    assert fx.get_gpu_algorithm_names() == ("baseline_gpu", "tiling_gpu", "multi_layer_tiling_gpu")
    assert fx.is_gpu_algorithm("multi_layer_tiling_gpu")
    assert not fx.is_gpu_algorithm("baseline_cpu")


def test_argument_errors_are_status_codes():
    L = _lib.lib()
    import ctypes
    null = ctypes.c_void_p(0)
    dummy = ctypes.c_void_p(64)  # never dereferenced: validation fails first
    assert L.fwx_run(b"nope", dummy, 8, 0, 0, None) == _lib.ERR_UNKNOWN_ALGO
    assert b"Unsupported GPU algorithm: nope" in L.fwx_last_error()
    assert L.fwx_multi_layer_tiling_gpu(null, 8, 0, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_multi_layer_tiling_gpu(dummy, 0, 0, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_multi_layer_tiling_gpu(dummy, 32770, 0, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_multi_layer_tiling_gpu(dummy, 8, 48, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_multi_layer_tiling_gpu(dummy, 8, 0, 17, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_tiling_gpu(dummy, 8, 16, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_baseline_gpu(null, 8, None) == _lib.ERR_INVALID_ARG
    with pytest.raises(fx.FwxError):
        fx.generate_graph(0, 0.5, 1)
    with pytest.raises(fx.FwxError):
        fx.generate_graph(4, 1.5, 1)


def test_tuned_config_is_supported():
    for n in (1, 31, 64, 65, 512, 513, 1024, 4096, 8192, 32768):
        t, k = fx.tuned_config(n)
        assert t in (32, 64, 128) and 1 <= k <= 16


def test_generator_matches_reference_bits(gen_ref):
    # outputs of the reference's own generateGraph1D (tests/golden/gen_ref.json)
    for c in gen_ref:
        g = fx.generate_graph(c["n"], c["density"], c["seed"])
        assert (g.ravel() == np.array(c["graph"], dtype=np.int32)).all(), (c["n"], c["seed"])


def test_generator_weight_range_and_density():
    g = fx.generate_graph(500, 0.9, 42, 1, 1000)
    off = ~np.eye(500, dtype=bool)
    vals = g[off]
    present = vals != fx.INF
    assert (np.diag(g) == 0).all()
    assert vals[present].min() >= 1 and vals[present].max() <= 1000
    assert abs(present.mean() - 0.9) < 0.01


def test_has_negative_cycle():
    D = np.array([[0, 1], [1, -1]], dtype=np.int32)
    assert fx.has_negative_cycle(D)
    assert not fx.has_negative_cycle(np.zeros((3, 3), dtype=np.int32))


def run_cli(*args):
    return subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                          capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "bench_runner")), reason="CLIs not built")
def test_bench_runner_argv_rules():
    # src/bench_runner.cpp:25-92
    r = run_cli("bench_runner", "baseline_cpu")
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = run_cli("bench_runner", "baseline_cpu", 0, 0.5)
    assert r.returncode == 1 and "Invalid graph size (0)" in r.stderr
    r = run_cli("bench_runner", "baseline_cpu", "abc", 0.5)
    assert r.returncode == 1 and "Invalid N_SIZE provided (abc)" in r.stderr
    r = run_cli("bench_runner", "baseline_cpu", 8, 1.5)
    assert r.returncode == 1 and "Invalid density" in r.stderr
    r = run_cli("bench_runner", "baseline_cpu", 8, 0.5, 1, 9)
    assert r.returncode == 1 and "Invalid TILE_SIZE (9)" in r.stderr
    r = run_cli("bench_runner", "baseline_cpu", 8, 0.5, 1, 0, -1)
    assert r.returncode == 1 and "Invalid KAPPA (-1)" in r.stderr
    r = run_cli("bench_runner", "nope_algo", 8, 0.5, 1)
    assert r.returncode == 1 and "Unsupported CPU algorithm: nope_algo" in r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "bench_runner")), reason="CLIs not built")
def test_bench_runner_cpu_output_lines():
    r = run_cli("bench_runner", "baseline_cpu", 64, 0.9, 42, 0, 0)
    assert r.returncode == 0, r.stderr
    for line in ("BENCHMARK: baseline_cpu", "Graph size: 64 vertices", "Edge density: 90.0%",
                 "Algorithm type: CPU", "Seed: 42 (Reproducible)", "Execution time: "):
        assert line in r.stdout, line


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "tests_runner")), reason="CLIs not built")
def test_tests_runner_cpu_suite():
    r = run_cli("tests_runner", "baseline_cpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RESULTS: 10/10 passed" in r.stdout
    r = run_cli("tests_runner", "nope")
    assert r.returncode == 1 and "Unsupported CPU algorithm" in r.stderr
    r = run_cli("tests_runner")
    assert r.returncode == 1 and "Usage:" in r.stderr


def test_batched_and_diff_argument_errors():
    import ctypes
    L = _lib.lib()
    dummy = ctypes.c_void_p(64)
    plan = ctypes.c_void_p()
    sizes = np.array([4], dtype=np.int32)
    offs = np.array([0], dtype=np.int64)
    assert L.fwx_batch_plan_create(ctypes.byref(plan), _lib.ptr(sizes), _lib.ptr(offs), 0, dummy,
                                   None) == _lib.ERR_INVALID_ARG
    bad = np.array([0], dtype=np.int32)
    assert L.fwx_batch_plan_create(ctypes.byref(plan), _lib.ptr(bad), _lib.ptr(offs), 1, dummy,
                                   None) == _lib.ERR_INVALID_ARG
    neg = np.array([-4], dtype=np.int64)
    assert L.fwx_batch_plan_create(ctypes.byref(plan), _lib.ptr(sizes), _lib.ptr(neg), 1, dummy,
                                   None) == _lib.ERR_INVALID_ARG
    assert L.fwx_batch_plan_run(None, 1) == _lib.ERR_INVALID_ARG
    r = ctypes.c_int32()
    assert L.fwx_has_negative_cycle_device(None, 4, ctypes.byref(r), None) == _lib.ERR_INVALID_ARG
    assert L.fwx_generate_graph_device(None, 4, 0.5, 1, 1, 10, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_generate_graph_device(dummy, 4, 1.5, 1, 1, 10, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_generate_graph_device(dummy, 4, 0.5, 1, 10, 1, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_host_alloc_pinned(None, 4) == _lib.ERR_INVALID_ARG
    assert L.fwx_predecessors(None, dummy, dummy, 4, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_predecessors(dummy, dummy, dummy, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_copy_to_device_async(None, None, 4, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_certify_rows(None, 4, dummy, 1, dummy, dummy, ctypes.byref(r), None) == _lib.ERR_INVALID_ARG
    assert L.fwx_certify_rows(dummy, 4, dummy, 0, dummy, dummy, ctypes.byref(r), None) == _lib.ERR_INVALID_ARG
    assert L.fwx_certify_rows(dummy, 32770, dummy, 1, dummy, dummy, ctypes.byref(r), None) == _lib.ERR_INVALID_ARG
    assert L.fwx_certify_rows(dummy, 4, dummy, 1, dummy, dummy, None, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_batch_plan_run(None, 3) == _lib.ERR_INVALID_ARG
    assert L.fwx_batch_plan_persistent_info(None, None, None, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_batch_loop(_lib.ptr(sizes), _lib.ptr(offs), 0, dummy, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_diff_forward(dummy, dummy, None, 0, 1, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_diff_forward(dummy, dummy, None, 8193, 1, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_diff_forward(dummy, dummy, None, 8, 0, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_diff_backward(None, dummy, None, dummy, dummy, 8, 1, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_diff_backward(dummy, dummy, None, dummy, None, 8, 1, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_relaxation_form(None, 4, ctypes.byref(r), None) == _lib.ERR_INVALID_ARG
    assert L.fwx_relaxation_form(dummy, 4, None, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_event_create(None) == _lib.ERR_INVALID_ARG
    assert L.fwx_event_record(None, None) == _lib.ERR_INVALID_ARG
    assert L.fwx_event_synchronize(None) == _lib.ERR_INVALID_ARG
    assert L.fwx_event_elapsed_ms(None, dummy, dummy) == _lib.ERR_INVALID_ARG
    assert L.fwx_event_destroy(None) == _lib.OK
    # the allocation-free backward rejects sizes beyond its workspace before
    # touching the device
    assert L.fwx_diff_backward_noalloc(dummy, dummy, dummy, dummy, 8192, 300, None) == _lib.ERR_INVALID_ARG
    assert b"16 MiB workspace" in L.fwx_last_error()
    assert L.fwx_diff_backward_noalloc(None, dummy, dummy, dummy, 8, 1, None) == _lib.ERR_INVALID_ARG


def test_batch_layout_alignment():
    offs, total = fx.batch_layout([3, 5, 1, 64])
    assert list(offs) == [0, 12, 40, 44] and total == 44 + 4096
    assert all(o % 4 == 0 for o in offs)
    ns, dens, seeds = fx.mixed_batch()
    assert len(ns) == 512 and ns.min() >= 64 and ns.max() <= 2048
    assert 0.05 <= dens.min() and dens.max() <= 0.9
