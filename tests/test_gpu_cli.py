"""The rebuilt reference CLIs on the GPU: tests_runner's own suite (six KATs +
eight seeded graphs against fwBaselineCPU, src/tests_gpu.cu:156-185) for each
GPU algorithm, and bench_runner's output lines."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fw_cuda_amd", "bin")


def run(*args):
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


@pytest.mark.parametrize("algo,tile,kappa", [("baseline_gpu", 0, 0), ("tiling_gpu", 0, 0),
                                             ("tiling_gpu", 32, 0), ("multi_layer_tiling_gpu", 0, 0),
                                             ("multi_layer_tiling_gpu", 32, 4),
                                             ("multi_layer_tiling_gpu", 64, 8)])
def test_tests_runner_gpu_suite(algo, tile, kappa):
    r = run("tests_runner", algo, tile, kappa)
    assert r.returncode == 0
    assert "RESULTS: 14/14 passed" in r.stdout
    assert "[ FAILED ]" not in r.stdout + r.stderr


def test_bench_runner_gpu_lines():
    r = run("bench_runner", "multi_layer_tiling_gpu", 1024, 0.9, 42, 0, 0, 1, 1000)
    assert r.returncode == 0
    for line in ("BENCHMARK: multi_layer_tiling_gpu", "Algorithm type: GPU", "Warming up GPU...",
                 "Copying graph to GPU...", "Running multi_layer_tiling_gpu...", "Execution time: "):
        assert line in r.stdout, line
    # the reference accepts any tile in [0, n] and ignores it; so must the drop-in
    r = run("bench_runner", "tiling_gpu", 300, 0.5, 7, 16)
    assert r.returncode == 0 and "Execution time: " in r.stdout
