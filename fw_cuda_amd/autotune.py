"""Autotuner for (tile, kappa) per size class of the multi-layer engine.

For every size class it times the candidate configurations of
fwx_multi_layer_tiling_gpu on a representative graph (density 0.9, weights
U[1,1000]) on the current GPU and keeps the fastest.  The rest-pass block
size R (and with it the register micro-tile: R=128 -> 8x4 cells per thread
on 512 threads, 64 -> 4x4, 32 -> 2x2) follows from the stage width kappa*tile, so choosing
(tile, kappa) chooses the micro-tile too.

Writes
  fw_cuda_amd/csrc/fwx_tune_table.inc   -- the table compiled into libfwx
  profiles/autotune_mi355x.json          -- every timing behind it

The last size class tuned also covers every larger n up to MAX_N = 32769
(include/constants.h:4).

Run on the GPU, then rebuild:  python -m fw_cuda_amd.autotune [--max-n 16384]
"""
import argparse
import json
import os
import time

import numpy as np

from . import DeviceMatrix, generate_graph, run_device

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CLASSES = [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]
MAX_N = 32769
TILES = (32, 64, 128)
KAPPAS = (1, 2, 3, 4, 6, 8, 12, 16)


def candidates(n):
    out = []
    for t in TILES:
        N = (n + t - 1) // t
        for k in KAPPAS:
            if k > max(1, N):
                continue
            if t == 32 and k * t < 64 and n > 256:
                continue  # 32-wide rest blocks only pay off on tiny graphs
            out.append((t, k))
    return out


def time_config(dm, g, n, t, k, reps):
    best = float("inf")
    for _ in range(reps):
        dm.upload(g)
        t0 = time.perf_counter()
        run_device("multi_layer_tiling_gpu", dm.ptr, n, t, k)
        best = min(best, time.perf_counter() - t0)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-n", type=int, default=16384)
    ap.add_argument("--out-inc", default=os.path.join(HERE, "csrc", "fwx_tune_table.inc"))
    ap.add_argument("--out-json", default=os.path.join(ROOT, "profiles", "autotune_mi355x.json"))
    a = ap.parse_args()
    table, record = [], {}
    for n in CLASSES:
        if n > a.max_n:
            break
        g = generate_graph(n, 0.9, 42, 1, 1000)
        reps = 5 if n <= 2048 else 2
        res = {}
        with DeviceMatrix(n) as dm:
            dm.upload(g)
            run_device("multi_layer_tiling_gpu", dm.ptr, n, 64, 1)  # warm
            for t, k in candidates(n):
                res[f"{t}x{k}"] = time_config(dm, g, n, t, k, reps)
                print(f"n={n} tile={t} kappa={k}: {res[f'{t}x{k}'] * 1e3:.3f} ms", flush=True)
        best = min(res, key=res.get)
        t, k = (int(x) for x in best.split("x"))
        table.append((n, t, k))
        record[str(n)] = {"best": best, "seconds": res, "n3_per_s": n ** 3 / res[best]}
        print(f"n<={n}: tile={t} kappa={k} ({n ** 3 / res[best]:.3e} n^3/s)", flush=True)
    if table and table[-1][0] < MAX_N:
        table.append((MAX_N, table[-1][1], table[-1][2]))
    with open(a.out_inc, "w") as f:
        f.write("// Generated by fw_cuda_amd/autotune.py on the target GPU: {n_max, tile, kappa}.\n")
        for n, t, k in table:
            f.write(f"{{{n}, {t}, {k}}},\n")
    os.makedirs(os.path.dirname(a.out_json), exist_ok=True)
    with open(a.out_json, "w") as f:
        json.dump(record, f, indent=1)


if __name__ == "__main__":
    main()
