"""Autotuner of the engine's per-size-class configurations, on the target GPU.

Plain solves: for every size class it times fwx_multi_layer_tiling_gpu on a
representative graph (density 0.9, weights U[1,1000]) over the candidate
(tile, kappa) pairs and, where the stage width kappa*tile makes the rest pass
128 wide (those passes run in the pk16 form on such graphs), over the chunk
depth of the pk16 rest kernel (16 or 32, FWX_P16_KC), and keeps the
fastest.  The shape of the pk64 128-wide rest kernel, which runs when the
pk16 form does not apply (FWX_REST_VARIANT: 512 threads of 8x4 cells, 256 of
8x8, 256 of 8x8 on 16-deep chunks), stays at the 512-thread shape that won
every class when it was searched, before the pk16 form existed.  (64- and
32-wide rest blocks have one shape each: 4x4 and 2x2 cells on 256 threads.)

Batched plans: per size class of a graph inside a batch, (tile, kappa) and
the threads of the 128-wide products (512: 8x4 cells, 256: 8x8), chosen by
timing whole batches of the 512-graph mix (n log-uniform in [64, 2048], seed
11) with the step launches replayed from a hipGraph: coordinate descent over
the classes from the current table, two passes, each candidate table set
through FWX_BATCH_TABLE.

Writes (into --out-dir, default the source tree)
  fwx_tune_table.inc    -- {n_max, tile, kappa, rest shape, pk16 chunk, tiling tile, schedule} compiled into libfwx
  fwx_batch_table.inc   -- {n_max, tile, kappa, product threads}
  autotune_mi355x.json  -- every timing behind both

The last plain size class tuned also covers every larger n up to MAX_N =
32769 (include/constants.h:4).

Run on the GPU, copy the tables into fw_cuda_amd/csrc/, then rebuild:
    python -m fw_cuda_amd.autotune [--max-n 16384] [--what plain,batched] [--out-dir DIR]
"""
import argparse
import json
import os
import time

import numpy as np

from . import (BatchPlan, batch_layout, generate_graph, mixed_batch, run_device)
from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CLASSES = [64, 128, 256, 512, 1024, 1536, 2048, 4096, 8192, 16384]
MAX_N = 32769
TILES = (32, 64, 128)
KAPPAS = (1, 2, 3, 4, 6, 8, 12, 16)
REST_SHAPE = 2  # FWX_REST_VARIANT of a pk64 128-wide rest pass (512 threads, 8x4 cells)
P16_CHUNKS = (16, 32)  # FWX_P16_KC of a pk16 rest pass
SCHEDULES = (2, 3)  # FWX_LOOKAHEAD: side-band, side-band a stage ahead

# batched size classes (n_max) and their candidates
BATCH_CLASSES = [64, 127, 255, 511, 1023, MAX_N]
BATCH_CANDIDATES = [(32, 2), (32, 4), (64, 1), (64, 2), (64, 4), (128, 1), (128, 2), (128, 3), (128, 4)]


def rest_block(tile, kappa):
    span = tile * kappa
    return 128 if span % 128 == 0 else 64 if span % 64 == 0 else 32


def candidates(n):
    out = []
    for t in TILES:
        N = (n + t - 1) // t
        for k in KAPPAS:
            if k > max(1, N):
                continue
            if t == 32 and k == 1 and n > 4096:
                continue  # (32, 1) runs the fused small-n rounds, up to n = 4096
            if t == 32 and k * t < 64 and k > 1 and n > 256:
                continue  # 32-wide rest blocks only pay off on tiny graphs
            for kc in (P16_CHUNKS if rest_block(t, k) == 128 and N > k else (0,)):
                for la in (SCHEDULES if N > k and not (t == 32 and k == 1) else (2,)):
                    out.append((t, k, kc, la))
    return out


def time_config(src, work, n, t, k, reps):
    """Best of `reps` solves, each on a device-to-device restore of the
    input, after one untimed solve: back-to-back GPU work as in a running
    application.  (Restoring from the host instead leaves the GPU idle for
    the length of a PCIe copy before every timed solve; at n = 8192 that
    measured every configuration 20-35 % slower and in a different order.)"""
    import torch
    from . import _lib
    best = float("inf")
    for r in range(reps + 1):
        work.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_device("multi_layer_tiling_gpu", _lib.ptr(work), n, t, k)
        if r:
            best = min(best, time.perf_counter() - t0)
    return best


def tune_class(n):
    """Every candidate of size class n, in this process: {key: seconds}."""
    import torch
    src = torch.from_numpy(generate_graph(n, 0.9, 42, 1, 1000)).cuda()
    work = torch.empty_like(src)
    reps = 5 if n <= 2048 else 3
    res = {}
    saved = {v: os.environ.get(v) for v in ("FWX_REST_VARIANT", "FWX_P16_KC", "FWX_LOOKAHEAD")}
    try:
        os.environ["FWX_REST_VARIANT"] = str(REST_SHAPE)
        for t, k, kc, la in candidates(n):
            if kc:
                os.environ["FWX_P16_KC"] = str(kc)
            os.environ["FWX_LOOKAHEAD"] = str(la)
            key = f"{t}x{k}/kc{kc}/la{la}"
            res[key] = time_config(src, work, n, t, k, reps)
            print(f"n={n} tile={t} kappa={k} pk16 chunk={kc} schedule={la}: {res[key] * 1e3:.3f} ms", flush=True)
    finally:
        for v, val in saved.items():
            if val is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = val
    return res


def tune_plain(max_n):
    """Each size class in a process of its own.  In one process that had
    already captured and replayed the smaller classes' solves as hipGraphs,
    the n = 8192 solves with 128-wide tiles measured 24 ms instead of the
    17.7 ms of a fresh process, so a class is never timed after another
    class's graphs."""
    import subprocess
    import sys
    import tempfile
    table, record = [], {}
    for n in CLASSES:
        if n > max_n:
            break
        with tempfile.NamedTemporaryFile(suffix=".json") as tf:
            subprocess.run([sys.executable, "-u", "-m", "fw_cuda_amd.autotune", "--one", str(n), "--one-out",
                            tf.name], check=True, cwd=ROOT)
            with open(tf.name) as f:
                res = json.load(f)
        best = min(res, key=res.get)
        tk, kc, la = best.split("/")
        t, k = (int(x) for x in tk.split("x"))
        kc, la = int(kc[2:]), int(la[2:])
        # tiling_gpu (kappa 1) with tile 0: the fastest kappa-1 candidate's tile
        best1 = min((key for key in res if key.split("/")[0].endswith("x1")), key=res.get)
        t1 = int(best1.split("x")[0])
        table.append((n, t, k, REST_SHAPE, kc, t1, la))
        record[str(n)] = {"best": best, "seconds": res, "n3_per_s": n ** 3 / res[best]}
        print(f"n<={n}: tile={t} kappa={k} pk16 chunk={kc} schedule={la} ({n ** 3 / res[best]:.3e} n^3/s)",
              flush=True)
    if table and table[-1][0] < MAX_N:
        table.append((MAX_N,) + table[-1][1:])
    return table, record


def table_env(tab):
    return ",".join(f"{n}:{t}:{k}:{r}" for n, t, k, r in tab)


def tune_batched(start, passes=2, reps=3):
    """Coordinate descent over BATCH_CLASSES on the 512-graph mix: for each
    class in turn, every candidate (tile, kappa, threads) with the other
    classes fixed; keeps a change only if the batch gets faster."""
    import torch
    ns, dens, seeds = mixed_batch()
    offs, total = batch_layout(ns)
    flat = np.zeros(total, dtype=np.int32)
    for n, p, sd, o in zip(ns, dens, seeds, offs):
        flat[o:o + int(n) * int(n)] = generate_graph(int(n), float(p), int(sd), 1, 1000).ravel()
    src = torch.from_numpy(flat).cuda()
    work = torch.empty_like(src)
    ref = None
    saved = os.environ.get("FWX_BATCH_TABLE")
    record = {"classes": BATCH_CLASSES, "evaluations": []}

    def measure(tab):
        nonlocal ref
        os.environ["FWX_BATCH_TABLE"] = table_env(tab)
        plan = BatchPlan(ns, offs, _lib.ptr(work))
        try:
            work.copy_(src)
            plan.run("graph")  # capture
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(reps):
                work.copy_(src)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                plan.run("graph")
                best = min(best, time.perf_counter() - t0)
            if ref is None:
                ref = work.clone()
            elif not torch.equal(ref, work):
                raise AssertionError(f"batched table {table_env(tab)} changed the results")
            return best, plan.stats()
        finally:
            plan.destroy()

    try:
        cur = [tuple(e) for e in start]
        cur_s, st = measure(cur)
        record["evaluations"].append({"table": table_env(cur), "ms": cur_s * 1e3, "plan": st})
        print(f"batched start {table_env(cur)}: {cur_s * 1e3:.3f} ms", flush=True)
        for _ in range(passes):
            for ci, nmax in enumerate(BATCH_CLASSES):
                lo = BATCH_CLASSES[ci - 1] + 1 if ci else 1
                upper = nmax if nmax < MAX_N else 2048
                for t, k in BATCH_CANDIDATES:
                    if k > 1 and (k - 1) * t >= upper:
                        continue  # the stage would be wider than every graph of the class
                    for thr in ((512, 256) if rest_block(t, k) == 128 and lo >= 128 else (512,)):
                        cand = list(cur)
                        cand[ci] = (nmax, t, k, thr)
                        if tuple(cand) == tuple(cur):
                            continue
                        s, st = measure(cand)
                        record["evaluations"].append({"table": table_env(cand), "ms": s * 1e3, "plan": st})
                        print(f"  class n<={nmax}: ({t}, {k}, {thr}) {s * 1e3:.3f} ms", flush=True)
                        if s < cur_s * 0.998:
                            cur, cur_s = cand, s
                print(f"class n<={nmax} -> {cur[ci]} ({cur_s * 1e3:.3f} ms)", flush=True)
    finally:
        if saved is None:
            os.environ.pop("FWX_BATCH_TABLE", None)
        else:
            os.environ["FWX_BATCH_TABLE"] = saved
    record["best"] = {"table": table_env(cur), "ms": cur_s * 1e3, "graphs_per_s": len(ns) / cur_s}
    return cur, record


def read_table(path):
    out = []
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            out.append(tuple(int(x) for x in line.strip("{},").split(",")))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-n", type=int, default=16384)
    ap.add_argument("--what", default="plain,batched")
    ap.add_argument("--out-dir", default=None)
    ap.add_argument("--one", type=int, default=0, help=argparse.SUPPRESS)  # one plain class (subprocess)
    ap.add_argument("--one-out", default=None, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.one:
        with open(a.one_out, "w") as f:
            json.dump(tune_class(a.one), f)
        return
    what = set(a.what.split(","))
    src_dir = os.path.join(HERE, "csrc")
    out_dir = a.out_dir or src_dir
    json_path = os.path.join(a.out_dir or os.path.join(ROOT, "profiles"), "autotune_mi355x.json")
    os.makedirs(out_dir, exist_ok=True)
    record = {}
    prev = os.path.join(ROOT, "profiles", "autotune_mi355x.json")
    for path in (prev, json_path):
        if os.path.exists(path):
            with open(path) as f:
                record = json.load(f)
    if "plain" in what:
        table, rec = tune_plain(a.max_n)
        # classes above --max-n keep their current entries and timings
        kept = [tuple(e[:5]) + (16,) * (5 - len(e)) + ((e[5] if len(e) > 5 else e[1]), (e[6] if len(e) > 6 else 2))
                for e in read_table(os.path.join(src_dir, "fwx_tune_table.inc")) if e[0] > a.max_n]
        if kept:
            table = [e for e in table if e[0] <= a.max_n] + kept
        record.setdefault("plain", {}).update(rec)
        with open(os.path.join(out_dir, "fwx_tune_table.inc"), "w") as f:
            f.write("// Generated by fw_cuda_amd/autotune.py on the target GPU: "
                    "{n_max, tile, kappa, rest shape, pk16 chunk, tiling tile, schedule}.\n")
            for e in table:
                f.write("{%d, %d, %d, %d, %d, %d, %d},\n" % tuple(e))
    if "batched" in what:
        start = read_table(os.path.join(src_dir, "fwx_batch_table.inc"))
        start = [next((e for e in start if e[0] >= c), start[-1]) for c in BATCH_CLASSES]
        start = [(c,) + tuple(e[1:]) for c, e in zip(BATCH_CLASSES, start)]
        table, rec = tune_batched(start)
        record["batched"] = rec
        with open(os.path.join(out_dir, "fwx_batch_table.inc"), "w") as f:
            f.write("// Generated by fw_cuda_amd/autotune.py on the target GPU: "
                    "{n_max, tile, kappa, threads of 128-wide products}.\n")
            for e in table:
                f.write("{%d, %d, %d, %d},\n" % e)
    with open(json_path, "w") as f:
        json.dump(record, f, indent=1)


if __name__ == "__main__":
    main()
