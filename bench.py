"""Benchmark of the plain Floyd-Warshall hot path on one MI355X.

A step is one in-place multi_layer_tiling_gpu solve of the n=8192 workload
(density 0.9, weights U[1,1000], seed 42), preceded by a device-to-device
restore of the input matrix (both resident in HBM).  The timed steps run
with no profiler markers; the same number of steps then runs again with HIP
events around every launch group, which time the dominant kernel (the rest
pass).  The JSON line reports n^3/s (relaxations per second) as a percentage
of the metric's 2-op VALU min-plus peak, the rest kernel against the VALU
instruction ceiling of the relaxation form it runs in (its roofline), and the
reference's own CPU fwBaselineCPU timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192] [--kappa 0] [--tile 0]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "n³/s at n=8192 int32 (% VALU min-plus peak); batched graphs/s; fwd+bwd ms"

# MI355X VALU min-plus peak of the metric: 256 CUs x 64 int32 lanes/clk x
# 2.4 GHz / 2 ops (add + min) per relaxation.
CUS, LANES, CLK = 256, 64, 2.4e9
PEAK_2OP = CUS * LANES * CLK / 2.0


def instruction_ceiling(cus, form):
    """Relaxations/s the dominant kernel's form can issue at the published
    clock: pk16 is half a VALU lane-op per relaxation (one 64-bit add and one
    v_pk_minimum3_f16 per 4 relaxations: CUs x 64 lanes x 2.4 GHz x 2 =
    7.86e13 at 256 CUs), pk64 one lane-op (3.93e13), min3 1.5 (2.62e13)."""
    return cus * LANES * CLK * {2: 2.0, 1: 1.0}.get(form, 2.0 / 3.0)


FORM_NAMES = {0: "min3", 1: "pk64", 2: "pk16"}
FORM_BASIS = {0: "1.5 VALU lane-ops per relaxation (x 2/3)", 1: "one VALU lane-op per relaxation",
              2: "half a VALU lane-op per relaxation (x 2)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--algo", default="multi_layer_tiling_gpu")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--kappa", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--cpu-n-warm", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--no-n32768", action="store_true")
    ap.add_argument("--no-n1024", action="store_true")
    return ap.parse_args()


def bench_batched(steps):
    """512 graphs, n log-uniform in [64, 2048] (seed 11), density U[0.05, 0.9],
    weights U[1, 1000]: graphs/s of one hipGraph replay of the merged step
    launches, of the one persistent cooperative launch (from its hipGraph and
    direct), and of the per-graph launch loop (one plain solve per graph)."""
    import numpy as np
    import torch
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib
    ns, dens, seeds = fx.mixed_batch()
    offs, total = fx.batch_layout(ns)
    flat = np.zeros(total, dtype=np.int32)
    for n, p, sd, o in zip(ns, dens, seeds, offs):
        flat[o:o + int(n) * int(n)] = fx.generate_graph(int(n), float(p), int(sd), 1, 1000).ravel()
    src = torch.from_numpy(flat).cuda()
    work = torch.empty_like(src)
    plan = fx.BatchPlan(ns, offs, _lib.ptr(work))

    def timed(mode):
        work.copy_(src)
        plan.run(mode)  # capture + first run
        torch.cuda.synchronize()
        t = 0.0
        for _ in range(steps):
            work.copy_(src)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.run(mode)
            t += time.perf_counter() - t0
        return t / steps, work.clone()

    steps_s, steps_out = timed("graph")
    pers = {}
    for mode in ("persistent_graph", "persistent"):
        p_s, p_out = timed(mode)
        pers[mode] = {"graphs_per_s": len(ns) / p_s, "ms_per_batch": p_s * 1e3,
                      "equals_step_replay": bool(torch.equal(p_out, steps_out))}
        del p_out
    work.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fx.batch_loop(ns, offs, _lib.ptr(work))
    loop_s = time.perf_counter() - t0
    same = bool(torch.equal(steps_out, work))
    st = plan.stats()
    pinfo = plan.persistent_info()
    plan.destroy()
    relax = float((ns.astype(np.float64) ** 3).sum())
    pers["kernel"] = ("one cooperative launch of k_batch_coop: %d resident 512-thread blocks, %d grid-wide "
                      "steps, %d tasks" % (pinfo["grid"], pinfo["steps"], pinfo["tasks"]))
    return {"graphs": int(len(ns)), "graphs_per_s": len(ns) / steps_s, "ms_per_batch": steps_s * 1e3,
            "mode": "hipGraph replay of the merged step launches (the fastest mode)",
            "persistent": pers,
            "loop_graphs_per_s": len(ns) / loop_s, "loop_ms_per_batch": loop_s * 1e3,
            "speedup_vs_loop": loop_s / steps_s, "n3_per_s": relax / steps_s,
            "plan": st, "replay_equals_loop": same}


def bench_pipeline(n):
    """Input pipeline at the workload's n: the reference-compatible host
    generator (std::mt19937, one thread) vs the Philox device generator, and
    a pageable vs a pinned host-to-device copy of the n x n matrix."""
    import ctypes
    import numpy as np
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib
    L = _lib.lib()
    out = {"n": n}
    t0 = time.perf_counter()
    h = fx.generate_graph(n, 0.9, 42, 1, 1000)
    out["host_generator_s"] = time.perf_counter() - t0
    with fx.DeviceMatrix(n) as dm:
        fx.generate_graph_device(dm.ptr, n, 0.9, 42)  # warm
        t0 = time.perf_counter()
        fx.generate_graph_device(dm.ptr, n, 0.9, 42)
        out["device_generator_s"] = time.perf_counter() - t0
        nbytes = 4.0 * n * n
        t0 = time.perf_counter()
        _lib.check(L.fwx_copy_to_device(dm.ptr, _lib.ptr(h), n))
        dt = time.perf_counter() - t0
        out["pageable_h2d_GBps"] = nbytes / dt / 1e9
        hp = ctypes.c_void_p()
        _lib.check(L.fwx_host_alloc_pinned(ctypes.byref(hp), n))
        try:
            buf = np.ctypeslib.as_array(ctypes.cast(hp, ctypes.POINTER(ctypes.c_int32)), shape=(n, n))
            buf[:] = h
            _lib.check(L.fwx_copy_to_device_async(dm.ptr, hp, n, None))
            _lib.check(L.fwx_stream_synchronize(None))
            t0 = time.perf_counter()
            _lib.check(L.fwx_copy_to_device_async(dm.ptr, hp, n, None))
            _lib.check(L.fwx_stream_synchronize(None))
            dt = time.perf_counter() - t0
            out["pinned_h2d_GBps"] = nbytes / dt / 1e9
        finally:
            _lib.check(L.fwx_host_free_pinned(hp))
    out["generator_speedup"] = out["host_generator_s"] / out["device_generator_s"]
    # predecessor matrix of the workload (one extra n^3 pass after the solve)
    with fx.DeviceMatrix(n) as dw, fx.DeviceMatrix(n) as dd, fx.DeviceMatrix(n) as dp:
        dw.upload(h)
        dd.upload(h)
        fx.run_device("multi_layer_tiling_gpu", dd.ptr, n, 0, 0)
        _lib.check(L.fwx_predecessors(dw.ptr, dd.ptr, dp.ptr, n, None))  # warm
        t0 = time.perf_counter()
        _lib.check(L.fwx_predecessors(dw.ptr, dd.ptr, dp.ptr, n, None))
        dt = time.perf_counter() - t0
        out["predecessors_s"] = dt
        out["predecessors_n3_per_s"] = float(n) ** 3 / dt
    return out


def bench_n1024(reps=20):
    """n=1024, density 0.9, weights U[1,1000], seed 42 (the reference-compatible
    host generator), for each GPU algorithm: per-call time with HIP events on
    the launch stream, mean and min of `reps` calls each on a fresh device copy
    of the input (a call returns after its stream is idle, as the reference's
    entry points do); the reference's protocol (src/timer.cu:43-55: one warmup
    call, then one event-timed call on the already-solved matrix); and every
    row of the result certified on the device against the input."""
    import torch
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib
    n = 1024
    dev = torch.device("cuda:0")
    src = torch.from_numpy(fx.generate_graph(n, 0.9, 42, 1, 1000)).to(dev)
    work = torch.empty_like(src)
    st = torch.cuda.current_stream(dev)
    sp = st.cuda_stream
    srcv = torch.arange(n, dtype=torch.int32, device=dev)
    okv = torch.zeros(n, dtype=torch.int32, device=dev)
    out = {"n": n}
    for algo in ("multi_layer_tiling_gpu", "tiling_gpu", "baseline_gpu"):
        work.copy_(src)
        fx.run_device(algo, _lib.ptr(work), n, 0, 0, sp)  # first call: plan and graph capture
        ms = []
        for _ in range(reps):
            work.copy_(src)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            fx.run_device(algo, _lib.ptr(work), n, 0, 0, sp)
            e1.record(st)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        failed = fx.certify_rows_device(_lib.ptr(src), n, _lib.ptr(srcv), n, _lib.ptr(work), _lib.ptr(okv), sp)
        work.copy_(src)
        fx.run_device(algo, _lib.ptr(work), n, 0, 0, sp)  # the protocol's warmup call solves the matrix
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fx.run_device(algo, _lib.ptr(work), n, 0, 0, sp)
        e1.record(st)
        e1.synchronize()
        mean = sum(ms) / len(ms)
        out[algo] = {"ms_mean": mean, "ms_min": min(ms), "n3_per_s": float(n) ** 3 / (mean * 1e-3),
                     "reference_protocol_ms": e0.elapsed_time(e1), "rows_certified": n, "rows_failed": failed}
    tile, kappa = fx.tuned_config(n)
    out["multi_layer_tiling_gpu"].update({"tile": tile, "kappa": kappa})
    return out


def n32768_input():
    """The n=32768 workload drawn by the reference-compatible host generator
    (one thread, ~7-9 s; run beside the other GPU legs) and its time."""
    import fw_cuda_amd as fx
    t0 = time.perf_counter()
    h = fx.generate_graph(32768, 0.1, 7, 1, 1000)
    return h, time.perf_counter() - t0


def bench_n32768(h, gen_s):
    """n=32768 (4 GiB int32), density 0.1, weights U[1,1000], seed 7 (the
    reference-compatible host generator): one multi_layer_tiling_gpu and one
    tiling_gpu solve (tuned configurations), n^3/s each, both results equal,
    and every row of the multi-layer result certified on the device against
    the input."""
    import torch
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib
    n = 32768
    dev = torch.device("cuda:0")
    w = torch.from_numpy(h).to(dev)
    del h
    stream = torch.cuda.current_stream(dev).cuda_stream
    out = {"n": n, "host_generator_s": gen_s}
    res = {}
    for name in ("multi_layer_tiling_gpu", "tiling_gpu"):
        d = w.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fx.run_device(name, _lib.ptr(d), n, 0, 0, stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        tile, kappa = fx.tuned_config(n)
        out[name] = {"seconds": dt, "n3_per_s": float(n) ** 3 / dt, "tile": tile,
                     "kappa": kappa if name == "multi_layer_tiling_gpu" else 1}
        res[name] = d
    out["equal"] = bool(torch.equal(res["multi_layer_tiling_gpu"], res["tiling_gpu"]))
    del res["tiling_gpu"]
    srcv = torch.arange(n, dtype=torch.int32, device=dev)
    okv = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    failed = fx.certify_rows_device(_lib.ptr(w), n, _lib.ptr(srcv), n, _lib.ptr(res["multi_layer_tiling_gpu"]),
                                    _lib.ptr(okv), stream)
    out["certificate"] = {"rows_certified": n, "rows_failed": failed, "seconds": time.perf_counter() - t0}
    return out


def bench_diff(steps):
    """Differentiable fp32 FW, weights U[0.5, 10], density 0.3, seed 3:
    forward + backward of L = sum(G * D), batch of 64 graphs at n=512 and one
    graph at n=4096."""
    import numpy as np
    import torch
    from fw_cuda_amd.autograd import floyd_warshall_diff
    out = {}
    for name, (b, n) in {"b64_n512": (64, 512), "b1_n4096": (1, 4096)}.items():
        rng = np.random.default_rng(3)
        W = np.where(rng.random((b, n, n)) < 0.3, rng.uniform(0.5, 10.0, (b, n, n)), np.inf).astype(np.float32)
        G = torch.from_numpy(rng.standard_normal((b, n, n)).astype(np.float32)).cuda()
        Wt = torch.from_numpy(W).cuda()

        def once():
            w = Wt.clone().requires_grad_(True)
            D = floyd_warshall_diff(w)
            (D * G).sum().backward()
            return w.grad

        once()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            once()
        torch.cuda.synchronize()
        out[name] = {"fwd_bwd_ms": (time.perf_counter() - t0) * 1e3 / steps}
    return out


def bench_baseline(L, work, src, n, stream):
    """baseline_gpu (one HBM sweep per k) once on the workload, each per-k
    launch bracketed by HIP events: n^3/s and the sweep kernel's HBM rate for
    the 4 n^2 bytes every sweep must read (it writes only improved cells);
    also the rate under the 8 B/cell/k model that counts a write of every
    cell, which exceeds what the kernel moves."""
    import ctypes
    import torch
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib
    work.copy_(src)
    torch.cuda.synchronize()
    _lib.check(L.fwx_profile_begin())
    c0 = time.perf_counter()
    fx.run_device("baseline_gpu", _lib.ptr(work), n, 0, 0, stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - c0
    ms = (ctypes.c_double * _lib.PROF_NCLASS)()
    wk = (ctypes.c_double * _lib.PROF_NCLASS)()
    grp = (ctypes.c_int64 * _lib.PROF_NCLASS)()
    _lib.check(L.fwx_profile_end(ms, wk, grp))
    k_ms, k_bytes, k_n = ms[_lib.PROF_BASELINE], wk[_lib.PROF_BASELINE], grp[_lib.PROF_BASELINE]
    gbps = k_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    return {"n3_per_s": float(n) ** 3 / dt, "seconds": dt, "kernel": "k_baseline_step",
            "launches": int(k_n), "avg_launch_us": k_ms * 1e3 / max(1, k_n),
            "hbm_read_GBps": gbps, "frac_of_8TBps": gbps / 8000.0, "model_8B_per_cell_GBps": 2 * gbps}


def cpu_baseline(n, n_warm):
    """The reference's fwBaselineCPU (oracle/_ref, built from the reference's
    own sources) or, if that is absent, the oracle's restatement, on one core:
    one cold call over an n x n sample of the workload (the reported value),
    plus the reference's own protocol (src/timer.cu:16-33: one warmup call,
    then one timed call on the already-solved matrix) at n_warm, and both at
    n=1024 (the n=1024 workload's CPU figure)."""
    import numpy as np
    import fw_cuda_amd as fx
    from oracle import pyoracle as po
    kind = "reference" if po.ref_available() else "port"
    run = po.ref_fw_baseline_cpu if kind == "reference" else po.fw_serial

    def once(g):
        t0 = time.perf_counter()
        out = run(g)
        return time.perf_counter() - t0, out

    dt, _ = once(fx.generate_graph(n, 0.9, 42, 1, 1000))
    _, solved = once(fx.generate_graph(n_warm, 0.9, 42, 1, 1000))
    dt_warm, _ = once(solved)
    dt1, solved1 = once(fx.generate_graph(1024, 0.9, 42, 1, 1000))
    dt1_warm, _ = once(solved1)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": float(n) ** 3 / dt, "unit": "relaxations/s", "cores": 1, "kind": kind,
            "sample": f"one cold fwBaselineCPU call, n={n}, density 0.9, U[1,1000], seed 42, "
                      f"{dt:.2f} s on 1 thread of {cpu} (nproc={os.cpu_count()})",
            "reference_protocol": {"n": n_warm, "seconds": dt_warm,
                                   "value": float(n_warm) ** 3 / dt_warm,
                                   "note": "timed call on the matrix solved by the warmup call"},
            "n1024": {"n": 1024, "cold_seconds": dt1, "reference_protocol_seconds": dt1_warm,
                        "value": 1024.0 ** 3 / dt1}}


def traffic_from_profiles(n, tile, kappa, kernel):
    """HBM bytes per rest-kernel launch from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE runs (calibrated against a copy kernel), if they were taken
    on this configuration and kernel; else None."""
    p = os.path.join(ROOT, "profiles", "rest_kernel_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    if (t.get("n"), t.get("tile"), t.get("kappa")) != (n, tile, kappa) or \
            not t.get("kernel", "").startswith(kernel.split("<")[0] + "<"):
        return None
    return t.get("bytes_per_launch")


def rest_kernel_name(tile, kappa, form, packed=False):
    """The rest-pass kernel the library launches for this configuration: block
    R = 128 if the stage width kappa*tile is a multiple of 128, else 64, else
    32; R = 128 runs the pk16 kernel in the pk16 form (on packed stage panels
    when the solve packed them), else the 512-thread kernel."""
    span = tile * kappa
    R = 128 if span % 128 == 0 else 64 if span % 64 == 0 else 32
    if R == 128:
        if form == 2:
            return "k_rest16p<128>" if packed else "k_rest16<128>"
        return "k_rest512<128>"
    return f"k_rest<{R}>"


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fw_cuda_amd as fx
    from fw_cuda_amd import _lib

    n = a.n
    tile, kappa = fx.tuned_config(n)
    tile = a.tile or tile
    kappa = a.kappa or kappa
    dev = torch.device("cuda:0")
    h = fx.generate_graph(n, 0.9, 42, 1, 1000)
    src = torch.from_numpy(h).to(dev)
    work = torch.empty_like(src)
    stream = torch.cuda.current_stream(dev).cuda_stream
    L = _lib.lib()

    def step():
        work.copy_(src)
        fx.run_device(a.algo, _lib.ptr(work), n, tile, kappa, stream)

    for _ in range(a.warmup):
        step()
    # the relaxation form the solves ran in (fwx_last_solve_form) fixes the
    # dominant kernel's VALU instruction ceiling
    form, exact_finish = fx.last_solve_form() if a.warmup else (_lib.FORM_PK64, 0)
    packed = fx.last_solve_packed() if a.warmup else False
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    ceiling = instruction_ceiling(cus, form)
    # timed region: the steps alone, no profiler markers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    value = float(n) ** 3 / (ms_per_step * 1e-3)

    # the same number of steps again with HIP events bracketing every launch
    # group on the stream it runs on: the rest kernel's mean launch duration
    _lib.check(L.fwx_profile_begin())
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    p1 = time.perf_counter()
    ms = (ctypes.c_double * _lib.PROF_NCLASS)()
    wk = (ctypes.c_double * _lib.PROF_NCLASS)()
    grp = (ctypes.c_int64 * _lib.PROF_NCLASS)()
    _lib.check(L.fwx_profile_end(ms, wk, grp))
    ms_per_step_prof = (p1 - p0) * 1e3 / a.steps
    rest_ms, rest_work, rest_launches = ms[_lib.PROF_REST], wk[_lib.PROF_REST], grp[_lib.PROF_REST]
    lead_ms = ms[_lib.PROF_LEAD]
    band_ms = ms[_lib.PROF_BAND]
    achieved = rest_work / (rest_ms * 1e-3) if rest_ms > 0 else 0.0

    # exactness: every row of the last timed solve certified on the device
    # against the input (fwx_certify_rows, n^3 relaxations, outside the
    # timed region)
    srcv = torch.arange(n, dtype=torch.int32, device=dev)
    okv = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    failed = fx.certify_rows_device(_lib.ptr(src), n, _lib.ptr(srcv), n, _lib.ptr(work), _lib.ptr(okv), stream)
    cert_s = time.perf_counter() - c0
    check = {"rows_certified": n, "rows_failed": failed, "seconds": cert_s,
             "n3_per_s": float(n) ** 3 / cert_s,
             "verdict": "every row exact" if failed == 0 else "MISMATCH"}
    del srcv, okv

    # the same rest kernel with the CUs to itself: two profiled steps of the
    # single-stream schedule (outside the timed region)
    alone = None
    if a.algo == "multi_layer_tiling_gpu":
        prev = os.environ.get("FWX_LOOKAHEAD")
        os.environ["FWX_LOOKAHEAD"] = "0"
        try:
            step()
            torch.cuda.synchronize()
            _lib.check(L.fwx_profile_begin())
            for _ in range(2):
                step()
            torch.cuda.synchronize()
            ms0 = (ctypes.c_double * _lib.PROF_NCLASS)()
            wk0 = (ctypes.c_double * _lib.PROF_NCLASS)()
            grp0 = (ctypes.c_int64 * _lib.PROF_NCLASS)()
            _lib.check(L.fwx_profile_end(ms0, wk0, grp0))
        finally:
            if prev is None:
                os.environ.pop("FWX_LOOKAHEAD", None)
            else:
                os.environ["FWX_LOOKAHEAD"] = prev
        r_ms, r_wk, r_n = ms0[_lib.PROF_REST], wk0[_lib.PROF_REST], grp0[_lib.PROF_REST]
        if r_ms > 0:
            ach = r_wk / (r_ms * 1e-3)
            alone = {"schedule": "single stream (FWX_LOOKAHEAD=0), 2 steps outside the timed region",
                     "launches": int(r_n), "avg_launch_ms": r_ms / max(1, r_n), "achieved": ach,
                     "frac": ach / ceiling, "x_2op_peak": ach / PEAK_2OP}

    compare = {}
    if not a.no_compare:
        for name, algo, t, k in (("tiling_gpu_refshape", "tiling_gpu_refshape", 0, 0),
                                 ("tiling_gpu", "tiling_gpu", 0, 0)):
            work.copy_(src)
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            fx.run_device(algo, _lib.ptr(work), n, t, k, stream)
            torch.cuda.synchronize()
            dt = time.perf_counter() - c0
            compare[name] = {"n3_per_s": float(n) ** 3 / dt, "seconds": dt}
        compare["baseline_gpu"] = bench_baseline(L, work, src, n, stream)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "relaxations/s (n^3/s)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"plain FW, n={n}, density 0.9, int32 weights U[1,1000], seed 42; "
                               f"{a.algo} tile={tile} kappa={kappa}",
                   "n": n, "tile": tile, "kappa": kappa, "algorithm": a.algo},
        "pct_valu_minplus_peak": 100.0 * value / PEAK_2OP,
        "frac_of_instruction_ceiling": value / ceiling,
        "roofline": {"bound": "valu", "kernel": rest_kernel_name(tile, kappa, form, packed), "achieved": achieved,
                     "peak": ceiling, "unit": "relaxations/s", "frac": achieved / ceiling,
                     "form": FORM_NAMES.get(form, str(form)), "exact_finish": exact_finish,
                     "peak_basis": f"{FORM_NAMES.get(form, form)} form: {cus} CUs x 64 lanes x 2.4 GHz, "
                                   f"{FORM_BASIS.get(form, '')}; the metric's 2-op peak is "
                                   f"{PEAK_2OP:.4g}",
                     "x_2op_peak": achieved / PEAK_2OP,
                     "traffic": traffic_from_profiles(n, tile, kappa, rest_kernel_name(tile, kappa, form, packed)),
                     "measured_over": f"{a.steps} profiled steps after the timed region "
                                      f"({ms_per_step_prof:.3f} ms per step with the markers)",
                     "launches": int(rest_launches), "avg_launch_ms": rest_ms / max(1, rest_launches),
                     "work_per_step": rest_work / a.steps,
                     "share_of_step": rest_ms / a.steps / ms_per_step_prof,
                     "lead_ms_per_step": lead_ms / a.steps,
                     "band_ms_per_step": band_ms / a.steps,
                     "kernel_alone": alone},
        "parity": check,
        "comparators": compare,
    }
    # The legs bound by launch latency or host work run first, alone: the
    # differentiable leg, the n=1024 leg and the input pipeline (whose host
    # generator it times).  Then the two host-only legs (the CPU baseline on
    # one core, and the single-threaded host generator of the n=32768 input)
    # start on threads of their own beside the batched leg, which is bound by
    # GPU throughput; both spend their time inside one ctypes call, which
    # releases the interpreter lock.  (Run beside the host threads, the
    # differentiable leg measured 10 % slower.)
    if not a.no_diff:
        out["differentiable"] = bench_diff(3)
        print("bench: differentiable leg done", file=sys.stderr, flush=True)
    if not a.no_n1024:
        out["n1024"] = bench_n1024()
        print("bench: n=1024 leg done", file=sys.stderr, flush=True)
    if not a.no_pipeline:
        out["input_pipeline"] = bench_pipeline(n)
    import concurrent.futures as cf
    pool = cf.ThreadPoolExecutor(max_workers=2)
    cpu_job = None if a.no_cpu else pool.submit(cpu_baseline, a.cpu_n, a.cpu_n_warm)
    gen_job = None if a.no_n32768 else pool.submit(n32768_input)
    if not a.no_batched:
        out["batched"] = bench_batched(3)
        print("bench: batched leg done", file=sys.stderr, flush=True)
    if gen_job is not None:
        h32, gen_s = gen_job.result()
        out["n32768"] = bench_n32768(h32, gen_s)
        del h32
        print("bench: n=32768 leg done", file=sys.stderr, flush=True)
    if cpu_job is not None:
        out["cpu_baseline"] = cpu_job.result()
    pool.shutdown()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
