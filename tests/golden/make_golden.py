"""Writes the golden fixtures under tests/golden/.

kats.json   -- the reference's six known-answer tests, as data (inputs and
               expected matrices from src/tests.cpp:59-132 and their GPU twins
               in src/tests_gpu.cu:76-148).
gen_ref.json -- outputs of the reference's own generateGraph1D
               (src/graph_generator.cpp:48-80) and fwBaselineCPU
               (src/fw_baseline_cpu.cpp:3-16), built by oracle/Makefile into
               oracle/_ref/libfwref.so, for a few small (n, density, seed).

Run from the repo root after `make -C oracle`:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle  # noqa: E402

INF = pyoracle.INF
HERE = os.path.dirname(os.path.abspath(__file__))

KATS = [
    {"name": "Single node (1x1)", "n": 1, "input": [0], "expected": [0]},
    {"name": "No edges (2x2)", "n": 2, "input": [0, INF, INF, 0], "expected": [0, INF, INF, 0]},
    {"name": "Simple path optimization (4x4)", "n": 4,
     "input": [0, 1, INF, 100, INF, 0, 1, INF, INF, INF, 0, 1, INF, INF, INF, 0],
     "expected": [0, 1, 2, 3, INF, 0, 1, 2, INF, INF, 0, 1, INF, INF, INF, 0]},
    {"name": "Disconnected components (4x4)", "n": 4,
     "input": [0, 1, INF, INF, 1, 0, INF, INF, INF, INF, 0, 2, INF, INF, 2, 0],
     "expected": [0, 1, INF, INF, 1, 0, INF, INF, INF, INF, 0, 2, INF, INF, 2, 0]},
    {"name": "Negative weights (3x3)", "n": 3,
     "input": [0, 5, 10, 0, 0, -1, INF, INF, 0],
     "expected": [0, 5, 4, 0, 0, -1, INF, INF, 0]},
    # passes when any diagonal entry of the result is negative (src/tests.cpp:113-132)
    {"name": "Negative cycle detection (3x3)", "n": 3,
     "input": [0, 2, INF, INF, 0, 1, INF, -3, 0], "expected": None, "check": "negative_cycle"},
]

GEN_CASES = [(1, 0.5, 1), (5, 0.1, 7), (8, 0.5, 42), (16, 0.9, 123), (33, 0.3, 99), (64, 0.9, 42)]


def main():
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump({"INF": INF, "kats": KATS}, f, indent=1)
    if not pyoracle.ref_available():
        print("oracle/_ref not built: gen_ref.json not regenerated")
        return
    cases = []
    for n, p, seed in GEN_CASES:
        g = pyoracle.ref_generate_graph(n, p, seed)
        d = pyoracle.ref_fw_baseline_cpu(g)
        cases.append({"n": n, "density": p, "seed": seed,
                      "graph": g.ravel().tolist(), "fw": d.ravel().tolist()})
    with open(os.path.join(HERE, "gen_ref.json"), "w") as f:
        json.dump({"source": "reference generateGraph1D + fwBaselineCPU (libstdc++ 11.4)",
                   "cases": cases}, f)
    print(f"wrote {len(KATS)} KATs and {len(cases)} generator cases")


if __name__ == "__main__":
    main()
