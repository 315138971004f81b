"""Measure the fused-sweep shape per precision and register size on the GPU
and write the tuning table that new registers and the request server read
(``tune_table.txt``, format in dgquest_amd/csrc/host/tune.hpp).

For every size n and precision, each candidate shape (k, R, W) plans and runs
a layered U3 + CZ circuit of depth 8 (the gate mix of the benchmark
workloads) through the QuEST API; the GPU time of its sweeps is taken with
HIP events on the register's stream (best of `reps` after one warm run) and
the fastest shape is kept.  This replaces the reference's offline sampler and
closed-form time model (sampler/main.cpp, src/cpp/time_predictor.cpp:127-395):
the device itself is the model.

    python -m dgquest_amd.autotune --out tune_table.txt
"""
from __future__ import annotations

import argparse
import sys
import time

# (k, R, W): tile = 6 + R + W qubits, k = target qubits above the six coalesced
# ones per sweep; only shapes with built kernels (engine.hip SweepTable)
CANDIDATES = [(3, 3, 0), (4, 4, 0), (5, 5, 0), (5, 4, 1), (6, 5, 1), (6, 4, 2), (7, 5, 2), (6, 3, 3),
              (7, 4, 3), (8, 5, 3)]
FP32_ONLY = {(8, 5, 3)}


def measure(quest, eng, env, n, shape, circ, reps):
    k, R, W = shape
    q = quest.createQureg(n, env)
    quest.dgq_set_tuning(q, k, R, W, 4096)
    stream = quest.dgq_stream(q)
    a, b = eng.event(), eng.event()
    best = None
    for i in range(reps + 1):
        quest.initZeroState(q)
        quest.apply_circuit(q, circ)
        eng.record(a, stream)
        quest.dgq_flush(q)
        eng.record(b, stream)
        ms = eng.elapsed_ms(a, b)
        if i:
            best = ms if best is None else min(best, ms)
    sweeps = quest.plan_stats(q)["sweeps"]
    quest.destroyQureg(q, env)
    return best, sweeps


def tune(sizes, precs=(2, 1), depth=8, reps=2, log=sys.stderr):
    from . import circuits
    from . import quest as Q
    from .engine import Engine

    eng = Engine()
    rows = []
    for prec in precs:
        quest = Q.load(prec)
        env = quest.createQuESTEnv()
        for n in sizes:
            circ = circuits.layered_u3_cz(n, depth, 1)
            res = {}
            for shape in CANDIDATES:
                if 6 + shape[1] + shape[2] > n or (prec == 2 and shape in FP32_ONLY):
                    continue
                res[shape] = measure(quest, eng, env, n, shape, circ, reps)
            if not res:
                continue
            best = min(res, key=lambda s: res[s][0])
            rows.append((prec, n, best, res[best][0]))
            print(f"prec {prec} n {n}: " + ", ".join(f"k{s[0]}R{s[1]}W{s[2]}={v[0]:.3f}ms/{v[1]}sw"
                                                     for s, v in res.items()) + f" -> {best}", file=log, flush=True)
    return rows


def write(rows, path):
    with open(path, "w") as f:
        f.write("# fused-sweep shape per precision and register size, measured by dgquest_amd/autotune.py\n")
        f.write(f"# on {time.strftime('%Y-%m-%d')}: layered U3+CZ depth 8, best of 2 after a warm run\n")
        f.write("# prec n k R W max_ops ms_of_tuning_circuit\n")
        for prec, n, (k, R, W), ms in rows:
            f.write(f"{prec} {n} {k} {R} {W} 4096 {ms:.4f}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="tune_table.txt")
    ap.add_argument("--sizes", default="12,14,16,18,20,22,24,26,28,30")
    ap.add_argument("--precs", default="2,1")
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args()
    rows = tune([int(x) for x in a.sizes.split(",")], tuple(int(x) for x in a.precs.split(",")), reps=a.reps)
    write(rows, a.out)
    print(f"wrote {len(rows)} entries to {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
