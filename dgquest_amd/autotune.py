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

With --cost it also fits, per precision and shape, the device-time model
that searchQuregTaskPartitions uses to pick a shape for each recorded
circuit (tune.hpp): for several circuit families and sizes, lanes rotating
and fixed, the flush of the circuit is timed and its plan's features are
read from a recording register (per block: bytes moved, phases after the
first, device ops; the launch count), and the four coefficients are fitted by
non-negative least squares.  The fit replaces the reference's sampled
per-gate model (src/cpp/time_predictor.cpp:127-338, sampler/main.cpp) with
per-sweep terms measured on the device.
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


# shapes of the cost model per precision (the FP32-only 14-qubit tile included)
COST_SHAPES = {2: [(3, 3, 0), (4, 4, 0), (5, 4, 1), (6, 4, 2), (6, 3, 3), (7, 4, 3)],
               1: [(4, 4, 0), (5, 4, 1), (6, 4, 2), (6, 3, 3), (7, 4, 3), (7, 5, 2), (8, 5, 3)]}


def cost_circuits(n):
    """Circuit families of the cost fit: the benchmark's layered U3+CZ (two
    depths), a hardware-efficient ansatz, GHZ + QFT and, when n is a multiple
    of 4, a 4 x n/4 grid circuit."""
    from . import circuits

    out = [("layered24", circuits.layered_u3_cz(n, 24, 2)), ("layered8", circuits.layered_u3_cz(n, 8, 1)),
           ("hea10", circuits.hea(n, 10, 3)), ("ghz_qft", circuits.ghz_qft(n))]
    if n % 4 == 0:
        out.append(("grid12", circuits.grid_circuit(4, n // 4, 12, 4)))
    return out


def plan_features(quest, n, circ, shape, rot):
    """[GB moved, GB x phases after the first, GB x device ops, launches] of
    the plan a flush of `circ` from |0..0> makes with this shape and lane
    mode, summed over its blocks (read from a recording register)."""
    import numpy as np

    k, R, W = shape
    r = quest.lib.dgq_create_recorder(n)
    try:
        quest.dgq_set_lane_rotation(r, rot)
        quest.apply_circuit(r, circ)
        b, p, _, _ = quest.recorder_plan(r, k, R, W, 4096)
    finally:
        quest.lib.dgq_destroy_recorder(r)
    amp = 16 if quest.prec == 2 else 8
    f = np.zeros(4)
    for blk in b:
        gb = (2 if blk["init_kind"] == 0 else 1) * amp * float(1 << n) * 1e-9
        ops = int(p["op_count"][blk["phase_begin"]:blk["phase_begin"] + blk["phase_count"]].sum())
        f += [gb, gb * (int(blk["phase_count"]) - 1), gb * ops, 1.0]
    return f


def measure_flush(quest, eng, env, n, shape, rot, circ, reps):
    """Least GPU ms of the flush of `circ` (events on the register's stream)
    over `reps` runs after a warm one; the plan is made once and reused."""
    k, R, W = shape
    q = quest.createQureg(n, env)
    try:
        quest.dgq_set_tuning(q, k, R, W, 4096)
        quest.dgq_set_lane_rotation(q, rot)
        stream = quest.dgq_stream(q)
        a, b = eng.event(), eng.event()
        best = None
        for i in range(reps + 1):
            quest.initZeroState(q)
            quest.apply_circuit(q, circ)
            quest.dgq_sync(q)
            eng.record(a, stream)
            quest.dgq_flush(q)
            eng.record(b, stream)
            ms = eng.elapsed_ms(a, b)
            if i:
                best = ms if best is None else min(best, ms)
    finally:
        quest.destroyQureg(q, env)
    return best


def fit_costs(sizes, precs=(2, 1), reps=2, log=sys.stderr):
    """Per (prec, shape): the non-negative least-squares fit of measured flush
    times on plan features, and its largest relative error."""
    import numpy as np
    from scipy.optimize import nnls

    from . import quest as Q
    from .engine import Engine

    eng = Engine()
    out = []
    for prec in precs:
        quest = Q.load(prec)
        env = quest.createQuESTEnv()
        for shape in COST_SHAPES[prec]:
            A, y, tags = [], [], []
            for n in sizes:
                if 6 + shape[1] + shape[2] > n:
                    continue
                for name, circ in cost_circuits(n):
                    for rot in (1, 0):
                        f = plan_features(quest, n, circ, shape, rot)
                        ms = measure_flush(quest, eng, env, n, shape, rot, circ, reps)
                        A.append(f)
                        y.append(ms)
                        tags.append(f"{name}/n{n}/rot{rot}")
            if not A:
                continue
            A, y = np.array(A), np.array(y)
            # relative least squares: weight each row by 1 / measured time
            x, _ = nnls(A / y[:, None], np.ones(len(y)))
            pred = A @ x
            rel = np.abs(pred - y) / y
            out.append((prec, shape, x, float(rel.max()), float(np.median(rel)), len(y)))
            print(f"cost prec {prec} k{shape[0]}R{shape[1]}W{shape[2]}: sweep {x[0]:.5f} phase {x[1]:.5f} "
                  f"op {x[2]:.3e} launch {x[3]:.4f}; rel err median {np.median(rel):.3f} max {rel.max():.3f} "
                  f"({tags[int(rel.argmax())]})", file=log, flush=True)
    return out


def write(rows, path, costs=()):
    with open(path, "w") as f:
        f.write("# fused-sweep shape per precision and register size, measured by dgquest_amd/autotune.py\n")
        f.write(f"# on {time.strftime('%Y-%m-%d')}: layered U3+CZ depth 8, best of 2 after a warm run\n")
        f.write("# prec n k R W max_ops ms_of_tuning_circuit\n")
        for prec, n, (k, R, W), ms in rows:
            f.write(f"{prec} {n} {k} {R} {W} 4096 {ms:.4f}\n")
        if costs:
            f.write("# device-time model per shape (tune.hpp): ms of a sweep = launch + GB moved x\n")
            f.write("# (sweep + phase x (phases - 1) + op x device ops); fitted on timed flushes of\n")
            f.write("# layered U3+CZ, HEA, GHZ+QFT and grid circuits, lanes rotating and fixed\n")
            f.write("# cost prec k R W sweep_ms_per_GB phase_ms_per_GB op_ms_per_GB launch_ms  # fit: rel err\n")
            for prec, (k, R, W), x, rmax, rmed, m in costs:
                f.write(f"cost {prec} {k} {R} {W} {x[0]:.6f} {x[1]:.6f} {x[2]:.4e} {x[3]:.5f}"
                        f"  # {m} flushes, median {rmed:.3f}, max {rmax:.3f}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="tune_table.txt")
    ap.add_argument("--sizes", default="12,14,16,18,20,22,24,26,28,30")
    ap.add_argument("--precs", default="2,1")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--cost", action="store_true", help="also fit the per-shape device-time model")
    ap.add_argument("--cost-sizes", default="24,28,30")
    a = ap.parse_args()
    precs = tuple(int(x) for x in a.precs.split(","))
    rows = tune([int(x) for x in a.sizes.split(",")], precs, reps=a.reps)
    costs = fit_costs([int(x) for x in a.cost_sizes.split(",")], precs, reps=a.reps) if a.cost else ()
    write(rows, a.out, costs)
    print(f"wrote {len(rows)} entries and {len(costs)} cost models to {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
