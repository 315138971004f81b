#!/usr/bin/env python3
"""Benchmark of the plain fused-sweep method on one MI355X.

Workload: the 30-qubit layered random circuit in complex FP64 (16 GiB state),
depth 24 -- per layer U3(theta, phi, lambda) on every qubit with angles
uniform on [0, 2pi), then CZ on a seeded perfect matching -- seed 2, state
starting at |0...0>.  One step = one whole execution of that circuit through
the QuEST API: initZeroState, the 1080 recorded gates, flush (fused sweeps)
on the GPU.  The state stays resident in HBM across steps.

value = algorithmic HBM bytes of all timed steps / timed wall time, where a
sweep moves 2 x 16 x 2^30 bytes (one read and one write of every amplitude)
and the first sweep of a step, which generates |0...0> instead of reading,
moves half of that.

The autotuner times every candidate shape with lane rotation on and off and
keeps the least GPU time per step.  Rotation (the default for API flushes)
needs fewer sweeps per circuit, each doing more gates, so it lowers GB/s while
it shortens the step; config.fixed_lanes_best gives the best fixed-lane plan
of the same run (qubits 0..5 in every tile, as in the reference's blocks).

roofline: the sweep kernel's algorithmic bytes per launch over its average
launch duration, timed with HIP events on the register's own stream.
cpu_baseline: the CPU oracle (oracle/oracle.c, OpenMP, gate by gate) on a
bounded sample of the same circuit on this host.

serving: a Poisson stream of circuits on 10..30 qubits served in FP32 by the
request server (continuous batching under a 200 GB HBM budget, 1% FP64
recheck) against the plain one-at-a-time FP64 loop on the same arrivals:
circuits/s of both (arrivals at 2000/s, above either loop's capacity, so both
are measured work-bound), their ratio, p50/p99 latency, and latency at half
the served capacity.  paged_bv: a Bernstein-Vazirani register forced into pinned
host pieces by a small device budget.  paged_layered: a layered circuit paged
the same way, whose sweeps give the copies something to overlap with.
resident_bv: Bernstein-Vazirani on a 34-qubit FP64 register held in HBM
(256 GiB), the largest that fits one MI355X.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "30q FP64 HBM GB/s (frac of peak); served circuits/s + p99 at fixed fidelity"
HBM_PEAK_GBS = 8000.0

# Fusion shapes tried by the autotuner: (k, R, W); k = targets above the six
# coalesced qubits per sweep, tile = 6 + R + W qubits.
CANDIDATES = [(3, 3, 0), (4, 4, 0), (5, 4, 1), (6, 4, 2), (6, 3, 3), (7, 4, 3)]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=24)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--shape", type=str, default="", help="k,R,W to skip the autotuner")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--paged-bv", type=int, default=31,
                    help="qubits of the host-paged Bernstein-Vazirani run (0 = skip; 31 is the largest FP64 "
                         "register whose pieces a host granting 64 GiB of pinned memory can hold)")
    ap.add_argument("--paged-budget-gib", type=float, default=4.0)
    ap.add_argument("--paged-layered", type=str, default="30,6,4",
                    help="n,depth,budget_GiB of the host-paged layered U3+CZ run ('' = skip)")
    ap.add_argument("--serve", type=int, default=400, help="requests of the serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="Poisson arrival rate of the throughput run, requests/s (above capacity: both loops "
                         "are measured work-bound; latency is measured again at half the served capacity)")
    ap.add_argument("--serve-nmax", type=int, default=30)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-dump", type=str, default="", help="write per-request results (JSON lines) here")
    ap.add_argument("--serve-replay-only", action="store_true",
                    help="run only the served replay of the serving mix (to trace its launch groups) and exit")
    ap.add_argument("--resident-bv", type=int, default=34,
                    help="qubits of the Bernstein-Vazirani run held resident in HBM (0 = skip)")
    ap.add_argument("--precision-qubits", type=int, default=32,
                    help="grid-circuit size of the FP32 MFMA precision-form run (0 = skip)")
    ap.add_argument("--precision-k", type=str, default="5,6,7")
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per sweep launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (profiles/sweep_traffic_30q.txt, line 'traffic_bytes_per_launch: X')."""
    p = os.path.join(ROOT, "profiles", "sweep_traffic_30q.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def host_info():
    """The host the bench ran on: CPU model, logical CPUs, CPUs this process
    may use, and total RAM (from /proc)."""
    model, mem_kib = None, None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal:"):
                mem_kib = int(line.split()[1])
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "mem_total_GiB": round(mem_kib / (1 << 20), 1) if mem_kib else None}


def cpu_baseline(circ, n, seconds):
    """Oracle gate by gate on a 2^n FP64 state until `seconds` have passed."""
    import numpy as np

    from oracle.binding import Oracle, build

    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build(ref=False)
    orc = Oracle()
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1.0
    done = 0
    t0 = time.perf_counter()
    for g in circ:
        orc.gate(state, g[0], g[1], g[2])
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    gate_bytes = 2 * 16 * (1 << n)  # one read + one write of the state per gate
    return {
        "value": round(done * gate_bytes / dt / 1e9, 2),
        "unit": "GB/s",
        "cores": orc.threads(),
        "kind": "port",
        "sample": f"first {done} of {len(circ)} gates of the same {n}-qubit circuit, FP64, gate by gate "
                  f"(one state read+write per gate) in {dt:.1f} s; whole circuit extrapolates to "
                  f"{dt / done * len(circ):.0f} s",
    }


def paged_register(q64, env, n, budget_bytes):
    """An n-qubit FP64 register forced into pinned host pieces by a device
    budget of budget_bytes (DGQ_HBM_BUDGET is read when a register is created)."""
    old = os.environ.get("DGQ_HBM_BUDGET")
    os.environ["DGQ_HBM_BUDGET"] = str(int(budget_bytes))
    try:
        return q64.createQureg(n, env)
    finally:
        if old is None:
            del os.environ["DGQ_HBM_BUDGET"]
        else:
            os.environ["DGQ_HBM_BUDGET"] = old


def paged_stats(q64, q):
    """(piece qubits, log2 pieces, stats dict) of a paged register's last flush."""
    import ctypes as C

    L = q64.lib
    L.dgq_paged_info.restype = C.c_int
    L.dgq_paged_info.argtypes = [type(q), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    nl, K = C.c_int(), C.c_int()
    st = (C.c_double * 10)()
    if L.dgq_paged_info(q, C.byref(nl), C.byref(K), st) != 0:
        raise RuntimeError("register is not paged")
    keys = ["sections", "remaps", "pieces_run", "pieces_skipped", "copy_ms", "hidden_ms", "compute_ms", "wall_ms",
            "bytes_in", "bytes_out"]
    return nl.value, K.value, dict(zip(keys, list(st)))


def overlap_figures(st):
    """Copy/compute overlap of a paged flush from its stream timeline: copy
    time that ran while a sweep ran, over the shorter of the two totals (1 =
    the shorter one hidden entirely), and over all copy time."""
    copy, comp, hid = st["copy_ms"], st["compute_ms"], st["hidden_ms"]
    short = min(copy, comp)
    return {"copy_ms": round(copy, 1), "compute_ms": round(comp, 1), "hidden_copy_ms": round(hid, 1),
            "overlap_of_shorter": round(hid / short, 4) if short > 0 else None,
            "hidden_fraction_of_copy": round(hid / copy, 4) if copy > 0 else None,
            "link_GBps": round((st["bytes_in"] + st["bytes_out"]) / (st["wall_ms"] * 1e-3) / 1e9, 1)
            if st["wall_ms"] > 0 else None}


def paged_layered(q64, env, n, depth, budget_bytes, seed=2):
    """A compute-carrying paged flush: the layered U3+CZ circuit on n qubits
    in pinned host pieces, paged through a device budget of budget_bytes.
    Every section runs fused sweeps on the pieces while the next pieces copy
    in and the previous ones copy out.  Checked against the same circuit on a
    resident register: P(q = 0) for every qubit and the total probability."""
    from dgquest_amd import circuits

    circ = circuits.layered_u3_cz(n, depth, seed)
    q = paged_register(q64, env, n, budget_bytes)
    t0 = time.perf_counter()
    q64.apply_circuit(q, circ)
    q64.dgq_flush(q)
    flush_s = time.perf_counter() - t0
    nl, K, st = paged_stats(q64, q)
    p_paged = [q64.calcProbOfOutcome(q, i, 0) for i in range(n)]
    tot_paged = q64.calcTotalProb(q)
    q64.destroyQureg(q, env)
    r = q64.createQureg(n, env)
    q64.apply_circuit(r, circ)
    p_res = [q64.calcProbOfOutcome(r, i, 0) for i in range(n)]
    q64.destroyQureg(r, env)
    out = {
        "workload": f"{n}-qubit layered U3+CZ circuit, depth {depth}, seed {seed}, FP64 ({16 << n >> 30} GiB in "
                    f"pinned host memory), device budget {budget_bytes >> 30} GiB",
        "pieces": 1 << K, "piece_qubits": nl, "gates": len(circ),
        "flush_s": round(flush_s, 3), "sections": int(st["sections"]), "qubit_moves": int(st["remaps"]),
        "pieces_run": int(st["pieces_run"]),
        "max_p0_diff_vs_resident": float(max(abs(a - b) for a, b in zip(p_paged, p_res))),
        "total_prob": tot_paged,
    }
    out.update(overlap_figures(st))
    return out


def paged_bv(q64, env, n_data, budget_bytes, seed=3):
    """Config-3 shape: Bernstein-Vazirani on n_data + 1 qubits held in pinned
    host pieces and paged through a device budget of `budget_bytes`.  Returns
    the recovered-string check and the copy/compute overlap of the flush."""
    import ctypes as C

    from dgquest_amd import circuits

    circ, hidden = circuits.bernstein_vazirani(n_data, seed)
    # createQureg exits when it cannot pin every piece (QuEST semantics): pin
    # the state's size once first and report a refusal instead
    need = 16 << (n_data + 1)
    L = q64.lib
    L.dq_host_alloc.restype = C.c_int
    L.dq_host_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.dq_host_free.argtypes = [C.c_void_p]
    p = C.c_void_p()
    if L.dq_host_alloc(C.byref(p), need) != 0:
        return {"skipped": f"the host refused {need >> 30} GiB of pinned memory"}
    L.dq_host_free(p)
    t_alloc = time.perf_counter()
    q = paged_register(q64, env, n_data + 1, budget_bytes)
    alloc_s = time.perf_counter() - t_alloc
    paged_stats(q64, q)
    t0 = time.perf_counter()
    q64.apply_circuit(q, circ)
    q64.dgq_flush(q)
    flush_s = time.perf_counter() - t0
    nl, K, st = paged_stats(q64, q)
    q64.seed([seed])
    got = 0
    t1 = time.perf_counter()
    for i in range(n_data):
        if q64.measure(q, i):
            got |= 1 << i
    measure_s = time.perf_counter() - t1
    q64.destroyQureg(q, env)
    out = {
        "workload": f"{n_data + 1}-qubit Bernstein-Vazirani (seed {seed}), FP64 "
                    f"({16 << (n_data + 1) >> 30} GiB in pinned host memory), device budget {budget_bytes >> 30} GiB",
        "pieces": 1 << K, "piece_qubits": nl,
        "recovered_hidden_string": got == hidden,
        "pinned_alloc_s": round(alloc_s, 3), "flush_s": round(flush_s, 3), "measure_all_s": round(measure_s, 3),
        "sections": int(st["sections"]), "qubit_moves": int(st["remaps"]),
        "pieces_run": int(st["pieces_run"]), "pieces_skipped_zero": int(st["pieces_skipped"]),
    }
    out.update(overlap_figures(st))
    return out


def resident_bv(q64, env, eng, n, seed=3):
    """Bernstein-Vazirani on the largest FP64 register one MI355X holds in HBM
    (34 qubits: 256 GiB of its 288 GiB): the flush, then a measurement of every
    data qubit (deterministic outcomes: one P(0) pass serves them all)."""
    import ctypes as C

    from dgquest_amd import circuits

    need = 16 << n
    free = eng.info()["free_mem"]
    # createQureg exits on an allocation failure (QuEST semantics) and puts a
    # register above the device budget (free memory less 4 GiB) into pinned
    # host pieces instead: check both before creating it
    if free - (4 << 30) < need:
        return {"skipped": f"{free >> 30} GiB free on the device, {need >> 30} GiB needed + 4 GiB margin"}
    p = C.c_void_p()
    if eng.lib.dq_malloc(C.byref(p), need) != 0:
        return {"skipped": f"a trial allocation of {need >> 30} GiB failed"}
    eng.lib.dq_free(p)
    circ, hidden = circuits.bernstein_vazirani(n - 1, seed)
    t0 = time.perf_counter()
    q = q64.createQureg(n, env)
    t1 = time.perf_counter()
    q64.apply_circuit(q, circ)
    q64.dgq_flush(q)
    q64.dgq_sync(q)
    t2 = time.perf_counter()
    st = q64.plan_stats(q)
    q64.seed([seed])
    got = 0
    for i in range(n - 1):
        if q64.measure(q, i):
            got |= 1 << i
    t3 = time.perf_counter()
    tp = q64.calcTotalProb(q)
    q64.destroyQureg(q, env)
    return {
        "workload": f"{n}-qubit Bernstein-Vazirani (seed {seed}), FP64, {need >> 30} GiB resident in HBM",
        "recovered_hidden_string": got == hidden, "total_prob": tp,
        "create_s": round(t1 - t0, 3), "flush_s": round(t2 - t1, 3), "sweeps": st["sweeps"],
        "measure_all_s": round(t3 - t2, 3),
    }


def serving(q64, args):
    """The serving workload: a Poisson stream of circuits on 10..30 qubits (40%
    hardware-efficient ansatz, 30% GHZ + QFT, 30% layered U3 + CZ), each
    returning sum <Z_i> and 1024 seeded samples.  Served: FP32 states, 1% of
    requests rechecked in FP64 on the device, continuous batching under a
    200 GB HBM budget.  Against: the plain one-at-a-time FP64 loop over the
    same arrivals.  Both loops are native (dgs_replay / dgs_run_plain); the
    requests are recorded through the QuEST API before the clock starts."""
    import numpy as np

    from dgquest_amd import circuits
    from dgquest_amd import serve as S

    mix = circuits.serving_mix(args.serve, seed=5, rate=args.serve_rate, n_max=args.serve_nmax, build=False)
    recs = [S.record(q64, d["n"], circuits.mix_circuit(d)) for d in mix]
    rs = S.RequestSet(recs, [1024] * len(mix), [d["shot_seed"] for d in mix])
    arrive = [d["arrive_s"] for d in mix]
    plain = S.run_plain(q64, rs, arrive)
    # the same one-at-a-time loop with FP32 states: separates what batching
    # and overlap gain from what the lower precision gains
    plain32 = S.run_plain(q64, rs, arrive, prec=S.DQ_F32)
    budget = int(args.serve_budget_gb * 1e9)
    with S.Server(q64, prec=S.DQ_F32, hbm_budget=budget, recheck_fraction=0.01) as srv:
        served = srv.replay(rs, arrive)
        st = srv.stats()
        # latency at half the measured served capacity, same request kinds
        cap = S.latency_stats(served)["circuits_per_s"]
        mix2 = circuits.serving_mix(args.serve, seed=6, rate=0.5 * cap, n_max=args.serve_nmax, build=False)
        recs2 = [S.record(q64, d["n"], circuits.mix_circuit(d)) for d in mix2]
        rs2 = S.RequestSet(recs2, [1024] * len(mix2), [d["shot_seed"] for d in mix2])
        at_load = srv.replay(rs2, [d["arrive_s"] for d in mix2])
        # accuracy pass (not timed): the FP64 recheck forced on a spread of
        # the same requests -- every 20th, plus the five largest registers
        pick = sorted(set(range(0, len(mix), 20)) |
                      set(sorted(range(len(mix)), key=lambda i: (-mix[i]["n"], i))[:5]))
        rs3 = S.RequestSet([recs[i] for i in pick], [1024] * len(pick), [mix[i]["shot_seed"] for i in pick],
                           [S.DGS_RECHECK] * len(pick))
        checked = srv.replay(rs3)
    if args.serve_dump:
        with open(args.serve_dump, "w") as f:
            for tag, rows, mx in (("plain", plain, mix), ("served", served, mix), ("at_load", at_load, mix2)):
                for d, r in zip(mx, rows):
                    f.write(json.dumps({"run": tag, "kind": d["kind"], "n": d["n"], "depth": d["depth"],
                                        **{k: r[k] for k in ("status", "t_arrive_ms", "t_start_ms", "t_done_ms",
                                                             "batch_size", "rechecked", "fidelity", "fidelity_raw",
                                                             "total_prob", "sum_z", "shot_mismatches")}}) + "\n")
    bad = [r for r in plain + plain32 + served + at_load + checked if r["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} requests failed: {bad[0].get('error')}")
    ls, lp, ll = S.latency_stats(served), S.latency_stats(plain), S.latency_stats(at_load)
    lp32 = S.latency_stats(plain32)
    # service time of each request in the one-at-a-time loops (always
    # backlogged: start = the previous request's end), by register size
    svc64 = [r["t_done_ms"] - r["t_start_ms"] for r in plain]
    svc32 = [r["t_done_ms"] - r["t_start_ms"] for r in plain32]
    nq = [d["n"] for d in mix]
    big = 23  # registers above the server's batch limit (22 qubits) run one per launch group
    buckets = {}
    for lo, hi in ((10, 16), (17, 22), (23, 26), (27, 30)):
        idx = [i for i in range(len(mix)) if lo <= nq[i] <= hi]
        if idx:
            buckets[f"{lo}-{hi}"] = {"requests": len(idx),
                                     "fp64_ms_mean": round(sum(svc64[i] for i in idx) / len(idx), 3),
                                     "fp32_ms_mean": round(sum(svc32[i] for i in idx) / len(idx), 3),
                                     "fp64_ms_sum": round(sum(svc64[i] for i in idx), 1)}
    t64 = sum(svc64)
    t32_big = sum(svc32[i] for i in range(len(mix)) if nq[i] >= big)
    rech = [r for r in served + at_load + checked if r["rechecked"]]
    dz = [abs(a["sum_z"] - b["sum_z"]) / max(1, a["n_qubits"]) for a, b in zip(served, plain)]
    r3 = lambda x: None if x is None else round(x, 3)
    fid_bar = 1.0 - 1e-6
    min_fid = min((r["fidelity"] for r in rech), default=None)
    below = [r for r in rech if r["fidelity"] < fid_bar]
    return {
        "workload": f"{len(mix)} requests, Poisson arrivals at {args.serve_rate:g}/s, n log-uniform on "
                    f"[10, {args.serve_nmax}], 40% HEA (depth 4-16) / 30% GHZ+QFT / 30% U3+CZ (depth 8-24), "
                    f"sum<Z> + 1024 seeded samples each, seed 5",
        "served_fp32": {"circuits_per_s": r3(ls["circuits_per_s"]), "p50_ms": r3(ls["p50_ms"]),
                        "p99_ms": r3(ls["p99_ms"]), "makespan_s": r3(ls["makespan_s"]),
                        "launch_groups": int(st["batches"]), "batched_registers": int(st["batched_registers"]),
                        "peak_arena_GB": round(st["peak_arena_bytes"] / 1e9, 2), "memory_waits": int(st["memory_waits"]),
                        "host_plan_ms": round(st["plan_ms"], 1)},
        "plain_fp64_loop": {"circuits_per_s": r3(lp["circuits_per_s"]), "p50_ms": r3(lp["p50_ms"]),
                            "p99_ms": r3(lp["p99_ms"]), "makespan_s": r3(lp["makespan_s"])},
        "plain_fp32_loop": {"circuits_per_s": r3(lp32["circuits_per_s"]), "p50_ms": r3(lp32["p50_ms"]),
                            "p99_ms": r3(lp32["p99_ms"]), "makespan_s": r3(lp32["makespan_s"])},
        "plain_service_ms_by_size": buckets,
        # an upper bound of served / plain FP64: every register up to the
        # batch limit free (perfect batching), every larger one at its FP32
        # one-at-a-time service time
        "bound_vs_plain_fp64": r3(t64 / t32_big) if t32_big > 0 else None,
        "speedup_vs_plain": r3(ls["circuits_per_s"] / lp["circuits_per_s"]),
        "speedup_vs_plain_fp32": r3(ls["circuits_per_s"] / lp32["circuits_per_s"]),
        "at_half_capacity": {"rate_per_s": r3(0.5 * cap), "circuits_per_s": r3(ll["circuits_per_s"]),
                             "p50_ms": r3(ll["p50_ms"]), "p99_ms": r3(ll["p99_ms"])},
        "fp64_recheck": {"fidelity_definition": "|<a|b>|^2 / (<a|a> <b|b>), a = FP64 run, b = served FP32 state",
                         "requests": len(rech), "requests_n_ge_25": sum(1 for r in rech if r["n_qubits"] >= 25),
                         "fidelity_bar": fid_bar, "min_fidelity": min_fid,
                         "bar_met": bool(rech) and not below, "requests_below_bar": len(below),
                         "min_fidelity_raw": min((r["fidelity_raw"] for r in rech), default=None),
                         "max_fidelity_raw": max((r["fidelity_raw"] for r in rech), default=None),
                         "max_abs_err": max((r["max_abs_err"] for r in rech), default=None),
                         "max_norm_drift_fp32": max((abs(r["total_prob"] - 1.0) for r in rech), default=None),
                         "max_sum_z_diff_per_qubit": max((abs(r["sum_z"] - r["sum_z_fp64"]) / r["n_qubits"]
                                                          for r in rech), default=None),
                         "sample_mismatches": int(sum(r["shot_mismatches"] for r in rech)),
                         "samples_compared": int(1024 * len(rech))},
        "max_sum_z_diff_per_qubit_vs_fp64_loop": float(max(dz)) if dz else None,
        "qubit_hist": {int(k): int(v) for k, v in zip(*np.unique([d["n"] for d in mix], return_counts=True))},
    }


def api_default_step(q64, env, eng, circ, n, best_ms):
    """The same circuit through the API with no shape given: the program
    records it and calls searchQuregTaskPartitions, as the reference's
    testcases do (testcases/bv-32.c:90), which plans every candidate shape
    and keeps the least predicted device time.  Reports the shape picked, its
    predicted and measured GPU ms per step, and the ratio to the bench's own
    best (autotuned, or the --shape given)."""
    import ctypes as C

    q = q64.createQureg(n, env)
    stream = q64.dgq_stream(q)
    e0, e1 = eng.event(), eng.event()
    q64.apply_circuit(q, circ)
    q64.searchQuregTaskPartitions(q, env)
    L = q64.lib
    L.dgq_last_search.argtypes = [type(q), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    shape, pred = (C.c_int * 4)(), C.c_double()
    searched = L.dgq_last_search(q, shape, C.byref(pred)) == 0
    q64.dgq_flush(q)  # the first flush runs the plan the search kept
    ms = []
    for _ in range(2):
        q64.initZeroState(q)
        q64.apply_circuit(q, circ)
        eng.record(e0, stream)
        q64.dgq_flush(q)
        eng.record(e1, stream)
        ms.append(eng.elapsed_ms(e0, e1))
    q64.destroyQureg(q, env)
    m = min(ms)
    return {"searched": searched, "shape": f"k{shape[0]}_R{shape[1]}_W{shape[2]}" + ("" if shape[3] else "_fixed_lanes"),
            "predicted_ms": round(pred.value, 2), "gpu_ms_per_step": round(m, 2),
            "vs_bench_best": round(m / best_ms, 4)}


def serve_replay_only(q64, args):
    """The served replay of the serving mix alone (no plain loops, no recheck
    pass): what a kernel trace of the server's launch groups needs."""
    from dgquest_amd import circuits
    from dgquest_amd import serve as S

    mix = circuits.serving_mix(args.serve, seed=5, rate=args.serve_rate, n_max=args.serve_nmax, build=False)
    recs = [S.record(q64, d["n"], circuits.mix_circuit(d)) for d in mix]
    rs = S.RequestSet(recs, [1024] * len(mix), [d["shot_seed"] for d in mix])
    with S.Server(q64, prec=S.DQ_F32, hbm_budget=int(args.serve_budget_gb * 1e9), recheck_fraction=0.0) as srv:
        served = srv.replay(rs, [d["arrive_s"] for d in mix])
        st = srv.stats()
    ls = S.latency_stats(served)
    return {"served_replay": {"requests": len(mix), "n_max": args.serve_nmax, "circuits_per_s": ls["circuits_per_s"],
                              "p50_ms": ls["p50_ms"], "p99_ms": ls["p99_ms"], "launch_groups": int(st["batches"]),
                              "batched_registers": int(st["batched_registers"])}}


def precision_form(q64, args):
    """Config 4: a 4 x (n/4) grid circuit of depth 20 (sqrt X/Y/W per qubit per
    cycle, CZ on four edge patterns in rotation), seed 4.  FP64 butterflies on
    the GPU give the reference state; then an FP32 state runs the same circuit
    as dense 2^k x 2^k blocks on FP32 MFMA for each k.  Reported per k: time of
    the dense sweeps (HIP events on the register's stream), dense ops,
    GEMM flops (8 * 2^k * 2^n per op), MFMA utilisation against the 157.3
    TFLOP/s FP32 dense peak, HBM GB/s, fidelity |<psi64|psi32>|^2 (normalised)
    and max |delta amplitude|."""
    import ctypes as C

    from dgquest_amd import circuits
    from dgquest_amd import quest as Q
    from dgquest_amd.engine import Engine

    n = args.precision_qubits
    cols = n // 4
    circ = circuits.grid_circuit(4, cols, 20, 4)
    eng = Engine()
    q32 = Q.load(1)
    env = q64.createQuESTEnv()
    env32 = q32.createQuESTEnv()
    ref = q64.createQureg(n, env)
    e0, e1 = eng.event(), eng.event()
    s64 = q64.dgq_stream(ref)
    q64.apply_circuit(ref, circ)
    eng.record(e0, s64)
    q64.dgq_flush(ref)
    eng.record(e1, s64)
    ms64 = eng.elapsed_ms(e0, e1)
    sweeps64 = q64.plan_stats(ref)["sweeps"]
    p64 = C.c_int()
    d64 = q64.dgq_state_pointer(ref, C.byref(p64))
    rows = {}
    for k in [int(x) for x in args.precision_k.split(",")]:
        st = q32.createQureg(n, env32)
        if q32.dgq_set_dense(st, k) != 0:
            continue
        s32 = q32.dgq_stream(st)
        q32.apply_circuit(st, circ)
        eng.record(e0, s32)
        q32.dgq_flush(st)
        eng.record(e1, s32)
        ms = eng.elapsed_ms(e0, e1)
        ds = q32.dense_stats(st)
        p32 = C.c_int()
        d32 = q32.dgq_state_pointer(st, C.byref(p32))
        cmp = eng.compare_states(d64, p64.value, d32, p32.value, n)
        fid = abs(cmp["inner"]) ** 2 / (cmp["norm_a"] * cmp["norm_b"])
        hbm = 2 * 8 * (1 << n) * ds["sweeps"] / (ms * 1e-3) / 1e9
        rows[f"k{k}"] = {"ms": round(ms, 2), "sweeps": ds["sweeps"], "dense_ops": ds["ops"],
                         "gemm_TFLOP": round(ds["flops"] / 1e12, 2),
                         "TFLOPs": round(ds["flops"] / (ms * 1e-3) / 1e12, 2),
                         "mfma_utilisation": round(ds["flops"] / (ms * 1e-3) / 157.3e12, 4),
                         "hbm_GBps": round(hbm, 1), "fidelity": fid, "raw_overlap_sq": abs(cmp["inner"]) ** 2,
                         "max_abs_err": cmp["max_abs_diff"], "host_plan_ms": round(ds["flush_ms"], 1)}
        q32.destroyQureg(st, env32)
    q64.destroyQureg(ref, env)
    return {"workload": f"{n}-qubit 4x{cols} grid circuit, depth 20, sqrt X/Y/W + CZ patterns, seed 4",
            "fp64_reference_ms": round(ms64, 2), "fp64_sweeps": sweeps64, "fp32_dense": rows,
            "mfma_peak_TFLOPs": 157.3}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the state-vector engine runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)

    from dgquest_amd import circuits
    from dgquest_amd import quest as Q
    from dgquest_amd.engine import Engine

    q64 = Q.load(2)
    eng = Engine()
    env = q64.createQuESTEnv()
    if args.serve_replay_only:
        print(json.dumps(serve_replay_only(q64, args)), flush=True)
        return
    n = args.qubits
    circ = circuits.layered_u3_cz(n, args.depth, args.seed)
    q = q64.createQureg(n, env)
    stream = q64.dgq_stream(q)
    e0, e1 = eng.event(), eng.event()

    def one_step():
        q64.initZeroState(q)
        q64.apply_circuit(q, circ)
        q64.dgq_flush(q)

    def gpu_time_of_step():
        q64.initZeroState(q)
        q64.apply_circuit(q, circ)
        eng.record(e0, stream)
        q64.dgq_flush(q)
        eng.record(e1, stream)
        return eng.elapsed_ms(e0, e1)

    # ---- autotuner: fusion shape with the least GPU time for this circuit
    # every shape with lane rotation on and off (fixed lanes keep qubits 0..5
    # in every tile, as the reference's blocks do); the least GPU time wins
    tuning = {}
    if args.shape:
        k, R, W = (int(x) for x in args.shape.split(","))
        best = (k, R, W, 1)
        q64.dgq_set_tuning(q, k, R, W, 100000)
    else:
        for (k, R, W) in CANDIDATES:
            for rot in (1, 0):
                q64.dgq_set_tuning(q, k, R, W, 100000)
                q64.dgq_set_lane_rotation(q, rot)
                gpu_time_of_step()  # first run of a shape warms it
                ms = gpu_time_of_step()
                st = q64.plan_stats(q)
                step_bytes_c, _ = q64.flush_bytes(q)
                tuning[f"k{k}_R{R}_W{W}" + ("" if rot else "_fixed_lanes")] = {
                    "ms": round(ms, 2), "sweeps": st["sweeps"],
                    "hbm_GBps": round(step_bytes_c / (ms * 1e-3) / 1e9, 1)}
        key = lambda c: f"k{c[0]}_R{c[1]}_W{c[2]}" + ("" if c[3] else "_fixed_lanes")
        combos = [(k, R, W, rot) for (k, R, W) in CANDIDATES for rot in (1, 0)]
        best = min(combos, key=lambda c: tuning[key(c)]["ms"])
        best_fixed = min([c for c in combos if not c[3]], key=lambda c: tuning[key(c)]["ms"])
        q64.dgq_set_tuning(q, best[0], best[1], best[2], 100000)
    q64.dgq_set_lane_rotation(q, best[3])

    # ---- warmup
    for _ in range(args.warmup):
        one_step()
    q64.dgq_sync(q)

    # ---- timed region: K steps back to back; events bracket each step's sweeps
    evs = [(eng.event(), eng.event()) for _ in range(args.steps)]
    eng.sync()
    t0 = time.perf_counter()
    for a, b in evs:
        q64.initZeroState(q)
        q64.apply_circuit(q, circ)
        eng.record(a, stream)
        q64.dgq_flush(q)
        eng.record(b, stream)
    q64.dgq_sync(q)
    eng.sync()
    wall = time.perf_counter() - t0
    gpu_ms = sum(eng.elapsed_ms(a, b) for a, b in evs)

    st = q64.plan_stats(q)
    step_bytes, launches = q64.flush_bytes(q)
    tp = q64.calcTotalProb(q)
    if abs(tp - 1.0) > 1e-10:
        raise RuntimeError(f"total probability drifted: {tp}")
    total_bytes = step_bytes * args.steps
    value = total_bytes / wall / 1e9
    achieved = total_bytes / (gpu_ms * 1e-3) / 1e9
    q64.destroyQureg(q, env)
    api = api_default_step(q64, env, eng, circ, n, gpu_ms / args.steps)
    # before the other legs: the largest allocation, made while device memory
    # is least fragmented
    res_bv = resident_bv(q64, env, eng, args.resident_bv) if args.resident_bv else None

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded MT19937 circuit)",
        "config": {
            "workload": f"{n}-qubit layered random circuit, U3 on every qubit + CZ on a seeded perfect "
                        f"matching per layer, depth {args.depth}, seed {args.seed}, complex FP64 "
                        f"({16 << n >> 30} GiB state), plain fused-sweep method",
            "n_qubits": n,
            "depth": args.depth,
            "gates": len(circ),
            "fusion_k": best[0],
            "tile_R": best[1],
            "tile_W": best[2],
            "lane_rotation": bool(best[3]),
            "sweeps_per_step": st["sweeps"],
            "lds_exchanges_per_step": st["exchanges"],
            "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
            "autotune_ms_per_shape": tuning,
            "api_default": api,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "kernel": f"sweep_kernel<double,{best[1]},{best[2]}>",
            "bytes_per_launch": step_bytes / max(1, launches),
            "avg_launch_ms": round(gpu_ms / (args.steps * max(1, launches)), 4),
        },
    }
    if not args.shape:
        # the best plan that keeps qubits 0..5 in every tile, from the same
        # autotuning run: fewer gates per sweep, more sweeps per circuit
        fb = tuning[f"k{best_fixed[0]}_R{best_fixed[1]}_W{best_fixed[2]}_fixed_lanes"]
        out["config"]["fixed_lanes_best"] = {
            "shape": f"k{best_fixed[0]}_R{best_fixed[1]}_W{best_fixed[2]}", "gpu_ms_per_step": fb["ms"],
            "sweeps": fb["sweeps"], "hbm_GBps": fb["hbm_GBps"], "frac": round(fb["hbm_GBps"] / HBM_PEAK_GBS, 4)}
    # serving first among the extra legs: its plain loop allocates and frees a
    # register per request, and device allocations became ~10x slower for
    # small registers after the precision leg's 64 + 3 x 32 GiB registers had
    # been allocated and freed in the same process
    if res_bv is not None:
        out["resident_bv"] = res_bv
    if args.serve:
        out["serving"] = serving(q64, args)
    if args.paged_bv:
        out["paged_bv"] = paged_bv(q64, env, args.paged_bv - 1, int(args.paged_budget_gib * (1 << 30)))
    if args.paged_layered:
        pn, pd, pb = (float(x) for x in args.paged_layered.split(","))
        out["paged_layered"] = paged_layered(q64, env, int(pn), int(pd), int(pb * (1 << 30)))
    if args.precision_qubits:
        out["precision_form"] = precision_form(q64, args)
    out["host"] = host_info()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(circ, n, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
