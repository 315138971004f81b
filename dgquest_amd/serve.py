"""ctypes binding of the request server (include/dgq_serve.h).

Circuits become requests through a recording register of the QuEST API
(``dgq_create_recorder`` + the usual gate calls + ``dgq_recorder_export``), so
a served circuit is encoded exactly as the QuEST front end encodes it.  The
server itself -- planning threads, admission against the HBM budget,
batching, streams, read-out -- is native (dgquest_amd/csrc/host/serve.cpp);
Python only prepares requests and reads results.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Sequence

import numpy as np

DQ_F32, DQ_F64 = 1, 2
DQ_INIT_CLASSICAL, DQ_INIT_PLUS = 1, 2
DGS_RECHECK, DGS_NO_RECHECK = 1, 2

GATE_REC_DT = np.dtype([("type", "<i4"), ("pad", "<i4"), ("tmask", "<u8"), ("cmask", "<u8"), ("op", "<f8", 8)])
assert GATE_REC_DT.itemsize == 88


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("prec", C.c_int), ("hbm_budget", C.c_uint64), ("planner_threads", C.c_int),
                ("batch_lanes", C.c_int), ("large_lanes", C.c_int), ("lane_depth", C.c_int),
                ("batch_max_qubits", C.c_int), ("batch_max", C.c_int), ("recheck_fraction", C.c_double),
                ("k", C.c_int), ("R", C.c_int), ("W", C.c_int), ("max_ops", C.c_int)]


class Request(C.Structure):
    _fields_ = [("n_qubits", C.c_int), ("n_gates", C.c_uint32), ("gates", C.c_void_p), ("l2p", C.c_void_p),
                ("init_kind", C.c_int), ("init_index", C.c_uint64), ("init_value", C.c_double),
                ("shots", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("n_qubits", C.c_int), ("sum_z", C.c_double),
                ("total_prob", C.c_double), ("t_arrive_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double), ("rechecked", C.c_int), ("batch_size", C.c_int),
                ("fidelity", C.c_double), ("max_abs_err", C.c_double), ("sum_z_fp64", C.c_double),
                ("shot_mismatches", C.c_uint32), ("pad", C.c_uint32), ("fidelity_raw", C.c_double),
                ("norm_fp64", C.c_double)]


def _bind(lib):
    if getattr(lib, "_dgs_bound", False):
        return lib
    vp = C.c_void_p
    lib.dgs_config_default.argtypes = [C.POINTER(Config)]
    lib.dgs_config_default.restype = None
    lib.dgs_last_error.restype = C.c_char_p
    lib.dgs_create.argtypes = [C.POINTER(Config)]
    lib.dgs_create.restype = vp
    lib.dgs_destroy.argtypes = [vp]
    lib.dgs_destroy.restype = None
    lib.dgs_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    lib.dgs_wait.argtypes = [vp, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    lib.dgs_poll.argtypes = [vp, C.c_uint64]
    lib.dgs_replay.argtypes = [vp, C.POINTER(Request), C.c_uint32, C.POINTER(C.c_double), C.POINTER(Result),
                               C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    lib.dgs_run_plain.argtypes = [C.POINTER(Config), C.POINTER(Request), C.c_uint32, C.POINTER(C.c_double),
                                  C.POINTER(Result), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    lib.dgs_stats.argtypes = [vp, C.POINTER(C.c_double)]
    lib.dgs_uniforms.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]
    lib.dgs_arena_selftest.argtypes = [C.c_uint32, C.c_int]
    from .quest import Qureg

    lib.dgq_recorder_export.argtypes = [Qureg, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_ulonglong), C.POINTER(C.c_double)]
    lib.dgq_sample.argtypes = [Qureg, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_ulonglong)]
    lib._dgs_bound = True
    return lib


class Recorded:
    """One circuit as the QuEST front end records it: gate records (physical
    qubits), the final logical -> physical map and the initial state."""

    def __init__(self, n, gates, l2p, init_kind, init_index, init_value):
        self.n = n
        self.gates = gates
        self.l2p = np.ascontiguousarray(l2p, dtype=np.int32)
        self.init_kind = init_kind
        self.init_index = init_index
        self.init_value = init_value


def record(quest, n: int, circuit: Iterable, init: str | int = "zero") -> Recorded:
    """Record `circuit` (QuEST API gate tuples) on an n-qubit recording
    register and export it.  init: "zero", "plus" or a basis-state index."""
    lib = _bind(quest.lib)
    q = quest.lib.dgq_create_recorder(n)
    try:
        if init == "plus":
            quest.lib.initPlusState(q)
        elif isinstance(init, int) and init:
            quest.lib.initClassicalState(q, init)
        quest.apply_circuit(q, circuit)
        cnt = C.c_int()
        l2p = (C.c_int * n)()
        kind = C.c_int()
        idx = C.c_ulonglong()
        val = C.c_double()
        rc = lib.dgq_recorder_export(q, None, 0, C.byref(cnt), l2p, C.byref(kind), C.byref(idx), C.byref(val))
        if rc < 0:
            raise RuntimeError("dgq_recorder_export failed")
        gates = np.zeros(cnt.value, dtype=GATE_REC_DT)
        rc = lib.dgq_recorder_export(q, gates.ctypes.data, len(gates), C.byref(cnt), l2p, C.byref(kind),
                                     C.byref(idx), C.byref(val))
        if rc != 0:
            raise RuntimeError("dgq_recorder_export failed")
        return Recorded(n, gates, list(l2p), kind.value, idx.value, val.value)
    finally:
        quest.lib.dgq_destroy_recorder(q)


class RequestSet:
    """A C array of dgs_request over recorded circuits (keeps their arrays alive)."""

    def __init__(self, recs: Sequence[Recorded], shots: Sequence[int], seeds: Sequence[int], flags=None):
        self.recs = list(recs)
        self.arr = (Request * len(recs))()
        for i, r in enumerate(recs):
            q = self.arr[i]
            q.n_qubits = r.n
            q.n_gates = len(r.gates)
            q.gates = r.gates.ctypes.data if len(r.gates) else None
            q.l2p = r.l2p.ctypes.data
            q.init_kind = r.init_kind
            q.init_index = r.init_index
            q.init_value = r.init_value
            q.shots = int(shots[i])
            q.seed = int(seeds[i]) & 0xFFFFFFFF
            q.flags = 0 if flags is None else int(flags[i])
        self.shots = [int(s) for s in shots]
        self.ns = [r.n for r in recs]

    def __len__(self):
        return len(self.recs)


def _unpack(rs: RequestSet, res, samples, p0):
    out = []
    so = po = 0
    for i in range(len(rs)):
        r = res[i]
        d = {f: getattr(r, f) for f, _ in Result._fields_ if f != "pad"}
        s = rs.shots[i]
        n = rs.ns[i]
        d["samples"] = samples[so:so + s].copy()
        d["p0"] = p0[po:po + n + 1].copy()
        so += s
        po += n + 1
        out.append(d)
    return out


def default_config(quest, **kw) -> Config:
    lib = _bind(quest.lib)
    c = Config()
    lib.dgs_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Server:
    """A running request server (dgs_create .. dgs_destroy)."""

    def __init__(self, quest, **cfg):
        self.lib = _bind(quest.lib)
        self.cfg = default_config(quest, **cfg)
        self.h = self.lib.dgs_create(C.byref(self.cfg))
        if not self.h:
            raise RuntimeError("dgs_create: " + self.lib.dgs_last_error().decode())

    def replay(self, rs: RequestSet, arrive_s: Sequence[float] | None = None):
        n = len(rs)
        res = (Result * n)()
        samples = np.zeros(max(1, sum(rs.shots)), dtype=np.uint64)
        p0 = np.zeros(max(1, sum(x + 1 for x in rs.ns)), dtype=np.float64)
        arr = None
        if arrive_s is not None:
            a = np.ascontiguousarray(arrive_s, dtype=np.float64)
            arr = a.ctypes.data_as(C.POINTER(C.c_double))
        rc = self.lib.dgs_replay(self.h, rs.arr, n, arr, res, samples.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 p0.ctypes.data_as(C.POINTER(C.c_double)))
        out = _unpack(rs, res, samples, p0)
        if rc != 0:
            err = self.lib.dgs_last_error().decode()
            for d in out:
                d.setdefault("error", err)
        return out

    def stats(self):
        v = (C.c_double * 8)()
        self.lib.dgs_stats(self.h, v)
        keys = ["done", "batches", "batched_registers", "peak_arena_bytes", "arena_bytes", "memory_waits",
                "rechecked", "plan_ms"]
        return {k: v[i] for i, k in enumerate(keys)}

    def close(self):
        if self.h:
            self.lib.dgs_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_plain(quest, rs: RequestSet, arrive_s: Sequence[float] | None = None, **cfg):
    """The plain one-request-at-a-time loop (dgs_run_plain); FP64 unless prec= is given."""
    lib = _bind(quest.lib)
    cfg.setdefault("prec", DQ_F64)
    c = default_config(quest, **cfg)
    n = len(rs)
    res = (Result * n)()
    samples = np.zeros(max(1, sum(rs.shots)), dtype=np.uint64)
    p0 = np.zeros(max(1, sum(x + 1 for x in rs.ns)), dtype=np.float64)
    arr = None
    if arrive_s is not None:
        a = np.ascontiguousarray(arrive_s, dtype=np.float64)
        arr = a.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib.dgs_run_plain(C.byref(c), rs.arr, n, arr, res, samples.ctypes.data_as(C.POINTER(C.c_uint64)),
                           p0.ctypes.data_as(C.POINTER(C.c_double)))
    out = _unpack(rs, res, samples, p0)
    if rc != 0:
        err = lib.dgs_last_error().decode()
        for d in out:
            d.setdefault("error", err)
    return out


def uniforms(quest, seed: int, count: int) -> np.ndarray:
    """The sample uniforms a request with this seed uses (dgs_uniforms)."""
    lib = _bind(quest.lib)
    out = np.zeros(count, dtype=np.float64)
    rc = lib.dgs_uniforms(seed & 0xFFFFFFFF, count, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError(lib.dgs_last_error().decode())
    return out


def sample(quest, q, u: Sequence[float]) -> np.ndarray:
    """dgq_sample: basis-state samples of register q for uniforms u in [0, 1)."""
    lib = _bind(quest.lib)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(len(uu), dtype=np.uint64)
    rc = lib.dgq_sample(q, len(uu), uu.ctypes.data_as(C.POINTER(C.c_double)),
                        out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    if rc != 0:
        raise RuntimeError("dgq_sample failed")
    return out


def latency_stats(results: List[dict]):
    """circuits/s over the run and p50 / p99 latency (arrival -> result, ms)."""
    ok = [r for r in results if r["status"] == 0]
    if not ok:
        return {"circuits_per_s": 0.0, "p50_ms": None, "p99_ms": None, "makespan_s": None}
    lat = np.array([r["t_done_ms"] - r["t_arrive_ms"] for r in ok])
    t0 = min(r["t_arrive_ms"] for r in ok)
    t1 = max(r["t_done_ms"] for r in ok)
    span = (t1 - t0) / 1e3
    return {"circuits_per_s": len(ok) / span if span > 0 else float("inf"),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "makespan_s": span}
