"""ctypes binding of the QuEST C API served by the gfx950 engine.

The libraries are built in-tree (dgquest_amd/lib/) by ``dgquest_amd.build``:

* ``libQuEST_f64.so`` -- QuEST_PREC=2 (double amplitudes)
* ``libQuEST_f32.so`` -- QuEST_PREC=1 (float amplitudes)
* ``libdgq_engine.so`` -- the HIP kernels behind both (include/dgq_engine.h)

The Python names are the C names (``createQureg``, ``hadamard``,
``controlledPhaseShift`` ...), taking and returning the same values, so a
circuit written against the reference's include/QuEST.h reads the same here.
There is no CPU path: loading fails loudly when the libraries are missing and
every amplitude is computed by the GPU kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")


class QuESTEnv(C.Structure):
    _fields_ = [("value", C.c_void_p), ("rank", C.c_int)]


class Qureg(C.Structure):
    _fields_ = [
        ("isDensityMatrix", C.c_int),
        ("numQubitsRepresented", C.c_ulonglong),
        ("numQubitsInStateVec", C.c_ulonglong),
        ("numAmpsTotal", C.c_ulonglong),
        ("env", QuESTEnv),
        ("state", C.c_void_p),
    ]


def _complex_types(qreal):
    class Complex(C.Structure):
        _fields_ = [("real", qreal), ("imag", qreal)]

    class ComplexMatrix2(C.Structure):
        _fields_ = [("r0c0", Complex), ("r0c1", Complex), ("r1c0", Complex), ("r1c1", Complex)]

    class Vector(C.Structure):
        _fields_ = [("x", qreal), ("y", qreal), ("z", qreal)]

    return Complex, ComplexMatrix2, Vector


# argument kinds of each gate: q = qubit, qs = qubit list, r = real,
# c = complex (2 reals), m = 2x2 matrix (8 reals), v = vector (3 reals)
GATE_SIGNATURES = {
    "hadamard": ("q",), "pauliX": ("q",), "pauliY": ("q",), "pauliZ": ("q",),
    "sGate": ("q",), "tGate": ("q",), "SqX": ("q",), "SqY": ("q",), "SqW": ("q",),
    "phaseShift": ("q", "r"), "rotateX": ("q", "r"), "rotateY": ("q", "r"), "rotateZ": ("q", "r"),
    "u1Gate": ("q", "r"), "u2Gate": ("q", "r", "r"), "u3Gate": ("q", "r", "r", "r"),
    "controlledPhaseShift": ("q", "q", "r"), "controlledPhaseFlip": ("q", "q"),
    "controlledRotateX": ("q", "q", "r"), "controlledRotateY": ("q", "q", "r"),
    "controlledRotateZ": ("q", "q", "r"), "controlledNot": ("q", "q"), "controlledPauliY": ("q", "q"),
    "compactUnitary": ("q", "c", "c"), "controlledCompactUnitary": ("q", "q", "c", "c"),
    "unitary": ("q", "m"), "controlledUnitary": ("q", "q", "m"),
    "multiControlledUnitary": ("qs", "q", "m"),
    "multiControlledPhaseFlip": ("qs",), "multiControlledPhaseShift": ("qs", "r"),
    "rotateAroundAxis": ("q", "r", "v"), "controlledRotateAroundAxis": ("q", "q", "r", "v"),
    "fSim": ("q", "q", "r", "r"), "swapGate": ("q", "q"),
}

_NREALS = {"r": 1, "c": 2, "m": 8, "v": 3}


def _descriptor_dtypes():
    """numpy layouts of dq_block, dq_phase and dq_op (include/dgq_engine.h)."""
    import numpy as np

    block = np.dtype([("n_qubits", "<u4"), ("tile_bits", "<u4"), ("reg_bits", "<u4"), ("wave_bits", "<u4"),
                      ("init_kind", "<u4"), ("phase_begin", "<u4"), ("phase_count", "<u4"), ("pad0", "<u4"),
                      ("nontile_mask", "<u8"), ("init_index", "<u8"), ("init_value", "<f8"), ("pad1", "<u8"),
                      ("ld_reg", "<u8", 5), ("ld_wave", "<u8", 3), ("st_reg", "<u8", 5), ("st_wave", "<u8", 3)])
    phase = np.dtype([("reg_t", "<u4", 5), ("wave_t", "<u4", 3), ("lane_t", "<u4", 6), ("op_begin", "<u4"),
                      ("op_count", "<u4")])
    op = np.dtype([("kind", "<u4"), ("tbit", "<u4"), ("c_reg", "<u4"), ("v_reg", "<u4"), ("c_thr", "<u4"),
                   ("v_thr", "<u4"), ("c_wg", "<u8"), ("v_wg", "<u8"), ("m", "<f8", 8)])
    assert block.itemsize == 192 and phase.itemsize == 64 and op.itemsize == 104
    return block, phase, op


try:
    BLOCK_DT, PHASE_DT, OP_DT = _descriptor_dtypes()
except ImportError:  # numpy is only needed for the plan export
    BLOCK_DT = PHASE_DT = OP_DT = None


class QuEST:
    """One loaded precision of the QuEST API (``prec`` 2 = double, 1 = float)."""

    def __init__(self, prec: int = 2, libdir: str = LIBDIR):
        name = {2: "libQuEST_f64.so", 1: "libQuEST_f32.so"}[prec]
        path = os.path.join(libdir, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run dgquest_amd.build.build()")
        self.prec = prec
        self.path = path
        self.lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        self.qreal = C.c_double if prec == 2 else C.c_float
        self.Complex, self.ComplexMatrix2, self.Vector = _complex_types(self.qreal)
        L, qr = self.lib, self.qreal
        sig = {
            "createQuESTEnv": (QuESTEnv, []),
            "destroyQuESTEnv": (None, [QuESTEnv]),
            "seedQuEST": (None, [C.POINTER(C.c_ulong), C.c_int]),
            "seedQuESTDefault": (None, []),
            "getQuEST_PREC": (C.c_int, []),
            "createQureg": (Qureg, [C.c_int, QuESTEnv]),
            "destroyQureg": (None, [Qureg, QuESTEnv]),
            "searchQuregTaskPartitions": (None, [Qureg, QuESTEnv]),
            "getNumQubits": (C.c_int, [Qureg]),
            "getNumAmps": (C.c_int, [Qureg]),
            "initZeroState": (None, [Qureg]),
            "initPlusState": (None, [Qureg]),
            "initClassicalState": (None, [Qureg, C.c_longlong]),
            "initStateFromAmps": (None, [Qureg, C.POINTER(qr), C.POINTER(qr)]),
            "setAmps": (None, [Qureg, C.c_longlong, C.POINTER(qr), C.POINTER(qr), C.c_longlong]),
            "getAmp": (self.Complex, [Qureg, C.c_longlong]),
            "getRealAmp": (qr, [Qureg, C.c_longlong]),
            "getImagAmp": (qr, [Qureg, C.c_longlong]),
            "getProbAmp": (qr, [Qureg, C.c_longlong]),
            "calcProbOfOutcome": (qr, [Qureg, C.c_int, C.c_int]),
            "calcTotalProb": (qr, [Qureg]),
            "collapseToOutcome": (qr, [Qureg, C.c_int, C.c_int]),
            "measure": (C.c_int, [Qureg, C.c_int]),
            "measureWithStats": (C.c_int, [Qureg, C.c_int, C.POINTER(qr)]),
            "calcInnerProduct": (self.Complex, [Qureg, Qureg]),
            "calcFidelity": (qr, [Qureg, Qureg]),
        }
        for fname, (res, args) in sig.items():
            f = getattr(L, fname)
            f.restype = res
            f.argtypes = args
        kinds = {"q": [C.c_int], "qs": [C.POINTER(C.c_int), C.c_int], "r": [qr],
                 "c": [self.Complex], "m": [self.ComplexMatrix2], "v": [self.Vector]}
        for gname, ks in GATE_SIGNATURES.items():
            f = getattr(L, gname)
            f.restype = None
            at = [Qureg]
            for k in ks:
                at += kinds[k]
            f.argtypes = at
        # engine statistics of the last flush (host extension, see include/dgquest_ext.h)
        L.dgq_last_plan_stats.restype = C.c_int
        L.dgq_last_plan_stats.argtypes = [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.dgq_set_tuning.restype = C.c_int
        L.dgq_set_tuning.argtypes = [Qureg, C.c_int, C.c_int, C.c_int, C.c_int]
        L.dgq_set_lane_rotation.restype = C.c_int
        L.dgq_set_lane_rotation.argtypes = [Qureg, C.c_int]
        L.dgq_set_fusion.restype = C.c_int
        L.dgq_set_fusion.argtypes = [Qureg, C.c_int]
        L.dgq_stream.restype = C.c_void_p
        L.dgq_stream.argtypes = [Qureg]
        L.dgq_last_flush_bytes.restype = C.c_int
        L.dgq_last_flush_bytes.argtypes = [Qureg, C.POINTER(C.c_ulonglong), C.POINTER(C.c_int)]
        L.dgq_sync.restype = C.c_int
        L.dgq_sync.argtypes = [Qureg]
        L.dgq_flush.restype = C.c_int
        L.dgq_flush.argtypes = [Qureg]
        L.dgq_read_state.restype = C.c_int
        L.dgq_read_state.argtypes = [Qureg, C.c_ulonglong, C.c_ulonglong, C.POINTER(C.c_double)]
        L.dgq_create_recorder.restype = Qureg
        L.dgq_create_recorder.argtypes = [C.c_int]
        L.dgq_destroy_recorder.restype = None
        L.dgq_destroy_recorder.argtypes = [Qureg]
        L.dgq_recorder_plan.restype = C.c_int
        L.dgq_recorder_plan.argtypes = [Qureg, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]

        L.dgq_state_pointer.restype = C.c_void_p
        L.dgq_state_pointer.argtypes = [Qureg, C.POINTER(C.c_int)]
        L.dgq_set_dense.restype = C.c_int
        L.dgq_set_dense.argtypes = [Qureg, C.c_int]
        L.dgq_last_dense_stats.restype = C.c_int
        L.dgq_last_dense_stats.argtypes = [Qureg, C.POINTER(C.c_double)]
        L.dgq_recorder_plan_dense.restype = C.c_int
        L.dgq_recorder_plan_dense.argtypes = [Qureg, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                              C.c_void_p, C.c_void_p]

    # --- thin pass-throughs -------------------------------------------------
    def __getattr__(self, name):
        return getattr(self.lib, name)

    # --- helpers ---------------------------------------------------------------
    def complex(self, re: float, im: float = 0.0):
        return self.Complex(re, im)

    def matrix(self, m: Sequence[float]):
        cx = self.Complex
        return self.ComplexMatrix2(cx(m[0], m[1]), cx(m[2], m[3]), cx(m[4], m[5]), cx(m[6], m[7]))

    def apply(self, q: Qureg, name: str, qubits: Sequence[int], params: Sequence[float] = ()):
        """Apply one gate given as (name, qubits, flattened real params)."""
        ks = GATE_SIGNATURES[name]
        args = [q]
        qi = 0
        pi = 0
        for k in ks:
            if k == "q":
                args.append(int(qubits[qi]))
                qi += 1
            elif k == "qs":
                nctrl = len(qubits) - (1 if "m" in ks else 0)
                arr = (C.c_int * nctrl)(*[int(x) for x in qubits[qi:qi + nctrl]])
                args += [arr, nctrl]
                qi += nctrl
            else:
                n = _NREALS[k]
                vals = [float(x) for x in params[pi:pi + n]]
                pi += n
                if k == "r":
                    args.append(vals[0])
                elif k == "c":
                    args.append(self.Complex(vals[0], vals[1]))
                elif k == "m":
                    args.append(self.matrix(vals))
                else:
                    args.append(self.Vector(*vals))
        getattr(self.lib, name)(*args)

    def apply_circuit(self, q: Qureg, circuit: Iterable):
        for g in circuit:
            self.apply(q, g[0], g[1], g[2] if len(g) > 2 else ())

    def seed(self, keys: Sequence[int]):
        arr = (C.c_ulong * len(keys))(*keys)
        self.lib.seedQuEST(arr, len(keys))

    def state_vector(self, q: Qureg, start: int = 0, count: int | None = None):
        """Amplitudes start..start+count-1 in logical order as a complex128 numpy array."""
        import numpy as np

        n = int(q.numQubitsRepresented)
        if count is None:
            count = (1 << n) - start
        out = np.empty(2 * count, dtype=np.float64)
        rc = self.lib.dgq_read_state(q, start, count, out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError("dgq_read_state failed")
        return out.view(np.complex128)

    def flush_bytes(self, q: Qureg):
        """(algorithmic HBM bytes, sweep launches) of the last flush."""
        b = C.c_ulonglong()
        l = C.c_int()
        self.lib.dgq_last_flush_bytes(q, C.byref(b), C.byref(l))
        return b.value, l.value

    def recorder_plan(self, q: Qureg, k=6, R=4, W=2, max_ops=4096, fuse=True):
        """The plan of what the recording register q holds, as numpy structured
        arrays (blocks, phases, ops) in the layout of include/dgq_engine.h,
        plus the logical -> physical qubit map."""
        import numpy as np

        n = int(q.numQubitsRepresented)
        counts = (C.c_int * 3)()
        l2p = (C.c_int * n)()
        rc = self.lib.dgq_recorder_plan(q, k, R, W, max_ops, int(fuse), None, 0, None, 0, None, 0, counts, l2p)
        if rc < 0:
            raise RuntimeError("dgq_recorder_plan failed")
        b = np.zeros(counts[0], dtype=BLOCK_DT)
        p = np.zeros(counts[1], dtype=PHASE_DT)
        o = np.zeros(counts[2], dtype=OP_DT)
        rc = self.lib.dgq_recorder_plan(q, k, R, W, max_ops, int(fuse), b.ctypes.data, len(b), p.ctypes.data, len(p),
                                        o.ctypes.data, len(o), counts, l2p)
        if rc != 0:
            raise RuntimeError("dgq_recorder_plan failed")
        return b, p, o, list(l2p)

    def dense_stats(self, q: Qureg):
        """Of the last flush: sweeps, dense ops, device GEMM flops, host ms."""
        v = (C.c_double * 4)()
        self.lib.dgq_last_dense_stats(q, v)
        return {"sweeps": int(v[0]), "ops": int(v[1]), "flops": v[2], "flush_ms": v[3]}

    def recorder_plan_dense(self, q: Qureg, k: int):
        """Dense plan of a recording register: (op qubits [ops x k], matrices [ops x 2^k x 2^k] complex)."""
        import numpy as np

        n_ops = C.c_int()
        n_blocks = C.c_int()
        rc = self.lib.dgq_recorder_plan_dense(q, k, 0, C.byref(n_ops), C.byref(n_blocks), None, None)
        if rc < 0:
            raise RuntimeError("dgq_recorder_plan_dense failed")
        K = 1 << k
        qs = np.zeros(n_ops.value * k, dtype=np.int32)
        mats = np.zeros(n_ops.value * K * K * 2, dtype=np.float64)
        rc = self.lib.dgq_recorder_plan_dense(q, k, n_ops.value, C.byref(n_ops), C.byref(n_blocks), qs.ctypes.data,
                                              mats.ctypes.data)
        if rc != 0:
            raise RuntimeError("dgq_recorder_plan_dense failed")
        return qs.reshape(-1, k), mats.view(np.complex128).reshape(-1, K, K), n_blocks.value

    def plan_stats(self, q: Qureg):
        """Statistics of the last flush: sweeps, phases, exchanges, ops, gates_in, k, R, W + flush ms."""
        ints = (C.c_int * 8)()
        dbl = (C.c_double * 1)()
        self.lib.dgq_last_plan_stats(q, ints, dbl)
        keys = ["sweeps", "phases", "exchanges", "ops", "gates_in", "k", "R", "W"]
        d = {k: ints[i] for i, k in enumerate(keys)}
        d["flush_ms"] = dbl[0]
        return d


_loaded = {}


def load(prec: int = 2) -> QuEST:
    """Load (once) and return the QuEST API of the given precision."""
    if prec not in _loaded:
        _loaded[prec] = QuEST(prec)
    return _loaded[prec]
