"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.  The product
(dgquest_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_MT = os.path.join(HERE, "_ref", "libmt19937ar_ref.so")


def build(ref: bool = True) -> None:
    """Compile the oracle (and, when the reference sources are present, the
    reference's own mt19937ar.c into oracle/_ref/)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if ref and os.path.exists("/root/reference/src/mt19937ar.c"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        self.lib = L
        dp = C.POINTER(C.c_double)
        L.oc_apply.argtypes = [dp, C.c_int, C.c_int, C.c_uint64, C.c_uint64, dp]
        L.oc_init_classical.argtypes = [dp, C.c_int, C.c_int64]
        L.oc_init_plus.argtypes = [dp, C.c_int]
        L.oc_prob_of_outcome.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.oc_prob_of_outcome.restype = C.c_double
        L.oc_total_prob.argtypes = [dp, C.c_int]
        L.oc_total_prob.restype = C.c_double
        L.oc_mt_init_by_array.argtypes = [C.POINTER(C.c_uint64), C.c_int]
        L.oc_mt_genrand_int32.restype = C.c_uint32
        L.oc_mt_genrand_real1.restype = C.c_double
        L.oc_mt_genrand_res53.restype = C.c_double
        L.oc_collapse.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.oc_collapse.restype = C.c_double
        L.oc_measure.argtypes = [dp, C.c_int, C.c_int, dp]
        L.oc_measure.restype = C.c_int
        L.oc_gate.argtypes = [dp, C.c_int, C.c_char_p, C.POINTER(C.c_int), C.c_int, dp]
        L.oc_gate.restype = C.c_int
        L.oc_num_threads.restype = C.c_int

    # state helpers -------------------------------------------------------------
    @staticmethod
    def _ptr(state: np.ndarray):
        assert state.dtype == np.complex128 and state.flags["C_CONTIGUOUS"]
        return state.ctypes.data_as(C.POINTER(C.c_double))

    def zero_state(self, n: int) -> np.ndarray:
        s = np.zeros(1 << n, dtype=np.complex128)
        s[0] = 1.0
        return s

    def gate(self, state: np.ndarray, name: str, qubits, params=()) -> None:
        n = int(state.size).bit_length() - 1
        q = (C.c_int * max(1, len(qubits)))(*[int(x) for x in qubits])
        p = (C.c_double * max(1, len(params)))(*[float(x) for x in params])
        rc = self.lib.oc_gate(self._ptr(state), n, name.encode(), q, len(qubits), p)
        if rc != 0:
            raise ValueError(f"oracle: unknown gate {name}")

    def run(self, state: np.ndarray, circuit) -> np.ndarray:
        for g in circuit:
            self.gate(state, g[0], g[1], g[2] if len(g) > 2 else ())
        return state

    def apply_record(self, state: np.ndarray, gtype: int, tmask: int, cmask: int, op=(0.0,) * 8):
        n = int(state.size).bit_length() - 1
        arr = (C.c_double * 8)(*[float(x) for x in op])
        self.lib.oc_apply(self._ptr(state), n, gtype, tmask, cmask, arr)

    def prob_of_outcome(self, state, q, outcome):
        n = int(state.size).bit_length() - 1
        return self.lib.oc_prob_of_outcome(self._ptr(state), n, q, outcome)

    def total_prob(self, state):
        n = int(state.size).bit_length() - 1
        return self.lib.oc_total_prob(self._ptr(state), n)

    def seed(self, keys):
        arr = (C.c_uint64 * len(keys))(*keys)
        self.lib.oc_mt_init_by_array(arr, len(keys))

    def measure(self, state, q):
        n = int(state.size).bit_length() - 1
        p = C.c_double()
        o = self.lib.oc_measure(self._ptr(state), n, q, C.byref(p))
        return o, p.value

    def collapse(self, state, q, outcome):
        n = int(state.size).bit_length() - 1
        return self.lib.oc_collapse(self._ptr(state), n, q, outcome)

    def threads(self) -> int:
        return self.lib.oc_num_threads()


class ReferenceMT:
    """The reference's own src/mt19937ar.c, compiled into oracle/_ref/."""

    def __init__(self, path: str = REF_MT):
        L = C.CDLL(path)
        L.init_by_array.argtypes = [C.POINTER(C.c_ulong), C.c_int]
        L.init_genrand.argtypes = [C.c_ulong]
        L.genrand_int32.restype = C.c_ulong
        L.genrand_real1.restype = C.c_double
        L.genrand_res53.restype = C.c_double
        self.lib = L

    def init_by_array(self, keys):
        arr = (C.c_ulong * len(keys))(*keys)
        self.lib.init_by_array(arr, len(keys))

    def int32(self):
        return self.lib.genrand_int32()

    def real1(self):
        return self.lib.genrand_real1()

    def res53(self):
        return self.lib.genrand_res53()
