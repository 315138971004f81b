"""Registers held resident in HBM: deterministic measurements and the largest
FP64 register one MI355X holds.

Deterministic measurements reuse one P(0) pass (their collapses stay pending
until something needs the state); a 50/50 measurement after them must
execute the pending collapses first.  Outcomes, probabilities and the final
state equal the oracle's for the same seed (FP64 bar 1e-12).  The 34-qubit
case (256 GiB) checks the Bernstein-Vazirani string exactly and the total
probability; it runs only when the device has the memory free.
"""
import ctypes as C
import time

import numpy as np
import pytest

from dgquest_amd import circuits

pytestmark = pytest.mark.gpu

TOL = 1e-12


def test_resident_deterministic_measurements_then_random_one(quest, env, oracle):
    n_data = 19
    circ, s = circuits.bernstein_vazirani(n_data, 5)
    n = n_data + 1
    q = quest.createQureg(n, env)
    quest.apply_circuit(q, circ)
    ref = np.zeros(1 << n, dtype=np.complex128)
    ref[0] = 1.0
    ref = oracle.run(ref, circ)
    keys = [17, 4]
    quest.seed(keys)
    oracle.seed(keys)
    got = 0
    t0 = time.perf_counter()
    for i in range(n_data):
        o = quest.measure(q, i)
        ro, _ = oracle.measure(ref, i)
        assert o == ro
        got |= o << i
    t_det = time.perf_counter() - t0
    assert got == s
    p = C.c_double()
    o = quest.measureWithStats(q, n_data, C.byref(p))
    ro, rp = oracle.measure(ref, n_data)
    assert o == ro
    assert abs(p.value - rp) <= TOL and abs(rp - 0.5) <= TOL
    print(f"{n_data} deterministic measurements in {t_det * 1e3:.1f} ms", flush=True)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    quest.destroyQureg(q, env)


def test_state_write_invalidates_cached_probabilities(quest, env):
    """setAmps right after a P(0) pass (nothing pending, so no flush in
    between): the next measurement must see the written state, not the
    probabilities cached before the write."""
    n = 8
    q = quest.createQureg(n, env)
    quest.initClassicalState(q, 1)  # qubit 0 is 1 with certainty
    assert abs(quest.calcProbOfOutcome(q, 0, 1) - 1.0) <= TOL
    re = (quest.qreal * 1)(1.0)
    im = (quest.qreal * 1)(0.0)
    zero = (quest.qreal * 1)(0.0)
    quest.setAmps(q, 1, zero, zero, 1)
    quest.setAmps(q, 0, re, im, 1)  # now |0...0>: qubit 0 is 0 with certainty
    assert quest.measure(q, 0) == 0
    quest.destroyQureg(q, env)


def test_bv_34_qubits_resident(quest, env):
    from dgquest_amd.engine import Engine

    engine = Engine()
    n = 34
    need = 16 << n
    free = engine.info()["free_mem"]
    if free - (4 << 30) < need:
        pytest.skip(f"{free >> 30} GiB free on the device; the 34-qubit FP64 state needs {need >> 30} GiB")
    p = C.c_void_p()
    if engine.lib.dq_malloc(C.byref(p), need) != 0:
        pytest.skip("a trial allocation of 256 GiB failed")
    engine.lib.dq_free(p)
    circ, hidden = circuits.bernstein_vazirani(n - 1, 3)
    print(f"34-qubit BV, {need >> 30} GiB resident", flush=True)
    q = quest.createQureg(n, env)
    quest.apply_circuit(q, circ)
    quest.seed([3])
    got = 0
    for i in range(n - 1):
        if quest.measure(q, i):
            got |= 1 << i
    assert got == hidden
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    quest.destroyQureg(q, env)
