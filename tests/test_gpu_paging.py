"""GPU: registers paged through the device from pinned host memory.

A small device budget (DGQ_HBM_BUDGET, read when a register is created)
forces a register into 2^K host pieces; every API path then runs through
the paging executor (sections, high <-> piece qubit moves folded into the
copies, zero-piece skipping).  Bar: the oracle's state to 1e-12, measurement
outcomes identical for the same seed, the hidden string of a
Bernstein-Vazirani circuit recovered exactly."""
import contextlib
import os

import numpy as np
import pytest

from dgquest_amd import circuits
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


@contextlib.contextmanager
def budget(nbytes):
    old = os.environ.get("DGQ_HBM_BUDGET")
    os.environ["DGQ_HBM_BUDGET"] = str(int(nbytes))
    try:
        yield
    finally:
        if old is None:
            del os.environ["DGQ_HBM_BUDGET"]
        else:
            os.environ["DGQ_HBM_BUDGET"] = old


def paged_register(quest, env, n, nbytes):
    with budget(nbytes):
        q = quest.createQureg(n, env)
    info = paged_info(quest, q)
    assert info is not None, "register was expected to be paged"
    return q, info


def paged_info(quest, q):
    import ctypes as C

    nl, K = C.c_int(), C.c_int()
    st = (C.c_double * 10)()
    quest.lib.dgq_paged_info.restype = C.c_int
    quest.lib.dgq_paged_info.argtypes = [type(q), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    rc = quest.lib.dgq_paged_info(q, C.byref(nl), C.byref(K), st)
    if rc != 0:
        return None
    keys = ["sections", "remaps", "pieces_run", "pieces_skipped", "copy_ms", "hidden_ms", "compute_ms", "wall_ms",
            "bytes_in", "bytes_out"]
    d = {k: st[i] for i, k in enumerate(keys)}
    d["nl"], d["K"] = nl.value, K.value
    return d


def oracle_state(oracle, n, circ, init=None):
    s = np.zeros(1 << n, dtype=np.complex128)
    if init == "plus":
        s[:] = 1.0 / np.sqrt(1 << n)
    elif isinstance(init, int):
        s[init] = 1.0
    else:
        s[0] = 1.0
    return oracle.run(s, circ)


@pytest.mark.parametrize("n,seed", [(14, 1), (15, 2), (16, 3)])
def test_paged_random_circuits(quest, env, oracle, n, seed):
    q, info = paged_register(quest, env, n, 64 << 10)
    rng = np.random.default_rng(seed)
    circ = H.random_circuit(rng, n, 150)
    quest.apply_circuit(q, circ)
    sv = quest.state_vector(q)
    st = paged_info(quest, q)
    print(f"n={n} nl={info['nl']} K={info['K']} stats={st}", flush=True)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(sv - ref).max() <= TOL
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    assert st["remaps"] >= 1  # gates on the top qubits forced moves
    quest.destroyQureg(q, env)


def test_paged_layered_against_resident(quest, env, oracle):
    """The same layered circuit, paged and resident, agree (and match the oracle)."""
    n = 16
    circ = circuits.layered_u3_cz(n, 8, 5)
    q, _ = paged_register(quest, env, n, 64 << 10)
    quest.apply_circuit(q, circ)
    sv_p = quest.state_vector(q)
    r = quest.createQureg(n, env)
    quest.apply_circuit(r, circ)
    sv_r = quest.state_vector(r)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(sv_p - ref).max() <= TOL
    assert np.abs(sv_p - sv_r).max() <= TOL
    quest.destroyQureg(q, env)
    quest.destroyQureg(r, env)


@pytest.mark.parametrize("init", ["plus", 12345])
def test_paged_init_and_probabilities(quest, env, oracle, init):
    n = 15
    q, _ = paged_register(quest, env, n, 64 << 10)
    if init == "plus":
        quest.initPlusState(q)
    else:
        quest.initClassicalState(q, init)
    rng = np.random.default_rng(7)
    circ = H.random_circuit(rng, n, 60)
    quest.apply_circuit(q, circ)
    ref = oracle_state(oracle, n, circ, init=init)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    for t in range(n):
        assert abs(quest.calcProbOfOutcome(q, t, 0) - oracle.prob_of_outcome(ref, t, 0)) <= TOL
    quest.destroyQureg(q, env)


def test_paged_measurement_bit_exact(quest, env, oracle):
    n = 14
    q, _ = paged_register(quest, env, n, 64 << 10)
    rng = np.random.default_rng(11)
    circ = H.random_circuit(rng, n, 50)
    quest.apply_circuit(q, circ)
    ref = oracle_state(oracle, n, circ)
    keys = [3, 0xC0FFEE]
    quest.seed(keys)
    oracle.seed(keys)
    for t in [int(x) for x in rng.permutation(n)]:
        assert quest.measure(q, t) == oracle.measure(ref, t)[0]
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    quest.destroyQureg(q, env)


def test_paged_zero_pieces_skipped_and_set_amps(quest, env, oracle):
    """From |0..0>, gates on the low qubits touch one piece only; setAmps and
    initStateFromAmps write through the qubit map."""
    import ctypes as C

    n = 16
    q, info = paged_register(quest, env, n, 64 << 10)
    circ = [("hadamard", (i,), ()) for i in range(info["nl"])]
    quest.apply_circuit(q, circ)
    sv = quest.state_vector(q)
    st = paged_info(quest, q)
    assert st["pieces_skipped"] >= (1 << info["K"]) - 1
    assert np.abs(sv - oracle_state(oracle, n, circ)).max() <= TOL
    # swap relabels, then setAmps through the map
    quest.swapGate(q, 0, n - 1)
    ref = oracle_state(oracle, n, circ + [("swapGate", (0, n - 1), ())])
    dp = C.POINTER(C.c_double)
    new = np.array([0.1 + 0.2j, -0.3j, 0.25])
    quest.setAmps(q, 1000, np.ascontiguousarray(new.real).ctypes.data_as(dp),
                  np.ascontiguousarray(new.imag).ctypes.data_as(dp), 3)
    ref[1000:1003] = new
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    quest.destroyQureg(q, env)


def test_paged_bernstein_vazirani(quest, env):
    """Config-3 shape at 20 qubits (19 data + ancilla), pieces of 2^12: the
    measured data qubits give the hidden string."""
    n_data = 19
    circ, s = circuits.bernstein_vazirani(n_data, 3)
    q, info = paged_register(quest, env, n_data + 1, 1 << 20)
    quest.apply_circuit(q, circ)
    quest.seed([3])
    got = 0
    for i in range(n_data):
        if quest.measure(q, i):
            got |= 1 << i
    st = paged_info(quest, q)
    print(f"nl={info['nl']} K={info['K']} stats={st}", flush=True)
    assert got == s


def test_paged_deterministic_measurements_then_random_one(quest, env, oracle):
    """Deterministic measurements of a paged register reuse one P(0) pass (their
    collapses stay pending); the ancilla's 50/50 measurement after them must
    execute the pending collapses first.  Outcomes, probabilities and the final
    state equal the oracle's for the same seed."""
    import time

    n_data = 13
    circ, s = circuits.bernstein_vazirani(n_data, 5)
    n = n_data + 1
    q, _ = paged_register(quest, env, n, 64 << 10)
    quest.apply_circuit(q, circ)
    ref = oracle_state(oracle, n, circ)
    keys = [17, 4]
    quest.seed(keys)
    oracle.seed(keys)
    got = 0
    t0 = time.perf_counter()
    for i in range(n_data):
        o = quest.measure(q, i)
        ro, rp = oracle.measure(ref, i)
        assert o == ro
        got |= o << i
    t_det = time.perf_counter() - t0
    assert got == s
    import ctypes as C

    p = C.c_double()
    o = quest.measureWithStats(q, n_data, C.byref(p))
    ro, rp = oracle.measure(ref, n_data)
    assert o == ro
    assert abs(p.value - rp) <= TOL and abs(rp - 0.5) <= TOL
    print(f"13 deterministic measurements in {t_det * 1e3:.1f} ms", flush=True)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("n,seed", [(15, 4), (16, 5)])
def test_paged_fp32_register(quest32, env32, oracle, n, seed):
    """FP32 amplitudes (8-byte elements) through the paging executor: the
    oracle's FP64 state within FP32 rounding (2e-6 per amplitude over 150
    gates), total probability within 1e-5, and the same circuit resident in
    FP32 within the same bound."""
    rng = np.random.default_rng(seed)
    circ = H.random_circuit(rng, n, 150)
    q, info = paged_register(quest32, env32, n, 32 << 10)
    quest32.apply_circuit(q, circ)
    sv_p = quest32.state_vector(q)
    st = paged_info(quest32, q)
    print(f"FP32 n={n} nl={info['nl']} K={info['K']} stats={st}", flush=True)
    r = quest32.createQureg(n, env32)
    quest32.apply_circuit(r, circ)
    sv_r = quest32.state_vector(r)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(sv_p - ref).max() <= 2e-6
    assert np.abs(sv_p - sv_r).max() <= 2e-6
    assert abs(quest32.calcTotalProb(q) - 1.0) <= 1e-5
    assert st["remaps"] >= 1
    quest32.destroyQureg(q, env32)
    quest32.destroyQureg(r, env32)
