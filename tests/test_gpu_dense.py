"""GPU: dense sweeps on FP32 MFMA (the precision form).

An FP32 register in dense mode (dgq_set_dense) runs its flushes as dense
2^k x 2^k blocks on MFMA inside 13-qubit tiles.  Bar, against the FP64
oracle: max |delta amplitude| <= 1e-5 and fidelity >= 1 - 1e-6 (normalised
overlap), for random circuits over the whole gate set and grid circuits, for
k = 5, 6, 7; and a register that switches between butterflies and dense
sweeps stays correct.
"""
import numpy as np
import pytest

from dgquest_amd import circuits as CC
from tests import helpers as H

pytestmark = pytest.mark.gpu


def oracle_state(oracle, n, circ):
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return oracle.run(s, circ)


def check(sv, ref, tag):
    err = np.abs(sv - ref).max()
    fid = abs(np.vdot(ref, sv)) ** 2 / np.vdot(sv, sv).real
    print(f"{tag}: err={err:.3e} fid={fid:.9f}", flush=True)
    assert err <= 1e-5 and fid >= 1 - 1e-6


@pytest.mark.parametrize("k", [5, 6, 7])
@pytest.mark.parametrize("n", [12, 13, 16])
def test_dense_random(quest32, env32, oracle, k, n):
    rng = np.random.default_rng(100 * k + n)
    circ = H.random_circuit(rng, n, 150)
    q = quest32.createQureg(n, env32)
    assert quest32.dgq_set_dense(q, k) == 0
    quest32.apply_circuit(q, circ)
    sv = quest32.state_vector(q)
    st = quest32.dense_stats(q)
    check(sv, oracle_state(oracle, n, circ), f"random n={n} k={k} {st}")
    assert st["ops"] > 0 and st["flops"] > 0
    quest32.destroyQureg(q, env32)


@pytest.mark.parametrize("k", [5, 6, 7])
def test_dense_grid(quest32, env32, oracle, k):
    n = 20
    circ = CC.grid_circuit(4, 5, 12, 4)
    q = quest32.createQureg(n, env32)
    assert quest32.dgq_set_dense(q, k) == 0
    quest32.apply_circuit(q, circ)
    check(quest32.state_vector(q), oracle_state(oracle, n, circ), f"grid 4x5 k={k} {quest32.dense_stats(q)}")
    quest32.destroyQureg(q, env32)


def test_dense_mixed_with_butterflies(quest32, env32, oracle):
    n = 15
    rng = np.random.default_rng(9)
    c1 = H.random_circuit(rng, n, 60)
    c2 = H.random_circuit(rng, n, 60)
    c3 = H.random_circuit(rng, n, 60)
    q = quest32.createQureg(n, env32)
    quest32.initPlusState(q)
    quest32.apply_circuit(q, c1)          # butterflies
    assert quest32.dgq_set_dense(q, 6) == 0  # flushes c1 first
    quest32.apply_circuit(q, c2)          # dense
    quest32.state_vector(q, 0, 1)
    assert quest32.dgq_set_dense(q, 0) == 0
    quest32.apply_circuit(q, c3)          # butterflies again
    s = np.full(1 << n, 1 / np.sqrt(1 << n), dtype=np.complex128)
    ref = oracle.run(s, c1 + c2 + c3)
    check(quest32.state_vector(q), ref, "mixed")
    quest32.destroyQureg(q, env32)


def test_dense_refuses_fp64_and_small(quest, env, quest32, env32):
    q = quest.createQureg(14, env)
    assert quest.dgq_set_dense(q, 6) == -1
    quest.destroyQureg(q, env)
    q = quest32.createQureg(10, env32)
    assert quest32.dgq_set_dense(q, 6) == -1
    assert quest32.dgq_set_dense(q, 5) == 0
    quest32.destroyQureg(q, env32)


def test_inner_product_and_fidelity(quest, env, oracle):
    """calcInnerProduct / calcFidelity (declared by the reference, defined
    here) against numpy on the oracle's states, with relabelled qubits."""
    n = 14
    rng = np.random.default_rng(3)
    c1 = H.random_circuit(rng, n, 80)
    c2 = H.random_circuit(rng, n, 80)
    a = quest.createQureg(n, env)
    b = quest.createQureg(n, env)
    quest.apply_circuit(a, c1)
    quest.apply_circuit(b, c2)
    ra, rb = oracle_state(oracle, n, c1), oracle_state(oracle, n, c2)
    ip = quest.calcInnerProduct(a, b)
    want = np.vdot(ra, rb)
    assert abs(complex(ip.real, ip.imag) - want) <= 1e-12
    assert abs(quest.calcFidelity(a, b) - abs(np.vdot(rb, ra)) ** 2) <= 1e-12
    assert abs(quest.calcFidelity(a, a) - 1.0) <= 1e-12
    quest.destroyQureg(a, env)
    quest.destroyQureg(b, env)


def test_cross_precision_compare(quest, env, quest32, env32):
    """The same circuit in FP64 and in FP32 dense mode, compared on the device."""
    from dgquest_amd.engine import Engine
    import ctypes as C

    n = 16
    circ = CC.grid_circuit(4, 4, 10, 7)
    a = quest.createQureg(n, env)
    b = quest32.createQureg(n, env32)
    assert quest32.dgq_set_dense(b, 6) == 0
    quest.apply_circuit(a, circ)
    quest32.apply_circuit(b, circ)
    pa, pb = C.c_int(), C.c_int()
    da = quest.dgq_state_pointer(a, C.byref(pa))
    db = quest32.dgq_state_pointer(b, C.byref(pb))
    r = Engine().compare_states(da, pa.value, db, pb.value, n)
    sa, sb = quest.state_vector(a), quest32.state_vector(b)
    assert abs(r["inner"] - np.vdot(sa, sb)) <= 1e-9
    assert abs(r["max_abs_diff"] - np.abs(sa - sb).max()) <= 1e-12
    assert abs(r["inner"]) ** 2 >= 1 - 2e-6
    quest.destroyQureg(a, env)
    quest32.destroyQureg(b, env32)
