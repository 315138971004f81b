"""CPU: the dense-sweep planner (precision form) against the oracle.

The planner groups recorded gates into dense ops of k qubits and forms each
op's 2^k x 2^k matrix on the host.  Here the exported ops are applied in
numpy (a tensor contraction per op, independent of the engine) and the result
must equal the oracle's gate-by-gate state to 1e-12, for random circuits over
the whole gate set and for the grid circuits of the precision workload.
"""
import numpy as np
import pytest

from dgquest_amd import circuits as CC
from tests import helpers as H


def apply_dense(psi, n, qs, U):
    """psi[index] with qubit q = bit q; U acts on the k-bit index sum_j bit(qs[j]) << j."""
    k = len(qs)
    axes = [n - 1 - int(q) for q in qs]
    order = [a for a in range(n) if a not in axes] + [axes[j] for j in reversed(range(k))]
    t = np.transpose(psi.reshape([2] * n), order).reshape(-1, 1 << k)
    t = t @ U.T
    return np.transpose(t.reshape([2] * n), np.argsort(order)).reshape(-1)


def dense_plan(quest, n, circ, k):
    r = quest.dgq_create_recorder(n)
    quest.apply_circuit(r, circ)
    qs, mats, nb = quest.recorder_plan_dense(r, k)
    quest.dgq_destroy_recorder(r)
    return qs, mats, nb


@pytest.fixture(scope="module")
def q64():
    from dgquest_amd import quest as Q

    return Q.load(2)


@pytest.mark.parametrize("k", [5, 6, 7])
@pytest.mark.parametrize("seed", [1, 2])
def test_dense_plan_random_circuits(q64, oracle, k, seed):
    n = 12
    rng = np.random.default_rng(seed)
    # swapGate only relabels qubits (the plan then acts on physical positions);
    # the qubit map is covered by the register tests, so it is left out here
    circ = [g for g in H.random_circuit(rng, n, 200) if g[0] != "swapGate"]
    qs, mats, nb = dense_plan(q64, n, circ, k)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    ref = oracle.run(psi.copy(), circ)
    out = psi.copy()
    for q, U in zip(qs, mats):
        assert len(set(q.tolist())) == k
        out = apply_dense(out, n, q, U)
    assert np.abs(out - ref).max() <= 1e-12
    # every op matrix is unitary (the circuit has no projections)
    for U in mats:
        assert np.abs(U @ U.conj().T - np.eye(len(U))).max() <= 1e-12


@pytest.mark.parametrize("k", [5, 6, 7])
def test_dense_plan_grid_circuit(q64, oracle, k):
    n = 16
    circ = CC.grid_circuit(4, 4, 10, 4)
    qs, mats, nb = dense_plan(q64, n, circ, k)
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1
    ref = oracle.run(s.copy(), circ)
    out = s.copy()
    for q, U in zip(qs, mats):
        out = apply_dense(out, n, q, U)
    assert np.abs(out - ref).max() <= 1e-12
    print(f"k={k}: {len(qs)} dense ops in {nb} sweeps for {len(circ)} gates")


def test_dense_plan_rejects_small_registers(q64):
    import ctypes as C

    r = q64.dgq_create_recorder(10)
    n_ops = C.c_int()
    assert q64.lib.dgq_recorder_plan_dense(r, 6, 0, C.byref(n_ops), None, None, None) == -1  # T = 10 < 6 + 5
    assert q64.lib.dgq_recorder_plan_dense(r, 5, 0, C.byref(n_ops), None, None, None) == 0
    assert q64.lib.dgq_recorder_plan_dense(r, 4, 0, C.byref(n_ops), None, None, None) == -1
    q64.dgq_destroy_recorder(r)


def test_grid_circuit_shape():
    c = CC.grid_circuit(4, 8, 8, 4)
    singles = [g for g in c if g[0] != "controlledPhaseFlip"]
    assert len(singles) == 32 * 8
    # never the same gate twice in a row on a qubit
    last = {}
    for name, (q,), _ in singles:
        assert last.get(q) != name
        last[q] = name
    cz = [g[1] for g in c if g[0] == "controlledPhaseFlip"]
    assert all(abs(a - b) in (1, 8) for a, b in cz)
    # pattern sizes: 16 horizontal even pairs, 12 horizontal odd, 16 vertical even, 8 vertical odd
    per_cycle = [16, 12, 16, 8]
    assert len(cz) == sum(per_cycle[i % 4] for i in range(8))
