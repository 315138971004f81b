"""The planner on the CPU: plans made by a recording register (no GPU) are
executed by the numpy model of the descriptor semantics (tests/plan_model.py)
and compared with the oracle.  A planner mistake -- a wrong factor of a
sheared rotation, a pending monomial or flipped condition lost, a bad
layout -- shows here without a GPU; the GPU tests then check that the
kernels implement the descriptors.

Bar: max |amplitude difference| <= 1e-12 (FP64)."""
import math

import numpy as np
import pytest

from dgquest_amd import circuits
from tests import helpers as H
from tests.plan_model import plan_and_run

TOL = 1e-12
K_SHEAR, K_RMAT, K_DIAGR = 8, 6, 7


def oracle_state(oracle, n, circ, init=None):
    s = np.zeros(1 << n, dtype=np.complex128)
    if init == "plus":
        s[:] = 1.0 / math.sqrt(1 << n)
    elif isinstance(init, int):
        s[init] = 1.0
    else:
        s[0] = 1.0
    return oracle.run(s, circ)


@pytest.mark.parametrize("n", [1, 3, 6, 7, 8, 9, 10, 12])
@pytest.mark.parametrize("fuse", [True, False])
def test_random_circuits(quest, oracle, n, fuse):
    rng = np.random.default_rng(4000 + n)
    circ = H.random_circuit(rng, n, 80)
    psi, _ = plan_and_run(quest, n, circ, fuse=fuse)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(psi - ref).max() <= TOL


@pytest.mark.parametrize("shape", [(1, 1, 0, 64), (3, 3, 0, 64), (4, 4, 0, 8), (5, 3, 2, 4096), (6, 4, 2, 4096),
                                   (6, 5, 1, 4096), (7, 4, 3, 4096), (7, 5, 2, 4096), (2, 2, 1, 3), (4, 2, 2, 16),
                                   (6, 3, 3, 4096)])
def test_shapes(quest, oracle, shape):
    n = 14
    rng = np.random.default_rng(99)
    circ = H.random_circuit(rng, n, 150)
    k, R, W, mo = shape
    psi, _ = plan_and_run(quest, n, circ, k=k, R=R, W=W, max_ops=mo)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(psi - ref).max() <= TOL


@pytest.mark.parametrize("n,depth,seed", [(12, 8, 1), (14, 6, 2)])
def test_layered_u3_cz_uses_shears(quest, oracle, n, depth, seed):
    """The benchmark's gate mix: every U3 but the last on each qubit becomes a
    SHEAR (4 FMAs per pair); the result still matches the oracle."""
    circ = circuits.layered_u3_cz(n, depth, seed)
    psi, (blocks, phases, ops) = plan_and_run(quest, n, circ)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(psi - ref).max() <= TOL
    kinds = ops["kind"]
    n_shear = int((kinds == K_SHEAR).sum())
    n_rmat = int((kinds == K_RMAT).sum())
    assert n_shear + n_rmat == n * depth
    assert n_rmat <= n  # only the last rotation on a qubit


def test_ghz_qft_closed_form(quest):
    n = 12
    circ = circuits.ghz_qft(n)
    psi, _ = plan_and_run(quest, n, circ)
    assert np.abs(psi - circuits.ghz_qft_closed_form(n)).max() <= TOL


@pytest.mark.parametrize("init", ["plus", 77])
def test_init_kinds(quest, oracle, init):
    n = 11
    rng = np.random.default_rng(17)
    circ = H.random_circuit(rng, n, 60)
    psi, _ = plan_and_run(quest, n, circ, init=init)
    assert np.abs(psi - oracle_state(oracle, n, circ, init=init)).max() <= TOL


def test_pending_x_through_controls_and_diagonals(quest, oracle):
    """Rotations near pi (pending X), then controls, phases and CZs on the
    same qubits: the flipped conditions must follow the pending X."""
    n = 9
    rng = np.random.default_rng(5)
    circ = []
    for layer in range(6):
        for q in range(n):
            circ.append(("rotateY", (q,), (math.pi - 0.01 * (q + 1) + 0.3 * layer,)))
        qs = [int(x) for x in rng.permutation(n)]
        circ.append(("controlledPhaseFlip", (qs[0], qs[1]), ()))
        circ.append(("controlledNot", (qs[2], qs[3]), ()))
        circ.append(("controlledPhaseShift", (qs[4], qs[5]), (0.7,)))
        circ.append(("controlledRotateZ", (qs[6], qs[7]), (1.1,)))
        circ.append(("multiControlledPhaseFlip", (qs[1], qs[3], qs[5]), ()))
        circ.append(("pauliX", (qs[8],), ()))
    psi, _ = plan_and_run(quest, n, circ)
    assert np.abs(psi - oracle_state(oracle, n, circ)).max() <= TOL


def test_swap_relabel(quest, oracle):
    n = 10
    rng = np.random.default_rng(8)
    circ = H.random_circuit(rng, n, 30) + [("swapGate", (0, 9), ()), ("swapGate", (2, 7), ())]
    circ += H.random_circuit(rng, n, 30)
    psi, _ = plan_and_run(quest, n, circ)
    assert np.abs(psi - oracle_state(oracle, n, circ)).max() <= TOL


def test_planted_mistake_is_caught(quest, oracle):
    """The model is a real check: a wrong shear coefficient changes the state."""
    n = 8
    circ = circuits.layered_u3_cz(n, 3, 4)
    from tests import plan_model

    q = quest.dgq_create_recorder(n)
    quest.apply_circuit(q, circ)
    blocks, phases, ops, _ = quest.recorder_plan(q)
    quest.dgq_destroy_recorder(q)
    i = int(np.nonzero(ops["kind"] == K_SHEAR)[0][0])
    ops[i]["m"][1] += 1e-9
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    plan_model.run_plan(psi, n, blocks, phases, ops)
    assert np.abs(psi - oracle_state(oracle, n, circ)).max() > TOL


def test_shape_fit_per_precision():
    """Requested shapes become instantiated kernels: (8, 5, 3) keeps its
    14-qubit tile in FP32 (an FP32-only kernel) and falls back to W = 2 in
    FP64."""
    from dgquest_amd import quest as Q

    for prec, want in ((1, (14, 5, 3)), (2, (13, 5, 2))):
        lib = Q.load(prec)
        r = lib.dgq_create_recorder(20)
        lib.apply_circuit(r, circuits.layered_u3_cz(20, 2, 1))
        b, _, _, _ = lib.recorder_plan(r, k=8, R=5, W=3)
        assert {(int(x["tile_bits"]), int(x["reg_bits"]), int(x["wave_bits"])) for x in b} == {want}
        lib.dgq_destroy_recorder(r)


@pytest.mark.parametrize("which", ["layered", "random", "ghz_qft", "grid"])
@pytest.mark.parametrize("shape", [(4, 4, 0, 4096), (5, 4, 1, 4096), (6, 3, 3, 4096)])
def test_rotating_lanes(quest, oracle, which, shape):
    """Registers wider than the tile: blocks hand lane positions 0..5 over to
    the qubits needed next (stores that move qubits, the fusion state moving
    with them).  The plan ends with a permuted qubit map and matches the
    oracle once read through it."""
    n = 16 if which != "grid" else 15
    if which == "layered":
        circ = circuits.layered_u3_cz(n, 8, 3)
    elif which == "random":
        circ = H.random_circuit(np.random.default_rng(77), n, 200)
    elif which == "ghz_qft":
        circ = circuits.ghz_qft(n)
    else:
        circ = circuits.grid_circuit(3, 5, 10, 4)
    k, R, W, mo = shape
    q = quest.dgq_create_recorder(n)
    quest.apply_circuit(q, circ)
    _, _, _, l2p = quest.recorder_plan(q, k, R, W, mo, True)
    quest.dgq_destroy_recorder(q)
    assert l2p != list(range(n))  # lanes were handed over
    psi, _ = plan_and_run(quest, n, circ, k=k, R=R, W=W, max_ops=mo)
    ref = oracle_state(oracle, n, circ)
    assert np.abs(psi - ref).max() <= TOL


def test_lane_rotation_switch(quest, oracle):
    """dgq_set_lane_rotation(q, 0) keeps qubits 0..5 in every tile: the plan
    leaves the qubit map alone and needs at least as many sweeps as the
    rotating plan; both match the oracle."""
    from tests.plan_model import run_plan

    n = 16
    circ = circuits.layered_u3_cz(n, 8, 3)
    ref = oracle_state(oracle, n, circ)
    sweeps = {}
    for rot in (0, 1):
        q = quest.dgq_create_recorder(n)
        quest.dgq_set_lane_rotation(q, rot)
        quest.apply_circuit(q, circ)
        blocks, phases, ops, l2p = quest.recorder_plan(q, 6, 3, 3, 4096, True)
        quest.dgq_destroy_recorder(q)
        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1.0
        run_plan(psi, n, blocks, phases, ops)
        idx = np.arange(1 << n, dtype=np.int64)
        phys = np.zeros_like(idx)
        for lq in range(n):
            phys |= ((idx >> lq) & 1) << l2p[lq]
        assert np.abs(psi[phys] - ref).max() <= TOL
        assert (l2p == list(range(n))) == (rot == 0)
        sweeps[rot] = len(blocks)
    assert sweeps[1] < sweeps[0]


@pytest.mark.parametrize("which,n", [("layered", 30), ("hea", 30), ("ghz_qft", 30), ("grid", 28), ("random", 24),
                                     ("bv", 31)])
def test_rotation_never_adds_sweeps(quest, which, n):
    """Plans only (no execution) of the workload families at their benchmark
    sizes: with lane rotation on, never more sweeps than with it off, and
    strictly fewer on the layered, grid and random circuits."""
    if which == "layered":
        circ = circuits.layered_u3_cz(n, 24, 2)
    elif which == "hea":
        circ = circuits.hea(n, 10, 7)
    elif which == "ghz_qft":
        circ = circuits.ghz_qft(n)
    elif which == "grid":
        circ = circuits.grid_circuit(4, 7, 20, 4)
    elif which == "random":
        circ = H.random_circuit(np.random.default_rng(5), n, 600)
    else:
        circ = circuits.bernstein_vazirani(n - 1, 3)[0]
    sweeps = {}
    for rot in (0, 1):
        q = quest.dgq_create_recorder(n)
        quest.dgq_set_lane_rotation(q, rot)
        quest.apply_circuit(q, circ)
        blocks, _, _, _ = quest.recorder_plan(q, 7, 4, 3, 4096, True)
        quest.dgq_destroy_recorder(q)
        sweeps[rot] = len(blocks)
    assert sweeps[1] <= sweeps[0]
    if which in ("layered", "grid", "random"):
        assert sweeps[1] < sweeps[0]
