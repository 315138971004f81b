"""GPU parity: the HIP path (through the QuEST C API) against the CPU oracle.

Bar (FP64): max |amplitude difference| <= 1e-12, total probability 1 within
1e-12, measurement outcomes identical for the same seed.
"""
import math

import numpy as np
import pytest

from dgquest_amd import circuits
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run_gpu(quest, env, n, circuit, tune=None, init=None):
    q = quest.createQureg(n, env)
    if init == "plus":
        quest.initPlusState(q)
    elif isinstance(init, int):
        quest.initClassicalState(q, init)
    if tune is not None:
        quest.dgq_set_tuning(q, *tune)
    quest.apply_circuit(q, circuit)
    sv = quest.state_vector(q)
    stats = quest.plan_stats(q)
    return q, sv, stats


def run_oracle(oracle, n, circuit, init=None):
    s = np.zeros(1 << n, dtype=np.complex128)
    if init == "plus":
        s[:] = 1.0 / math.sqrt(1 << n)
    elif isinstance(init, int):
        s[init] = 1.0
    else:
        s[0] = 1.0
    return oracle.run(s, circuit)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 16])
def test_random_circuits_all_gates(quest, env, oracle, n):
    rng = np.random.default_rng(1000 + n)
    circ = H.random_circuit(rng, n, 60 if n < 14 else 200)
    q, sv, st = run_gpu(quest, env, n, circ)
    ref = run_oracle(oracle, n, circ)
    err = np.abs(sv - ref).max()
    print(f"n={n} err={err:.3e} stats={st}", flush=True)
    assert err <= TOL
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("n", [7, 12, 17])
def test_diagonal_fusion_on_off(quest, env, oracle, n, fuse):
    """Real-rotation fusion and the plain complex path give the oracle's state."""
    rng = np.random.default_rng(300 + n)
    circ = H.random_circuit(rng, n, 120)
    q = quest.createQureg(n, env)
    quest.dgq_set_fusion(q, fuse)
    quest.apply_circuit(q, circ)
    sv = quest.state_vector(q)
    ref = run_oracle(oracle, n, circ)
    err = np.abs(sv - ref).max()
    print(f"n={n} fuse={fuse} err={err:.3e} stats={quest.plan_stats(q)}")
    assert err <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("shape", [(1, 1, 0, 64), (3, 3, 0, 64), (4, 4, 0, 8), (5, 3, 2, 4096), (6, 4, 2, 4096),
                                   (6, 5, 1, 4096), (7, 4, 3, 4096), (7, 5, 2, 4096), (2, 2, 1, 3), (4, 2, 2, 16),
                                   (6, 3, 3, 4096)])
def test_tile_shapes(quest, env, oracle, shape):
    """Every kernel shape (k, R, W, max_ops) on the same 16-qubit circuit."""
    n = 16
    rng = np.random.default_rng(77)
    circ = H.random_circuit(rng, n, 150)
    q, sv, st = run_gpu(quest, env, n, circ, tune=shape)
    ref = run_oracle(oracle, n, circ)
    err = np.abs(sv - ref).max()
    print(f"shape={shape} err={err:.3e} stats={st}", flush=True)
    assert err <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("init", ["plus", 5, 0])
def test_init_states(quest, env, oracle, init):
    n = 12
    rng = np.random.default_rng(5)
    circ = H.random_circuit(rng, n, 40)
    q, sv, _ = run_gpu(quest, env, n, circ, init=init)
    ref = run_oracle(oracle, n, circ, init=init)
    assert np.abs(sv - ref).max() <= TOL
    # init alone, no gates
    quest.initPlusState(q) if init == "plus" else quest.initClassicalState(q, init)
    sv2 = quest.state_vector(q)
    ref2 = run_oracle(oracle, n, [], init=init)
    assert np.abs(sv2 - ref2).max() == 0.0
    quest.destroyQureg(q, env)


def test_init_from_amps_and_set_amps(quest, env, oracle):
    import ctypes as C

    n = 10
    rng = np.random.default_rng(9)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    q = quest.createQureg(n, env)
    re = np.ascontiguousarray(v.real)
    im = np.ascontiguousarray(v.imag)
    dp = C.POINTER(C.c_double)
    quest.initStateFromAmps(q, re.ctypes.data_as(dp), im.ctypes.data_as(dp))
    circ = H.random_circuit(rng, n, 50)
    quest.apply_circuit(q, circ)
    ref = oracle.run(v.copy(), circ)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    # setAmps after a relabelling swap: the engine must move qubits home first
    quest.swapGate(q, 1, 8)
    quest.swapGate(q, 0, 9)
    ref = oracle.run(ref, [("swapGate", (1, 8), ()), ("swapGate", (0, 9), ())])
    new = rng.normal(size=7) + 1j * rng.normal(size=7)
    quest.setAmps(q, 100, np.ascontiguousarray(new.real).ctypes.data_as(dp),
                  np.ascontiguousarray(new.imag).ctypes.data_as(dp), 7)
    ref[100:107] = new
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    quest.destroyQureg(q, env)


def test_ghz_qft_20_closed_form(quest, env, oracle):
    """Config 1: 20-qubit GHZ + textbook QFT; the full vector against the
    closed form and the oracle, total probability 1."""
    n = 20
    circ = circuits.ghz_qft(n)
    q, sv, st = run_gpu(quest, env, n, circ)
    cf = circuits.ghz_qft_closed_form(n)
    ref = run_oracle(oracle, n, circ)
    print(f"stats={st} err_cf={np.abs(sv - cf).max():.3e} err_oracle={np.abs(sv - ref).max():.3e}")
    assert np.abs(sv - cf).max() <= TOL
    assert np.abs(sv - ref).max() <= TOL
    assert abs(quest.calcTotalProb(q) - 1.0) <= TOL
    for t in range(n):
        p1 = quest.calcProbOfOutcome(q, t, 1)
        assert abs(p1 - oracle.prob_of_outcome(ref, t, 1)) <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("n,depth,seed", [(14, 6, 1), (20, 8, 2), (22, 4, 3)])
def test_layered_u3_cz(quest, env, oracle, n, depth, seed):
    circ = circuits.layered_u3_cz(n, depth, seed)
    q, sv, st = run_gpu(quest, env, n, circ)
    ref = run_oracle(oracle, n, circ)
    err = np.abs(sv - ref).max()
    print(f"n={n} err={err:.3e} stats={st}")
    assert err <= TOL
    quest.destroyQureg(q, env)


def test_measurement_bit_exact(quest, env, oracle):
    """Same seed -> same outcomes and same post-measurement state."""
    n = 12
    rng = np.random.default_rng(21)
    for trial in range(4):
        circ = H.random_circuit(rng, n, 40)
        q = quest.createQureg(n, env)
        quest.apply_circuit(q, circ)
        ref = run_oracle(oracle, n, circ)
        keys = [trial + 1, 0xBEEF]
        quest.seed(keys)
        oracle.seed(keys)
        order = [int(x) for x in rng.permutation(n)]
        for t in order:
            o_gpu = quest.measure(q, t)
            o_ref, _ = oracle.measure(ref, t)
            assert o_gpu == o_ref
        assert np.abs(quest.state_vector(q) - ref).max() <= TOL
        quest.destroyQureg(q, env)


def test_collapse_and_probabilities(quest, env, oracle):
    n = 9
    rng = np.random.default_rng(3)
    circ = H.random_circuit(rng, n, 30)
    q = quest.createQureg(n, env)
    quest.apply_circuit(q, circ)
    ref = run_oracle(oracle, n, circ)
    for t in range(n):
        for o in (0, 1):
            assert abs(quest.calcProbOfOutcome(q, t, o) - oracle.prob_of_outcome(ref, t, o)) <= TOL
    p = quest.collapseToOutcome(q, 4, 1)
    pr = oracle.collapse(ref, 4, 1)
    assert abs(p - pr) <= TOL
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    amp = quest.getAmp(q, 37)
    assert abs(complex(amp.real, amp.imag) - ref[37]) <= TOL
    assert abs(quest.getProbAmp(q, 37) - abs(ref[37]) ** 2) <= TOL
    quest.destroyQureg(q, env)


def test_swap_relabel_matches_textbook_swap(quest, env, oracle):
    n = 13
    rng = np.random.default_rng(8)
    circ = H.random_circuit(rng, n, 20)
    circ += [("swapGate", (0, 12), ()), ("swapGate", (3, 7), ())] + H.random_circuit(rng, n, 40)
    q, sv, _ = run_gpu(quest, env, n, circ)
    ref = run_oracle(oracle, n, circ)
    assert np.abs(sv - ref).max() <= TOL
    quest.destroyQureg(q, env)


def test_f32_fidelity(quest32, env32, oracle):
    """FP32 build: fidelity against the FP64 oracle >= 1 - 1e-6."""
    n = 16
    circ = circuits.layered_u3_cz(n, 10, 7)
    q = quest32.createQureg(n, env32)
    quest32.apply_circuit(q, circ)
    sv = quest32.state_vector(q)
    ref = run_oracle(oracle, n, circ)
    fid = abs(np.vdot(ref, sv)) ** 2
    err = np.abs(sv - ref).max()
    print(f"f32 fidelity 1-{1 - fid:.3e} max err {err:.3e}")
    assert fid >= 1 - 1e-6
    assert err <= 1e-5
    quest32.destroyQureg(q, env32)


def test_30q_layered_norm_and_inverse(quest, env):
    """Config-2 size (30 qubits, 16 GiB): norm stays 1, and applying the
    circuit's inverse returns |0...0> (size-independent properties)."""
    n = 30
    circ = circuits.layered_u3_cz(n, 4, 2)
    q = quest.createQureg(n, env)
    quest.apply_circuit(q, circ)
    tp = quest.calcTotalProb(q)
    assert abs(tp - 1.0) <= 1e-12
    inv = []
    for g in reversed(circ):
        name, qs, p = g
        if name == "u3Gate":
            th, ph, la = p
            inv.append(("u3Gate", qs, (-th, -la, -ph)))
        else:
            inv.append(g)
    quest.apply_circuit(q, inv)
    a0 = quest.getAmp(q, 0)
    assert abs(a0.real - 1.0) <= 1e-12 and abs(a0.imag) <= 1e-12
    assert abs(quest.calcProbOfOutcome(q, 29, 0) - 1.0) <= 1e-12
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("shape", [(4, 4, 0, 4096), (6, 4, 2, 4096), (7, 5, 2, 4096), (7, 4, 3, 4096),
                                   (6, 3, 3, 4096), (8, 5, 3, 4096)])
def test_tile_shapes_fp32(quest32, env32, oracle, shape):
    """FP32 kernels (split exchanges, and the FP32-only 14-bit tile) on a
    17-qubit circuit: max |delta amplitude| <= 1e-5 and fidelity >= 1 - 1e-6
    against the FP64 oracle."""
    n = 17
    rng = np.random.default_rng(78)
    circ = H.random_circuit(rng, n, 150)
    q, sv, st = run_gpu(quest32, env32, n, circ, tune=shape)
    ref = run_oracle(oracle, n, circ)
    err = np.abs(sv - ref).max()
    fid = abs(np.vdot(ref, sv)) ** 2 / np.vdot(sv, sv).real
    print(f"fp32 shape={shape} err={err:.3e} fid={fid:.9f} stats={st}", flush=True)
    assert err <= 1e-5 and fid >= 1 - 1e-6
    quest32.destroyQureg(q, env32)


@pytest.mark.parametrize("rot", [1, 0])
@pytest.mark.parametrize("n,which", [(16, "random"), (18, "layered"), (16, "grid")])
def test_lane_rotation_on_off(quest, env, oracle, rot, n, which):
    """Registers wider than the tile (12 qubits here): flushes whose sweeps
    hand lane positions 0..5 over to other qubits (rotation on) and flushes
    that keep qubits 0..5 in every tile (off) both match the oracle, read
    through the register's qubit map; measurement outcomes from one seed then
    match too, and so does the collapsed state."""
    if which == "random":
        circ = H.random_circuit(np.random.default_rng(600 + n), n, 200)
    elif which == "layered":
        circ = circuits.layered_u3_cz(n, 6, 5)
    else:
        circ = circuits.grid_circuit(4, 4, 10, 4)
    q = quest.createQureg(n, env)
    quest.dgq_set_tuning(q, 6, 3, 3, 4096)
    quest.dgq_set_lane_rotation(q, rot)
    quest.apply_circuit(q, circ)
    ref = run_oracle(oracle, n, circ)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    keys = [n, rot, 7]
    quest.seed(keys)
    oracle.seed(keys)
    for t in [int(x) for x in np.random.default_rng(n).permutation(n)][:6]:
        assert quest.measure(q, t) == oracle.measure(ref, t)[0]
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    quest.destroyQureg(q, env)


@pytest.mark.parametrize("tune", [None, (6, 3, 3, 4096)])
def test_random_22q_default_plan(quest, env, oracle, tune):
    """A 22-qubit random circuit over the whole gate set (relabelling swaps,
    fSim, multi-controlled gates) with the register's default plan (lane
    rotation on; the tuning table's shape, or a 12-qubit tile): amplitudes
    within 1e-12 of the oracle, then seeded measurements and the collapsed
    state."""
    n = 22
    circ = H.random_circuit(np.random.default_rng(2222), n, 300)
    q = quest.createQureg(n, env)
    if tune is not None:
        quest.dgq_set_tuning(q, *tune)
    quest.apply_circuit(q, circ)
    ref = run_oracle(oracle, n, circ)
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    keys = [22, 5]
    quest.seed(keys)
    oracle.seed(keys)
    for t in (21, 0, 13, 6):
        assert quest.measure(q, t) == oracle.measure(ref, t)[0]
    assert np.abs(quest.state_vector(q) - ref).max() <= TOL
    quest.destroyQureg(q, env)


def _inverse_layered(circ):
    inv = []
    for name, qs, p in reversed(circ):
        if name == "u3Gate":
            th, ph, la = p
            inv.append(("u3Gate", qs, (-th, -la, -ph)))
        else:
            inv.append((name, qs, p))
    return inv


# The 30-qubit benchmark's own plan: fusion width 7 on a 13-qubit tile of
# 4 register and 3 wave bits, one block per sweep (max_ops large), lane
# rotation on (the register default).
BENCH_TUNE = (7, 4, 3, 100000)


def test_30q_bench_plan_norm_and_inverse(quest, env):
    """The benchmark circuit itself (30 qubits, depth 24, seed 2) with the
    benchmark's plan: total probability 1 within 1e-12, then the circuit's
    inverse, planned the same way, returns |0...0> within 1e-12."""
    n = 30
    circ = circuits.layered_u3_cz(n, 24, 2)
    q = quest.createQureg(n, env)
    quest.dgq_set_tuning(q, *BENCH_TUNE)
    quest.dgq_set_lane_rotation(q, 1)
    quest.apply_circuit(q, circ)
    tp = quest.calcTotalProb(q)
    st = quest.plan_stats(q)
    print(f"30q bench plan: total prob 1{tp - 1:+.2e} stats={st}", flush=True)
    assert abs(tp - 1.0) <= 1e-12
    quest.apply_circuit(q, _inverse_layered(circ))
    a0 = quest.getAmp(q, 0)
    print(f"after inverse: a0 = {a0}", flush=True)
    assert abs(a0.real - 1.0) <= 1e-12 and abs(a0.imag) <= 1e-12
    for t in (0, 17, 29):
        assert abs(quest.calcProbOfOutcome(q, t, 0) - 1.0) <= 1e-12
    quest.destroyQureg(q, env)


def test_26q_bench_plan_vs_oracle(quest, env, oracle):
    """The benchmark's circuit family and plan (depth 24, 13-qubit tile, k=7,
    lane rotation) at 26 qubits, where 13 qubits index tiles: every amplitude
    within 1e-12 of the oracle, read through the register's qubit map."""
    n = 26
    circ = circuits.layered_u3_cz(n, 24, 2)
    q = quest.createQureg(n, env)
    quest.dgq_set_tuning(q, *BENCH_TUNE)
    quest.dgq_set_lane_rotation(q, 1)
    quest.apply_circuit(q, circ)
    st = quest.plan_stats(q)
    sv = quest.state_vector(q)
    ref = run_oracle(oracle, n, circ)
    err = float(np.abs(sv - ref).max())
    print(f"26q bench plan: err={err:.3e} stats={st}", flush=True)
    assert err <= TOL
    quest.destroyQureg(q, env)
