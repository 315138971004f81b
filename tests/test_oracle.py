"""CPU checks of the oracle against fixtures and mathematics (no GPU).

Pins: the reference's own MT19937 (golden streams, and the live build in
oracle/_ref when present), the published mt19937ar known answers, a dense
numpy brute force independent of the oracle code (tests/golden/states.npz),
closed forms (GHZ + QFT, Bernstein-Vazirani), and invariants (norm, inverse
circuit).  One test plants a mistake in the oracle to show the pins catch it.
"""
import json
import math
import os

import numpy as np
import pytest

from dgquest_amd import circuits
from dgquest_amd.circuits import MT19937
from tests import helpers as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_mt_published_known_answers(oracle):
    # mt19937ar.out: init_by_array({0x123, 0x234, 0x345, 0x456}, 4)
    oracle.seed([0x123, 0x234, 0x345, 0x456])
    got = [oracle.lib.oc_mt_genrand_int32() for _ in range(5)]
    assert got == [1067595299, 955945823, 477289528, 4107218783, 4228976476]


def test_mt_oracle_matches_reference_streams(oracle):
    kat = load("mt19937_kat.json")
    for s in kat["seeds"]:
        oracle.seed(s["keys"])
        assert [oracle.lib.oc_mt_genrand_int32() for _ in range(32)] == s["int32"]
        assert [oracle.lib.oc_mt_genrand_real1() for _ in range(16)] == s["real1"]
        assert [oracle.lib.oc_mt_genrand_res53() for _ in range(16)] == s["res53"]
    lg = kat["long"]
    oracle.seed(lg["keys"])
    stream = [oracle.lib.oc_mt_genrand_int32() for _ in range(lg["from"] + len(lg["int32"]))]
    assert stream[lg["from"]:] == lg["int32"]


def test_mt_python_generator_matches_reference_streams():
    kat = load("mt19937_kat.json")
    for s in kat["seeds"]:
        g = MT19937(s["keys"])
        assert [g.genrand_int32() for _ in range(32)] == s["int32"]
        assert [g.genrand_real1() for _ in range(16)] == s["real1"]
        assert [g.genrand_res53() for _ in range(16)] == s["res53"]


def test_mt_live_reference_build(oracle):
    from oracle.binding import REF_MT, ReferenceMT

    if not os.path.exists(REF_MT):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref = ReferenceMT()
    for keys in ([7], [1, 2, 3], [0xFFFFFFFF, 0]):
        ref.init_by_array(keys)
        oracle.seed(keys)
        assert [ref.int32() for _ in range(2000)] == [oracle.lib.oc_mt_genrand_int32() for _ in range(2000)]


def _cases():
    return load("circuits.json")


@pytest.mark.parametrize("i", range(81))
def test_oracle_matches_dense_brute_force(oracle, i):
    cases = _cases()
    if i >= len(cases):
        pytest.skip("fewer cases")
    st = np.load(os.path.join(GOLD, "states.npz"))
    cs = cases[i]
    v = st[f"in_{i}"].copy()
    out = oracle.run(v, [(g[0], tuple(g[1]), tuple(g[2])) for g in cs["circuit"]])
    assert np.abs(out - st[f"out_{i}"]).max() <= 1e-13, cs["name"]


def test_planted_mistake_is_caught(oracle):
    """Swap PAULI_Y for its conjugate (a sign error) inside the oracle and
    check that the brute-force fixture rejects it."""
    cases = _cases()
    st = np.load(os.path.join(GOLD, "states.npz"))
    i = next(k for k, c in enumerate(cases) if c["name"].startswith("pauliY"))
    cs = cases[i]
    v = st[f"in_{i}"].copy()
    t = cs["circuit"][0][1][0]
    oracle.apply_record(v, 13, 1 << t, 0)  # type 13 = Y conjugate: the planted mistake
    assert np.abs(v - st[f"out_{i}"]).max() > 1e-3


def test_measurement_outcomes_match_reference_mt(oracle):
    cases = _cases()
    st = np.load(os.path.join(GOLD, "states.npz"))
    for m in load("measure.json"):
        cs = cases[m["case"]]
        v = st[f"in_{m['case']}"].copy()
        oracle.run(v, [(g[0], tuple(g[1]), tuple(g[2])) for g in cs["circuit"]])
        oracle.seed(m["keys"])
        got = [oracle.measure(v, q)[0] for q in m["order"]]
        assert got == m["outcomes"]


@pytest.mark.parametrize("n", [6, 10, 14])
def test_ghz_qft_closed_form(oracle, n):
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    oracle.run(v, circuits.ghz_qft(n))
    assert np.abs(v - circuits.ghz_qft_closed_form(n)).max() <= 1e-13
    assert abs(oracle.total_prob(v) - 1) <= 1e-13


def test_ghz_qft_closed_form_matches_fft():
    # the closed form itself, against numpy's FFT of the GHZ vector
    n = 10
    N = 1 << n
    ghz = np.zeros(N, dtype=complex)
    ghz[0] = ghz[-1] = 1 / math.sqrt(2)
    f = np.fft.ifft(ghz) * math.sqrt(N)  # exp(+2 pi i x y / N)
    i = np.arange(N)
    rev = np.zeros(N, dtype=np.int64)
    for b in range(n):
        rev |= ((i >> b) & 1) << (n - 1 - b)
    assert np.abs(circuits.ghz_qft_closed_form(n) - f[rev]).max() <= 1e-13


def test_bernstein_vazirani_recovers_string(oracle):
    c, s = circuits.bernstein_vazirani(12, 3)
    n = 13
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    oracle.run(v, c)
    # data register in |s>, ancilla in |->: amplitudes +-1/sqrt2 at s and s + 2^12
    assert abs(abs(v[s]) - 1 / math.sqrt(2)) <= 1e-13
    assert abs(abs(v[s | (1 << 12)]) - 1 / math.sqrt(2)) <= 1e-13
    for q in range(12):
        assert abs(oracle.prob_of_outcome(v, q, (s >> q) & 1) - 1) <= 1e-13


@pytest.mark.parametrize("n", [3, 5, 7])
def test_tensor_oracle_matches_dense_brute_force(n):
    """The tensordot oracle (one axis per qubit) agrees with the Kronecker
    brute force on random circuits over the whole gate set."""
    rng = np.random.default_rng(300 + n)
    circ = H.random_circuit(rng, n, 60)
    assert np.abs(H.tensor_run(n, circ) - H.dense_run(n, circ)).max() <= 1e-12


def test_tensor_oracle_planted_mistake():
    """A transposed 2x2 matrix in the tensordot oracle is caught by the brute
    force (the target-axis placement is not symmetric in the operands)."""
    rng = np.random.default_rng(7)
    n = 4
    circ = [g for g in H.random_circuit(rng, n, 40, allow={"u3Gate", "controlledNot", "unitary"})]
    good = H.tensor_run(n, circ)
    psi = np.zeros((2,) * n, dtype=complex)
    psi[(0,) * n] = 1
    for g in circ:
        if g[0] == "u3Gate":
            H._apply_1q(psi, n, g[1][0], H.mat1("u3Gate", g[2]).T)
        else:
            H.tensor_gate(psi, n, g)
    assert np.abs(psi.reshape(-1) - H.dense_run(n, circ)).max() > 1e-3
    assert np.abs(good - H.dense_run(n, circ)).max() <= 1e-12


@pytest.mark.parametrize("which", ["random", "layered", "grid"])
def test_oracle_matches_tensor_oracle_20q(oracle, which):
    """Two independent oracles at 20 qubits: the C per-gate loop and the numpy
    tensordot oracle, on a random circuit over the whole gate set, the
    layered U3 + CZ workload (4 layers) and a 4 x 5 grid circuit."""
    n = 20
    if which == "random":
        circ = H.random_circuit(np.random.default_rng(2020), n, 60)
    elif which == "layered":
        circ = circuits.layered_u3_cz(n, 4, 2)
    else:
        circ = circuits.grid_circuit(4, 5, 6, 4)
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    got = oracle.run(v, circ)
    want = H.tensor_run(n, circ)
    assert np.abs(got - want).max() <= 1e-12
    assert abs(np.vdot(got, got).real - 1.0) <= 1e-12


@pytest.mark.parametrize("n", [8, 12])
def test_layered_circuit_norm_and_inverse(oracle, n):
    c = circuits.layered_u3_cz(n, 6, 9)
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    oracle.run(v, c)
    assert abs(oracle.total_prob(v) - 1) <= 1e-12
    inv = []
    for g in reversed(c):
        if g[0] == "u3Gate":
            th, ph, la = g[2]
            inv.append(("u3Gate", g[1], (-th, -la, -ph)))
        else:
            inv.append(g)
    oracle.run(v, inv)
    assert abs(v[0] - 1) <= 1e-12


def test_layered_generator_structure():
    c = circuits.layered_u3_cz(30, 24, 2)
    assert len(c) == 24 * (30 + 15)
    for L in range(24):
        layer = c[L * 45:(L + 1) * 45]
        assert [g[1][0] for g in layer[:30]] == list(range(30))
        pairs = [g[1] for g in layer[30:]]
        assert sorted(q for p in pairs for q in p) == list(range(30))
        for g in layer[:30]:
            assert all(0 <= a < 2 * math.pi for a in g[2])


def test_grid_generator_no_repeat():
    c = circuits.grid_random(4, 8, 20, 4)
    last = {}
    for g in c:
        if g[0] in ("SqX", "SqY", "SqW"):
            q = g[1][0]
            assert last.get(q) != g[0]
            last[q] = g[0]


@pytest.mark.parametrize("n", [3, 6, 9, 11])
def test_ghz_qft_testcase_closed_form(oracle, n):
    """The closed form used to check the reference's GHZ_QFT-30 program run
    through the drop-in: pinned against the oracle (every amplitude) and the
    dense brute force at small n; P(q = 1) = 1/2 for every qubit."""
    c = circuits.ghz_qft_testcase(n)
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    v = oracle.run(s, c)
    cf = circuits.ghz_qft_testcase_amps(n, np.arange(1 << n))
    assert np.abs(v - cf).max() <= 1e-13
    if n <= 9:
        assert np.abs(H.dense_run(n, c) - cf).max() <= 1e-12
    p = np.abs(cf) ** 2
    idx = np.arange(1 << n)
    for q in range(n):
        assert abs(p[((idx >> q) & 1) == 1].sum() - 0.5) <= 1e-13
    # the mistyped pi matters: with the true pi the amplitudes differ
    assert np.abs(circuits.ghz_qft_testcase_amps(n, idx, math.pi) - cf).max() > 1e-6
