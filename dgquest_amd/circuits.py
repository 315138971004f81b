"""Synthetic circuits of the benchmark workloads, as (name, qubits, params) lists.

Every random choice comes from MT19937 seeded with init_by_array([seed]):
angles are 2*pi*genrand_res53(), shuffles are Fisher-Yates with
genrand_int32() % (i + 1).  The same seed gives the same circuit on every
machine, so GPU runs and the CPU oracle see identical gate lists.

Gate tuples use the QuEST API names; see ``dgquest_amd.quest.GATE_SIGNATURES``.
"""
from __future__ import annotations

import math
from typing import List, Tuple

Gate = Tuple[str, tuple, tuple]


class MT19937:
    """Matsumoto-Nishimura MT19937 (mt19937ar), pure Python."""

    N, M = 624, 397

    def __init__(self, seed_keys=(5489,)):
        self.mt = [0] * self.N
        self.mti = self.N + 1
        self.init_by_array(list(seed_keys))

    def init_genrand(self, s):
        mt = self.mt
        mt[0] = s & 0xFFFFFFFF
        for i in range(1, self.N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mti = self.N

    def init_by_array(self, key):
        self.init_genrand(19650218)
        mt, N = self.mt, self.N
        i, j = 1, 0
        for _ in range(max(N, len(key))):
            mt[i] = ((mt[i] ^ (((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525) & 0xFFFFFFFF)) + key[j] + j) & 0xFFFFFFFF
            i += 1
            j += 1
            if i >= N:
                mt[0] = mt[N - 1]
                i = 1
            if j >= len(key):
                j = 0
        for _ in range(N - 1):
            mt[i] = ((mt[i] ^ (((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941) & 0xFFFFFFFF)) - i) & 0xFFFFFFFF
            i += 1
            if i >= N:
                mt[0] = mt[N - 1]
                i = 1
        mt[0] = 0x80000000

    def genrand_int32(self):
        mt, N, M = self.mt, self.N, self.M
        if self.mti >= N:
            for kk in range(N):
                y = (mt[kk] & 0x80000000) | (mt[(kk + 1) % N] & 0x7FFFFFFF)
                mt[kk] = mt[(kk + M) % N] ^ (y >> 1) ^ (0x9908B0DF if (y & 1) else 0)
            self.mti = 0
        y = mt[self.mti]
        self.mti += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    def genrand_real1(self):
        return self.genrand_int32() * (1.0 / 4294967295.0)

    def genrand_res53(self):
        a = self.genrand_int32() >> 5
        b = self.genrand_int32() >> 6
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0)


def ghz_qft(n: int) -> List[Gate]:
    """GHZ then a textbook QFT: H(0), CNOT chain 0->1->..->n-1; for each qubit j
    (qubit 0 first) H(j) then controlled-phase pi/2^m from qubit j+m; then the
    qubit-order reversal by swaps."""
    c: List[Gate] = [("hadamard", (0,), ())]
    for i in range(n - 1):
        c.append(("controlledNot", (i, i + 1), ()))
    for j in range(n):
        c.append(("hadamard", (j,), ()))
        for m in range(1, n - j):
            c.append(("controlledPhaseShift", (j, j + m), (math.pi / (1 << m),)))
    for i in range(n // 2):
        c.append(("swapGate", (i, n - 1 - i), ()))
    return c


def ghz_qft_closed_form(n: int):
    """a_i = 2^{-(n+1)/2} (1 + exp(-2 pi i rev(i) / 2^n)), rev = n-bit reversal.

    Qubit 0 plays the most significant input bit of the textbook circuit, so
    the output amplitude of index i belongs to y = rev(i)."""
    import numpy as np

    N = 1 << n
    i = np.arange(N, dtype=np.int64)
    rev = np.zeros(N, dtype=np.int64)
    for b in range(n):
        rev |= ((i >> b) & 1) << (n - 1 - b)
    return (1.0 + np.exp(-2j * np.pi * rev / N)) / math.sqrt(2 * N)


# The angle constant of the reference's GHZ_QFT testcases
# (testcases/GHZ_QFT-30.c:6): pi mistyped as 3.1415265358979...
PI_GHZ_QFT_TESTCASE = 3.141526535897932384626


def ghz_qft_testcase(n: int, pi_: float = PI_GHZ_QFT_TESTCASE) -> List[Gate]:
    """The circuit of the reference's GHZ_QFT-<n> programs
    (testcases/GHZ_QFT-30.c:25-43): GHZ (H(0), CNOT chain), then H(0) and, for
    i = 1 .. n-2, controlled-RZ(pi_ / 2^(i-j)) from each j < i onto i followed
    by H(i); qubit n-1 is left as GHZ made it."""
    c: List[Gate] = [("hadamard", (0,), ())]
    for i in range(1, n):
        c.append(("controlledNot", (i - 1, i), ()))
    c.append(("hadamard", (0,), ()))
    for i in range(1, n - 1):
        for j in range(i):
            c.append(("controlledRotateZ", (j, i), (pi_ * 0.5 ** (i - j),)))
        c.append(("hadamard", (i,), ()))
    return c


def ghz_qft_testcase_amps(n: int, idx, pi_: float = PI_GHZ_QFT_TESTCASE):
    """Amplitudes of ghz_qft_testcase(n) at basis indices idx, in closed form.

    In the branch |b...b> of the GHZ state every target of a controlled RZ is
    still the basis state |b> when the rotation acts, so each RZ only puts a
    phase exp(i s theta / 2) (s = -1 for b = 0, +1 for b = 1) on the |1> part
    of its control.  The branch ends as the product over j <= n-2 of
    (|0> + (-1)^b exp(i phi_j) |1>) / sqrt(2), phi_j = s * sum_{i=j+1}^{n-2}
    pi_ / 2^(i-j) / 2, times |b> on qubit n-1.  So a(y) = 2^{-n/2} * product
    over the set bits j <= n-2 of y of (-1)^b exp(i phi_j), b = bit n-1 of y."""
    import numpy as np

    idx = np.asarray(idx, dtype=np.int64)
    half = np.array([sum(pi_ * 0.5 ** (i - j) for i in range(j + 1, n - 1)) / 2.0 for j in range(n - 1)])
    b = (idx >> (n - 1)) & 1
    s = np.where(b == 1, 1.0, -1.0)
    out = np.full(idx.shape, 2.0 ** (-n / 2), dtype=np.complex128)
    for j in range(n - 1):
        on = ((idx >> j) & 1) == 1
        out = np.where(on, out * np.where(b == 1, -1.0, 1.0) * np.exp(1j * s * half[j]), out)
    return out


def layered_u3_cz(n: int, depth: int, seed: int) -> List[Gate]:
    """Per layer: U3(theta, phi, lambda) on every qubit (angles U[0, 2pi)), then
    CZ on a seeded perfect matching (shuffle, paired consecutively)."""
    rng = MT19937([seed])
    c: List[Gate] = []
    twopi = 2.0 * math.pi
    for _ in range(depth):
        for q in range(n):
            th, ph, la = (twopi * rng.genrand_res53() for _ in range(3))
            c.append(("u3Gate", (q,), (th, ph, la)))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.genrand_int32() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for k in range(0, n - 1, 2):
            c.append(("controlledPhaseFlip", (perm[k], perm[k + 1]), ()))
    return c


def bernstein_vazirani(n_data: int, seed: int):
    """n_data data qubits + ancilla (qubit n_data) in |->: X,H on the ancilla, H on
    data, CNOT data->ancilla for each set bit of a seeded hidden string
    (density 0.5), H on data.  Returns (circuit, hidden_string)."""
    rng = MT19937([seed])
    s = 0
    for q in range(n_data):
        if rng.genrand_int32() & 1:
            s |= 1 << q
    anc = n_data
    c: List[Gate] = [("pauliX", (anc,), ()), ("hadamard", (anc,), ())]
    for q in range(n_data):
        c.append(("hadamard", (q,), ()))
    for q in range(n_data):
        if (s >> q) & 1:
            c.append(("controlledNot", (q, anc), ()))
    for q in range(n_data):
        c.append(("hadamard", (q,), ()))
    return c, s


def grid_random(rows: int, cols: int, depth: int, seed: int) -> List[Gate]:
    """2D-grid random circuit: per cycle a sqrt(X)/sqrt(Y)/sqrt(W) per qubit chosen
    uniformly (never the same gate twice in a row on a qubit), then CZ on one
    of four nearest-neighbour edge patterns in rotation."""
    rng = MT19937([seed])
    n = rows * cols
    names = ("SqX", "SqY", "SqW")
    last = [-1] * n
    c: List[Gate] = []

    def q(r, k):
        return r * cols + k

    patterns = []
    # 0: horizontal edges starting at even column, 1: odd column,
    # 2: vertical edges starting at even row, 3: odd row
    for par in (0, 1):
        patterns.append([(q(r, k), q(r, k + 1)) for r in range(rows) for k in range(par, cols - 1, 2)])
    for par in (0, 1):
        patterns.append([(q(r, k), q(r + 1, k)) for r in range(par, rows - 1, 2) for k in range(cols)])
    for cyc in range(depth):
        for i in range(n):
            while True:
                g = rng.genrand_int32() % 3
                if g != last[i]:
                    break
            last[i] = g
            c.append((names[g], (i,), ()))
        for a, b in patterns[cyc % 4]:
            c.append(("controlledPhaseFlip", (a, b), ()))
    return c


def hea(n: int, depth: int, seed: int) -> List[Gate]:
    """Hardware-efficient ansatz: per layer RY(theta) then RZ(phi) on every
    qubit (angles U[0, 2pi)), then a ring of CNOTs i -> i+1 (and n-1 -> 0)."""
    rng = MT19937([seed])
    c: List[Gate] = []
    twopi = 2.0 * math.pi
    for _ in range(depth):
        for q in range(n):
            c.append(("rotateY", (q,), (twopi * rng.genrand_res53(),)))
        for q in range(n):
            c.append(("rotateZ", (q,), (twopi * rng.genrand_res53(),)))
        if n >= 2:
            for q in range(n if n > 2 else 1):
                c.append(("controlledNot", (q, (q + 1) % n), ()))
    return c


def serving_mix(count: int, seed: int = 5, rate: float = 100.0, n_min: int = 10, n_max: int = 30,
                build: bool = True):
    """A request stream of the serving workload: Poisson arrivals at `rate`
    per second; qubit counts log-uniform on [n_min, n_max] (n = floor of
    exp(U[ln n_min, ln(n_max + 1)))); 40% hardware-efficient ansatz (depth
    4..16), 30% GHZ + QFT, 30% layered U3 + CZ (depth 8..24).  Returns a list
    of dicts (kind, n, depth, circuit_seed, shot_seed, arrive_s, circuit);
    with build=False the circuits are left out (see mix_circuit)."""
    rng = MT19937([seed])
    lo, hi = math.log(n_min), math.log(n_max + 1)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(1.0 - rng.genrand_res53()) / rate
        n = min(n_max, int(math.exp(lo + (hi - lo) * rng.genrand_res53())))
        u = rng.genrand_res53()
        cseed = rng.genrand_int32()
        sseed = rng.genrand_int32()
        if u < 0.4:
            kind, depth = "hea", 4 + rng.genrand_int32() % 13
        elif u < 0.7:
            kind, depth = "ghz_qft", 0
        else:
            kind, depth = "u3_cz", 8 + rng.genrand_int32() % 17
        d = {"kind": kind, "n": n, "depth": depth, "circuit_seed": cseed, "shot_seed": sseed, "arrive_s": t}
        if build:
            d["circuit"] = mix_circuit(d)
        out.append(d)
    return out


def mix_circuit(d) -> List[Gate]:
    """The circuit of one serving_mix entry."""
    if d["kind"] == "hea":
        return hea(d["n"], d["depth"], d["circuit_seed"])
    if d["kind"] == "ghz_qft":
        return ghz_qft(d["n"])
    return layered_u3_cz(d["n"], d["depth"], d["circuit_seed"])


def grid_circuit(rows: int, cols: int, depth: int, seed: int) -> List[Gate]:
    """Random circuit on a rows x cols grid (qubit = r * cols + c): per cycle a
    sqrt(X), sqrt(Y) or sqrt(W) on every qubit, drawn uniformly from the seed
    and never the same gate twice in a row on a qubit; then CZ on one of four
    nearest-neighbour edge patterns in rotation (cycle i uses pattern i % 4:
    horizontal pairs from even columns, from odd columns, vertical pairs from
    even rows, from odd rows)."""
    rng = MT19937([seed])
    names = ("SqX", "SqY", "SqW")
    last = [-1] * (rows * cols)
    c: List[Gate] = []
    for cyc in range(depth):
        for q in range(rows * cols):
            if last[q] < 0:
                g = rng.genrand_int32() % 3
            else:
                g = (last[q] + 1 + rng.genrand_int32() % 2) % 3
            last[q] = g
            c.append((names[g], (q,), ()))
        pat = cyc % 4
        if pat < 2:
            for r in range(rows):
                for col in range(pat, cols - 1, 2):
                    c.append(("controlledPhaseFlip", (r * cols + col, r * cols + col + 1), ()))
        else:
            for r in range(pat - 2, rows - 1, 2):
                for col in range(cols):
                    c.append(("controlledPhaseFlip", (r * cols + col, (r + 1) * cols + col), ()))
    return c
