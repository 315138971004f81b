"""GPU: qubit moves (permutation sweeps) are exact.

Every amplitude is set to its own index, the qubit order is permuted with
swapGate (a relabelling only), then setAmps materialises the map with
permutation sweeps; afterwards every amplitude must be exactly the expected
one.  The shapes include sweeps that move qubits between wave bits inside one
phase (loads and stores of a tile then touch other waves' positions, so they
must be ordered inside the workgroup), in both precisions (indices up to
2^23 are exact in float)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(7, 4, 3), (6, 3, 3), (6, 4, 2), (5, 2, 3), (4, 4, 0)]


def permuted_index_state(quest, env, n, shape, swaps, prec):
    N = 1 << n
    q = quest.createQureg(n, env)
    quest.dgq_set_tuning(q, *shape, 4096)
    dt = np.float64 if prec == 2 else np.float32
    ptr = C.POINTER(C.c_double if prec == 2 else C.c_float)
    re = np.arange(N, dtype=dt)
    im = np.zeros(N, dtype=dt)
    quest.initStateFromAmps(q, re.ctypes.data_as(ptr), im.ctypes.data_as(ptr))
    l2p = list(range(n))
    for a, b in swaps:
        quest.swapGate(q, a, b)
        l2p[a], l2p[b] = l2p[b], l2p[a]
    quest.setAmps(q, 0, re.ctypes.data_as(ptr), im.ctypes.data_as(ptr), 0)  # materialises the map
    got = quest.state_vector(q).real.astype(np.int64)
    quest.destroyQureg(q, env)
    # logical index i holds the old amplitude whose bit l2p[b] is bit b of i
    idx = np.arange(N, dtype=np.int64)
    want = np.zeros(N, dtype=np.int64)
    for b in range(n):
        want |= ((idx >> b) & 1) << l2p[b]
    return got, want


@pytest.mark.parametrize("prec", [2, 1])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("n", [14, 20, 23])
def test_reversal_exact(quest, env, quest32, env32, prec, shape, n):
    qq, ee = (quest, env) if prec == 2 else (quest32, env32)
    swaps = [(i, n - 1 - i) for i in range(n // 2)]
    got, want = permuted_index_state(qq, ee, n, shape, swaps, prec)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{len(bad)} misplaced, first {bad[:4]}"


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_permutation_exact(quest, env, seed):
    n = 21
    rng = np.random.default_rng(seed)
    swaps = [tuple(int(x) for x in rng.choice(n, 2, replace=False)) for _ in range(12)]
    for shape in [(7, 4, 3), (6, 3, 3)]:
        got, want = permuted_index_state(quest, env, n, shape, swaps, 2)
        assert np.array_equal(got, want), shape
