"""GPU: the allocation-free P(0) read-out (dq_prob_zero_all_noalloc) against
the workspace entry point it is a variant of (dq_prob_zero_all).

Two calls in a row, the second on a smaller register, must each equal the
workspace entry point bit for bit (up to 30 qubits both run the same
launches); a 31-qubit FP32 register takes the chunked path and must agree to
1e-12.  Sizes above the limit are refused.
"""
import ctypes as C

import numpy as np
import pytest

from dgquest_amd import circuits
from dgquest_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _lib():
    L = Engine().lib
    vp = C.c_void_p
    L.dq_prob_ws_bytes.restype = C.c_size_t
    L.dq_prob_ws_bytes.argtypes = [C.c_int]
    L.dq_prob_zero_all.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), vp, C.c_size_t, vp]
    L.dq_prob_zero_all_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    return L


def _ck(L, rc):
    assert rc == 0, L.dq_last_error().decode()


def _both(L, api, q, n):
    """(workspace entry point, allocation-free entry point) results for q."""
    prec = C.c_int()
    d_state = api.dgq_state_pointer(q, C.byref(prec))
    assert d_state
    wsb = L.dq_prob_ws_bytes(n)
    ws, d_out = C.c_void_p(), C.c_void_p()
    _ck(L, L.dq_malloc(C.byref(ws), wsb))
    _ck(L, L.dq_malloc(C.byref(d_out), 64 * 8))  # the caller's output buffer
    try:
        ref = (C.c_double * (n + 1))()
        _ck(L, L.dq_prob_zero_all(d_state, n, prec.value, ref, ws, wsb, None))
        _ck(L, L.dq_prob_zero_all_noalloc(d_state, n, prec.value, d_out, None))
        got = (C.c_double * (n + 1))()
        _ck(L, L.dq_memcpy_async(C.addressof(got), d_out, (n + 1) * 8, 2, None))
        _ck(L, L.dq_stream_sync(None))
        # sizes beyond the limit are refused before any work
        assert L.dq_prob_zero_all_noalloc(d_state, 37, prec.value, d_out, None) != 0
        assert b"limit" in L.dq_last_error()
    finally:
        L.dq_free(ws)
        L.dq_free(d_out)
    return np.array(ref[:]), np.array(got[:])


def test_noalloc_equals_workspace_entry_point(quest, env):
    L = _lib()
    for n, depth, seed in ((22, 3, 7), (17, 4, 8)):  # second call smaller
        q = quest.createQureg(n, env)
        quest.apply_circuit(q, circuits.layered_u3_cz(n, depth, seed))
        ref, got = _both(L, quest, q, n)
        quest.destroyQureg(q, env)
        print(f"noalloc n={n}: total {got[n]:.15f}", flush=True)
        assert np.array_equal(ref, got), (n, np.abs(ref - got).max())
        assert abs(got[n] - 1.0) <= 1e-12
        assert np.all((got[:n] > 0) & (got[:n] < 1))


def test_noalloc_chunked_31q_fp32(quest32, env32):
    L = _lib()
    n = 31
    q = quest32.createQureg(n, env32)
    circ = [("hadamard", (i,), ()) for i in range(n)] + [("rotateY", (i,), (0.05 * (i + 1),)) for i in range(n)]
    quest32.apply_circuit(q, circ)
    ref, got = _both(L, quest32, q, n)
    quest32.destroyQureg(q, env32)
    print(f"noalloc n={n} fp32: max diff {np.abs(ref - got).max():.3e}", flush=True)
    assert np.abs(ref - got).max() <= 1e-12
    # product state: P(q=0) = (1 - sin(theta_q)) / 2 after H then RY(theta_q)
    theta = 0.05 * (np.arange(n) + 1)
    assert np.abs(got[:n] - (1 - np.sin(theta)) / 2).max() <= 1e-5
