"""CPU: the sweep kernel's device programs against the op semantics.

The engine turns each phase's ops (dq_op, include/dgq_engine.h) into a device
program for its assembly interpreter: register conditions resolved into op
variants, runs of shears merged, diagonal ops folded into their run's table,
a table fused with the shears after it (TSH ops), workgroup guards.  Here
dq_build_program returns that program on the host, a numpy model of the
interpreter (op codes and sizes from gen_ops_asm.py, handler semantics as in
its generated assembly) runs it on every thread of a tile, and the result
must equal the ops applied one at a time (the dq_op semantics), for the
plans of random circuits over the whole gate set, every tile shape, both
precisions, and workgroup bases that make guarded ops both run and skip.
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

from tests import helpers as H

K_MAT, K_NEG, K_PHASE, K_SWAP, K_RMAT, K_DIAGR, K_SHEAR = 1, 2, 3, 5, 6, 7, 8
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gen():
    spec = importlib.util.spec_from_file_location("gen_ops_asm", os.path.join(ROOT, "dgquest_amd", "csrc", "gen_ops_asm.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


GEN = _gen()


@pytest.fixture(scope="module")
def eng():
    from dgquest_amd.engine import Engine

    L = Engine().lib
    L.dq_build_program.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]
    return L


def build(eng, prec, R, ops):
    n = C.c_size_t()
    assert eng.dq_build_program(prec, R, ops.ctypes.data, len(ops), None, 0, C.byref(n)) == 0, \
        eng.dq_last_error().decode()
    buf = np.zeros(n.value, dtype=np.uint8)
    assert eng.dq_build_program(prec, R, ops.ctypes.data, len(ops), buf.ctypes.data, len(buf), C.byref(n)) == 0
    return buf.tobytes()


def coef(o, prec):
    """An op's 8 coefficients as stored (FP32 ops keep floats from the start of m)."""
    raw = o["m"].tobytes()
    return np.frombuffer(raw, dtype=np.float64 if prec == 2 else np.float32)[:8].astype(np.float64)


def run_ops(A, ops, prec, R, base):
    """dq_op semantics on a tile: A[thread, register index]."""
    NT, NR = A.shape
    tid = np.arange(NT)[:, None]
    j = np.arange(NR)[None, :]
    k = 0
    while k < len(ops):
        o = ops[k]
        kind = int(o["kind"])
        wg_ok = (base & int(o["c_wg"])) == int(o["v_wg"])
        cond = ((j & int(o["c_reg"])) == int(o["v_reg"])) & ((tid & int(o["c_thr"])) == int(o["v_thr"])) & wg_ok
        m = coef(o, prec)
        if kind == K_DIAGR:
            table = np.empty(NR, dtype=np.complex128)
            for e in range(NR):
                mm = coef(ops[k + 1 + e // 4], prec)
                table[e] = complex(mm[2 * (e % 4)], mm[2 * (e % 4) + 1])
            if wg_ok:
                A *= table[None, :]
            k += 1 + int(o["tbit"])
            continue
        if kind in (K_NEG, K_PHASE):
            f = -1.0 if kind == K_NEG else complex(m[0], m[1])
            A[np.broadcast_to(cond, A.shape)] *= f
            k += 1
            continue
        b = int(o["tbit"])
        for j0 in range(NR):
            if (j0 >> b) & 1:
                continue
            j1 = j0 | (1 << b)
            c = cond[:, j0]
            x, y = A[c, j0].copy(), A[c, j1].copy()
            if kind == K_MAT:
                u = [complex(m[2 * i], m[2 * i + 1]) for i in range(4)]
                A[c, j0], A[c, j1] = u[0] * x + u[1] * y, u[2] * x + u[3] * y
            elif kind == K_RMAT:
                A[c, j0], A[c, j1] = m[0] * x + m[1] * y, m[2] * x + m[3] * y
            elif kind == K_SWAP:
                A[c, j0], A[c, j1] = y, x
            elif kind == K_SHEAR:
                x = x + m[0] * y
                A[c, j0], A[c, j1] = x, y + m[1] * x
            else:
                raise ValueError(kind)
        k += 1
    return A


def run_program(A, prog, prec, R, base):
    """The interpreter's semantics on a tile (gen_ops_asm.py handlers)."""
    NT, NR = A.shape
    ts = 8 if prec == 2 else 4
    ft = np.float64 if prec == 2 else np.float32
    codes = GEN.op_codes(R)
    tid = np.arange(NT)
    at = 0
    skip_next = False
    seen = set()
    while True:
        code, mask, cthr, vthr = np.frombuffer(prog, dtype=np.uint32, count=4, offset=at)
        name, params = codes[int(code)]
        size = GEN.op_size(name, R, ts)
        seen.add(name)
        cf = np.frombuffer(prog, dtype=ft, count=(size - 16) // ts, offset=at + 16).astype(np.float64) \
            if size > 16 else np.zeros(0)
        if skip_next:
            skip_next = False
            at += size
            continue
        if name == "END":
            return seen
        lanes = (tid & int(cthr)) == int(vthr)

        def table_of(n_entries):
            return np.array([complex(cf[2 * i], cf[2 * i + 1]) for i in range(n_entries)])

        def shears(smask, slots):
            for B in range(R):
                if not (smask >> B) & 1:
                    continue
                u, l = slots[2 * B], slots[2 * B + 1]
                for j0 in range(NR):
                    if (j0 >> B) & 1:
                        continue
                    j1 = j0 | (1 << B)
                    A[:, j0] = A[:, j0] + u * A[:, j1]
                    A[:, j1] = A[:, j1] + l * A[:, j0]

        kind = name.split("_")[0]
        if name == "WG":
            c_wg, v_wg = np.frombuffer(prog, dtype=np.uint64, count=2, offset=at + 16)
            skip_next = (base & int(c_wg)) != int(v_wg)
        elif name == "SHEARS":
            shears(int(mask), cf)
        elif kind in ("TABLE", "TSH"):
            variant = name.split("_")[1]
            if variant == "HALF":
                (r,) = params
                js = [j for j in range(NR) if (j >> r) & 1]
                tab = table_of(len(js))
                for i, j in enumerate(js):
                    A[:, j] *= tab[i]
                nt = len(js)
            else:
                tab = table_of(NR)
                sel = int(cthr) if (kind == "TSH" and variant == "GEN") else (int(mask) if variant == "GEN" else (1 << NR) - 1)
                for j in range(NR):
                    if (sel >> j) & 1:
                        A[:, j] *= tab[j]
                nt = NR
            if kind == "TSH":
                shears(int(mask), cf[2 * nt:])
        elif kind in ("NEG", "PHASE"):
            variant = name.split("_", 1)[1]
            if variant == "ALL":
                js = range(NR)
            elif variant == "GEN":
                js = [j for j in range(NR) if (int(mask) >> j) & 1]
            elif variant == "B":
                r, v = params
                js = [j for j in range(NR) if ((j >> r) & 1) == v]
            else:
                r1, r2, v1, v2 = params
                js = [j for j in range(NR) if ((j >> r1) & 1) == v1 and ((j >> r2) & 1) == v2]
            f = -1.0 if kind == "NEG" else complex(cf[0], cf[1])
            for j in js:
                A[lanes, j] *= f
        elif kind in ("MAT", "SWAP", "RMAT"):
            (B,) = params
            for j0 in range(NR):
                if (j0 >> B) & 1:
                    continue
                if kind != "RMAT" and not (int(mask) >> j0) & 1:
                    continue
                j1 = j0 | (1 << B)
                c = lanes if kind != "RMAT" else np.ones(NT, dtype=bool)
                x, y = A[c, j0].copy(), A[c, j1].copy()
                if kind == "MAT":
                    u = [complex(cf[2 * i], cf[2 * i + 1]) for i in range(4)]
                    A[c, j0], A[c, j1] = u[0] * x + u[1] * y, u[2] * x + u[3] * y
                elif kind == "RMAT":
                    A[c, j0], A[c, j1] = cf[0] * x + cf[1] * y, cf[2] * x + cf[3] * y
                else:
                    A[c, j0], A[c, j1] = y, x
        else:
            raise ValueError(name)
        at += size


def phases_of(lib, n, circ, shape):
    k, R, W = shape
    r = lib.dgq_create_recorder(n)
    lib.apply_circuit(r, circ)
    blocks, phases, ops, _ = lib.recorder_plan(r, k=k, R=R, W=W)
    lib.dgq_destroy_recorder(r)
    out = []
    for b in blocks:
        Rb, Wb = int(b["reg_bits"]), int(b["wave_bits"])
        if Rb == 0:
            continue
        for p in range(int(b["phase_begin"]), int(b["phase_begin"] + b["phase_count"])):
            ph = phases[p]
            out.append((Rb, Wb, ops[int(ph["op_begin"]): int(ph["op_begin"] + ph["op_count"])].copy(),
                        int(b["nontile_mask"])))
    return out


@pytest.mark.parametrize("prec", [2, 1])
@pytest.mark.parametrize("shape", [(7, 4, 3), (6, 3, 3), (5, 4, 1), (6, 4, 2), (4, 4, 0), (3, 2, 1), (7, 5, 2)])
def test_programs_match_ops(quest, quest32, eng, prec, shape):
    lib = quest if prec == 2 else quest32
    rng = np.random.default_rng(1000 + 10 * shape[0] + shape[1] + prec)
    n = 14
    circ = H.random_circuit(rng, n, 160)
    tol = 1e-12 if prec == 2 else 2e-5
    seen = set()
    for R, W, ops, nontile in phases_of(lib, n, circ, shape):
        if len(ops) == 0:
            continue
        prog = build(eng, prec, R, ops)
        for trial in range(2):
            base = int(rng.integers(0, 1 << 62)) & nontile
            A0 = rng.normal(size=(64 << W, 1 << R)) + 1j * rng.normal(size=(64 << W, 1 << R))
            want = run_ops(A0.copy(), ops, prec, R, base)
            got = A0.copy()
            seen |= run_program(got, prog, prec, R, base)
            assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())
    # the rewrites this test is for took place
    assert "TSH_FULL" in seen or "TSH_HALF" in seen or "TSH_GEN" in seen


def test_cz_between_register_bits_folds_into_the_table(quest, eng):
    """A layered U3 + CZ circuit: its register-bit CZs leave no NEG op behind
    a table in the same diagonal run."""
    from dgquest_amd import circuits

    n = 14
    circ = circuits.layered_u3_cz(n, 6, 3)
    neg_bb = 0
    for R, W, ops, _ in phases_of(quest, n, circ, (7, 4, 3)):
        if len(ops) == 0:
            continue
        prog = build(eng, 2, R, ops)
        at = 0
        codes = GEN.op_codes(R)
        while True:
            code = int(np.frombuffer(prog, dtype=np.uint32, count=1, offset=at)[0])
            name, _ = codes[code]
            if name == "END":
                break
            neg_bb += name.startswith("NEG_BB")
            at += GEN.op_size(name, R, 8)
    raw_bb = 0
    for R, W, ops, _ in phases_of(quest, n, circ, (7, 4, 3)):
        for o in ops:
            if int(o["kind"]) == K_NEG and bin(int(o["c_reg"])).count("1") == 2 and not int(o["c_thr"]) and not int(o["c_wg"]):
                raw_bb += 1
    assert raw_bb > 0 and neg_bb < raw_bb
