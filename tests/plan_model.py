"""A numpy model of the engine's descriptor semantics (include/dgq_engine.h).

It executes a plan -- blocks, phases, ops, as exported by dgq_recorder_plan --
on a complex128 state the way the sweep and small kernels define it, one op
at a time over the whole state.  This is test infrastructure: it lets the
planner's output (fusion, sheared rotations, pending monomials, condition
flips, layouts) be checked against the oracle on any host, with no GPU.
"""
import numpy as np

K_MAT, K_NEG, K_PHASE, K_SWAP, K_RMAT, K_DIAGR, K_SHEAR = 1, 2, 3, 5, 6, 7, 8
INIT_LOAD, INIT_CLASSICAL, INIT_PLUS = 0, 1, 2


def _lg(x):
    x = int(x)
    assert x and not (x & (x - 1)), "mask not one-hot"
    return x.bit_length() - 1


def _bits(idx, q):
    return (idx >> q) & 1


def _pair_op(psi, idx, cond, tq, kind, m):
    lower = cond & (_bits(idx, tq) == 0)
    i0 = idx[lower]
    i1 = i0 | (1 << tq)
    x = psi[i0].copy()
    y = psi[i1].copy()
    if kind == K_MAT:
        u = [complex(m[2 * i], m[2 * i + 1]) for i in range(4)]
        psi[i0] = u[0] * x + u[1] * y
        psi[i1] = u[2] * x + u[3] * y
    elif kind == K_RMAT:
        psi[i0] = m[0] * x + m[1] * y
        psi[i1] = m[2] * x + m[3] * y
    elif kind == K_SWAP:
        psi[i0] = y
        psi[i1] = x
    elif kind == K_SHEAR:
        x = x + m[0] * y
        psi[i0] = x
        psi[i1] = y + m[1] * x
    else:
        raise ValueError(kind)


def run_plan(psi, n, blocks, phases, ops):
    """Apply the plan to psi (complex128, length 2^n) in place."""
    N = 1 << n
    idx = np.arange(N, dtype=np.int64)
    for b in blocks:
        if b["init_kind"] == INIT_CLASSICAL:
            psi[:] = 0
            psi[int(b["init_index"])] = 1
        elif b["init_kind"] == INIT_PLUS:
            psi[:] = float(b["init_value"])
        R, W = int(b["reg_bits"]), int(b["wave_bits"])
        pb, pc = int(b["phase_begin"]), int(b["phase_count"])
        if R == 0:  # small form: ops in global-index form
            for p in range(pb, pb + pc):
                ph = phases[p]
                for k in range(int(ph["op_begin"]), int(ph["op_begin"] + ph["op_count"])):
                    o = ops[k]
                    cond = (idx & int(o["c_wg"])) == int(o["v_wg"])
                    _apply(psi, idx, cond, o, int(o["tbit"]), None, None)
            continue
        ph0 = phases[pb]
        tb2q = {}
        for l in range(6):
            tb2q[_lg(ph0["lane_t"][l])] = l
        for r in range(R):
            tb2q[_lg(ph0["reg_t"][r])] = _lg(b["ld_reg"][r])
        for w in range(W):
            tb2q[_lg(ph0["wave_t"][w])] = _lg(b["ld_wave"][w])
        for p in range(pb, pb + pc):
            ph = phases[p]
            regq = [tb2q[_lg(ph["reg_t"][r])] for r in range(R)]
            thrq = [tb2q[_lg(ph["lane_t"][l])] for l in range(6)] + [tb2q[_lg(ph["wave_t"][w])] for w in range(W)]
            k = int(ph["op_begin"])
            end = k + int(ph["op_count"])
            while k < end:
                o = ops[k]
                cond = (idx & int(o["c_wg"])) == int(o["v_wg"])
                for r in range(R):
                    if (int(o["c_reg"]) >> r) & 1:
                        cond &= _bits(idx, regq[r]) == ((int(o["v_reg"]) >> r) & 1)
                for t in range(6 + W):
                    if (int(o["c_thr"]) >> t) & 1:
                        cond &= _bits(idx, thrq[t]) == ((int(o["v_thr"]) >> t) & 1)
                if int(o["kind"]) == K_DIAGR:
                    nc = int(o["tbit"])
                    table = np.empty(1 << R, dtype=np.complex128)
                    for e in range(1 << R):
                        mm = ops[k + 1 + e // 4]["m"]
                        table[e] = complex(mm[2 * (e % 4)], mm[2 * (e % 4) + 1])
                    j = np.zeros(N, dtype=np.int64)
                    for r in range(R):
                        j |= _bits(idx, regq[r]) << r
                    psi[cond] *= table[j[cond]]
                    k += nc + 1
                    continue
                tq = regq[int(o["tbit"])] if int(o["kind"]) in (K_MAT, K_SWAP, K_RMAT, K_SHEAR) else None
                _apply(psi, idx, cond, o, tq, regq, thrq)
                k += 1
        # store: the last layout's bits go to st_* qubits (a permutation when they differ)
        last = phases[pb + pc - 1]
        store = {}
        for l in range(6):
            store[_lg(last["lane_t"][l])] = l
        for r in range(R):
            store[_lg(last["reg_t"][r])] = _lg(b["st_reg"][r])
        for w in range(W):
            store[_lg(last["wave_t"][w])] = _lg(b["st_wave"][w])
        moves = {tb2q[tb]: store[tb] for tb in tb2q}
        if any(s != d for s, d in moves.items()):
            dst = idx.copy()
            for s, d in moves.items():
                dst &= ~(1 << d)
            for s, d in moves.items():
                dst |= _bits(idx, s) << d
            out = np.empty_like(psi)
            out[dst] = psi
            psi[:] = out


def _apply(psi, idx, cond, o, tq, regq, thrq):
    kind = int(o["kind"])
    m = o["m"]
    if kind in (K_MAT, K_SWAP, K_RMAT, K_SHEAR):
        _pair_op(psi, idx, cond, tq, kind, m)
    elif kind == K_NEG:
        psi[cond] *= -1
    elif kind == K_PHASE:
        psi[cond] *= complex(m[0], m[1])
    else:
        raise ValueError(f"op kind {kind}")


def plan_and_run(quest, n, circuit, k=6, R=4, W=2, max_ops=4096, fuse=True, init=None, start=None):
    """Record `circuit` on a recording register, plan it and run the plan on
    the model.  Returns the state in logical qubit order and the plan."""
    q = quest.dgq_create_recorder(n)
    try:
        if init == "plus":
            quest.initPlusState(q)
        elif isinstance(init, int):
            quest.initClassicalState(q, init)
        elif init == "load":
            pass
        quest.apply_circuit(q, circuit)
        blocks, phases, ops, l2p = quest.recorder_plan(q, k, R, W, max_ops, fuse)
    finally:
        quest.dgq_destroy_recorder(q)
    psi = np.zeros(1 << n, dtype=np.complex128)
    if start is not None:
        psi[:] = start
    else:
        psi[0] = 1.0
    run_plan(psi, n, blocks, phases, ops)
    # logical qubit q sits at physical position l2p[q]
    if any(l2p[i] != i for i in range(n)):
        idx = np.arange(1 << n, dtype=np.int64)
        phys = np.zeros_like(idx)
        for lq in range(n):
            phys |= ((idx >> lq) & 1) << l2p[lq]
        psi = psi[phys]
    return psi, (blocks, phases, ops)
