"""Shared test helpers: random circuits over the whole gate set, and a numpy
brute force (dense 2^n x 2^n operators built by Kronecker products) that is
independent of both the oracle and the engine."""
import math

import numpy as np

from dgquest_amd.quest import GATE_SIGNATURES

ONE_Q = ["hadamard", "pauliX", "pauliY", "pauliZ", "sGate", "tGate", "SqX", "SqY", "SqW"]
ONE_Q_R = ["phaseShift", "rotateX", "rotateY", "rotateZ", "u1Gate"]
TWO_Q = ["controlledPhaseFlip", "controlledNot", "controlledPauliY"]
TWO_Q_R = ["controlledPhaseShift", "controlledRotateX", "controlledRotateY", "controlledRotateZ"]


def random_unitary_params(rng):
    """A 2x2 unitary e^{i a} Rz(b) Ry(c) Rz(d), flattened (re, im) row-major."""
    a, b, c, d = rng.uniform(0, 2 * math.pi, 4)
    rz = lambda t: np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])
    ry = np.array([[math.cos(c / 2), -math.sin(c / 2)], [math.sin(c / 2), math.cos(c / 2)]])
    u = np.exp(1j * a) * (rz(b) @ ry @ rz(d))
    return [float(x) for z in u.reshape(-1) for x in (z.real, z.imag)]


def random_compact_params(rng):
    v = rng.normal(size=4)
    v /= np.linalg.norm(v)
    return [float(x) for x in v]


def random_gate(rng, n, allow=None):
    kinds = []
    kinds += [("1q", g) for g in ONE_Q] + [("1qr", g) for g in ONE_Q_R]
    kinds += [("u2", "u2Gate"), ("u3", "u3Gate"), ("un", "unitary"), ("cu", "compactUnitary"),
              ("ax", "rotateAroundAxis")]
    if n >= 2:
        kinds += [("2q", g) for g in TWO_Q] + [("2qr", g) for g in TWO_Q_R]
        kinds += [("cun", "controlledUnitary"), ("ccu", "controlledCompactUnitary"),
                  ("cax", "controlledRotateAroundAxis"), ("fsim", "fSim"), ("swap", "swapGate")]
    if n >= 3:
        kinds += [("mcu", "multiControlledUnitary"), ("mcf", "multiControlledPhaseFlip"),
                  ("mcs", "multiControlledPhaseShift")]
    if allow is not None:
        kinds = [k for k in kinds if k[1] in allow]
    kind, name = kinds[rng.integers(len(kinds))]
    ang = lambda: float(rng.uniform(-2 * math.pi, 2 * math.pi))
    qs = [int(x) for x in rng.permutation(n)]
    if kind == "1q":
        return (name, (qs[0],), ())
    if kind == "1qr":
        return (name, (qs[0],), (ang(),))
    if kind == "u2":
        return (name, (qs[0],), (ang(), ang()))
    if kind == "u3":
        return (name, (qs[0],), (ang(), ang(), ang()))
    if kind == "un":
        return (name, (qs[0],), tuple(random_unitary_params(rng)))
    if kind == "cu":
        return (name, (qs[0],), tuple(random_compact_params(rng)))
    if kind == "ax":
        return (name, (qs[0],), (ang(), *[float(x) for x in rng.normal(size=3)]))
    if kind == "2q":
        return (name, (qs[0], qs[1]), ())
    if kind == "2qr":
        return (name, (qs[0], qs[1]), (ang(),))
    if kind == "cun":
        return (name, (qs[0], qs[1]), tuple(random_unitary_params(rng)))
    if kind == "ccu":
        return (name, (qs[0], qs[1]), tuple(random_compact_params(rng)))
    if kind == "cax":
        return (name, (qs[0], qs[1]), (ang(), *[float(x) for x in rng.normal(size=3)]))
    if kind == "fsim":
        return (name, (qs[0], qs[1]), (ang(), ang()))
    if kind == "swap":
        return (name, (qs[0], qs[1]), ())
    nc = int(rng.integers(1, min(4, n - 1) + 1))
    if kind == "mcu":
        return (name, tuple(qs[:nc + 1]), tuple(random_unitary_params(rng)))
    if kind == "mcf":
        return (name, tuple(qs[:nc + 1]), ())
    return (name, tuple(qs[:nc + 1]), (ang(),))


def random_circuit(rng, n, count, allow=None):
    return [random_gate(rng, n, allow) for _ in range(count)]


# ---- dense brute force (independent of oracle and engine) -------------------
I2 = np.eye(2, dtype=complex)
P0 = np.diag([1, 0]).astype(complex)
P1 = np.diag([0, 1]).astype(complex)


def op_on(n, mats):
    """Kronecker product with mats[q] on qubit q (qubit 0 = least significant)."""
    out = np.array([[1.0 + 0j]])
    for q in range(n - 1, -1, -1):
        out = np.kron(out, mats.get(q, I2))
    return out


def controlled(n, controls, target, u):
    """sum over control patterns: u on target only when all controls are 1."""
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=complex)
    k = len(controls)
    for pat in range(1 << k):
        mats = {c: (P1 if (pat >> i) & 1 else P0) for i, c in enumerate(controls)}
        if pat == (1 << k) - 1 and target is not None:
            mats[target] = u
        full += op_on(n, mats)
    return full


def mat1(name, p):
    s2 = 1 / math.sqrt(2)
    if name == "hadamard":
        return np.array([[1, 1], [1, -1]]) * s2
    if name == "pauliX":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if name == "pauliY":
        return np.array([[0, -1j], [1j, 0]])
    if name == "pauliZ":
        return np.diag([1, -1]).astype(complex)
    if name == "sGate":
        return np.diag([1, 1j])
    if name == "tGate":
        return np.diag([1, np.exp(1j * math.pi / 4)])
    if name == "SqX":
        return np.array([[1, -1j], [-1j, 1]]) * s2
    if name == "SqY":
        return np.array([[1, -1], [1, 1]]) * s2
    if name == "SqW":
        return np.array([[s2, -0.5 - 0.5j], [0.5 - 0.5j, s2]])
    if name in ("phaseShift", "u1Gate"):
        return np.diag([1, np.exp(1j * p[0])])
    if name in ("rotateX", "controlledRotateX"):
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return np.array([[c, -1j * s], [-1j * s, c]])
    if name in ("rotateY", "controlledRotateY"):
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return np.array([[c, -s], [s, c]])
    if name in ("rotateZ", "controlledRotateZ"):
        return np.diag([np.exp(-0.5j * p[0]), np.exp(0.5j * p[0])])
    if name == "u2Gate":
        phi, lam = p
        return np.array([[1, -np.exp(1j * lam)], [np.exp(1j * phi), np.exp(1j * (phi + lam))]]) * s2
    if name == "u3Gate":
        th, phi, lam = p
        return np.array([[math.cos(th / 2), -np.exp(1j * lam) * math.sin(th / 2)],
                         [np.exp(1j * phi) * math.sin(th / 2), np.exp(1j * (phi + lam)) * math.cos(th / 2)]])
    if name in ("unitary", "controlledUnitary", "multiControlledUnitary"):
        return np.array(p[0::2]).reshape(2, 2) + 1j * np.array(p[1::2]).reshape(2, 2)
    if name in ("compactUnitary", "controlledCompactUnitary"):
        a, b = p[0] + 1j * p[1], p[2] + 1j * p[3]
        return np.array([[a, -np.conj(b)], [b, np.conj(a)]])
    if name in ("rotateAroundAxis", "controlledRotateAroundAxis"):
        th, x, y, z = p
        v = np.array([x, y, z]) / np.linalg.norm([x, y, z])
        X = np.array([[0, 1], [1, 0]]); Y = np.array([[0, -1j], [1j, 0]]); Z = np.diag([1, -1])
        return math.cos(th / 2) * np.eye(2) - 1j * math.sin(th / 2) * (v[0] * X + v[1] * Y + v[2] * Z)
    raise KeyError(name)


def dense_gate(n, g):
    name, qs, p = g[0], g[1], (g[2] if len(g) > 2 else ())
    if name in ONE_Q or name in ONE_Q_R or name in ("u2Gate", "u3Gate", "unitary", "compactUnitary",
                                                   "rotateAroundAxis"):
        return op_on(n, {qs[0]: mat1(name, p)})
    if name == "controlledPhaseFlip":
        return np.diag([(-1.0 if ((i >> qs[0]) & 1 and (i >> qs[1]) & 1) else 1.0)
                        for i in range(1 << n)]).astype(complex)
    if name == "controlledPhaseShift":
        return np.diag([(np.exp(1j * p[0]) if ((i >> qs[0]) & 1 and (i >> qs[1]) & 1) else 1.0)
                        for i in range(1 << n)]).astype(complex)
    if name == "multiControlledPhaseFlip":
        return np.diag([(-1.0 if all((i >> q) & 1 for q in qs) else 1.0) for i in range(1 << n)]).astype(complex)
    if name == "multiControlledPhaseShift":
        return np.diag([(np.exp(1j * p[0]) if all((i >> q) & 1 for q in qs) else 1.0)
                        for i in range(1 << n)]).astype(complex)
    if name == "controlledNot":
        return controlled(n, [qs[0]], qs[1], mat1("pauliX", ()))
    if name == "controlledPauliY":
        return controlled(n, [qs[0]], qs[1], mat1("pauliY", ()))
    if name in ("controlledRotateX", "controlledRotateY", "controlledRotateZ", "controlledUnitary",
                "controlledCompactUnitary", "controlledRotateAroundAxis"):
        return controlled(n, [qs[0]], qs[1], mat1(name, p))
    if name == "multiControlledUnitary":
        return controlled(n, list(qs[:-1]), qs[-1], mat1(name, p))
    if name == "swapGate":
        a, b = qs
        perm = np.zeros((1 << n, 1 << n), dtype=complex)
        for i in range(1 << n):
            j = i
            if ((i >> a) & 1) != ((i >> b) & 1):
                j = i ^ (1 << a) ^ (1 << b)
            perm[j, i] = 1
        return perm
    if name == "fSim":
        a, b = qs
        th, phi = p
        out = np.eye(1 << n, dtype=complex)
        for sub in (("controlledPhaseShift", (a, b), (-phi,)), ("controlledNot", (a, b), ()),
                    ("controlledRotateX", (b, a), (2 * th,)), ("controlledNot", (a, b), ())):
            out = dense_gate(n, sub) @ out
        return out
    raise KeyError(name)


def dense_run(n, circuit, state=None):
    if state is None:
        state = np.zeros(1 << n, dtype=complex)
        state[0] = 1
    for g in circuit:
        state = dense_gate(n, g) @ state
    return state


# ---- tensor oracle: the state as an n-axis tensor, one gate at a time --------
# Independent of the oracle's C loops and of the engine: each gate acts by a
# numpy tensordot on the target axis of the sub-tensor where its controls are
# 1 (qubit q = axis n - 1 - q of the C-ordered reshape).  Runs at 20+ qubits,
# where the Kronecker brute force above cannot.
def _axis(n, q):
    return n - 1 - q


def _apply_1q(psi, n, target, u, controls=()):
    sl = [slice(None)] * n
    for c in controls:
        sl[_axis(n, c)] = 1
    sub = psi[tuple(sl)]
    ax = _axis(n, target) - sum(1 for c in controls if _axis(n, c) < _axis(n, target))
    out = np.moveaxis(np.tensordot(u, sub, axes=([1], [ax])), 0, ax)
    psi[tuple(sl)] = out


def _phase_where_all_one(psi, n, qs, f):
    sl = [slice(None)] * n
    for q in qs:
        sl[_axis(n, q)] = 1
    psi[tuple(sl)] *= f


def tensor_gate(psi, n, g):
    name, qs, p = g[0], g[1], (g[2] if len(g) > 2 else ())
    if name in ONE_Q or name in ONE_Q_R or name in ("u2Gate", "u3Gate", "unitary", "compactUnitary",
                                                   "rotateAroundAxis"):
        _apply_1q(psi, n, qs[0], mat1(name, p))
    elif name in ("controlledPhaseFlip", "multiControlledPhaseFlip"):
        _phase_where_all_one(psi, n, qs, -1.0)
    elif name in ("controlledPhaseShift", "multiControlledPhaseShift"):
        _phase_where_all_one(psi, n, qs, np.exp(1j * p[0]))
    elif name == "controlledNot":
        _apply_1q(psi, n, qs[1], mat1("pauliX", ()), (qs[0],))
    elif name == "controlledPauliY":
        _apply_1q(psi, n, qs[1], mat1("pauliY", ()), (qs[0],))
    elif name in ("controlledRotateX", "controlledRotateY", "controlledRotateZ", "controlledUnitary",
                  "controlledCompactUnitary", "controlledRotateAroundAxis"):
        _apply_1q(psi, n, qs[1], mat1(name, p), (qs[0],))
    elif name == "multiControlledUnitary":
        _apply_1q(psi, n, qs[-1], mat1(name, p), tuple(qs[:-1]))
    elif name == "swapGate":
        a, b = _axis(n, qs[0]), _axis(n, qs[1])
        psi[...] = np.swapaxes(psi, a, b).copy()
    elif name == "fSim":
        a, b = qs
        th, phi = p
        for sub in (("controlledPhaseShift", (a, b), (-phi,)), ("controlledNot", (a, b), ()),
                    ("controlledRotateX", (b, a), (2 * th,)), ("controlledNot", (a, b), ())):
            tensor_gate(psi, n, sub)
    else:
        raise KeyError(name)


def tensor_run(n, circuit, state=None):
    if state is None:
        state = np.zeros(1 << n, dtype=complex)
        state[0] = 1
    psi = state.astype(complex).reshape((2,) * n).copy()
    for g in circuit:
        tensor_gate(psi, n, g)
    return psi.reshape(-1)
