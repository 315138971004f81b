// planner.cpp -- see planner.hpp.
#include "planner.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>

#include "dag.hpp"

namespace dgq {

static const double kRsqrt2 = 0.7071067811865475244008443621;
static const int kSmallMax = 6;  // n <= kSmallMax: whole-state-in-LDS kernel

TuneParams fit_shape(const TuneParams &tp, int n, int prec) {
    TuneParams t = tp;
    if (t.R < 1) t.R = 1;
    if (t.R > DQ_MAX_REG_BITS) t.R = DQ_MAX_REG_BITS;
    if (t.W < 0) t.W = 0;
    if (t.W > DQ_MAX_WAVE_BITS) t.W = DQ_MAX_WAVE_BITS;
    while (6 + t.R + t.W > n && t.W > 0) t.W--;
    while (6 + t.R + t.W > n && t.R > 1) t.R--;
    // instantiated shapes: W = 0 for R 1..5, W 1..2 for R 2..5, W = 3 for R 3..4
    // (and R = 5 in FP32: the 14-bit tile)
    if (t.W >= 1 && t.R < 2) t.R = 2;
    if (t.W == 3 && t.R > 4 && prec != DQ_F32) t.W = 2;
    if (t.W == 3 && t.R < 3) t.W = 2;
    if (t.k > t.R + t.W) t.k = t.R + t.W;
    if (t.k < 1) t.k = 1;
    if (t.max_ops < 1) t.max_ops = 1;
    return t;
}

namespace {

// Planner-side layout: tile bit of each register, lane and wave bit.
struct Layout {
    uint8_t reg[8];
    uint8_t lane[6];
    uint8_t wave[4];
};

struct Roles {
    int kind[DQ_MAX_TILE_BITS];  // 0 reg, 1 lane, 2 wave
    int idx[DQ_MAX_TILE_BITS];
};

Roles roles_of(const Layout &L, int R, int W) {
    Roles r;
    for (int i = 0; i < R; i++) { r.kind[L.reg[i]] = 0; r.idx[L.reg[i]] = i; }
    for (int i = 0; i < 6; i++) { r.kind[L.lane[i]] = 1; r.idx[L.lane[i]] = i; }
    for (int i = 0; i < W; i++) { r.kind[L.wave[i]] = 2; r.idx[L.wave[i]] = i; }
    return r;
}

void put(dq_op &o, int i, double v, int prec) {
    if (prec == DQ_F64) o.m[i] = v;
    else reinterpret_cast<float *>(o.m)[i] = (float)v;
}

struct TileCtx {
    int n, T, R, W, prec;
    int q2t[64];
    Roles roles;
};

// Translate a qubit mask into register / thread / tile-index bit masks.
void split_mask(const TileCtx &c, uint64_t qmask, uint32_t &reg, uint32_t &thr, uint64_t &wg) {
    reg = 0; thr = 0; wg = 0;
    while (qmask) {
        const int q = ctz64(qmask);
        qmask &= qmask - 1;
        const int b = c.q2t[q];
        if (b < 0) { wg |= 1ull << q; continue; }
        const int k = c.roles.kind[b], i = c.roles.idx[b];
        if (k == 0) reg |= 1u << i;
        else if (k == 1) thr |= 1u << i;
        else thr |= 1u << (6 + i);
    }
}

using cd = std::complex<double>;

// Ops of one phase with double coefficients; rounded to the state precision
// when appended to the plan.
struct OpList {
    std::vector<dq_op> ops;
    std::vector<std::array<double, 8>> coef;
    void add(uint32_t kind, uint32_t tbit, const dq_op &cond, const double *m, int nm) {
        dq_op o = cond;
        o.kind = kind;
        o.tbit = tbit;
        std::array<double, 8> a{};
        for (int i = 0; i < nm; i++) a[i] = m[i];
        ops.push_back(o);
        coef.push_back(a);
    }
    void flush_to(std::vector<dq_op> &out, int prec) {
        for (size_t i = 0; i < ops.size(); i++) {
            dq_op o = ops[i];
            for (int e = 0; e < 8; e++) put(o, e, coef[i][e], prec);
            out.push_back(o);
        }
        ops.clear();
        coef.clear();
    }
};

dq_op cond_of(const TileCtx &c, bool small, uint64_t cmask, uint64_t cval) {
    dq_op o;
    memset(&o, 0, sizeof o);
    if (small) {
        o.c_wg = cmask;
        o.v_wg = cval;
    } else {
        split_mask(c, cmask, o.c_reg, o.c_thr, o.c_wg);
        split_mask(c, cval, o.v_reg, o.v_thr, o.v_wg);
    }
    return o;
}

// register bit holding qubit q in the current layout, or -1
int reg_of(const TileCtx &c, int q) {
    const int b = c.q2t[q];
    if (b < 0 || c.roles.kind[b] != 0) return -1;
    return c.roles.idx[b];
}

// Fusion state of one flush.  Invariant: logical state = g * (prod_q P_q) *
// device state, with P_q = diag(p0[q], p1[q]) * X^x[q] a monomial of unit
// moduli.  Single-qubit diagonals only update P_q; a diagonal or a control on
// q only sees x[q] (it flips the condition value); the next non-diagonal gate
// on q absorbs P_q.  A pending X (x[q] = 1) only lives between two
// non-diagonal gates on q: the last one of the flush on each qubit leaves a
// plain diagonal, applied at the end of the flush.
struct Fuse {
    bool on = false;
    std::vector<cd> p0, p1;
    std::vector<uint8_t> x;
    std::vector<int> left;  // non-diagonal gates still to come per target qubit
    cd g = 1.0;
    explicit Fuse(int n = 0) : p0(n, 1.0), p1(n, 1.0), x(n, 0), left(n, 0) {}
    // The qubit at position a moves to to[a] (a permutation of the positions).
    void permute(const std::vector<int> &to) {
        std::vector<cd> a0(p0.size()), a1(p1.size());
        std::vector<uint8_t> ax(x.size());
        std::vector<int> al(left.size());
        for (size_t a = 0; a < to.size() && a < p0.size(); a++) {
            a0[to[a]] = p0[a];
            a1[to[a]] = p1[a];
            ax[to[a]] = x[a];
            al[to[a]] = left[a];
        }
        p0.swap(a0);
        p1.swap(a1);
        x.swap(ax);
        left.swap(al);
    }
    bool trivial(int q) const { return p0[q] == cd(1.0) && p1[q] == cd(1.0) && !x[q]; }
    uint64_t flips(uint64_t mask) const {
        uint64_t f = 0;
        for (uint64_t m = mask; m; m &= m - 1) {
            const int q = ctz64(m);
            if (q < (int)x.size() && x[q]) f |= 1ull << q;
        }
        return f;
    }
};

// Pending factored gates of one phase (at most one per qubit): their right
// diagonals diag(d0, d1) are applied together as one register-indexed table,
// then the real parts (SHEAR, or RMAT for the last gate on a qubit).
struct Batch {
    std::vector<int> reg;
    std::vector<uint32_t> kind;
    std::vector<std::array<double, 4>> m;
    std::vector<cd> d0, d1;
    uint64_t qmask = 0;
    void add(int r, uint32_t k, const double *c, int nc, cd a0, cd a1, int q) {
        std::array<double, 4> v{};
        for (int i = 0; i < nc; i++) v[i] = c[i];
        reg.push_back(r);
        kind.push_back(k);
        m.push_back(v);
        d0.push_back(a0);
        d1.push_back(a1);
        qmask |= 1ull << q;
    }
};

void emit_table(OpList &L, const TileCtx &c, const std::vector<cd> &table) {
    const int n_ent = 1 << c.R;
    const int n_cont = (n_ent + 3) / 4;
    dq_op z;
    memset(&z, 0, sizeof z);
    L.add(DQ_OP_DIAGR, (uint32_t)n_cont, z, nullptr, 0);
    for (int k = 0; k < n_cont; k++) {
        double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int e = 0; e < 4 && 4 * k + e < n_ent; e++) {
            m[2 * e] = table[4 * k + e].real();
            m[2 * e + 1] = table[4 * k + e].imag();
        }
        L.add(0, 0, z, m, 8);
    }
}

// The table of a batch holds diag(1, d1 / d0) per gate (the device skips
// entries that are exactly 1, half the registers for one gate); the d0 go to
// the global factor, folded into a table once it drifts far from 1 (the
// device state is the logical one over the global factor).
void flush_batch(Batch &b, OpList &L, const TileCtx &c, Fuse &F) {
    if (b.reg.empty()) return;
    const int n_ent = 1 << c.R;
    for (size_t i = 0; i < b.reg.size(); i++) {
        F.g *= b.d0[i];
        b.d1[i] /= b.d0[i];
    }
    const double ag = std::abs(F.g);
    std::vector<cd> table(n_ent, 1.0);
    if (ag < 0x1p-20 || ag > 0x1p20) {
        for (auto &t : table) t = F.g;
        F.g = 1.0;
    }
    for (size_t i = 0; i < b.reg.size(); i++)
        for (int j = 0; j < n_ent; j++)
            if ((j >> b.reg[i]) & 1) table[j] *= b.d1[i];
    bool any = false;
    for (const cd &t : table) any |= t != cd(1.0);
    if (any) emit_table(L, c, table);
    dq_op z;
    memset(&z, 0, sizeof z);
    for (size_t i = 0; i < b.reg.size(); i++)
        L.add(b.kind[i], (uint32_t)b.reg[i], z, b.m[i].data(), b.kind[i] == DQ_OP_SHEAR ? 2 : 4);
    b = Batch();
}

// Apply the diagonal part of P_q now (x[q] must be 0).
void apply_pending(Fuse &F, int q, OpList &L, const TileCtx &c) {
    if (F.p0[q] == cd(1.0) && F.p1[q] == cd(1.0)) return;
    const int r = reg_of(c, q);
    if (r >= 0) {
        std::vector<cd> table(1 << c.R);
        F.g *= F.p0[q];
        for (int j = 0; j < (1 << c.R); j++) table[j] = ((j >> r) & 1) ? F.p1[q] / F.p0[q] : cd(1.0);
        emit_table(L, c, table);
    } else {
        F.g *= F.p0[q];
        const cd ph = F.p1[q] / F.p0[q];
        const double m[2] = {ph.real(), ph.imag()};
        const uint64_t qm = 1ull << q;
        L.add(DQ_OP_PHASE, 0, cond_of(c, false, qm, qm), m, 2);
    }
    F.p0[q] = F.p1[q] = 1.0;
}

// Apply everything pending (end of a flush): register qubits in one table,
// the others as phase ops, the global factor last.
void apply_all_pending(Fuse &F, OpList &L, const TileCtx &c, int n) {
    const int n_ent = 1 << c.R;
    std::vector<cd> table(n_ent, 1.0);
    bool any = false;
    for (int q = 0; q < n; q++) {
        if (F.trivial(q)) continue;
        const int r = reg_of(c, q);
        if (r < 0) {
            apply_pending(F, q, L, c);
            continue;
        }
        any = true;
        for (int j = 0; j < n_ent; j++) table[j] *= ((j >> r) & 1) ? F.p1[q] : F.p0[q];
        F.p0[q] = F.p1[q] = 1.0;
    }
    if (F.g != cd(1.0)) {
        any = true;
        for (auto &t : table) t *= F.g;
        F.g = 1.0;
    }
    if (any) emit_table(L, c, table);
}

// U = diag(l0, l1) * M * diag(1, r1) with M real.  Returns false when the
// factorisation would drop more than rounding (a matrix that is unitary only
// to the validation tolerance); the caller then keeps the complex matrix.
bool real_factor(const cd U[4], cd &l0, cd &l1, cd &r1, double M[4]) {
    auto unit = [](cd z) { const double a = std::abs(z); return a > 0 ? z / a : cd(1.0); };
    const double tiny = 1e-300;
    if (std::abs(U[1]) <= tiny && std::abs(U[2]) <= tiny) {
        l0 = unit(U[0]); l1 = unit(U[3]); r1 = 1.0;
    } else if (std::abs(U[0]) <= tiny && std::abs(U[3]) <= tiny) {
        l0 = unit(U[1]); l1 = unit(U[2]); r1 = 1.0;
    } else {
        l0 = unit(U[0]); l1 = unit(U[2]); r1 = unit(U[1]) / l0;
    }
    const cd m[4] = {std::conj(l0) * U[0], std::conj(l0) * U[1] * std::conj(r1), std::conj(l1) * U[2],
                     std::conj(l1) * U[3] * std::conj(r1)};
    for (int i = 0; i < 4; i++) {
        if (std::abs(m[i].imag()) > 4e-16 * (1.0 + std::abs(m[i]))) return false;
        M[i] = m[i].real();
    }
    return true;
}

// Sheared form of a 2x2 unitary M (times unit-modulus monomials):
//   M = diag(l0, l1) X^x * L(t) * U(u) * diag(d0, d1),
//   U(u) = [[1, u], [0, 1]], L(t) = [[1, 0], [t, 1]], |t| <= 1, |u| <= 1/2.
// From M = diag(l0, l1) R(th) diag(1, r) with R(th) = [[c, -s], [s, c]]:
//   c >= s:  R = L(s/c) U(-s c) diag(c, 1/c)
//   s >  c:  R = diag(-1, 1) X R(th - pi/2), the same form with (c, s) -> (s, -c).
// The device applies diag(d0, d1) (in a register table), then the two shears
// (4 FMAs per amplitude pair, in place); diag(l0, l1) X^x stays pending.
// kind: 0 general, 1 M is diagonal, 2 M is anti-diagonal (no device op then).
// Returns false if rebuilding M from the factors misses by more than rounding.
bool shear_factor(const cd M[4], int &kind, cd &l0, cd &l1, int &x, double &u, double &t, cd &d0, cd &d1) {
    const double c = std::abs(M[0]), s = std::abs(M[2]);
    x = 0; u = t = 0; d0 = d1 = 1.0;
    if (s == 0.0 && std::abs(M[1]) == 0.0) {
        kind = 1; l0 = M[0]; l1 = M[3];
        return std::abs(std::abs(l0) - 1.0) < 1e-14 && std::abs(std::abs(l1) - 1.0) < 1e-14;
    }
    if (c == 0.0 && std::abs(M[3]) == 0.0) {
        kind = 2; l0 = M[1]; l1 = M[2]; x = 1;
        return std::abs(std::abs(l0) - 1.0) < 1e-14 && std::abs(std::abs(l1) - 1.0) < 1e-14;
    }
    kind = 0;
    if (c == 0.0 || s == 0.0) return false;  // not unitary
    l0 = M[0] / c;
    l1 = M[2] / s;
    const cd r = (c >= s) ? M[3] / (l1 * c) : -M[1] / (l0 * s);
    double cc = c, ss = s;
    if (c >= s) {
        x = 0;
    } else {  // diag(l0, l1) diag(-1, 1) X: pending -l0 on the |0> row
        x = 1;
        l0 = -l0;
        cc = s;
        ss = -c;
    }
    t = ss / cc;
    u = -ss * cc;
    d0 = cc;
    d1 = r / cc;
    // rebuild: diag(l0, l1) X^x L(t) U(u) diag(d0, d1)
    const cd S[4] = {d0, u * d1, t * d0, (1.0 + t * u) * d1};
    cd R4[4];
    if (!x) {
        R4[0] = l0 * S[0]; R4[1] = l0 * S[1]; R4[2] = l1 * S[2]; R4[3] = l1 * S[3];
    } else {
        R4[0] = l0 * S[2]; R4[1] = l0 * S[3]; R4[2] = l1 * S[0]; R4[3] = l1 * S[1];
    }
    for (int i = 0; i < 4; i++)
        if (std::abs(R4[i] - M[i]) > 8e-16 * (1.0 + std::abs(M[i]))) return false;
    return true;
}

void emit_general(OpList &L, const TileCtx &c, bool small, uint32_t kind, uint32_t tbit, uint64_t cm, uint64_t cv,
                  const double *m) {
    L.add(kind, tbit, cond_of(c, small, cm, cv), m, 8);
}

// Diagonal gates with two or more qubits (or a control): emitted as they come,
// condition values flipped on qubits with a pending X.
void emit_diagonal(const Gate &g, OpList &L, Fuse &F, const TileCtx &c, bool small) {
    const uint64_t cm = g.cmask;
    const uint64_t fl = F.flips(g.cmask | g.tmask);
    switch (g.type) {
        case OP_ONE: return;
        case OP_PAULI_Z: L.add(DQ_OP_NEG, 0, cond_of(c, small, cm, cm ^ fl), nullptr, 0); return;
        case OP_PHASE_SHIFT: {
            const double p[2] = {std::cos(g.op[0]), std::sin(g.op[0])};
            L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm, cm ^ fl), p, 2);
            return;
        }
        case OP_ROTATE_Z: {  // controlled: diag(e^{-ia}, e^{ia}) on the target where the controls are 1
            const double a = g.op[0];
            const uint64_t t = g.tmask;
            const double p0[2] = {std::cos(a), -std::sin(a)}, p1[2] = {std::cos(a), std::sin(a)};
            L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm | t, cm ^ fl), p0, 2);
            L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm | t, (cm | t) ^ fl), p1, 2);
            return;
        }
        case OP_COLLAPSE: {  // keep the outcome branch scaled by 1/sqrt(P), zero the other
            const int outcome = (int)g.op[0];
            const uint64_t t = g.tmask;
            const double zero[2] = {0, 0};
            L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm | t, (cm | (outcome ? 0 : t)) ^ fl), zero, 2);
            if (cm == 0 && F.on) {
                F.g *= g.op[1];
            } else {
                const double sc[2] = {g.op[1], 0};
                L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm | t, (cm | (outcome ? t : 0)) ^ fl), sc, 2);
            }
            return;
        }
        default: {  // phase gates e^{i k pi/4} where the mask qubits are all 1 (QuEST_gpu_kernel.cu:17-64)
            static const double pr[8] = {1, kRsqrt2, 0, -kRsqrt2, -1, -kRsqrt2, 0, kRsqrt2};
            static const double pi[8] = {0, kRsqrt2, 1, kRsqrt2, 0, -kRsqrt2, -1, -kRsqrt2};
            const double p[2] = {pr[g.type], pi[g.type]};
            L.add(DQ_OP_PHASE, 0, cond_of(c, small, cm, cm ^ fl), p, 2);
            return;
        }
    }
}

// Single-qubit diagonal without control?  Returns its qubit and factors.
bool single_diag(const Gate &g, int &q, cd &d0, cd &d1) {
    if (g.type == OP_ROTATE_Z) {
        if (g.cmask) return false;
        q = ctz64(g.tmask);
        d0 = std::polar(1.0, -g.op[0]);
        d1 = std::polar(1.0, g.op[0]);
        return true;
    }
    if (g.type > OP_PHASE_SHIFT || popc64(g.cmask) != 1) return false;
    q = ctz64(g.cmask);
    d0 = 1.0;
    if (g.type == OP_PHASE_SHIFT) d1 = std::polar(1.0, g.op[0]);
    else d1 = std::polar(1.0, std::atan(1.0) * g.type);
    if (g.type == OP_PAULI_Z) d1 = -1.0;
    if (g.type == OP_S_GATE) d1 = cd(0, 1);
    if (g.type == OP_S_GATE_CONJ) d1 = cd(0, -1);
    return true;
}

// U * P_q as a complex 2x2 (P_q = diag(p0, p1) X^x)
void absorb(const double m[8], const Fuse &F, int t, cd M[4]) {
    const cd U[4] = {cd(m[0], m[1]), cd(m[2], m[3]), cd(m[4], m[5]), cd(m[6], m[7])};
    M[0] = U[0] * F.p0[t]; M[1] = U[1] * F.p1[t];
    M[2] = U[2] * F.p0[t]; M[3] = U[3] * F.p1[t];
    if (F.x[t]) { std::swap(M[0], M[1]); std::swap(M[2], M[3]); }
}

void flat(const cd M[4], double m[8]) {
    for (int i = 0; i < 4; i++) { m[2 * i] = M[i].real(); m[2 * i + 1] = M[i].imag(); }
}

bool is_x(const cd M[4]) { return M[0] == cd(0.0) && M[3] == cd(0.0) && M[1] == cd(1.0) && M[2] == cd(1.0); }

// Ops of one phase.  With fusion on, uncontrolled single-qubit gates become
// sheared rotations plus pending monomials (see Fuse and shear_factor).
void pack_phase(const std::vector<Gate> &gates, Fuse &F, OpList &L, const TileCtx &c, bool small) {
    Batch b;
    dq_op z;
    memset(&z, 0, sizeof z);
    for (const Gate &g : gates) {
        const uint64_t touched = g.tmask | g.cmask;
        if (gate_is_diagonal(g.type)) {
            int q;
            cd d0, d1;
            if (F.on && single_diag(g, q, d0, d1)) {
                F.p0[q] *= d0;
                F.p1[q] *= d1;
                continue;
            }
            if (b.qmask & touched) flush_batch(b, L, c, F);
            emit_diagonal(g, L, F, c, small);
            continue;
        }
        const int t = ctz64(g.tmask);
        const uint32_t tb = small ? (uint32_t)t : (uint32_t)reg_of(c, t);
        double m[8];
        gate_matrix(g, m);
        if (!F.on) {
            emit_general(L, c, small, g.type == OP_PAULI_X ? DQ_OP_SWAP : DQ_OP_MAT, tb, g.cmask, g.cmask, m);
            continue;
        }
        if (b.qmask & touched) flush_batch(b, L, c, F);
        const bool last = --F.left[t] <= 0;
        if (g.cmask) {
            // C-U * P_t = P_t * C-(P_t^-1 U P_t); controls with a pending X flip
            const uint64_t cv = g.cmask ^ F.flips(g.cmask);
            if (last && F.x[t]) {  // the last gate on t leaves no pending X
                L.add(DQ_OP_SWAP, tb, z, nullptr, 0);
                F.x[t] = 0;
            }
            cd U[4] = {cd(m[0], m[1]), cd(m[2], m[3]), cd(m[4], m[5]), cd(m[6], m[7])};
            const cd a0 = F.p0[t], a1 = F.p1[t];
            cd V[4] = {U[0], U[1] * a1 / a0, U[2] * a0 / a1, U[3]};  // D^-1 U D
            if (F.x[t]) { std::swap(V[0], V[3]); std::swap(V[1], V[2]); }  // X V X
            double mv[8];
            flat(V, mv);
            emit_general(L, c, small, is_x(V) ? DQ_OP_SWAP : DQ_OP_MAT, tb, g.cmask, cv, mv);
            continue;
        }
        if (g.type == OP_PAULI_X) {  // X diag(p0, p1) X^x = diag(p1, p0) X^(x+1)
            std::swap(F.p0[t], F.p1[t]);
            F.x[t] ^= 1;
            if (last && F.x[t]) {
                L.add(DQ_OP_SWAP, tb, z, nullptr, 0);
                F.x[t] = 0;
            }
            continue;
        }
        cd M[4];
        absorb(m, F, t, M);
        if (!last) {
            int kind, x;
            cd l0, l1, d0, d1;
            double u, tt;
            if (shear_factor(M, kind, l0, l1, x, u, tt, d0, d1)) {
                F.p0[t] = l0;
                F.p1[t] = l1;
                F.x[t] = (uint8_t)x;
                if (kind == 0) {
                    const double sh[2] = {u, tt};
                    b.add((int)tb, DQ_OP_SHEAR, sh, 2, d0, d1, t);
                }
                continue;
            }
        } else {
            cd l0, l1, r1;
            double R4[4];
            if (real_factor(M, l0, l1, r1, R4)) {
                F.p0[t] = l0;
                F.p1[t] = l1;
                F.x[t] = 0;
                b.add((int)tb, DQ_OP_RMAT, R4, 4, 1.0, r1, t);
                continue;
            }
        }
        // fall back to the complex matrix (absorbs the pending monomial)
        double mu[8];
        flat(M, mu);
        F.p0[t] = F.p1[t] = 1.0;
        F.x[t] = 0;
        emit_general(L, c, small, DQ_OP_MAT, tb, 0, 0, mu);
    }
    flush_batch(b, L, c, F);
}

Layout standard_layout(int T, int R, int W) {
    Layout L;
    memset(&L, 0, sizeof L);
    for (int l = 0; l < 6; l++) L.lane[l] = (uint8_t)l;
    for (int r = 0; r < R; r++) L.reg[r] = (uint8_t)(6 + r);
    for (int w = 0; w < W; w++) L.wave[w] = (uint8_t)(6 + R + w);
    (void)T;
    return L;
}

bool lanes_identity(const Layout &L) {
    for (int l = 0; l < 6; l++)
        if (L.lane[l] != l) return false;
    return true;
}

// Layout whose register bits hold the tile bits of `reg_tbits` (<= R of
// them).  Lanes prefer tile bits 0..5 so that phases without low targets
// keep the coalesced lane order and need no exchange at the ends.
Layout layout_for(uint32_t reg_tbits, int T, int R, int W) {
    Layout L;
    memset(&L, 0, sizeof L);
    uint32_t used = 0;
    int nr = 0;
    for (int b = 0; b < T && nr < R; b++)
        if ((reg_tbits >> b) & 1) { L.reg[nr++] = (uint8_t)b; used |= 1u << b; }
    for (int b = 6; b < T && nr < R; b++)
        if (!((used >> b) & 1)) { L.reg[nr++] = (uint8_t)b; used |= 1u << b; }
    for (int b = 0; b < 6 && nr < R; b++)
        if (!((used >> b) & 1)) { L.reg[nr++] = (uint8_t)b; used |= 1u << b; }
    int nl = 0;
    for (int b = 0; b < T && nl < 6; b++)
        if (!((used >> b) & 1)) { L.lane[nl++] = (uint8_t)b; used |= 1u << b; }
    int nw = 0;
    for (int b = 0; b < T && nw < W; b++)
        if (!((used >> b) & 1)) { L.wave[nw++] = (uint8_t)b; used |= 1u << b; }
    return L;
}

// Layout of the next phase as pairwise swaps of A: register bits that stay
// targets keep their slots; each new target takes the slot of a register bit
// that is no longer one, which moves to the target's old lane or wave position.
// An exchange between such layouts needs no barrier between its passes
// (owned-slot exchanges).
Layout swap_layout(const Layout &A, uint32_t reg_tbits, int T, int R, int W) {
    Layout L = A;
    uint32_t in_regs = 0;
    for (int r = 0; r < R; r++) in_regs |= 1u << A.reg[r];
    int free_slots[DQ_MAX_REG_BITS], nf = 0;
    for (int r = 0; r < R; r++)
        if (!((reg_tbits >> A.reg[r]) & 1)) free_slots[nf++] = r;
    int k = 0;
    for (int b = 0; b < T && k < nf; b++) {
        if (!((reg_tbits >> b) & 1) || ((in_regs >> b) & 1)) continue;
        const int r = free_slots[k++];
        const uint8_t y = A.reg[r];
        for (int l = 0; l < 6; l++)
            if (L.lane[l] == b) L.lane[l] = y;
        for (int w = 0; w < W; w++)
            if (L.wave[w] == b) L.wave[w] = y;
        L.reg[r] = (uint8_t)b;
    }
    return L;
}

dq_phase phase_of(const Layout &L, int R, int W) {
    dq_phase ph;
    memset(&ph, 0, sizeof ph);
    for (int r = 0; r < R; r++) ph.reg_t[r] = 1u << L.reg[r];
    for (int w = 0; w < W; w++) ph.wave_t[w] = 1u << L.wave[w];
    for (int l = 0; l < 6; l++) ph.lane_t[l] = 1u << L.lane[l];
    return ph;
}

// Qubit masks of the register and wave bits at the two ends of a block.
void set_end_masks(dq_block &blk, const Layout &first, const Layout &last, const uint8_t *load_q,
                   const uint8_t *store_q, int R, int W) {
    for (int r = 0; r < R; r++) {
        blk.ld_reg[r] = 1ull << load_q[first.reg[r]];
        blk.st_reg[r] = 1ull << store_q[last.reg[r]];
    }
    for (int w = 0; w < W; w++) {
        blk.ld_wave[w] = 1ull << load_q[first.wave[w]];
        blk.st_wave[w] = 1ull << store_q[last.wave[w]];
    }
}

// Layout with the given lanes (tile bits); register bits from `prefer`'s
// where they are not lanes, then the lowest other tile bits; waves the rest.
Layout layout_with_lanes(const int *lane_tb, const Layout &prefer, int T, int R, int W) {
    Layout L;
    memset(&L, 0, sizeof L);
    uint32_t used = 0;
    for (int l = 0; l < 6; l++) { L.lane[l] = (uint8_t)lane_tb[l]; used |= 1u << lane_tb[l]; }
    int nr = 0;
    for (int r = 0; r < R; r++)
        if (!((used >> prefer.reg[r]) & 1)) { L.reg[nr++] = prefer.reg[r]; used |= 1u << prefer.reg[r]; }
    for (int b = 0; b < T && nr < R; b++)
        if (!((used >> b) & 1)) { L.reg[nr++] = (uint8_t)b; used |= 1u << b; }
    int nw = 0;
    for (int b = 0; b < T && nw < W; b++)
        if (!((used >> b) & 1)) { L.wave[nw++] = (uint8_t)b; used |= 1u << b; }
    return L;
}

// One fused block.  store_pos (or null: every tile qubit stays): the qubit
// position each tile bit is stored to, a permutation of the tile's own
// positions that keeps 0..5 among them; the block then ends in a layout whose
// lane l holds the tile bit stored to position l.
void emit_block(const std::vector<Gate> &bg, uint64_t tmask, int n, int nl, const TuneParams &t, int prec,
                int init_kind, uint64_t init_index, double init_value, Fuse &F, bool final, Plan &P,
                const uint8_t *store_pos = nullptr) {
    const int T = 6 + t.R + t.W, R = t.R, W = t.W;
    // tile qubits: the block's targets (always containing 0..5), padded with
    // the lowest other qubits (src/cpp/QuEST_gpu_kernel.cu:505-510 pads low to
    // high); all below nl (the piece's own qubits)
    uint64_t tile = tmask | 0x3Full;
    for (int q = 6; q < nl && popc64(tile) < T; q++) tile |= 1ull << q;
    TileCtx c;
    c.n = n; c.T = T; c.R = R; c.W = W; c.prec = prec;
    for (int q = 0; q < 64; q++) c.q2t[q] = -1;
    dq_block blk;
    memset(&blk, 0, sizeof blk);
    uint8_t tile_q[DQ_MAX_TILE_BITS];
    {
        int b = 0;
        for (int q = 0; q < n; q++)
            if ((tile >> q) & 1) { tile_q[b] = (uint8_t)q; c.q2t[q] = b; b++; }
    }
    blk.n_qubits = (uint32_t)nl;
    blk.tile_bits = (uint32_t)T;
    blk.reg_bits = (uint32_t)R;
    blk.wave_bits = (uint32_t)W;
    blk.init_kind = (uint32_t)init_kind;
    blk.init_index = init_index;
    blk.init_value = init_value;
    blk.nontile_mask = ((1ull << nl) - 1) & ~tile;

    // split the block into register phases of at most R targets each, in
    // both selection orders unless one is forced; the fewer phases win (ties:
    // the default order)
    std::vector<std::vector<Gate>> sets;
    std::vector<uint32_t> set_tbits;
    {
        auto split = [&](SelectOrder o, std::vector<std::vector<Gate>> &ss, std::vector<uint32_t> &tbs) {
            Dag pd(n, false);
            for (const auto &g : bg) pd.add(g);
            while (!pd.empty()) {
                std::vector<Gate> pg;
                uint64_t pm = 0;
                pd.select(pg, pm, R, (size_t)-1, false, o);
                uint32_t tb = 0;
                for (uint64_t m = pm; m; m &= m - 1) tb |= 1u << c.q2t[ctz64(m)];
                ss.push_back(std::move(pg));
                tbs.push_back(tb);
            }
        };
        split(select_order(), sets, set_tbits);
        if (!select_order_forced() && sets.size() > 1) {  // one phase cannot be beaten
            std::vector<std::vector<Gate>> s2;
            std::vector<uint32_t> t2;
            split(SelectOrder::DepthFirst, s2, t2);
            if (s2.size() < sets.size()) { sets.swap(s2); set_tbits.swap(t2); }
        }
        if (sets.empty()) { sets.emplace_back(); set_tbits.push_back(0); }
    }
    std::vector<Layout> lays;
    std::vector<const std::vector<Gate> *> opsets;
    static const std::vector<Gate> kNone;
    for (size_t i = 0; i < sets.size(); i++) {
        // after the first phase, pairwise swaps of the phase before
        Layout L = i > 0 ? swap_layout(lays.back(), set_tbits[i], T, R, W) : layout_for(set_tbits[i], T, R, W);
        if (i == 0 && !lanes_identity(L)) { lays.push_back(standard_layout(T, R, W)); opsets.push_back(&kNone); }
        lays.push_back(L);
        opsets.push_back(&sets[i]);
    }
    uint8_t store_q[DQ_MAX_TILE_BITS];
    for (int b = 0; b < T; b++) store_q[b] = store_pos ? store_pos[b] : tile_q[b];
    int lane_tb[6];
    bool lanes_moved = false;
    for (int l = 0; l < 6; l++) lane_tb[l] = -1;
    for (int b = 0; b < T; b++)
        if (store_q[b] < 6) lane_tb[store_q[b]] = b;
    for (int l = 0; l < 6; l++) lanes_moved |= lane_tb[l] != l;
    if (!lanes_moved) {
        if (!lanes_identity(lays.back())) { lays.push_back(standard_layout(T, R, W)); opsets.push_back(&kNone); }
    } else {
        bool same = true;
        for (int l = 0; l < 6; l++) same &= lays.back().lane[l] == lane_tb[l];
        // the last phase can take the store's lanes itself when it targets
        // none of them and is not the block's first (whose lanes the load
        // fixes); otherwise one more exchange, into a phase without ops
        uint32_t lane_set = 0;
        for (int l = 0; l < 6; l++) lane_set |= 1u << lane_tb[l];
        const uint32_t last_targets = set_tbits.empty() ? 0u : set_tbits.back();
        if (!same && lays.size() > 1 && opsets.back() != &kNone && !(last_targets & lane_set)) {
            lays.back() = layout_with_lanes(lane_tb, lays.back(), T, R, W);
            same = true;
        }
        if (!same) {
            lays.push_back(layout_with_lanes(lane_tb, lays.back(), T, R, W));
            opsets.push_back(&kNone);
        }
    }

    blk.phase_begin = (uint32_t)P.phases.size();
    blk.phase_count = (uint32_t)lays.size();
    set_end_masks(blk, lays.front(), lays.back(), tile_q, store_q, R, W);
    OpList L;
    for (size_t p = 0; p < lays.size(); p++) {
        dq_phase ph = phase_of(lays[p], R, W);
        ph.op_begin = (uint32_t)P.ops.size();
        c.roles = roles_of(lays[p], R, W);
        pack_phase(*opsets[p], F, L, c, false);
        if (final && p + 1 == lays.size()) apply_all_pending(F, L, c, n);
        L.flush_to(P.ops, prec);
        ph.op_count = (uint32_t)P.ops.size() - ph.op_begin;
        P.phases.push_back(ph);
    }
    P.blocks.push_back(blk);
    P.st.sweeps++;
    P.st.phases += (int)lays.size();
    P.st.exchanges += (int)lays.size() - 1;
}

}  // namespace

namespace {

// Positions after a block that hands lanes `from` over to `next` (masks of
// qubits by their start-of-plan label; pos: label -> position): each qubit
// leaving the lanes swaps positions with one entering them, both taken in
// position order.
void hand_over_lanes(uint64_t from, uint64_t next, std::vector<int> &pos) {
    std::vector<int> out, in;
    for (uint64_t m = from & ~next; m; m &= m - 1) out.push_back(ctz64(m));
    for (uint64_t m = next & ~from; m; m &= m - 1) in.push_back(ctz64(m));
    auto by_pos = [&](int a, int b) { return pos[a] < pos[b]; };
    std::sort(out.begin(), out.end(), by_pos);
    std::sort(in.begin(), in.end(), by_pos);
    for (size_t j = 0; j < out.size() && j < in.size(); j++) std::swap(pos[out[j]], pos[in[j]]);
}

uint64_t to_positions(uint64_t labels, const std::vector<int> &pos) {
    uint64_t m = 0;
    for (; labels; labels &= labels - 1) m |= 1ull << pos[ctz64(labels)];
    return m;
}

}  // namespace

void plan_circuit(const std::vector<Gate> &gates, int n, int prec, int init_kind, uint64_t init_index,
                  double init_value, const TuneParams &tp, Plan &P, int n_local, bool rotate_lanes) {
    if (n_local < 0 || n_local > n) n_local = n;
    P.blocks.clear();
    P.phases.clear();
    P.ops.clear();
    P.st = PlanStats();
    P.st.gates_in = (int)gates.size();

    if (n_local <= kSmallMax) {
        Dag d(n, true);
        for (const auto &g : gates) d.add(g);
        const std::vector<Gate> gl = d.gates();
        if (gl.empty() && init_kind == DQ_INIT_LOAD) return;
        TileCtx c;
        memset(&c, 0, sizeof c);
        c.prec = prec;
        dq_block blk;
        memset(&blk, 0, sizeof blk);
        blk.n_qubits = (uint32_t)n_local;
        blk.tile_bits = (uint32_t)n_local;
        blk.init_kind = (uint32_t)init_kind;
        blk.init_index = init_index;
        blk.init_value = init_value;
        blk.phase_begin = (uint32_t)P.phases.size();
        blk.phase_count = 1;
        dq_phase ph;
        memset(&ph, 0, sizeof ph);
        ph.op_begin = 0;
        Fuse F(0);
        OpList L;
        pack_phase(gl, F, L, c, true);
        L.flush_to(P.ops, prec);
        ph.op_count = (uint32_t)P.ops.size();
        P.phases.push_back(ph);
        P.blocks.push_back(blk);
        P.st.sweeps = 1;
        P.st.phases = 1;
        P.st.ops = (int)P.ops.size();
        return;
    }

    const TuneParams t = fit_shape(tp, n_local, prec);
    Fuse F(n);
    F.on = tp.fuse_diagonals;
    // fused blocks in both selection orders unless one is forced; the fewer
    // sweeps win (ties: the default order)
    std::vector<std::vector<Gate>> bgs;
    std::vector<uint64_t> tms;
    auto partition = [&](SelectOrder o, std::vector<std::vector<Gate>> &out, std::vector<uint64_t> &masks) {
        Dag d(n, true);
        for (const auto &g : gates) d.add(g);
        while (!d.empty()) {
            std::vector<Gate> bg;
            uint64_t tm = 0x3Full;
            d.select(bg, tm, 6 + t.k, (size_t)t.max_ops, false, o);
            out.push_back(std::move(bg));
            masks.push_back(tm);
        }
    };
    partition(select_order(), bgs, tms);
    if (!select_order_forced() && bgs.size() > 1) {  // one block cannot be beaten
        std::vector<std::vector<Gate>> b2;
        std::vector<uint64_t> m2;
        partition(SelectOrder::DepthFirst, b2, m2);
        if (b2.size() < bgs.size()) { bgs.swap(b2); tms.swap(m2); }
    }

    // rotating lanes: blocks over qubit labels (positions at the start of the
    // plan); lanes[i] = the labels at positions 0..5 during block i
    const int T = 6 + t.R + t.W;
    std::vector<std::vector<Gate>> rbg;
    std::vector<uint64_t> rtile, rlanes;
    if (rotate_lanes && n_local == n && n > T && bgs.size() > 1) {
        Dag d(n, true);
        for (const auto &g : gates) d.add(g);
        std::vector<int> pos(n), at(n);
        for (int q = 0; q < n; q++) pos[q] = at[q] = q;
        uint64_t lanes = 0x3Full;
        while (!d.empty() && rbg.size() < bgs.size()) {
            std::vector<Gate> bg;
            uint64_t tm = lanes;
            d.select(bg, tm, 6 + t.k, (size_t)t.max_ops, false, select_order());
            for (int p = 6; p < n && popc64(tm) < T; p++) tm |= 1ull << at[p];
            rbg.push_back(std::move(bg));
            rtile.push_back(tm);
            rlanes.push_back(lanes);
            // next lanes: the six tile qubits whose next targeted gates come
            // first (ties: the current lanes, then the lower position)
            const std::vector<size_t> first = d.first_target_use();
            std::vector<int> cand;
            for (uint64_t m = tm; m; m &= m - 1) cand.push_back(ctz64(m));
            std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) {
                if (first[a] != first[b]) return first[a] < first[b];
                const bool la = (lanes >> a) & 1, lb = (lanes >> b) & 1;
                if (la != lb) return la;
                return pos[a] < pos[b];
            });
            uint64_t next = 0;
            for (int i = 0; i < 6; i++) next |= 1ull << cand[i];
            hand_over_lanes(lanes, next, pos);
            for (int q = 0; q < n; q++) at[pos[q]] = q;
            lanes = next;
        }
        if (!d.empty() || rbg.size() >= bgs.size()) { rbg.clear(); rtile.clear(); rlanes.clear(); }
    }

    if (!rbg.empty()) {
        // emit in the positions of each block; the fusion state moves with
        // the qubits
        for (const auto &bg : rbg)
            for (const Gate &g : bg)
                if (!gate_is_diagonal(g.type)) F.left[ctz64(g.tmask)]++;
        std::vector<int> pos(n);
        for (int q = 0; q < n; q++) pos[q] = q;
        for (size_t i = 0; i < rbg.size(); i++) {
            std::vector<Gate> hg = rbg[i];
            for (Gate &g : hg) {
                g.tmask = to_positions(g.tmask, pos);
                g.cmask = to_positions(g.cmask, pos);
            }
            const uint64_t ptile = to_positions(rtile[i], pos);
            const uint64_t next = i + 1 < rbg.size() ? rlanes[i + 1] : rlanes[i];
            std::vector<int> pos2 = pos;
            hand_over_lanes(rlanes[i], next, pos2);
            std::vector<int> at(n);
            for (int q = 0; q < n; q++) at[pos[q]] = q;
            uint8_t store_pos[DQ_MAX_TILE_BITS];
            int b = 0;
            for (int p = 0; p < n; p++)
                if ((ptile >> p) & 1) store_pos[b++] = (uint8_t)pos2[at[p]];
            emit_block(hg, ptile, n, n_local, t, prec, i == 0 ? init_kind : DQ_INIT_LOAD, init_index, init_value, F,
                       i + 1 == rbg.size(), P, next != rlanes[i] ? store_pos : nullptr);
            if (next != rlanes[i]) {
                std::vector<int> to(n);
                for (int q = 0; q < n; q++) to[pos[q]] = pos2[q];
                F.permute(to);
            }
            pos.swap(pos2);
        }
        bool ident = true;
        for (int q = 0; q < n; q++) ident &= pos[q] == q;
        if (!ident) P.moved = pos;
        P.st.ops = (int)P.ops.size();
        return;
    }

    if (bgs.empty() && init_kind != DQ_INIT_LOAD) {
        bgs.emplace_back();
        tms.push_back(0x3Full);
    }
    for (const auto &bg : bgs)
        for (const Gate &g : bg)
            if (!gate_is_diagonal(g.type)) F.left[ctz64(g.tmask)]++;
    for (size_t i = 0; i < bgs.size(); i++)
        emit_block(bgs[i], tms[i], n, n_local, t, prec, i == 0 ? init_kind : DQ_INIT_LOAD, init_index, init_value, F,
                   i + 1 == bgs.size(), P);
    P.st.ops = (int)P.ops.size();
}

bool plan_permutation(int n, int prec, const std::vector<int> &src, const std::vector<int> &dst,
                      const TuneParams &tp, Plan &P, std::string &err) {
    P.blocks.clear();
    P.phases.clear();
    P.ops.clear();
    P.st = PlanStats();
    if (n <= kSmallMax) { err = "permutation pass needs n > 6"; return false; }
    const TuneParams t = fit_shape(tp, n, prec);
    const int T = 6 + t.R + t.W, R = t.R, W = t.W;
    if (src.size() != dst.size()) { err = "src/dst size mismatch"; return false; }
    uint64_t moved = 0;
    for (size_t i = 0; i < src.size(); i++) {
        if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) { err = "qubit out of range"; return false; }
        moved |= (1ull << src[i]) | (1ull << dst[i]);
    }
    uint64_t tile = moved | 0x3Full;
    if (popc64(tile) > T) { err = "moved qubits do not fit one tile"; return false; }
    for (int q = 6; q < n && popc64(tile) < T; q++) tile |= 1ull << q;
    dq_block blk;
    memset(&blk, 0, sizeof blk);
    int q2t[64];
    for (int q = 0; q < 64; q++) q2t[q] = -1;
    uint8_t load_q[DQ_MAX_TILE_BITS], store_q[DQ_MAX_TILE_BITS];
    {
        int b = 0;
        for (int q = 0; q < n; q++)
            if ((tile >> q) & 1) { load_q[b] = (uint8_t)q; q2t[q] = b; b++; }
    }
    for (int b = 0; b < T; b++) store_q[b] = load_q[b];
    uint64_t dst_seen = 0;
    for (size_t i = 0; i < src.size(); i++) {
        store_q[q2t[src[i]]] = (uint8_t)dst[i];
        dst_seen |= 1ull << dst[i];
    }
    {   // the result must be a permutation of the tile qubits
        uint64_t s = 0;
        for (int b = 0; b < T; b++) s |= 1ull << store_q[b];
        if (s != tile || popc64(dst_seen) != (int)dst.size()) { err = "not a permutation"; return false; }
    }
    blk.n_qubits = (uint32_t)n;
    blk.tile_bits = (uint32_t)T;
    blk.reg_bits = (uint32_t)R;
    blk.wave_bits = (uint32_t)W;
    blk.init_kind = DQ_INIT_LOAD;
    blk.nontile_mask = ((1ull << n) - 1) & ~tile;
    blk.phase_begin = 0;
    std::vector<Layout> lays;
    lays.push_back(standard_layout(T, R, W));
    bool low_ok = true;
    for (int l = 0; l < 6; l++) low_ok &= store_q[l] == l;
    if (!low_ok) {
        // second layout: lane l holds the tile bit stored to qubit l
        Layout L;
        memset(&L, 0, sizeof L);
        uint32_t used = 0;
        for (int l = 0; l < 6; l++)
            for (int b = 0; b < T; b++)
                if (store_q[b] == l) { L.lane[l] = (uint8_t)b; used |= 1u << b; }
        int nr = 0, nw = 0;
        for (int b = 0; b < T; b++) {
            if ((used >> b) & 1) continue;
            if (nr < R) L.reg[nr++] = (uint8_t)b;
            else L.wave[nw++] = (uint8_t)b;
        }
        lays.push_back(L);
    }
    for (const auto &L : lays) {
        dq_phase ph = phase_of(L, R, W);
        ph.op_begin = 0;
        ph.op_count = 0;
        P.phases.push_back(ph);
    }
    set_end_masks(blk, lays.front(), lays.back(), load_q, store_q, R, W);
    blk.phase_count = (uint32_t)P.phases.size();
    P.blocks.push_back(blk);
    P.st.sweeps = 1;
    P.st.phases = (int)P.phases.size();
    P.st.exchanges = (int)P.phases.size() - 1;
    return true;
}

void append_plan(Plan &dst, const Plan &src) {
    const uint32_t pb = (uint32_t)dst.phases.size(), ob = (uint32_t)dst.ops.size();
    for (dq_block b : src.blocks) {
        b.phase_begin += pb;
        dst.blocks.push_back(b);
    }
    for (dq_phase ph : src.phases) {
        ph.op_begin += ob;
        dst.phases.push_back(ph);
    }
    dst.ops.insert(dst.ops.end(), src.ops.begin(), src.ops.end());
    dst.st.sweeps += src.st.sweeps;
    dst.st.phases += src.st.phases;
    dst.st.exchanges += src.st.exchanges;
    dst.st.ops += src.st.ops;
}

bool plan_restore_mapping(int n, int prec, std::vector<int> &l2p, const TuneParams &tp, std::vector<Plan> &passes,
                          std::vector<Gate> &gates, std::string &err) {
    auto bit = [](int q) { return 1ull << q; };
    std::vector<int> p2l(n);
    for (int q = 0; q < n; q++) p2l[l2p[q]] = q;
    auto identity = [&] {
        for (int q = 0; q < n; q++)
            if (l2p[q] != q) return false;
        return true;
    };
    if (n <= kSmallMax) {
        // small registers: textbook SWAPs (3 CNOTs) on the physical positions
        for (int p = 0; p < n; p++) {
            const int cur = l2p[p];  // physical position of logical qubit p
            if (cur == p) continue;
            for (int k = 0; k < 3; k++) {
                Gate g;
                g.type = OP_PAULI_X;
                g.tmask = (k == 1) ? bit(p) : bit(cur);
                g.cmask = (k == 1) ? bit(cur) : bit(p);
                gates.push_back(g);
            }
            const int other = p2l[p];
            p2l[cur] = other;
            l2p[other] = cur;
            p2l[p] = p;
            l2p[p] = p;
        }
        return true;
    }
    const TuneParams t = fit_shape(tp, n, prec);
    const int cap_high = t.R + t.W;
    for (int guard = 0; guard < 4 * n + 4 && !identity(); guard++) {
        // positions taking part in this pass: pairs (current, home) while the
        // tile has room for their high qubits (qubits 0..5 are always in it)
        uint64_t S = 0;
        int high = 0;
        for (int lq = 0; lq < n; lq++) {
            const int cur = l2p[lq];
            if (cur == lq) continue;
            const uint64_t add = (bit(cur) | bit(lq)) & ~S;
            const int ah = popc64(add & ~0x3Full);
            if (high + ah > cap_high) continue;
            S |= add;
            high += ah;
        }
        // inside S every qubit whose home is in S goes home; the others fill
        // the remaining positions
        std::vector<int> np2l = p2l;
        uint64_t placed = 0;
        for (uint64_t m = S; m; m &= m - 1) {
            const int p = ctz64(m);
            if ((S >> l2p[p]) & 1) {
                np2l[p] = p;
                placed |= bit(p);
            }
        }
        std::vector<int> rest_q, rest_p;
        for (uint64_t m = S; m; m &= m - 1) {
            const int lq = p2l[ctz64(m)];
            if (!((placed >> lq) & 1)) rest_q.push_back(lq);
        }
        for (uint64_t m = S & ~placed; m; m &= m - 1) rest_p.push_back(ctz64(m));
        if (rest_q.size() != rest_p.size()) {
            err = "internal error in qubit re-placement";
            return false;
        }
        for (size_t i = 0; i < rest_p.size(); i++) np2l[rest_p[i]] = rest_q[i];
        std::vector<int> src, dst;
        for (uint64_t m = S; m; m &= m - 1) {
            const int p = ctz64(m);
            for (uint64_t m2 = S; m2; m2 &= m2 - 1)
                if (np2l[ctz64(m2)] == p2l[p] && ctz64(m2) != p) {
                    src.push_back(p);
                    dst.push_back(ctz64(m2));
                }
        }
        if (src.empty()) break;
        Plan P;
        if (!plan_permutation(n, prec, src, dst, tp, P, err)) return false;
        passes.push_back(std::move(P));
        p2l = np2l;
        for (int p = 0; p < n; p++) l2p[p2l[p]] = p;
    }
    if (!identity()) {
        err = "qubit re-placement did not converge";
        return false;
    }
    return true;
}

}  // namespace dgq
