// dense.cpp -- see dense.hpp.
#include "dense.hpp"

#include <cmath>
#include <cstring>

#include "dag.hpp"

namespace dgq {

static const double kR2 = 0.7071067811865475244008443621;

void apply_gate_small(double *psi, int m, const Gate &g) {
    const uint64_t N = 1ull << m;
    const uint64_t cm = g.cmask;
    auto at = [&](uint64_t i, double &r, double &im) { r = psi[2 * i]; im = psi[2 * i + 1]; };
    auto put = [&](uint64_t i, double r, double im) { psi[2 * i] = r; psi[2 * i + 1] = im; };
    if (g.type < OP_PHASE_SHIFT) {
        // e^{i k pi/4} on amplitudes whose control-mask bits are all 1
        static const double pr[8] = {1, kR2, 0, -kR2, -1, -kR2, 0, kR2};
        static const double pi[8] = {0, kR2, 1, kR2, 0, -kR2, -1, -kR2};
        if (g.type == OP_ONE) return;
        const double c = pr[g.type], s = pi[g.type];
        for (uint64_t i = 0; i < N; i++) {
            if ((i & cm) != cm) continue;
            double r, x;
            at(i, r, x);
            put(i, c * r - s * x, c * x + s * r);
        }
        return;
    }
    if (g.type == OP_PHASE_SHIFT) {
        const double c = std::cos(g.op[0]), s = std::sin(g.op[0]);
        for (uint64_t i = 0; i < N; i++) {
            if ((i & cm) != cm) continue;
            double r, x;
            at(i, r, x);
            put(i, c * r - s * x, c * x + s * r);
        }
        return;
    }
    if (g.type == OP_COLLAPSE) {
        const int outcome = (int)g.op[0];
        const double scale = g.op[1];
        for (uint64_t i = 0; i < N; i++) {
            const int bit = (i & g.tmask) != 0;
            if (bit == outcome) put(i, psi[2 * i] * scale, psi[2 * i + 1] * scale);
            else put(i, 0.0, 0.0);
        }
        return;
    }
    if (g.type == OP_ROTATE_Z_FAKE) return;
    double u[8];
    if (g.type == OP_ROTATE_Z) {  // diag(e^{-i a}, e^{i a}), a = recorded half angle
        const double c = std::cos(g.op[0]), s = std::sin(g.op[0]);
        const double d[8] = {c, -s, 0, 0, 0, 0, c, s};
        memcpy(u, d, sizeof u);
    } else {
        gate_matrix(g, u);
    }
    for (uint64_t i0 = 0; i0 < N; i0++) {
        if (i0 & g.tmask) continue;
        if ((i0 & cm) != cm) continue;
        const uint64_t i1 = i0 | g.tmask;
        double r0, m0, r1, m1;
        at(i0, r0, m0);
        at(i1, r1, m1);
        put(i0, u[0] * r0 - u[1] * m0 + u[2] * r1 - u[3] * m1, u[0] * m0 + u[1] * r0 + u[2] * m1 + u[3] * r1);
        put(i1, u[4] * r0 - u[5] * m0 + u[6] * r1 - u[7] * m1, u[4] * m0 + u[5] * r0 + u[6] * m1 + u[7] * r1);
    }
}

namespace {

// qubit mask -> mask over the positions given by q (q[j] = qubit of bit j)
uint64_t to_local(uint64_t mask, const int *q, int k) {
    uint64_t r = 0;
    for (int j = 0; j < k; j++)
        if ((mask >> q[j]) & 1) r |= 1ull << j;
    return r;
}

// The product of `gs` (acting on qubits q[0..k-1]) as a 2^k x 2^k row-major
// complex matrix: column c is the image of basis state c.
void op_matrix(const std::vector<Gate> &gs, const int *q, int k, double *out) {
    const uint64_t K = 1ull << k;
    std::vector<double> col(2 * K);
    for (uint64_t c = 0; c < K; c++) {
        std::fill(col.begin(), col.end(), 0.0);
        col[2 * c] = 1.0;
        for (const Gate &g : gs) {
            Gate l = g;
            l.tmask = to_local(g.tmask, q, k);
            l.cmask = to_local(g.cmask, q, k);
            apply_gate_small(col.data(), k, l);
        }
        for (uint64_t r = 0; r < K; r++) {
            out[2 * (r * K + c)] = col[2 * r];
            out[2 * (r * K + c) + 1] = col[2 * r + 1];
        }
    }
}

}  // namespace

bool plan_dense(const std::vector<Gate> &gates, int n, int k, DensePlan &P, std::string &err) {
    P = DensePlan();
    P.k = k;
    if (k < 5 || k > 7) { err = "dense width k must be 5, 6 or 7"; return false; }
    const int T = n < 13 ? n : 13;
    if (T < k + 5) { err = "register too small for dense sweeps of this width"; return false; }
    const uint64_t K = 1ull << k;
    Dag d(n, true);
    for (const Gate &g : gates) d.add(g);
    while (!d.empty()) {
        std::vector<Gate> bg;
        uint64_t tm = 0x3Full;
        // depth-first selection: a dense block counts every qubit a gate
        // touches, and following chains of partners packs more gates per
        // block than the front-first scan (32-qubit grid, k = 7: 34 sweeps
        // and 52 dense ops against 45 and 66)
        d.select(bg, tm, T, (size_t)-1, true, SelectOrder::DepthFirst);
        if (bg.empty()) { err = "dense planning made no progress"; return false; }
        for (int q = 6; q < n && popc64(tm) < T; q++) tm |= 1ull << q;
        dq_dense_block blk;
        memset(&blk, 0, sizeof blk);
        blk.n_qubits = (uint32_t)n;
        blk.tile_bits = (uint32_t)T;
        blk.nontile_mask = ((n == 64) ? ~0ull : ((1ull << n) - 1)) & ~tm;
        int q2t[64];
        for (int q = 0; q < 64; q++) q2t[q] = -1;
        {
            int b = 0;
            for (int q = 0; q < n; q++)
                if ((tm >> q) & 1) { blk.tile_q[b] = (uint8_t)q; q2t[q] = b; b++; }
        }
        blk.op_begin = (uint32_t)P.ops.size();
        // dense ops of at most k qubits each, in dependency order
        Dag od(n, false);
        for (const Gate &g : bg) od.add(g);
        while (!od.empty()) {
            std::vector<Gate> og;
            uint64_t om = 0;
            od.select(og, om, k, (size_t)-1, true, SelectOrder::DepthFirst);
            if (og.empty()) { err = "dense op selection made no progress"; return false; }
            // pad the support to exactly k qubits with other tile qubits
            for (int b = 0; b < T && popc64(om) < k; b++)
                if (!((om >> blk.tile_q[b]) & 1)) om |= 1ull << blk.tile_q[b];
            int q[8];
            int j = 0;
            for (uint64_t m = om; m; m &= m - 1) q[j++] = ctz64(m);
            dq_dense_op op;
            memset(&op, 0, sizeof op);
            for (int i = 0; i < k; i++) {
                op.bits[i] = (uint32_t)q2t[q[i]];
                P.op_qubits.push_back(q[i]);
            }
            op.mat = P.ops.size();
            const size_t mo = P.mats.size();
            P.mats.resize(mo + 2 * K * K);
            op_matrix(og, q, k, P.mats.data() + mo);
            P.ops.push_back(op);
            P.flops += 8.0 * (double)K * std::ldexp(1.0, n);
        }
        blk.op_count = (uint32_t)P.ops.size() - blk.op_begin;
        P.blocks.push_back(blk);
    }
    return true;
}

}  // namespace dgq
