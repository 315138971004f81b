// planner.cpp -- see planner.hpp.
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "dag.hpp"

namespace dgq {

static const double kRsqrt2 = 0.7071067811865475244008443621;
static const int kSmallMax = 6;  // n <= kSmallMax: whole-state-in-LDS kernel

TuneParams fit_shape(const TuneParams &tp, int n) {
    TuneParams t = tp;
    if (t.R < 1) t.R = 1;
    if (t.R > DQ_MAX_REG_BITS) t.R = DQ_MAX_REG_BITS;
    if (t.W < 0) t.W = 0;
    if (t.W > 3) t.W = 3;
    while (6 + t.R + t.W > n && t.W > 0) t.W--;
    while (6 + t.R + t.W > n && t.R > 1) t.R--;
    // instantiated shapes: W = 0 for R 1..5, W 1..2 for R 2..5, W = 3 for R 3..4
    if (t.W >= 1 && t.R < 2) t.R = 2;
    if (t.W == 3 && t.R > 4) t.W = 2;
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

// Device op(s) for one gate in the current phase layout (or, for the small
// form, in global-index form).  Returns false for a gate that needs nothing.
bool pack_gate(const Gate &g, const TileCtx &c, bool small, dq_op &o) {
    memset(&o, 0, sizeof o);
    const int prec = c.prec;
    auto ctrl = [&](uint64_t qm) {
        if (small) o.c_wg = qm;
        else split_mask(c, qm, o.c_reg, o.c_thr, o.c_wg);
    };
    auto sel = [&](uint64_t qm) {
        if (small) o.s_wg = qm;
        else split_mask(c, qm, o.s_reg, o.s_thr, o.s_wg);
    };
    auto target = [&](uint64_t qm) -> uint32_t {
        const int q = ctz64(qm);
        if (small) return (uint32_t)q;
        const int b = c.q2t[q];
        return (uint32_t)c.roles.idx[b];  // planner guarantees a register role
    };
    switch (g.type) {
        case OP_ONE: return false;
        case OP_PAULI_Z: o.kind = DQ_OP_NEG; ctrl(g.cmask); return true;
        case OP_T_GATE: case OP_S_GATE: case OP_T_GATE_CONJ_NEG: case OP_T_GATE_NEG:
        case OP_S_GATE_CONJ: case OP_T_GATE_CONJ: {
            // e^{i k pi/4}, QuEST_gpu_kernel.cu:17-64
            static const double pr[8] = {1, kRsqrt2, 0, -kRsqrt2, -1, -kRsqrt2, 0, kRsqrt2};
            static const double pi[8] = {0, kRsqrt2, 1, kRsqrt2, 0, -kRsqrt2, -1, -kRsqrt2};
            o.kind = DQ_OP_PHASE;
            put(o, 0, pr[g.type], prec);
            put(o, 1, pi[g.type], prec);
            ctrl(g.cmask);
            return true;
        }
        case OP_PHASE_SHIFT:
            o.kind = DQ_OP_PHASE;
            put(o, 0, std::cos(g.op[0]), prec);
            put(o, 1, std::sin(g.op[0]), prec);
            ctrl(g.cmask);
            return true;
        case OP_ROTATE_Z:  // diag(e^{-i a}, e^{+i a}) on the target bit, a = half angle
            o.kind = DQ_OP_DIAG2;
            put(o, 0, std::cos(g.op[0]), prec);
            put(o, 1, -std::sin(g.op[0]), prec);
            put(o, 2, std::cos(g.op[0]), prec);
            put(o, 3, std::sin(g.op[0]), prec);
            ctrl(g.cmask);
            sel(g.tmask);
            return true;
        case OP_COLLAPSE: {
            const int outcome = (int)g.op[0];
            o.kind = DQ_OP_DIAG2;
            put(o, 0, outcome == 0 ? g.op[1] : 0.0, prec);
            put(o, 1, 0.0, prec);
            put(o, 2, outcome == 1 ? g.op[1] : 0.0, prec);
            put(o, 3, 0.0, prec);
            ctrl(g.cmask);
            sel(g.tmask);
            return true;
        }
        case OP_PAULI_X:
            o.kind = DQ_OP_SWAP;
            o.tbit = target(g.tmask);
            ctrl(g.cmask);
            return true;
        default: {
            double m[8];
            gate_matrix(g, m);
            o.kind = DQ_OP_MAT;
            o.tbit = target(g.tmask);
            for (int i = 0; i < 8; i++) put(o, i, m[i], prec);
            ctrl(g.cmask);
            return true;
        }
    }
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

void emit_block(const std::vector<Gate> &bg, uint64_t tmask, int n, const TuneParams &t, int prec, int init_kind,
                uint64_t init_index, double init_value, Plan &P) {
    const int T = 6 + t.R + t.W, R = t.R, W = t.W;
    // tile qubits: the block's targets (always containing 0..5), padded with
    // the lowest other qubits (src/cpp/QuEST_gpu_kernel.cu:505-510 pads low to high)
    uint64_t tile = tmask | 0x3Full;
    for (int q = 6; q < n && popc64(tile) < T; q++) tile |= 1ull << q;
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
    blk.n_qubits = (uint32_t)n;
    blk.tile_bits = (uint32_t)T;
    blk.reg_bits = (uint32_t)R;
    blk.wave_bits = (uint32_t)W;
    blk.init_kind = (uint32_t)init_kind;
    blk.init_index = init_index;
    blk.init_value = init_value;
    blk.nontile_mask = ((1ull << n) - 1) & ~tile;

    // split the block into register phases of at most R targets each
    std::vector<std::vector<Gate>> sets;
    std::vector<uint32_t> set_tbits;
    {
        Dag pd(n, false);
        for (const auto &g : bg) pd.add(g);
        while (!pd.empty()) {
            std::vector<Gate> pg;
            uint64_t pm = 0;
            pd = pd.select(pg, pm, R, (size_t)-1);
            uint32_t tb = 0;
            for (uint64_t m = pm; m; m &= m - 1) tb |= 1u << c.q2t[ctz64(m)];
            sets.push_back(std::move(pg));
            set_tbits.push_back(tb);
        }
        if (sets.empty()) { sets.emplace_back(); set_tbits.push_back(0); }
    }
    std::vector<Layout> lays;
    std::vector<const std::vector<Gate> *> opsets;
    static const std::vector<Gate> kNone;
    for (size_t i = 0; i < sets.size(); i++) {
        Layout L = layout_for(set_tbits[i], T, R, W);
        if (i == 0 && !lanes_identity(L)) { lays.push_back(standard_layout(T, R, W)); opsets.push_back(&kNone); }
        lays.push_back(L);
        opsets.push_back(&sets[i]);
    }
    if (!lanes_identity(lays.back())) { lays.push_back(standard_layout(T, R, W)); opsets.push_back(&kNone); }

    blk.phase_begin = (uint32_t)P.phases.size();
    blk.phase_count = (uint32_t)lays.size();
    set_end_masks(blk, lays.front(), lays.back(), tile_q, tile_q, R, W);
    for (size_t p = 0; p < lays.size(); p++) {
        dq_phase ph = phase_of(lays[p], R, W);
        ph.op_begin = (uint32_t)P.ops.size();
        c.roles = roles_of(lays[p], R, W);
        for (const auto &g : *opsets[p]) {
            dq_op o;
            if (pack_gate(g, c, false, o)) P.ops.push_back(o);
        }
        ph.op_count = (uint32_t)P.ops.size() - ph.op_begin;
        P.phases.push_back(ph);
    }
    P.blocks.push_back(blk);
    P.st.sweeps++;
    P.st.phases += (int)lays.size();
    P.st.exchanges += (int)lays.size() - 1;
}

}  // namespace

void plan_circuit(const std::vector<Gate> &gates, int n, int prec, int init_kind, uint64_t init_index,
                  double init_value, const TuneParams &tp, Plan &P) {
    P.blocks.clear();
    P.phases.clear();
    P.ops.clear();
    P.st = PlanStats();
    P.st.gates_in = (int)gates.size();

    if (n <= kSmallMax) {
        Dag d(n, true);
        for (const auto &g : gates) d.add(g);
        const std::vector<Gate> gl = d.gates();
        if (gl.empty() && init_kind == DQ_INIT_LOAD) return;
        TileCtx c;
        memset(&c, 0, sizeof c);
        c.prec = prec;
        dq_block blk;
        memset(&blk, 0, sizeof blk);
        blk.n_qubits = (uint32_t)n;
        blk.tile_bits = (uint32_t)n;
        blk.init_kind = (uint32_t)init_kind;
        blk.init_index = init_index;
        blk.init_value = init_value;
        blk.phase_begin = (uint32_t)P.phases.size();
        blk.phase_count = 1;
        dq_phase ph;
        memset(&ph, 0, sizeof ph);
        ph.op_begin = 0;
        for (const auto &g : gl) {
            dq_op o;
            if (pack_gate(g, c, true, o)) P.ops.push_back(o);
        }
        ph.op_count = (uint32_t)P.ops.size();
        P.phases.push_back(ph);
        P.blocks.push_back(blk);
        P.st.sweeps = 1;
        P.st.phases = 1;
        P.st.ops = (int)P.ops.size();
        return;
    }

    const TuneParams t = fit_shape(tp, n);
    Dag d(n, true);
    for (const auto &g : gates) d.add(g);
    bool first = true;
    while (!d.empty()) {
        std::vector<Gate> bg;
        uint64_t tm = 0x3Full;
        d = d.select(bg, tm, 6 + t.k, (size_t)t.max_ops);
        emit_block(bg, tm, n, t, prec, first ? init_kind : DQ_INIT_LOAD, init_index, init_value, P);
        first = false;
    }
    if (first && init_kind != DQ_INIT_LOAD) {
        emit_block(std::vector<Gate>(), 0x3Full, n, t, prec, init_kind, init_index, init_value, P);
    }
    P.st.ops = (int)P.ops.size();
}

bool plan_permutation(int n, int prec, const std::vector<int> &src, const std::vector<int> &dst,
                      const TuneParams &tp, Plan &P, std::string &err) {
    P.blocks.clear();
    P.phases.clear();
    P.ops.clear();
    P.st = PlanStats();
    if (n <= kSmallMax) { err = "permutation pass needs n > 6"; return false; }
    const TuneParams t = fit_shape(tp, n);
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

}  // namespace dgq
