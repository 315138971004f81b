// dag.cpp -- see dag.hpp.  Merge and selection rules follow the reference's
// src/cpp/DAG.cpp:22-259; the data structure is this engine's own (node
// pool with index links, explicit DFS stack instead of recursion).
#include "dag.hpp"

#include <cstdlib>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <utility>

namespace dgq {

static const double kRsqrt2 = 0.7071067811865475244008443621;

void gate_matrix(const Gate &g, double m[8]) {
    for (int i = 0; i < 8; i++) m[i] = 0.0;
    const double a = g.op[0];
    switch (g.type) {
        case OP_ROTATE_Z:
            m[0] = std::cos(a); m[1] = -std::sin(a);
            m[6] = std::cos(a); m[7] = std::sin(a);
            break;
        case OP_PAULI_X: m[2] = 1; m[4] = 1; break;
        case OP_PAULI_Y: m[3] = -1; m[5] = 1; break;
        case OP_PAULI_Y_CONJ: m[3] = 1; m[5] = -1; break;
        case OP_HADAMARD: m[0] = kRsqrt2; m[2] = kRsqrt2; m[4] = kRsqrt2; m[6] = -kRsqrt2; break;
        case OP_ROTATE_X:
            m[0] = std::cos(a); m[3] = -std::sin(a);
            m[5] = -std::sin(a); m[6] = std::cos(a);
            break;
        case OP_ROTATE_Y:
            m[0] = std::cos(a); m[2] = -std::sin(a);
            m[4] = std::sin(a); m[6] = std::cos(a);
            break;
        case OP_COMPACT_UNITARY:  // [[alpha, -conj(beta)], [beta, conj(alpha)]]
            m[0] = g.op[0]; m[1] = g.op[1];
            m[2] = -g.op[2]; m[3] = g.op[3];
            m[4] = g.op[2]; m[5] = g.op[3];
            m[6] = g.op[0]; m[7] = -g.op[1];
            break;
        case OP_UNITARY:
            for (int i = 0; i < 8; i++) m[i] = g.op[i];
            break;
        default:
            break;
    }
}

// C = A * B for 2x2 complex matrices stored (re,im) row-major.
static void mat_mul(double C[8], const double A[8], const double B[8]) {
    auto mul = [](double ar, double ai, double br, double bi, double &cr, double &ci) {
        cr = ar * br - ai * bi;
        ci = ar * bi + ai * br;
    };
    double t1r, t1i, t2r, t2i;
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) {
            mul(A[4 * r + 0], A[4 * r + 1], B[2 * c + 0], B[2 * c + 1], t1r, t1i);
            mul(A[4 * r + 2], A[4 * r + 3], B[4 + 2 * c + 0], B[4 + 2 * c + 1], t2r, t2i);
            C[4 * r + 2 * c + 0] = t1r + t2r;
            C[4 * r + 2 * c + 1] = t1i + t2i;
        }
}

Dag::Dag(int num_qubits, bool merge) : nq_(num_qubits), merge_(merge), stacks_(num_qubits) {}

void Dag::erase_node(int top) {
    Node &t = nodes_[top];
    for (auto &s : stacks_)
        if (!s.empty() && s.back() == top) s.pop_back();
    if (t.prev.empty()) {
        front_.erase(std::remove(front_.begin(), front_.end(), top), front_.end());
    } else {
        for (int p : t.prev) {
            auto &r = nodes_[p].rear;
            r.erase(std::remove(r.begin(), r.end(), top), r.end());
        }
    }
    t.alive = false;
    alive_count_--;
}

bool Dag::try_merge(int top, Node &nn) {
    Node &tn = nodes_[top];
    if (nn.g.tmask != tn.g.tmask || nn.g.cmask != tn.g.cmask) return false;
    // gate2 is the one with the larger type (ties: the earlier gate)
    const bool node_is_2 = nn.g.type > tn.g.type;
    const Gate &g1 = node_is_2 ? tn.g : nn.g;
    const Gate &g2 = node_is_2 ? nn.g : tn.g;
    if (g2.type < OP_PHASE_SHIFT) {
        const int t = (g1.type + g2.type) % 8;
        if (t == 0) {
            erase_node(top);
        } else {
            tn.g.type = t;
        }
        return true;
    }
    if (g2.type == OP_PHASE_SHIFT) {
        Gate m = g2;
        m.op[0] = g2.op[0] + (g1.type == OP_PHASE_SHIFT ? g1.op[0] : std::atan(1.0) * g1.type);
        tn.g = m;
        return true;
    }
    if (g2.type == OP_UNITARY) {
        // the later gate multiplies from the left
        double a[8], u[8], c[8];
        Gate m = g2;
        if (!node_is_2) {  // the earlier gate (tn) is the unitary
            gate_matrix(g1, a);
            for (int i = 0; i < 8; i++) u[i] = g2.op[i];
            mat_mul(c, a, u);
        } else {           // the new gate is the unitary
            gate_matrix(g1, a);
            for (int i = 0; i < 8; i++) u[i] = g2.op[i];
            mat_mul(c, u, a);
        }
        for (int i = 0; i < 8; i++) m.op[i] = c[i];
        tn.g = m;
        return true;
    }
    if (g1.type == g2.type) {
        if (g1.type >= OP_PAULI_X && g1.type <= OP_HADAMARD) {
            erase_node(top);
            return true;
        }
        if (g1.type == OP_ROTATE_X || g1.type == OP_ROTATE_Y || g1.type == OP_ROTATE_Z) {
            if (g1.op[0] == -g2.op[0]) {
                erase_node(top);
                return true;
            }
            Gate m = g1;
            m.op[0] = g1.op[0] + g2.op[0];
            tn.g = m;
            return true;
        }
    }
    return false;
}

void Dag::add(const Gate &g) {
    nodes_.emplace_back();
    const int id = (int)nodes_.size() - 1;
    {
        Node &nn = nodes_[id];
        nn.g = g;
        nn.barrier = g.tmask | g.cmask;
    }
    alive_count_++;
    for (uint64_t bm = nodes_[id].barrier; bm; bm &= bm - 1) {
        const int i = ctz64(bm);
        auto &st = stacks_[i];
        if (!st.empty()) {
            const int top = st.back();
            if (merge_ && nodes_[top].rear.empty() && try_merge(top, nodes_[id])) {
                nodes_.pop_back();
                alive_count_--;
                if (nodes_[top].alive) remerge(top);
                return;
            }
            nodes_[id].prev.push_back(top);
            nodes_[top].rear.push_back(id);
        }
        st.push_back(id);
    }
    if (nodes_[id].prev.empty()) front_.push_back(id);
}

// A node whose gate just changed (it absorbed a later gate) may now merge
// into its predecessor, as it would if it were added anew: when that
// predecessor is its immediate one on every qubit and has no other
// successor.  Merging eagerly here means the DAG left after a selection
// needs no rebuild (no pair of remaining gates can newly merge).
void Dag::remerge(int id) {
    while (id >= 0) {
        Node &b = nodes_[id];
        if (!b.alive || b.prev.empty() || !b.rear.empty()) return;
        const int a = b.prev[0];
        for (int p : b.prev)
            if (p != a) return;
        Node &an = nodes_[a];
        for (int r : an.rear)
            if (r != id) return;
        if (an.g.tmask != b.g.tmask || an.g.cmask != b.g.cmask) return;
        // unlink b (it is the top of all its stacks), then offer it to a
        const std::vector<int> saved_rear = an.rear;
        for (uint64_t m = b.barrier; m; m &= m - 1) stacks_[ctz64(m)].pop_back();
        an.rear.clear();
        b.alive = false;
        alive_count_--;
        Node nb;
        nb.g = b.g;
        nb.barrier = b.barrier;
        if (!try_merge(a, nb)) {  // no side effects: restore the link
            Node &b2 = nodes_[id];
            b2.alive = true;
            alive_count_++;
            nodes_[a].rear = saved_rear;
            for (uint64_t m = b2.barrier; m; m &= m - 1) stacks_[ctz64(m)].push_back(id);
            return;
        }
        id = nodes_[a].alive ? a : -1;
    }
}

// Node ids are a topological order: a gate's predecessors were added before
// it, and a merge keeps the earlier node.  One pass in id order takes every
// gate whose predecessors are taken and whose targets fit, so the target
// budget goes to the gates nearest the front on every qubit, where the
// depth-first walk spends it along the first front gate's chain of partners.
void Dag::scan(uint64_t &mask, int limit) {
    for (auto &nd : nodes_) {
        if (!nd.alive) continue;
        bool ready = true;
        for (int p : nd.prev)
            if (!nodes_[p].selected) { ready = false; break; }
        if (ready && popc64(mask | nd.tbits) <= limit) {
            nd.selected = true;
            mask |= nd.tbits;
        }
    }
}

SelectOrder select_order() {
    static const SelectOrder o = [] {
        const char *e = getenv("DGQ_SELECT");
        return (e && strcmp(e, "dfs") == 0) ? SelectOrder::DepthFirst : SelectOrder::Scan;
    }();
    return o;
}

bool select_order_forced() {
    static const bool f = getenv("DGQ_SELECT") != nullptr;
    return f;
}

void Dag::dfs(int start, uint64_t &mask, int limit) {
    // Same visiting order as a recursive depth-first walk.
    std::vector<std::pair<int, size_t>> stk;
    auto visit = [&](int v) {
        Node &nd = nodes_[v];
        nd.visited = true;
        if (popc64(mask | nd.tbits) <= limit) {
            nd.selected = true;
            mask |= nd.tbits;
            stk.emplace_back(v, 0);
        }
    };
    visit(start);
    while (!stk.empty()) {
        auto &top = stk.back();
        Node &u = nodes_[top.first];
        if (top.second < u.rear.size()) {
            const int v = u.rear[top.second++];
            Node &rv = nodes_[v];
            if (!rv.visited && --rv.remain == 0) visit(v);
        } else {
            stk.pop_back();
        }
    }
}

void Dag::select(std::vector<Gate> &out, uint64_t &target_mask, int limit, size_t max_gates, bool full_support,
                 SelectOrder order) {
    for (auto &nd : nodes_) {
        if (!nd.alive) continue;
        nd.tbits = full_support ? nd.barrier : gate_target_bits(nd.g);
        nd.remain = nd.prev.size();
        nd.visited = false;
        nd.selected = false;
    }
    if (order == SelectOrder::DepthFirst) {
        const std::vector<int> front = front_;
        for (int f : front) dfs(f, target_mask, limit);
    } else {
        scan(target_mask, limit);
    }
    // The taken gates (the first max_gates selected ones in insertion order)
    // are closed under predecessors, so on every qubit they form a prefix of
    // its chain: removing them leaves the links among the others as they are
    // (and, merges being eager, no two of the others can newly merge).
    size_t taken = 0;
    for (auto &nd : nodes_) {
        if (!nd.alive || !nd.selected || taken >= max_gates) continue;
        out.push_back(nd.g);
        nd.alive = false;
        nd.prev.clear();
        nd.rear.clear();
        taken++;
    }
    if (!taken) return;
    alive_count_ -= taken;
    front_.clear();
    for (size_t i = 0; i < nodes_.size(); i++) {
        Node &nd = nodes_[i];
        if (!nd.alive) continue;
        nd.prev.erase(std::remove_if(nd.prev.begin(), nd.prev.end(), [&](int p) { return !nodes_[p].alive; }),
                      nd.prev.end());
        if (nd.prev.empty()) front_.push_back((int)i);
    }
    for (auto &st : stacks_) {
        size_t k = 0;
        while (k < st.size() && !nodes_[st[k]].alive) k++;
        if (k) st.erase(st.begin(), st.begin() + (long)k);
    }
}

uint64_t Dag::front_targets() const {
    uint64_t m = 0;
    for (int f : front_)
        if (nodes_[f].alive) m |= gate_target_bits(nodes_[f].g);
    return m;
}

std::vector<size_t> Dag::first_target_use() const {
    std::vector<size_t> first((size_t)nq_, SIZE_MAX);
    for (size_t i = 0; i < nodes_.size(); i++) {
        if (!nodes_[i].alive) continue;
        for (uint64_t m = gate_target_bits(nodes_[i].g); m; m &= m - 1) {
            const int q = ctz64(m);
            if (q < nq_ && first[q] == SIZE_MAX) first[q] = i;
        }
    }
    return first;
}

std::vector<Gate> Dag::gates() const {
    std::vector<Gate> v;
    v.reserve(alive_count_);
    for (auto &nd : nodes_)
        if (nd.alive) v.push_back(nd.g);
    return v;
}

}  // namespace dgq
