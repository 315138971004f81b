// dag.hpp -- gate dependency DAG with peephole merges and greedy
// target-bounded block selection.
//
// Behaviour restated from the reference's src/cpp/DAG.cpp:
//   add()     per-qubit barrier stacks; a new gate merges into the latest gate
//             on its qubits when both masks are equal and that gate has no
//             successor yet (:22-217): phase types add modulo 8 and vanish at
//             0; phase shifts add angles; 2x2 unitaries multiply; equal X, Y,
//             Y*, H cancel; equal rotations add half angles or cancel.
//   select()  takes every gate whose non-diagonal targets keep
//             popcount(targetMask) <= limit; the taken set is closed under
//             predecessors and is emitted in insertion order, capped at
//             max_gates; everything else goes to a new DAG (:219-259).  The
//             reference visits gates depth-first from the front nodes; here
//             the default visits them in topological (id) order, which fills
//             the target budget from the front on every qubit and needs
//             fewer blocks (30-qubit layered U3+CZ: 30 sweeps instead of 35).
//             DGQ_SELECT=dfs restores the depth-first order.
// The reference builds that new DAG by re-adding the remaining gates, which
// also merges pairs that became mergeable after their link was made (a gate
// that absorbed a later gate can now merge into its predecessor).  Here such
// merges happen at once, when the gate changes, so the selected gates are
// simply removed in place.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "gates.hpp"

namespace dgq {

enum class SelectOrder { Scan, DepthFirst };
// The default order select() visits gates in: Scan unless DGQ_SELECT=dfs.
SelectOrder select_order();
// Whether DGQ_SELECT fixes the order (callers that try both orders and keep
// the better partition then use only that one).
bool select_order_forced();

class Dag {
public:
    explicit Dag(int num_qubits, bool merge = true);

    void add(const Gate &g);
    bool empty() const { return alive_count_ == 0; }
    size_t size() const { return alive_count_; }

    // Move the selected gates to `out` (appended) and remove them from this
    // DAG; grow `target_mask`.  full_support counts every qubit a gate
    // touches (controls and diagonal gates too), for dense blocks that form
    // one matrix over their qubits.
    void select(std::vector<Gate> &out, uint64_t &target_mask, int limit, size_t max_gates,
                bool full_support = false, SelectOrder order = select_order());

    // All surviving gates in insertion order (a valid topological order).
    std::vector<Gate> gates() const;

    // Non-diagonal targets of the gates that have no predecessor.
    uint64_t front_targets() const;

    // Per qubit: the smallest node id (a topological position) of a
    // remaining gate with a non-diagonal target on it; SIZE_MAX if none.
    std::vector<size_t> first_target_use() const;

private:
    struct Node {
        Gate g;
        std::vector<int> prev, rear;
        uint64_t barrier = 0;
        uint64_t tbits = 0;
        size_t remain = 0;
        bool visited = false, selected = false, alive = true;
    };
    bool try_merge(int top, Node &nn);
    void remerge(int id);
    void erase_node(int top);
    void dfs(int start, uint64_t &mask, int limit);
    void scan(uint64_t &mask, int limit);

    int nq_;
    bool merge_;
    size_t alive_count_ = 0;
    std::vector<Node> nodes_;
    std::vector<std::vector<int>> stacks_;
    std::vector<int> front_;
};

// 2x2 matrix of a non-diagonal gate (rows r0c0, r0c1, r1c0, r1c1 as re,im),
// as src/cpp/QuEST_backend.h:480-530 defines it.
void gate_matrix(const Gate &g, double m[8]);

}  // namespace dgq
