// planner.hpp -- turns a recorded gate list into fused sweeps for the engine.
//
// Replaces the reference's planning chain for the one-GPU case:
//   splitSectionLocalViewIntoFusedGates  src/cpp/QuEST_backend.cpp:118-141
//     (blocks: targetMask starts with the coalesced qubits and may grow to a
//      tile; here the coalesced set is qubits 0..5 and the tile is 6 + k)
//   launchKernel per-target regrouping   src/cpp/QuEST_gpu_kernel.cu:529-546
//     (here: phases, each at most R register targets)
//   localizeTargetBits / partitionFusedOperator / statevec_fused
//                                        src/cpp/QuEST_gpu_kernel.cu:361-527
//     (here: control and target qubits translated to register, lane, wave or
//      tile-index bits; sin/cos on the host)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "dgq_engine.h"
#include "gates.hpp"

namespace dgq {

// Shape of a fused sweep.  k: maximum number of target qubits above the six
// coalesced ones per block (the fusion width); R/W: register and wave bits
// of the kernel (tile = 6 + R + W qubits, k <= R + W); max_ops caps the gates
// per block so long blocks stay memory-bound.
struct TuneParams {
    int k = 6;
    int R = 4;
    int W = 2;
    int max_ops = 4096;
    bool fuse_diagonals = true;  // real rotations + deferred diagonals (see planner.cpp)
};

struct PlanStats {
    int sweeps = 0;      // fused blocks (one full read + write of the state each)
    int phases = 0;      // register phases over all blocks
    int exchanges = 0;   // LDS re-distributions over all blocks
    int ops = 0;         // device ops
    int gates_in = 0;    // recorded gates before merging
};

struct Plan {
    std::vector<dq_block> blocks;
    std::vector<dq_phase> phases;
    std::vector<dq_op> ops;
    PlanStats st;
    // Set when the plan moves qubits (rotating lanes): the qubit at physical
    // position a before the plan is at moved[a] after it.  Empty otherwise.
    std::vector<int> moved;
};

// Plan `gates` (physical qubit indices) on an n-qubit state.  The first block
// generates the state instead of reading it when init_kind is
// DQ_INIT_CLASSICAL / DQ_INIT_PLUS.
//
// n_local < n plans for a register held in pieces of 2^n_local amplitudes
// (the qubits n_local .. n-1 index the piece): every non-diagonal target must
// be below n_local; controls and diagonals on the piece qubits become
// workgroup conditions, resolved per piece through dq_job.base_hi.
//
// rotate_lanes (whole registers only): the six lane positions 0..5 may take
// other qubits between blocks.  Each block's store moves the six tile qubits
// whose next targeted gates come first into positions 0..5, so the next block
// has 13 useful tile qubits rather than 7 and the six fixed ones.  The
// partition is kept only when it needs fewer sweeps than fixed lanes; the
// caller composes out.moved into its qubit map.
void plan_circuit(const std::vector<Gate> &gates, int n, int prec, int init_kind, uint64_t init_index,
                  double init_value, const TuneParams &tp, Plan &out, int n_local = -1, bool rotate_lanes = false);

// A pass that only moves qubits: afterwards physical qubit dst[i] holds what
// physical qubit src[i] held.  All moved qubits must fit one tile.
bool plan_permutation(int n, int prec, const std::vector<int> &src, const std::vector<int> &dst,
                      const TuneParams &tp, Plan &out, std::string &err);

// Qubit moves that return every logical qubit to its own physical position
// (l2p[q] = physical position of logical qubit q; the identity afterwards).
// n > 6: permutation passes, one sweep each, appended to `passes`; n <= 6:
// SWAPs as CNOT triples on physical positions, appended to `gates` for the
// caller to plan.  False (with err) if planning fails.
bool plan_restore_mapping(int n, int prec, std::vector<int> &l2p, const TuneParams &tp, std::vector<Plan> &passes,
                          std::vector<Gate> &gates, std::string &err);

// Append src's blocks, phases and ops to dst (indices rebased).
void append_plan(Plan &dst, const Plan &src);

// Clamp a tuning choice to an n-qubit state (tile = 6 + R + W <= n).
TuneParams fit_shape(const TuneParams &tp, int n, int prec);

}  // namespace dgq
