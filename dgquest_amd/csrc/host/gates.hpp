// gates.hpp -- the recorded-gate model of the host.
//
// One record per API call, as in the reference's Gate
// (src/cpp/QuEST_backend.h:160-178): a type, a target mask, a control mask
// and an operand.  Type numbering follows src/cpp/QuEST_backend.h:89-114:
// types 0..7 are the phase gates e^{i k pi/4} (they keep their qubit in the
// control mask and have no target), rotations store half angles.
// OP_COLLAPSE is this engine's own: the projection done by measure().
#pragma once
#include <cstdint>
#include <vector>

namespace dgq {

enum GateType : int {
    OP_ONE = 0,
    OP_T_GATE = 1,
    OP_S_GATE = 2,
    OP_T_GATE_CONJ_NEG = 3,
    OP_PAULI_Z = 4,
    OP_T_GATE_NEG = 5,
    OP_S_GATE_CONJ = 6,
    OP_T_GATE_CONJ = 7,
    OP_PHASE_SHIFT = 8,
    OP_ROTATE_Z_FAKE = 9,
    OP_ROTATE_Z = 10,
    OP_PAULI_X = 11,
    OP_PAULI_Y = 12,
    OP_PAULI_Y_CONJ = 13,
    OP_HADAMARD = 14,
    OP_ROTATE_X = 15,
    OP_ROTATE_Y = 16,
    OP_COMPACT_UNITARY = 17,
    OP_UNITARY = 18,
    OP_COLLAPSE = 19,  // op[0] = outcome (0/1), op[1] = 1/sqrt(P(outcome))
};

// Operand layout in op[8]: angle at op[0]; compact unitary alpha = (op0, op1),
// beta = (op2, op3); 2x2 unitary r0c0, r0c1, r1c0, r1c1 = (op0,op1) .. (op6,op7).
struct Gate {
    double op[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int type = OP_ONE;
    uint64_t tmask = 0;
    uint64_t cmask = 0;
};

// Diagonal gates touch no amplitude pair; they never need their qubit in a
// register (src/cpp/DAG.cpp:224-228 counts only types >= OP_PAULI_X).
inline bool gate_is_diagonal(int type) { return type <= OP_ROTATE_Z || type == OP_COLLAPSE; }

inline uint64_t gate_target_bits(const Gate &g) { return gate_is_diagonal(g.type) ? 0 : g.tmask; }

inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }
inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

// Why a gate record that did not come through the API cannot be planned, or
// nullptr when it can (the shapes the API itself records):
//   phase types 0..8: no target, their qubits in the control mask;
//   OP_ROTATE_Z_FAKE: kernel-internal in the reference
//     (src/cpp/QuEST_backend.h:102), never a recorded gate;
//   every other type: exactly one target qubit, not among the controls;
//   OP_COLLAPSE: outcome op[0] is 0 or 1.
// The caller checks that both masks lie inside the register.
inline const char *gate_record_error(int type, uint64_t tmask, uint64_t cmask, const double *op) {
    if (type < OP_ONE || type > OP_COLLAPSE) return "unknown gate type";
    if (type == OP_ROTATE_Z_FAKE) return "type 9 (ROTATE_Z_FAKE) is not a recordable gate";
    if (type <= OP_PHASE_SHIFT) {
        if (tmask) return "a phase gate has no target (its qubits go in the control mask)";
        if (!cmask) return "a phase gate needs at least one qubit in its control mask";
        return nullptr;
    }
    if (!tmask || (tmask & (tmask - 1))) return "the gate needs exactly one target qubit";
    if (tmask & cmask) return "the target qubit is also a control";
    if (type == OP_COLLAPSE && !(op[0] == 0.0 || op[0] == 1.0)) return "a collapse outcome must be 0 or 1";
    return nullptr;
}

}  // namespace dgq
