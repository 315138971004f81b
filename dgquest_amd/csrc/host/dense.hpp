// dense.hpp -- planning of dense sweeps (the FP32 precision form).
//
// A flush becomes dense blocks: tiles of T = min(13, n) qubits (qubits 0..5
// always in, for coalesced loads) selected on the full qubit support of every
// gate, each split into dense ops whose support is at most k qubits.  An op's
// 2^k x 2^k complex matrix is the product of its gates, formed on the host in
// double precision; the device applies it to the FP32 state on MFMA
// (dq_apply_dense).  Block selection is the reference's DAG selection
// (src/cpp/DAG.cpp:219-259) counting controls and diagonals as well, since a
// dense op has no free diagonal qubits.
#pragma once
#include <string>
#include <vector>

#include "dgq_engine.h"
#include "gates.hpp"

namespace dgq {

struct DensePlan {
    int k = 6;
    std::vector<dq_dense_block> blocks;
    std::vector<dq_dense_op> ops;
    std::vector<int> op_qubits;   // k qubits per op: matrix-index bit j acts on op_qubits[k * op + j]
    std::vector<double> mats;     // per op: 2^k x 2^k row-major {re, im}
    double flops = 0.0;           // real flops of the device GEMMs: 8 * 2^k * 2^n per op
};

// False (with err) when the register is too small for dense sweeps of width
// k (k + 5 <= min(n, 13) is needed) or k is not 5, 6 or 7.
bool plan_dense(const std::vector<Gate> &gates, int n, int k, DensePlan &out, std::string &err);

// Apply one recorded gate to a state of m qubits held as interleaved
// {re, im} doubles, gate masks already in this state's bit positions.  The
// per-type arithmetic of src/cpp/QuEST_gpu_kernel.cu:17-177.
void apply_gate_small(double *psi, int m, const Gate &g);

}  // namespace dgq
