// tune.hpp -- the persisted fusion-shape table.
//
// Replaces the reference's time predictor and its offline sampler
// (src/cpp/time_predictor.cpp:127-395, sampler/main.cpp:7-13), which model
// per-gate and per-kernel times to choose a task partition.  Here the choice
// is the fused-sweep shape (k, R, W, max_ops) per precision and register
// size, measured directly on the device by dgquest_amd/autotune.py and
// stored as text, one entry per line:
//     <prec 1|2> <n> <k> <R> <W> <max_ops> <ms of the tuning circuit>
// '#' starts a comment.  A register takes the entry of its precision with the
// nearest n (the larger n on a tie); DGQ_TUNE=k,R,W,max_ops overrides the
// table, and without either the planner defaults hold.
//
// Lines of a second kind give the device-time model of one shape:
//     cost <prec> <k> <R> <W> <sweep> <phase> <op> <launch>
// with which searchQuregTaskPartitions predicts the device time of a plan
// (predict_plan_ms) and picks a shape per recorded circuit -- the role of the
// reference's predictor (src/cpp/time_predictor.cpp:127-395) in
// searchQuregTaskPartitions (src/cpp/QuEST_backend_distributed.cpp:86-121).
#pragma once
#include <string>
#include <vector>

#include "planner.hpp"

namespace dgq {

// The shape for an n-qubit register of precision prec (DQ_F32 / DQ_F64).
TuneParams tune_for(int prec, int n);

// Device time of one fused sweep (block) b of a plan, in ms:
//   launch + GB_b * (sweep + phase * (phases_b - 1) + op * ops_b)
// GB_b: bytes the block moves / 1e9 (a read and a write of the state, or only
// the write when the block generates its input); phases_b, ops_b: the block's
// register phases and device ops.  Fitted per shape by autotune.py from
// measured flushes.
struct CostModel {
    double sweep = 0.0;   // ms per GB moved
    double phase = 0.0;   // ms per GB moved, per phase after the first
    double op = 0.0;      // ms per GB moved, per device op
    double launch = 0.0;  // ms per launch
};

// The model of shape tp (k, R, W) in precision prec: its table line, else a
// built-in estimate (HBM-bound sweeps at 5.5 TB/s, measured phase and op
// costs of the 13-qubit FP64 tile scaled by the tile's size).
CostModel tune_cost(int prec, const TuneParams &tp);

// Shapes the per-circuit search tries: those with cost lines for prec, else
// the built-in candidate list.
std::vector<TuneParams> tune_candidates(int prec);

// Predicted device ms of a plan for a register of precision prec (resident,
// one job per launch).
double predict_plan_ms(const Plan &P, int prec, const CostModel &c);

// (Re)load the table from `path` (empty: DGQ_TUNE_TABLE, else tune_table.txt
// in the directory above the library's own).  Returns the number of entries,
// or -1 if the file cannot be read (the table is then empty).
int tune_load(const std::string &path);

}  // namespace dgq
