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
#pragma once
#include <string>

#include "planner.hpp"

namespace dgq {

// The shape for an n-qubit register of precision prec (DQ_F32 / DQ_F64).
TuneParams tune_for(int prec, int n);

// (Re)load the table from `path` (empty: DGQ_TUNE_TABLE, else tune_table.txt
// in the directory above the library's own).  Returns the number of entries,
// or -1 if the file cannot be read (the table is then empty).
int tune_load(const std::string &path);

}  // namespace dgq
