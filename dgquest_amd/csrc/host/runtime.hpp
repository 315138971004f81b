// runtime.hpp -- register state and executor behind the QuEST C API.
//
// One GPU, one process.  The state of a register lives in HBM for its whole
// life; gates are recorded (in physical qubit positions) and flushed as fused
// sweeps when a result is read.  Replaces the reference's QuregState /
// QuESTEnvValueDistributed pair (src/cpp/QuEST_backend.h:314-431) for the
// resident-state case; MPI, task partitions and the page pool are gone.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "dgq_engine.h"
#include "gates.hpp"
#include "planner.hpp"

namespace dgq {
struct PagedStore;
}

struct QuESTEnvValue {
    int device = 0;
};

struct QuregState {
    int n = 0;
    int prec = DQ_F64;
    void *d_state = nullptr;
    void *d_ws = nullptr;
    size_t ws_cap = 0;
    void *stream = nullptr;
    std::vector<dgq::Gate> pending;   // physical qubit positions
    int init_kind = DQ_INIT_LOAD;     // pending initialisation (fused into the next sweep)
    uint64_t init_index = 0;
    double init_value = 0.0;
    std::vector<int> l2p, p2l;        // logical <-> physical qubit
    bool probs_dirty = true;
    std::vector<double> p0;           // per physical qubit, total at [n]
    dgq::TuneParams tune;
    dgq::PlanStats last_stats;
    unsigned long long last_bytes = 0;  // algorithmic HBM bytes of the last flush
    int last_launches = 0;
    double last_flush_ms = 0.0;
    bool recorder = false;            // records and plans only, no device state
    int dense_k = 0;                  // > 0: flushes run as dense sweeps of this width (FP32)
    int rotate = -1;                  // lane rotation in plans: 1 on, 0 off, -1 the default (on; DGQ_ROTATE)
    int rotate_pick = -1;             // searchQuregTaskPartitions' choice, used while rotate is -1
    double last_dense_flops = 0.0;    // device GEMM flops of the last dense flush
    dgq::PagedStore *paged = nullptr; // state in pinned host pieces, paged through the device
    // The last plan made for this register and what it was made from.  A
    // flush whose pending gates, initialisation, shape and lane rotation equal
    // these runs the cached plan without planning again: a circuit recorded
    // anew after an init* call, or the plan searchQuregTaskPartitions picked.
    struct {
        bool valid = false;
        std::vector<dgq::Gate> gates;
        int init_kind = 0;
        uint64_t init_index = 0;
        double init_value = 0.0;
        dgq::TuneParams tune;
        bool rotate = false;
        dgq::Plan plan;
    } cache;
    double predicted_ms = 0.0;        // searchQuregTaskPartitions' predicted device ms (0: no search)
    // While det_base holds, p0 is exact for the state before
    // the pending gates, and those are det_pending collapses of deterministic
    // measurements whose dropped branches held probability det_slack in all
    // (see measureWithStats).  Anything that changes the state clears det_base.
    bool det_base = false;
    int det_pending = 0;
    double det_slack = 0.0;
};

namespace dgq {

[[noreturn]] void engine_fatal(const char *what, int rc);
void ensure_ws(QuregState *s, size_t bytes);
void flush(QuregState *s);
void ensure_probs(QuregState *s);
void materialize_mapping(QuregState *s);

// MT19937 (Matsumoto & Nishimura's mt19937ar), process-global like the
// reference's src/mt19937ar.c:57-58.
void mt_init_by_array(const unsigned long *key, int len);
double mt_genrand_real1();

}  // namespace dgq
