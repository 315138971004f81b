// quest.cpp -- the QuEST C API (include/QuEST.h) on the gfx950 engine.
//
// Front-end behaviour follows the reference:
//   argument validation, error codes and messages   src/QuEST_validation.c:19-263
//   API -> gate record encoding                     src/cpp/QuEST_distributed.cpp:118-364
//   u1/u2/u3/SqX/SqY/SqW matrices, fSim             src/QuEST.c:320-423
//   rotateAroundAxis -> compact unitary             src/QuEST_common.c:56-63,154-183
//   measurement outcome rule, seeding               src/QuEST_common.c:97-152
//   lazy execution, flushed by getAmp / calcProbOfOutcome
//                                                   src/cpp/QuEST_distributed.cpp:106-116
//   P(1) = 1 - P(0)                                 src/cpp/QuEST_backend_distributed.cpp:735-737
// The execution underneath is this engine's: one resident state per register
// in HBM, fused sweeps planned by planner.cpp, kernels behind dgq_engine.h.
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "QuEST.h"
#include "dag.hpp"
#include "paging.hpp"
#include "runtime.hpp"
#include "tune.hpp"
#include "dense.hpp"

using namespace dgq;

// ---------------------------------------------------------------------------
// validation (error codes and messages as src/QuEST_validation.c:19-80)
// ---------------------------------------------------------------------------
namespace {
enum ErrorCode {
    E_SUCCESS = 0,
    E_INVALID_NUM_QUBITS,
    E_INVALID_TARGET_QUBIT,
    E_INVALID_CONTROL_QUBIT,
    E_INVALID_STATE_INDEX,
    E_INVALID_NUM_AMPS,
    E_INVALID_OFFSET_NUM_AMPS,
    E_TARGET_IS_CONTROL,
    E_TARGET_IN_CONTROLS,
    E_TARGETS_NOT_UNIQUE,
    E_INVALID_NUM_CONTROLS,
    E_NON_UNITARY_MATRIX,
    E_NON_UNITARY_COMPLEX_PAIR,
    E_ZERO_VECTOR,
    E_SYS_TOO_BIG_TO_PRINT,
    E_COLLAPSE_STATE_ZERO_PROB,
    E_INVALID_QUBIT_OUTCOME,
    E_CANNOT_OPEN_FILE,
    E_SECOND_ARG_MUST_BE_STATEVEC,
    E_MISMATCHING_QUREG_DIMENSIONS,
    E_MISMATCHING_QUREG_TYPES,
    E_DEFINED_ONLY_FOR_STATEVECS,
    E_DEFINED_ONLY_FOR_DENSMATRS,
};

const char *error_message(ErrorCode c) {
    switch (c) {
        case E_INVALID_NUM_QUBITS: return "Invalid number of qubits. Must create >0.";
        case E_INVALID_TARGET_QUBIT: return "Invalid target qubit. Note qubits are zero indexed.";
        case E_INVALID_CONTROL_QUBIT: return "Invalid control qubit. Note qubits are zero indexed.";
        case E_INVALID_STATE_INDEX: return "Invalid state index. Must be >=0 and <2^numQubits.";
        case E_INVALID_NUM_AMPS: return "Invalid number of amplitudes. Must be >=0 and <=2^numQubits.";
        case E_INVALID_OFFSET_NUM_AMPS: return "More amplitudes given than exist in the statevector from the given starting index.";
        case E_TARGET_IS_CONTROL: return "Control qubit cannot equal target qubit.";
        case E_TARGET_IN_CONTROLS: return "Control qubits cannot include target qubit.";
        case E_TARGETS_NOT_UNIQUE: return "The two target qubits must be unique.";
        case E_INVALID_NUM_CONTROLS: return "Invalid number of control qubits. Must be >0 and <numQubits.";
        case E_NON_UNITARY_MATRIX: return "Matrix is not unitary.";
        case E_NON_UNITARY_COMPLEX_PAIR: return "Compact matrix formed by given complex numbers is not unitary.";
        case E_ZERO_VECTOR: return "Invalid axis vector. Must be non-zero.";
        case E_COLLAPSE_STATE_ZERO_PROB: return "Can't collapse to state with zero probability.";
        case E_INVALID_QUBIT_OUTCOME: return "Invalid measurement outcome -- must be either 0 or 1.";
        case E_DEFINED_ONLY_FOR_STATEVECS: return "Operation valid only for state-vectors.";
        case E_SECOND_ARG_MUST_BE_STATEVEC: return "Second argument must be a state-vector.";
        case E_MISMATCHING_QUREG_DIMENSIONS: return "Dimensions of the qubit registers don't match.";
        case E_MISMATCHING_QUREG_TYPES: return "Registers must both be state-vectors or both be density matrices.";
        default: return "Unknown error.";
    }
}

[[noreturn]] void exit_with_error(ErrorCode code, const char *func) {
    printf("!!!\n");
    printf("QuEST Error in function %s: %s\n", func, error_message(code));
    printf("!!!\n");
    printf("exiting..\n");
    fflush(stdout);
    exit(code);
}

void quest_assert(bool ok, ErrorCode code, const char *func) {
    if (!ok) exit_with_error(code, func);
}

void v_target(const Qureg &q, int t, const char *f) {
    quest_assert(t >= 0 && (u64)t < q.numQubitsRepresented, E_INVALID_TARGET_QUBIT, f);
}
void v_control(const Qureg &q, int c, const char *f) {
    quest_assert(c >= 0 && (u64)c < q.numQubitsRepresented, E_INVALID_CONTROL_QUBIT, f);
}
void v_control_target(const Qureg &q, int c, int t, const char *f) {
    v_target(q, t, f);
    v_control(q, c, f);
    quest_assert(c != t, E_TARGET_IS_CONTROL, f);
}
void v_unique_targets(const Qureg &q, int a, int b, const char *f) {
    v_target(q, a, f);
    v_target(q, b, f);
    quest_assert(a != b, E_TARGETS_NOT_UNIQUE, f);
}
void v_multi_controls(const Qureg &q, const int *cs, int nc, const char *f) {
    quest_assert(nc > 0 && (u64)nc <= q.numQubitsRepresented, E_INVALID_NUM_CONTROLS, f);
    for (int i = 0; i < nc; i++) v_control(q, cs[i], f);
}
void v_multi_controls_target(const Qureg &q, const int *cs, int nc, int t, const char *f) {
    v_target(q, t, f);
    v_multi_controls(q, cs, nc, f);
    for (int i = 0; i < nc; i++) quest_assert(cs[i] != t, E_TARGET_IN_CONTROLS, f);
}
void v_state_index(const Qureg &q, long long idx, const char *f) {
    const long long mx = 1LL << q.numQubitsRepresented;
    quest_assert(idx >= 0 && idx < mx, E_INVALID_STATE_INDEX, f);
}
void v_num_amps(const Qureg &q, long long start, long long num, const char *f) {
    v_state_index(q, start, f);
    quest_assert(num >= 0 && (u64)num <= q.numAmpsTotal, E_INVALID_NUM_AMPS, f);
    quest_assert((u64)(num + start) <= q.numAmpsTotal, E_INVALID_OFFSET_NUM_AMPS, f);
}
bool is_matrix_unitary(const double m[8]) {
    // src/QuEST_validation.c:110-128, tolerance REAL_EPS
    const double eps = REAL_EPS;
    if (std::fabs(m[0] * m[0] + m[1] * m[1] + m[4] * m[4] + m[5] * m[5] - 1) > eps) return false;
    if (std::fabs(m[2] * m[2] + m[3] * m[3] + m[6] * m[6] + m[7] * m[7] - 1) > eps) return false;
    if (std::fabs(m[0] * m[2] + m[1] * m[3] + m[4] * m[6] + m[5] * m[7]) > eps) return false;
    if (std::fabs(m[2] * m[1] - m[0] * m[3] + m[6] * m[5] - m[4] * m[7]) > eps) return false;
    return true;
}
void v_unitary(const double m[8], const char *f) { quest_assert(is_matrix_unitary(m), E_NON_UNITARY_MATRIX, f); }
void v_complex_pair(Complex a, Complex b, const char *f) {
    const double s = (double)a.real * a.real + (double)a.imag * a.imag + (double)b.real * b.real + (double)b.imag * b.imag;
    quest_assert(std::fabs(-1 + s) < REAL_EPS, E_NON_UNITARY_COMPLEX_PAIR, f);
}
void v_vector(Vector v, const char *f) {
    quest_assert(std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z) > REAL_EPS, E_ZERO_VECTOR, f);
}
void v_outcome(int o, const char *f) { quest_assert(o == 0 || o == 1, E_INVALID_QUBIT_OUTCOME, f); }
void v_statevec(const Qureg &q, const char *f) { quest_assert(!q.isDensityMatrix, E_DEFINED_ONLY_FOR_STATEVECS, f); }

void mat_from(const ComplexMatrix2 &u, double m[8]) {
    m[0] = u.r0c0.real; m[1] = u.r0c0.imag; m[2] = u.r0c1.real; m[3] = u.r0c1.imag;
    m[4] = u.r1c0.real; m[5] = u.r1c0.imag; m[6] = u.r1c1.real; m[7] = u.r1c1.imag;
}

// ---------------------------------------------------------------------------
// recording
// ---------------------------------------------------------------------------
QuregState *st(const Qureg &q) { return q.state; }

uint64_t phys_mask(const QuregState *s, uint64_t logical) {
    uint64_t m = 0;
    while (logical) {
        const int q = ctz64(logical);
        logical &= logical - 1;
        m |= 1ull << s->l2p[q];
    }
    return m;
}

void record(const Qureg &q, int type, uint64_t tmask_l, uint64_t cmask_l, const double *op = nullptr) {
    QuregState *s = st(q);
    Gate g;
    g.type = type;
    g.tmask = phys_mask(s, tmask_l);
    g.cmask = phys_mask(s, cmask_l);
    if (op)
        for (int i = 0; i < 8; i++) g.op[i] = op[i];
    s->pending.push_back(g);
    s->probs_dirty = true;
}

inline uint64_t bit(int q) { return 1ull << q; }

void rec_unitary(const Qureg &q, int t, uint64_t cm, const double m[8]) { record(q, OP_UNITARY, bit(t), cm, m); }

}  // namespace

// ---------------------------------------------------------------------------
// executor
// ---------------------------------------------------------------------------
namespace dgq {

void engine_fatal(const char *what, int rc) {
    fprintf(stderr, "dgquest engine error in %s (code %d): %s\n", what, rc, dq_last_error());
    fflush(stderr);
    exit(EXIT_FAILURE);
}

#define ENG(call)                                 \
    do {                                          \
        int rc_ = (call);                         \
        if (rc_ != DQ_OK) engine_fatal(#call, rc_); \
    } while (0)

void ensure_ws(QuregState *s, size_t bytes) {
    if (bytes <= s->ws_cap) return;
    if (s->d_ws) {
        ENG(dq_stream_sync(s->stream));  // kernels in flight may still read it
        ENG(dq_free(s->d_ws));
    }
    size_t cap = 1 << 20;
    while (cap < bytes) cap *= 2;
    ENG(dq_malloc(&s->d_ws, cap));
    s->ws_cap = cap;
}

static void run_plan(QuregState *s, const Plan &P) {
    if (P.blocks.empty()) return;
    const uint32_t nb = (uint32_t)P.blocks.size();
    std::vector<dq_job> jobs(nb);
    std::vector<uint32_t> begin(nb + 1);
    for (uint32_t i = 0; i < nb; i++) {
        jobs[i].state = s->d_state;
        jobs[i].block = i;
        jobs[i].pad = 0;
        jobs[i].base_hi = 0;
        begin[i] = i;
    }
    begin[nb] = nb;
    ensure_ws(s, dq_apply_ws_bytes(nb, (uint32_t)P.phases.size(), (uint32_t)P.ops.size(), nb));
    ENG(dq_apply_blocks(s->prec, P.blocks.data(), nb, P.phases.data(), (uint32_t)P.phases.size(),
                        P.ops.empty() ? nullptr : P.ops.data(), (uint32_t)P.ops.size(), jobs.data(), begin.data(), nb,
                        s->d_ws, s->ws_cap, s->stream));
}

// Rotating lanes in whole-register plans (DGQ_ROTATE=0 turns them off, for
// comparison).
static bool rotate_lanes() {
    static const bool on = [] {
        const char *e = getenv("DGQ_ROTATE");
        return !e || atoi(e) != 0;
    }();
    return on;
}

// Lane rotation of a register's plans: the caller's setting, else the
// per-circuit search's pick, else the default.
static bool same_shape(const TuneParams &a, const TuneParams &b) {
    return a.k == b.k && a.R == b.R && a.W == b.W && a.max_ops == b.max_ops && a.fuse_diagonals == b.fuse_diagonals;
}

static bool same_gate(const Gate &a, const Gate &b) {
    return a.type == b.type && a.tmask == b.tmask && a.cmask == b.cmask && !memcmp(a.op, b.op, sizeof a.op);
}

// Was the register's last plan made from exactly what it now holds?
static bool plan_is_cached(const QuregState *s, const TuneParams &t, bool rot) {
    const auto &c = s->cache;
    if (!c.valid || c.rotate != rot || !same_shape(c.tune, t) || c.init_kind != s->init_kind ||
        c.init_index != s->init_index || c.init_value != s->init_value || c.gates.size() != s->pending.size())
        return false;
    for (size_t i = 0; i < c.gates.size(); i++)
        if (!same_gate(c.gates[i], s->pending[i])) return false;
    return true;
}

static void cache_plan(QuregState *s, const TuneParams &t, bool rot, const Plan &P) {
    auto &c = s->cache;
    c.valid = true;
    c.gates = s->pending;
    c.init_kind = s->init_kind;
    c.init_index = s->init_index;
    c.init_value = s->init_value;
    c.tune = t;
    c.rotate = rot;
    c.plan = P;
}

static bool rotation_allowed(const QuregState *s) { return s->rotate >= 0 ? s->rotate != 0 : rotate_lanes(); }
static bool rotation_for(const QuregState *s) {
    if (s->rotate < 0 && s->rotate_pick >= 0) return s->rotate_pick != 0;
    return rotation_allowed(s);
}

void flush(QuregState *s) {
    if (s->recorder) engine_fatal("a recording register cannot execute or read its state", DQ_ERR_UNSUPPORTED);
    if (s->pending.empty() && s->init_kind == DQ_INIT_LOAD) return;
    s->det_base = false;
    if (s->paged) {
        const size_t nrec = s->pending.size();
        paged_flush(s);
        s->last_stats = PlanStats();
        s->last_stats.gates_in = (int)nrec;
        s->last_stats.sweeps = s->paged->st.sections;
        s->last_bytes = s->paged->st.bytes_in + s->paged->st.bytes_out;
        s->last_launches = s->paged->st.sections;
        s->last_flush_ms = s->paged->st.wall_ms;
        return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const unsigned long long state_bytes = (unsigned long long)(s->prec == DQ_F64 ? 16 : 8) << s->n;
    s->last_dense_flops = 0.0;
    if (s->pending.empty()) {
        ENG(dq_init_state(s->d_state, s->n, s->prec, s->init_kind, s->init_index, s->init_value, s->stream));
        s->last_stats = PlanStats();
        s->last_bytes = state_bytes;
        s->last_launches = 1;
    } else if (s->dense_k) {
        // precision form: the initial state is written first, then dense sweeps
        s->last_bytes = 0;
        if (s->init_kind != DQ_INIT_LOAD) {
            ENG(dq_init_state(s->d_state, s->n, s->prec, s->init_kind, s->init_index, s->init_value, s->stream));
            s->last_bytes += state_bytes;
        }
        DensePlan D;
        std::string err;
        if (!plan_dense(s->pending, s->n, s->dense_k, D, err)) engine_fatal(("dense planning: " + err).c_str(), DQ_ERR_ARG);
        const uint32_t nb = (uint32_t)D.blocks.size(), no = (uint32_t)D.ops.size();
        ensure_ws(s, dq_dense_ws_bytes(s->dense_k, nb, no, no));
        ENG(dq_apply_dense(s->dense_k, D.blocks.data(), nb, D.ops.data(), no, D.mats.data(), no, s->d_state, s->d_ws,
                           s->ws_cap, s->stream));
        s->last_stats = PlanStats();
        s->last_stats.sweeps = (int)nb;
        s->last_stats.ops = (int)no;
        s->last_stats.gates_in = (int)s->pending.size();
        s->last_launches = (int)nb;
        s->last_bytes += 2 * state_bytes * nb;
        s->last_dense_flops = D.flops;
    } else {
        Plan P;
        const bool rot = rotation_for(s);
        if (plan_is_cached(s, s->tune, rot)) {
            P = s->cache.plan;
        } else {
            plan_circuit(s->pending, s->n, s->prec, s->init_kind, s->init_index, s->init_value, s->tune, P, -1, rot);
            cache_plan(s, s->tune, rot, P);
        }
        run_plan(s, P);
        if (!P.moved.empty()) {  // the plan moved qubits: logical q now sits at moved[l2p[q]]
            for (int q = 0; q < s->n; q++) s->l2p[q] = P.moved[s->l2p[q]];
            for (int q = 0; q < s->n; q++) s->p2l[s->l2p[q]] = q;
        }
        s->last_stats = P.st;
        s->last_launches = (int)P.blocks.size();
        s->last_bytes = 0;
        for (const auto &b : P.blocks) s->last_bytes += (b.init_kind == DQ_INIT_LOAD ? 2 : 1) * state_bytes;
    }
    s->pending.clear();
    s->init_kind = DQ_INIT_LOAD;
    s->probs_dirty = true;
    s->last_flush_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void ensure_probs(QuregState *s) {
    flush(s);
    if (!s->probs_dirty) return;
    s->det_pending = 0;
    s->det_slack = 0.0;
    if (s->paged) {
        paged_probs(s);
        s->det_base = true;
        return;
    }
    ensure_ws(s, dq_prob_ws_bytes(s->n));
    s->p0.assign(s->n + 1, 0.0);
    ENG(dq_prob_zero_all(s->d_state, s->n, s->prec, s->p0.data(), s->d_ws, s->ws_cap, s->stream));
    s->probs_dirty = false;
    s->det_base = true;
}

// Move every logical qubit back to its own physical position with
// permutation sweeps (each moves at most one tile's worth of qubits).
void materialize_mapping(QuregState *s) {
    flush(s);
    const int n = s->n;
    bool ident = true;
    for (int q = 0; q < n; q++) ident &= s->l2p[q] == q;
    if (ident) return;
    std::vector<Plan> passes;
    std::vector<Gate> gates;
    std::string err;
    if (!plan_restore_mapping(n, s->prec, s->l2p, s->tune, passes, gates, err)) {
        fprintf(stderr, "dgquest: qubit re-placement failed: %s\n", err.c_str());
        exit(EXIT_FAILURE);
    }
    for (const Plan &P : passes) run_plan(s, P);
    for (int q = 0; q < n; q++) s->p2l[s->l2p[q]] = q;
    s->probs_dirty = true;
    s->det_base = false;
    if (!gates.empty()) {
        s->pending.insert(s->pending.end(), gates.begin(), gates.end());
        flush(s);
    }
}

}  // namespace dgq

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

void seedQuEST(unsigned long int *seedArray, int numSeeds) { mt_init_by_array(seedArray, numSeeds); }

void seedQuESTDefault(void) {
    // key = {milliseconds since the epoch, pid}, src/QuEST_common.c:127-142
    struct timeval tv;
    gettimeofday(&tv, NULL);
    unsigned long key[2];
    key[0] = (unsigned long)((double)tv.tv_sec * 1000 + (double)(tv.tv_usec / 1000));
    key[1] = (unsigned long)getpid();
    mt_init_by_array(key, 2);
}

QuESTEnv createQuESTEnv(void) {
    seedQuESTDefault();
    QuESTEnvValue *ev = new QuESTEnvValue();
    int dev = 0;
    if (const char *d = getenv("DGQ_DEVICE")) dev = atoi(d);
    ev->device = dev;
    ENG(dq_device_init(dev));
    // QuEST.conf ("minK maxK [K]") selected the reference's task partition; with
    // the state resident on one GPU there is none to choose, so it is accepted
    // and ignored when present, and not required.
    QuESTEnv env;
    env.value = ev;
    env.rank = 0;
    return env;
}

void destroyQuESTEnv(QuESTEnv env) { delete env.value; }

int getQuEST_PREC(void) { return (int)(sizeof(qreal) / 4); }

Qureg createQureg(int numQubits, QuESTEnv env) {
    quest_assert(numQubits > 0, E_INVALID_NUM_QUBITS, __func__);
    Qureg q;
    memset(&q, 0, sizeof q);
    q.isDensityMatrix = 0;
    q.numQubitsRepresented = (u64)numQubits;
    q.numQubitsInStateVec = (u64)numQubits;
    q.numAmpsTotal = 1ull << numQubits;
    q.env = env;
    QuregState *s = new QuregState();
    s->n = numQubits;
    s->prec = sizeof(qreal) == 8 ? DQ_F64 : DQ_F32;
    s->tune = tune_for(s->prec, numQubits);
    s->l2p.resize(numQubits);
    s->p2l.resize(numQubits);
    for (int i = 0; i < numQubits; i++) s->l2p[i] = s->p2l[i] = i;
    const size_t bytes = (size_t)(sizeof(qreal) * 2) << numQubits;
    const size_t budget = device_budget();
    if (bytes > budget) {
        // larger than the device may hold: pinned host pieces paged through it
        s->paged = paged_create(numQubits, s->prec, budget);
    } else {
        int rc = dq_malloc(&s->d_state, bytes);
        if (rc != DQ_OK) engine_fatal("createQureg: state allocation", rc);
    }
    ENG(dq_stream_create(&s->stream));
    q.state = s;
    initZeroState(q);
    return q;
}

// With DGQ_REPORT_FILE set, a register appends one JSON line about its final
// state to that file when it is destroyed: size, precision, total
// probability, P(qubit q = 1) for every logical qubit (1 - P(0), as
// calcProbOfOutcome), and the first (up to 16) amplitudes in logical index
// order, all at full precision.  It lets an unmodified QuEST program (which
// prints a few values to 6 decimals, or nothing) be checked afterwards.
static void report_register(QuregState *s, const char *path) {
    ensure_probs(s);
    const int n = s->n;
    const uint64_t na = std::min<uint64_t>(16, 1ull << n);
    std::vector<double> amps(2 * na);
    if (s->paged) {
        paged_read(s, 0, na, amps.data());
    } else {
        ensure_ws(s, 4096);
        ENG(dq_read_amps(s->d_state, n, s->prec, s->l2p.data(), 0, na, amps.data(), s->d_ws, s->ws_cap, s->stream));
    }
    FILE *f = fopen(path, "a");
    if (!f) {
        fprintf(stderr, "dgquest: cannot open DGQ_REPORT_FILE %s\n", path);
        return;
    }
    char comm[64] = "";
    if (FILE *c = fopen("/proc/self/comm", "r")) {
        if (fgets(comm, sizeof comm, c)) comm[strcspn(comm, "\n\"\\")] = 0;
        fclose(c);
    }
    fprintf(f, "{\"program\": \"%s\", \"n\": %d, \"prec\": %d, \"paged\": %s, \"total_prob\": %.17g, \"p1\": [", comm, n,
            s->prec, s->paged ? "true" : "false", s->p0[n]);
    for (int q = 0; q < n; q++) fprintf(f, "%s%.17g", q ? ", " : "", 1.0 - s->p0[s->l2p[q]]);
    fprintf(f, "], \"amps\": [");
    for (uint64_t i = 0; i < na; i++) fprintf(f, "%s[%.17g, %.17g]", i ? ", " : "", amps[2 * i], amps[2 * i + 1]);
    fprintf(f, "]}\n");
    fclose(f);
}

void destroyQureg(Qureg qureg, QuESTEnv env) {
    (void)env;
    QuregState *s = st(qureg);
    if (!s) return;
    if (s->recorder) {
        delete s;
        return;
    }
    if (const char *path = getenv("DGQ_REPORT_FILE"))
        if (*path) report_register(s, path);
    dq_stream_sync(s->stream);
    if (s->paged) paged_destroy(s->paged);
    else dq_free(s->d_state);
    if (s->d_ws) dq_free(s->d_ws);
    dq_stream_destroy(s->stream);
    delete s;
}

void searchQuregTaskPartitions(Qureg qureg, QuESTEnv env) {
    // The reference predicts, for the recorded circuit, the time of every task
    // partition K with its sampled model and keeps the fastest
    // (src/cpp/QuEST_backend_distributed.cpp:86-121,
    // src/cpp/time_predictor.cpp:343-395).  With the state resident on one GPU
    // the choice is the fused-sweep shape instead: the recorded gates are
    // planned with every candidate shape (k, R, W), lanes rotating and fixed,
    // each plan's device time is predicted from the measured per-shape cost
    // model (tune.hpp), and the least is kept.  The register keeps that shape
    // for later flushes; the next flush runs the winning plan without planning
    // again.  DGQ_TUNE (an explicit shape) skips the search; so does a
    // register with lane rotation switched off, for the rotation choice.  A
    // recording register (dgq_create_recorder) is searched the same way, on
    // the host only.
    (void)env;
    QuregState *s = st(qureg);
    if (!s || s->paged || s->dense_k || s->pending.empty() || getenv("DGQ_TUNE")) return;
    // lane rotation: the caller's setting if made, else both when allowed
    int rots[2] = {1, 0}, nrot = 2;
    if (s->rotate >= 0 || !rotate_lanes()) {
        rots[0] = rotation_allowed(s) ? 1 : 0;
        nrot = 1;
    }
    double best = 0.0;
    bool have = false;
    Plan pick;
    for (TuneParams c : tune_candidates(s->prec)) {
        c.max_ops = s->tune.max_ops;
        c.fuse_diagonals = s->tune.fuse_diagonals;
        const TuneParams f = fit_shape(c, s->n, s->prec);
        if (f.k != c.k || f.R != c.R || f.W != c.W) continue;  // the tile does not fit this register
        const CostModel cm = tune_cost(s->prec, f);
        for (int i = 0; i < nrot; i++) {
            Plan P;
            plan_circuit(s->pending, s->n, s->prec, s->init_kind, s->init_index, s->init_value, f, P, -1, rots[i] != 0);
            const double ms = predict_plan_ms(P, s->prec, cm);
            if (!have || ms < best) {
                have = true;
                best = ms;
                s->tune = f;
                s->rotate_pick = rots[i];
                pick = std::move(P);
            }
        }
    }
    if (have) {
        cache_plan(s, s->tune, rotation_for(s), pick);
        s->predicted_ms = best;
    }
}

int getNumQubits(Qureg qureg) { return (int)qureg.numQubitsRepresented; }

int getNumAmps(Qureg qureg) {
    v_statevec(qureg, __func__);
    return (int)qureg.numAmpsTotal;
}

// ---- initialisation: discards the recorded, unexecuted gates (the state is
// defined anew), and is fused into the next sweep ----------------------------
static void set_init(Qureg q, int kind, uint64_t idx, double val) {
    QuregState *s = st(q);
    s->pending.clear();
    for (int i = 0; i < s->n; i++) s->l2p[i] = s->p2l[i] = i;
    s->init_kind = kind;
    s->init_index = idx;
    s->init_value = val;
    s->probs_dirty = true;
    s->det_base = false;
}

void initZeroState(Qureg qureg) { set_init(qureg, DQ_INIT_CLASSICAL, 0, 0.0); }

void initPlusState(Qureg qureg) {
    set_init(qureg, DQ_INIT_PLUS, 0, 1.0 / std::sqrt((double)(1ULL << qureg.numQubitsRepresented)));
}

void initClassicalState(Qureg qureg, long long int stateInd) {
    v_state_index(qureg, stateInd, __func__);
    set_init(qureg, DQ_INIT_CLASSICAL, (uint64_t)stateInd, 0.0);
}

void initStateFromAmps(Qureg qureg, qreal *reals, qreal *imags) {
    v_statevec(qureg, __func__);
    QuregState *s = st(qureg);
    if (s->recorder) engine_fatal("a recording register cannot execute or read its state", DQ_ERR_UNSUPPORTED);
    set_init(qureg, DQ_INIT_LOAD, 0, 0.0);
    const uint64_t N = qureg.numAmpsTotal;
    const uint64_t chunk = 1ull << 22;
    std::vector<double> re, im;
    for (uint64_t off = 0; off < N; off += chunk) {
        const uint64_t c = (N - off < chunk) ? N - off : chunk;
        re.assign(reals + off, reals + off + c);
        im.assign(imags + off, imags + off + c);
        if (s->paged) {
            paged_write(s, off, c, re.data(), im.data());
            continue;
        }
        ensure_ws(s, 2 * ((c * sizeof(double) + 255) / 256 * 256));
        ENG(dq_write_amps(s->d_state, s->n, s->prec, off, c, re.data(), im.data(), s->d_ws, s->ws_cap, s->stream));
    }
}

void setAmps(Qureg qureg, long long int startInd, qreal *reals, qreal *imags, long long int numAmps) {
    v_statevec(qureg, __func__);
    v_num_amps(qureg, startInd, numAmps, __func__);
    QuregState *s = st(qureg);
    if (s->paged) flush(s);          // host pieces are written through the qubit map
    else materialize_mapping(s);     // flushes; afterwards logical index == physical index
    const uint64_t chunk = 1ull << 22;
    std::vector<double> re, im;
    for (uint64_t off = 0; off < (uint64_t)numAmps; off += chunk) {
        const uint64_t c = ((uint64_t)numAmps - off < chunk) ? (uint64_t)numAmps - off : chunk;
        re.assign(reals + off, reals + off + c);
        im.assign(imags + off, imags + off + c);
        if (s->paged) {
            paged_write(s, (uint64_t)startInd + off, c, re.data(), im.data());
            continue;
        }
        ensure_ws(s, 2 * ((c * sizeof(double) + 255) / 256 * 256));
        ENG(dq_write_amps(s->d_state, s->n, s->prec, (uint64_t)startInd + off, c, re.data(), im.data(), s->d_ws,
                          s->ws_cap, s->stream));
    }
    s->probs_dirty = true;
    s->det_base = false;
}

// ---- gates ---------------------------------------------------------------
void hadamard(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_HADAMARD, bit(t), 0); }
void pauliX(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_PAULI_X, bit(t), 0); }
void pauliY(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_PAULI_Y, bit(t), 0); }
void pauliZ(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_PAULI_Z, 0, bit(t)); }
void sGate(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_S_GATE, 0, bit(t)); }
void tGate(Qureg q, const int t) { v_target(q, t, __func__); record(q, OP_T_GATE, 0, bit(t)); }

void phaseShift(Qureg q, const int t, qreal angle) {
    v_target(q, t, __func__);
    double op[8] = {(double)angle};
    record(q, OP_PHASE_SHIFT, 0, bit(t), op);
}
void controlledPhaseShift(Qureg q, const int a, const int b, qreal angle) {
    v_control_target(q, a, b, __func__);
    double op[8] = {(double)angle};
    record(q, OP_PHASE_SHIFT, 0, bit(a) | bit(b), op);
}
void multiControlledPhaseShift(Qureg q, int *cs, int nc, qreal angle) {
    v_multi_controls(q, cs, nc, __func__);
    uint64_t m = 0;
    for (int i = 0; i < nc; i++) m |= bit(cs[i]);
    double op[8] = {(double)angle};
    record(q, OP_PHASE_SHIFT, 0, m, op);
}
void controlledPhaseFlip(Qureg q, const int a, const int b) {
    v_control_target(q, a, b, __func__);
    record(q, OP_PAULI_Z, 0, bit(a) | bit(b));
}
void multiControlledPhaseFlip(Qureg q, int *cs, int nc) {
    v_multi_controls(q, cs, nc, __func__);
    uint64_t m = 0;
    for (int i = 0; i < nc; i++) m |= bit(cs[i]);
    record(q, OP_PAULI_Z, 0, m);
}

static void rot(Qureg q, int type, int t, uint64_t cm, qreal angle) {
    double op[8] = {(double)angle * 0.5};
    record(q, type, bit(t), cm, op);
}
void rotateX(Qureg q, const int t, qreal a) { v_target(q, t, __func__); rot(q, OP_ROTATE_X, t, 0, a); }
void rotateY(Qureg q, const int t, qreal a) { v_target(q, t, __func__); rot(q, OP_ROTATE_Y, t, 0, a); }
void rotateZ(Qureg q, const int t, qreal a) { v_target(q, t, __func__); rot(q, OP_ROTATE_Z, t, 0, a); }
void controlledRotateX(Qureg q, const int c, const int t, qreal a) {
    v_control_target(q, c, t, __func__);
    rot(q, OP_ROTATE_X, t, bit(c), a);
}
void controlledRotateY(Qureg q, const int c, const int t, qreal a) {
    v_control_target(q, c, t, __func__);
    rot(q, OP_ROTATE_Y, t, bit(c), a);
}
void controlledRotateZ(Qureg q, const int c, const int t, qreal a) {
    v_control_target(q, c, t, __func__);
    rot(q, OP_ROTATE_Z, t, bit(c), a);
}

void controlledNot(Qureg q, const int c, const int t) {
    v_control_target(q, c, t, __func__);
    record(q, OP_PAULI_X, bit(t), bit(c));
}
void controlledPauliY(Qureg q, const int c, const int t) {
    v_control_target(q, c, t, __func__);
    record(q, OP_PAULI_Y, bit(t), bit(c));
}

void compactUnitary(Qureg q, const int t, Complex alpha, Complex beta) {
    v_target(q, t, __func__);
    v_complex_pair(alpha, beta, __func__);
    double op[8] = {alpha.real, alpha.imag, beta.real, beta.imag};
    record(q, OP_COMPACT_UNITARY, bit(t), 0, op);
}
void controlledCompactUnitary(Qureg q, const int c, const int t, Complex alpha, Complex beta) {
    v_control_target(q, c, t, __func__);
    v_complex_pair(alpha, beta, __func__);
    double op[8] = {alpha.real, alpha.imag, beta.real, beta.imag};
    record(q, OP_COMPACT_UNITARY, bit(t), bit(c), op);
}
void unitary(Qureg q, const int t, ComplexMatrix2 u) {
    v_target(q, t, __func__);
    double m[8];
    mat_from(u, m);
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void controlledUnitary(Qureg q, const int c, const int t, ComplexMatrix2 u) {
    v_control_target(q, c, t, __func__);
    double m[8];
    mat_from(u, m);
    v_unitary(m, __func__);
    rec_unitary(q, t, bit(c), m);
}
void multiControlledUnitary(Qureg q, int *cs, const int nc, const int t, ComplexMatrix2 u) {
    v_multi_controls_target(q, cs, nc, t, __func__);
    double m[8];
    mat_from(u, m);
    v_unitary(m, __func__);
    uint64_t cm = 0;
    for (int i = 0; i < nc; i++) cm |= bit(cs[i]);
    rec_unitary(q, t, cm, m);
}

static void axis_pair(qreal angle, Vector axis, double op[8]) {
    // src/QuEST_common.c:56-63
    const double mag = std::sqrt((double)axis.x * axis.x + (double)axis.y * axis.y + (double)axis.z * axis.z);
    const double ux = axis.x / mag, uy = axis.y / mag, uz = axis.z / mag;
    op[0] = std::cos(angle / 2.0);
    op[1] = -std::sin(angle / 2.0) * uz;
    op[2] = std::sin(angle / 2.0) * uy;
    op[3] = -std::sin(angle / 2.0) * ux;
}
void rotateAroundAxis(Qureg q, const int t, qreal angle, Vector axis) {
    v_target(q, t, __func__);
    v_vector(axis, __func__);
    double op[8] = {0};
    axis_pair(angle, axis, op);
    record(q, OP_COMPACT_UNITARY, bit(t), 0, op);
}
void controlledRotateAroundAxis(Qureg q, const int c, const int t, qreal angle, Vector axis) {
    v_control_target(q, c, t, __func__);
    v_vector(axis, __func__);
    double op[8] = {0};
    axis_pair(angle, axis, op);
    record(q, OP_COMPACT_UNITARY, bit(t), bit(c), op);
}

void u1Gate(Qureg q, int t, qreal lambda) {
    v_target(q, t, __func__);
    double m[8] = {1, 0, 0, 0, 0, 0, std::cos((double)lambda), std::sin((double)lambda)};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void u2Gate(Qureg q, int t, qreal phi, qreal lambda) {
    v_target(q, t, __func__);
    const double is2 = 1 / std::sqrt(2.0), p = phi, l = lambda;
    double m[8] = {is2, 0, -std::cos(l) * is2, -std::sin(l) * is2,
                   std::cos(p) * is2, std::sin(p) * is2, std::cos(p + l) * is2, std::sin(p + l) * is2};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void u3Gate(Qureg q, int t, qreal theta, qreal phi, qreal lambda) {
    v_target(q, t, __func__);
    const double ct = std::cos((double)theta / 2), s_t = std::sin((double)theta / 2), p = phi, l = lambda;
    double m[8] = {ct, 0, -std::cos(l) * s_t, -std::sin(l) * s_t,
                   std::cos(p) * s_t, std::sin(p) * s_t, std::cos(p + l) * ct, std::sin(p + l) * ct};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void SqX(Qureg q, int t) {
    v_target(q, t, __func__);
    const double is2 = 1 / std::sqrt(2.0);
    double m[8] = {is2, 0, 0, -is2, 0, -is2, is2, 0};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void SqY(Qureg q, int t) {
    v_target(q, t, __func__);
    const double is2 = 1 / std::sqrt(2.0);
    double m[8] = {is2, 0, -is2, 0, is2, 0, is2, 0};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void SqW(Qureg q, int t) {
    v_target(q, t, __func__);
    const double is2 = 1 / std::sqrt(2.0);
    double m[8] = {is2, 0, -0.5, -0.5, 0.5, -0.5, is2, 0};
    v_unitary(m, __func__);
    rec_unitary(q, t, 0, m);
}
void fSim(Qureg q, int t1, int t2, qreal theta, qreal phi) {
    controlledPhaseShift(q, t1, t2, -phi);
    controlledNot(q, t1, t2);
    controlledRotateX(q, t2, t1, theta * 2);
    controlledNot(q, t1, t2);
}

// SWAP relabels the logical -> physical map: no amplitude moves.
void swapGate(Qureg q, int a, int b) {
    v_unique_targets(q, a, b, __func__);
    QuregState *s = st(q);
    const int pa = s->l2p[a], pb = s->l2p[b];
    s->l2p[a] = pb;
    s->l2p[b] = pa;
    s->p2l[pa] = b;
    s->p2l[pb] = a;
}

// ---- overlaps of two registers ---------------------------------------------
// <a|b>, <a|a>, <b|b>, max |a_i - b_i| of two resident registers, both moved
// to logical qubit order first, on a's stream.
static void compare_regs(QuregState *a, QuregState *b, double out[5]) {
    if (a->paged || b->paged) engine_fatal("overlaps of paged registers", DQ_ERR_UNSUPPORTED);
    materialize_mapping(a);
    materialize_mapping(b);
    ENG(dq_stream_sync(b->stream));
    const size_t cw = (dq_compare_ws_bytes(a->n) + 255) / 256 * 256;
    ensure_ws(a, cw + 256);
    double *d_out = (double *)((unsigned char *)a->d_ws + cw);
    ENG(dq_compare_states(a->d_state, a->prec, b->d_state, b->prec, a->n, d_out, a->d_ws, cw, a->stream));
    ENG(dq_memcpy_async(out, d_out, 5 * sizeof(double), DQ_D2H, a->stream));
    ENG(dq_stream_sync(a->stream));
}

Complex calcInnerProduct(Qureg bra, Qureg ket) {
    v_statevec(bra, __func__);
    v_statevec(ket, __func__);
    quest_assert(bra.numQubitsRepresented == ket.numQubitsRepresented, E_MISMATCHING_QUREG_DIMENSIONS, __func__);
    double o[5];
    compare_regs(st(bra), st(ket), o);
    Complex c;
    c.real = (qreal)o[0];
    c.imag = (qreal)o[1];
    return c;
}

qreal calcFidelity(Qureg qureg, Qureg pureState) {
    quest_assert(qureg.numQubitsRepresented == pureState.numQubitsRepresented, E_MISMATCHING_QUREG_DIMENSIONS,
                 __func__);
    quest_assert(!pureState.isDensityMatrix, E_SECOND_ARG_MUST_BE_STATEVEC, __func__);
    v_statevec(qureg, __func__);
    double o[5];
    compare_regs(st(pureState), st(qureg), o);
    return (qreal)(o[0] * o[0] + o[1] * o[1]);
}

// ---- results ---------------------------------------------------------------
Complex getAmp(Qureg qureg, long long int index) {
    v_statevec(qureg, __func__);
    v_state_index(qureg, index, __func__);
    QuregState *s = st(qureg);
    flush(s);
    ensure_ws(s, 4096);
    double out[2];
    if (s->paged) paged_read(s, (uint64_t)index, 1, out);
    else ENG(dq_read_amps(s->d_state, s->n, s->prec, s->l2p.data(), (uint64_t)index, 1, out, s->d_ws, s->ws_cap, s->stream));
    Complex c;
    c.real = (qreal)out[0];
    c.imag = (qreal)out[1];
    return c;
}
qreal getRealAmp(Qureg q, long long int i) { return getAmp(q, i).real; }
qreal getImagAmp(Qureg q, long long int i) { return getAmp(q, i).imag; }
qreal getProbAmp(Qureg q, long long int i) {
    Complex c = getAmp(q, i);
    return c.real * c.real + c.imag * c.imag;
}

qreal calcProbOfOutcome(Qureg qureg, const int q, int outcome) {
    v_target(qureg, q, __func__);
    v_outcome(outcome, __func__);
    QuregState *s = st(qureg);
    ensure_probs(s);
    double p = s->p0[s->l2p[q]];
    if (outcome == 1) p = 1.0 - p;
    return (qreal)p;
}

qreal calcTotalProb(Qureg qureg) {
    QuregState *s = st(qureg);
    ensure_probs(s);
    return (qreal)s->p0[s->n];
}

static void record_collapse(Qureg qureg, int q, int outcome, double prob) {
    double op[8] = {(double)outcome, 1.0 / std::sqrt(prob)};
    record(qureg, OP_COLLAPSE, bit(q), 0, op);
}

qreal collapseToOutcome(Qureg qureg, const int q, int outcome) {
    v_target(qureg, q, __func__);
    v_outcome(outcome, __func__);
    const qreal prob = calcProbOfOutcome(qureg, q, outcome);
    quest_assert(prob > REAL_EPS, E_COLLAPSE_STATE_ZERO_PROB, __func__);
    record_collapse(qureg, q, outcome, prob);
    return prob;
}

// P(q = 0) from the register's last P(0) pass when every gate recorded since
// is the collapse of a deterministic measurement: each such collapse drops a
// branch of probability at most its slack and rescales the rest, which moves
// any P(q = 0) by at most 2 * slack in all.  Only a value that decides the
// outcome rule the same way anywhere in that range is taken.  A run of
// deterministic measurements then costs one P(0) pass and one flush of the
// collapses, instead of a flush and a pass over the whole state each (over
// the host link for a paged register).
static bool cached_zero_prob(QuregState *s, int q, double *zero_prob) {
    if (!s->det_base || s->init_kind != DQ_INIT_LOAD) return false;
    if (s->pending.size() != (size_t)s->det_pending) return false;
    const double p = s->p0[s->l2p[q]], m = 2.0 * s->det_slack;
    if (!(p + m < REAL_EPS || 1.0 - (p - m) < REAL_EPS)) return false;
    *zero_prob = p;
    return true;
}

int measureWithStats(Qureg qureg, int q, qreal *outcomeProb) {
    v_target(qureg, q, __func__);
    QuregState *s = st(qureg);
    double cached = 0.0;
    const qreal zeroProb = cached_zero_prob(s, q, &cached) ? (qreal)cached : calcProbOfOutcome(qureg, q, 0);
    // src/QuEST_common.c:97-115
    int outcome;
    const bool determined = zeroProb < REAL_EPS || 1 - zeroProb < REAL_EPS;
    if (zeroProb < REAL_EPS) outcome = 1;
    else if (1 - zeroProb < REAL_EPS) outcome = 0;
    else outcome = (mt_genrand_real1() > zeroProb);
    const qreal prob = (outcome == 0) ? zeroProb : 1 - zeroProb;
    if (outcomeProb) *outcomeProb = prob;
    const bool chain = s->det_base && s->init_kind == DQ_INIT_LOAD && s->pending.size() == (size_t)s->det_pending;
    record_collapse(qureg, q, outcome, prob);
    if (chain && determined) {
        s->det_pending++;
        s->det_slack += std::max(0.0, 1.0 - (double)prob);
    }
    return outcome;
}

int measure(Qureg qureg, int q) {
    v_target(qureg, q, __func__);
    qreal p;
    return measureWithStats(qureg, q, &p);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// engine controls (include/dgquest_ext.h)
// ---------------------------------------------------------------------------
#include "dgquest_ext.h"

extern "C" {

int dgq_set_tuning(Qureg qureg, int k, int R, int W, int max_ops) {
    QuregState *s = st(qureg);
    if (!s || k < 1 || R < 1 || W < 0 || max_ops < 1) return -1;
    flush(s);
    s->tune.k = k;
    s->tune.R = R;
    s->tune.W = W;
    s->tune.max_ops = max_ops;
    return 0;
}

int dgq_set_lane_rotation(Qureg qureg, int on) {
    QuregState *s = st(qureg);
    if (!s) return -1;
    if (!s->recorder) flush(s);  // a recording register only plans: nothing to flush
    s->rotate = on != 0;
    s->rotate_pick = -1;
    return 0;
}

int dgq_last_search(Qureg qureg, int shape[4], double *predicted_ms) {
    QuregState *s = st(qureg);
    if (!s || !shape || !predicted_ms) return -1;
    shape[0] = s->tune.k;
    shape[1] = s->tune.R;
    shape[2] = s->tune.W;
    shape[3] = rotation_for(s) ? 1 : 0;
    *predicted_ms = s->predicted_ms;
    return s->predicted_ms > 0.0 ? 0 : 1;
}

int dgq_set_fusion(Qureg qureg, int on) {
    QuregState *s = st(qureg);
    if (!s) return -1;
    flush(s);
    s->tune.fuse_diagonals = on != 0;
    return 0;
}

int dgq_last_plan_stats(Qureg qureg, int stats[8], double *flush_ms) {
    QuregState *s = st(qureg);
    if (!s || !stats) return -1;
    const TuneParams t = fit_shape(s->tune, s->n, s->prec);
    stats[0] = s->last_stats.sweeps;
    stats[1] = s->last_stats.phases;
    stats[2] = s->last_stats.exchanges;
    stats[3] = s->last_stats.ops;
    stats[4] = s->last_stats.gates_in;
    stats[5] = t.k;
    stats[6] = t.R;
    stats[7] = t.W;
    if (flush_ms) *flush_ms = s->last_flush_ms;
    return 0;
}

void *dgq_stream(Qureg qureg) {
    QuregState *s = st(qureg);
    return s ? s->stream : nullptr;
}

int dgq_last_flush_bytes(Qureg qureg, unsigned long long *bytes, int *launches) {
    QuregState *s = st(qureg);
    if (!s) return -1;
    if (bytes) *bytes = s->last_bytes;
    if (launches) *launches = s->last_launches;
    return 0;
}

int dgq_flush(Qureg qureg) {
    QuregState *s = st(qureg);
    if (!s) return -1;
    flush(s);
    return 0;
}

int dgq_sync(Qureg qureg) {
    QuregState *s = st(qureg);
    if (!s) return -1;
    ENG(dq_stream_sync(s->stream));
    return 0;
}

int dgq_read_state(Qureg qureg, unsigned long long start, unsigned long long count, double *out) {
    QuregState *s = st(qureg);
    if (!s || !out) return -1;
    const uint64_t N = 1ull << s->n;
    if (start > N || count > N - start) return -1;
    flush(s);
    if (s->paged) {
        paged_read(s, start, count, out);
        return 0;
    }
    const uint64_t chunk = 1ull << 22;
    for (uint64_t off = 0; off < count; off += chunk) {
        const uint64_t c = (count - off < chunk) ? count - off : chunk;
        ensure_ws(s, 256 + 64 * sizeof(int) + c * 2 * sizeof(double));
        ENG(dq_read_amps(s->d_state, s->n, s->prec, s->l2p.data(), start + off, c, out + 2 * off, s->d_ws,
                         s->ws_cap, s->stream));
    }
    return 0;
}

int dgq_plan_only(int n, int prec, const dgq_gate_rec *gates, int ngates, int k, int R, int W, int max_ops,
                  int stats[8]) {
    if (n < 1 || n > 40 || ngates < 0 || (ngates && !gates) || !stats || (prec != DQ_F32 && prec != DQ_F64)) return -1;
    std::vector<Gate> gl(ngates);
    const uint64_t full = (1ull << n) - 1;
    for (int i = 0; i < ngates; i++) {
        if ((gates[i].tmask & ~full) || (gates[i].cmask & ~full) ||
            gate_record_error(gates[i].type, gates[i].tmask, gates[i].cmask, gates[i].op))
            return -1;
        gl[i].type = gates[i].type;
        gl[i].tmask = gates[i].tmask;
        gl[i].cmask = gates[i].cmask;
        for (int j = 0; j < 8; j++) gl[i].op[j] = gates[i].op[j];
    }
    TuneParams tp;
    tp.k = k;
    tp.R = R;
    tp.W = W;
    tp.max_ops = max_ops;
    Plan P;
    plan_circuit(gl, n, prec, DQ_INIT_LOAD, 0, 0.0, tp, P, -1, rotate_lanes());
    const TuneParams t = fit_shape(tp, n, prec);
    stats[0] = P.st.sweeps;
    stats[1] = P.st.phases;
    stats[2] = P.st.exchanges;
    stats[3] = P.st.ops;
    stats[4] = P.st.gates_in;
    stats[5] = t.k;
    stats[6] = t.R;
    stats[7] = t.W;
    return 0;
}

Qureg dgq_create_recorder(int numQubits) {
    quest_assert(numQubits > 0 && numQubits <= 40, E_INVALID_NUM_QUBITS, __func__);
    Qureg q;
    memset(&q, 0, sizeof q);
    q.numQubitsRepresented = (u64)numQubits;
    q.numQubitsInStateVec = (u64)numQubits;
    q.numAmpsTotal = 1ull << numQubits;
    QuregState *s = new QuregState();
    s->n = numQubits;
    s->prec = sizeof(qreal) == 8 ? DQ_F64 : DQ_F32;
    s->tune = tune_for(s->prec, numQubits);
    s->recorder = true;
    s->l2p.resize(numQubits);
    s->p2l.resize(numQubits);
    for (int i = 0; i < numQubits; i++) s->l2p[i] = s->p2l[i] = i;
    s->init_kind = DQ_INIT_CLASSICAL;  // a new register is |0...0>
    q.state = s;
    return q;
}

void dgq_destroy_recorder(Qureg q) { destroyQureg(q, q.env); }

int dgq_recorder_plan(Qureg q, int k, int R, int W, int max_ops, int fuse, dq_block *blocks, int cap_blocks,
                      dq_phase *phases, int cap_phases, dq_op *ops, int cap_ops, int counts[3], int *l2p) {
    QuregState *s = q.state;
    if (!s || !s->recorder || !counts) return -1;
    TuneParams tp;
    tp.k = k;
    tp.R = R;
    tp.W = W;
    tp.max_ops = max_ops;
    tp.fuse_diagonals = fuse != 0;
    Plan P;
    plan_circuit(s->pending, s->n, s->prec, s->init_kind, s->init_index, s->init_value, tp, P, -1, rotation_for(s));
    counts[0] = (int)P.blocks.size();
    counts[1] = (int)P.phases.size();
    counts[2] = (int)P.ops.size();
    if (l2p)  // the map after the plan, as a flush would leave it
        for (int i = 0; i < s->n; i++) l2p[i] = P.moved.empty() ? s->l2p[i] : P.moved[s->l2p[i]];
    if (counts[0] > cap_blocks || counts[1] > cap_phases || counts[2] > cap_ops) return 1;
    if (counts[0]) memcpy(blocks, P.blocks.data(), P.blocks.size() * sizeof(dq_block));
    if (counts[1]) memcpy(phases, P.phases.data(), P.phases.size() * sizeof(dq_phase));
    if (counts[2]) memcpy(ops, P.ops.data(), P.ops.size() * sizeof(dq_op));
    return 0;
}

int dgq_recorder_export(Qureg q, dgq_gate_rec *gates, int cap, int *n_gates, int *l2p, int *init_kind,
                        unsigned long long *init_index, double *init_value) {
    QuregState *s = q.state;
    if (!s || !s->recorder || !n_gates) return -1;
    *n_gates = (int)s->pending.size();
    if (l2p)
        for (int i = 0; i < s->n; i++) l2p[i] = s->l2p[i];
    if (init_kind) *init_kind = s->init_kind;
    if (init_index) *init_index = s->init_index;
    if (init_value) *init_value = s->init_value;
    if ((int)s->pending.size() > cap) return 1;
    for (size_t i = 0; i < s->pending.size(); i++) {
        const Gate &g = s->pending[i];
        dgq_gate_rec &r = gates[i];
        memset(&r, 0, sizeof r);
        r.type = g.type;
        r.tmask = g.tmask;
        r.cmask = g.cmask;
        for (int j = 0; j < 8; j++) r.op[j] = g.op[j];
    }
    return 0;
}

int dgq_sample(Qureg q, int shots, const double *u, unsigned long long *out) {
    QuregState *s = q.state;
    if (!s || shots < 0 || (shots && (!u || !out))) return -1;
    if (s->recorder) engine_fatal("a recording register cannot execute or read its state", DQ_ERR_UNSUPPORTED);
    if (s->paged) engine_fatal("dgq_sample: sampling a paged register", DQ_ERR_UNSUPPORTED);
    for (int i = 0; i < shots; i++)
        if (!(u[i] >= 0.0 && u[i] < 1.0)) return -1;
    if (!shots) return 0;
    materialize_mapping(s);  // samples walk the CDF in logical index order
    const size_t res = (((size_t)shots * sizeof(uint64_t)) + 255) / 256 * 256;
    ensure_ws(s, res + dq_readout_ws_bytes(s->n, 1, (uint32_t)shots));
    uint64_t *d_out = (uint64_t *)s->d_ws;
    const void *st = s->d_state;
    ENG(dq_readout_many(s->prec, s->n, &st, 1, u, (uint32_t)shots, nullptr, d_out, (unsigned char *)s->d_ws + res,
                        s->ws_cap - res, s->stream));
    std::vector<uint64_t> idx((size_t)shots);
    ENG(dq_memcpy_async(idx.data(), d_out, (size_t)shots * sizeof(uint64_t), DQ_D2H, s->stream));
    ENG(dq_stream_sync(s->stream));
    for (int i = 0; i < shots; i++) out[i] = idx[i];
    return 0;
}

int dgq_tune_lookup(int prec, int n, int out[4]) {
    if ((prec != DQ_F32 && prec != DQ_F64) || n < 1 || n > 64 || !out) return -1;
    const TuneParams t = tune_for(prec, n);
    out[0] = t.k;
    out[1] = t.R;
    out[2] = t.W;
    out[3] = t.max_ops;
    return 0;
}

int dgq_tune_reload(const char *path) { return tune_load(path ? path : ""); }

int dgq_set_dense(Qureg q, int k) {
    QuregState *s = q.state;
    if (!s || k < 0) return -1;
    if (k == 0) {
        s->dense_k = 0;
        return 0;
    }
    const int T = s->n < 13 ? s->n : 13;
    if (s->prec != DQ_F32 || s->paged || k < 5 || k > 7 || T < k + 5) return -1;
    flush(s);  // gates recorded so far run in the current form
    s->dense_k = k;
    return 0;
}

int dgq_last_dense_stats(Qureg q, double out[4]) {
    QuregState *s = q.state;
    if (!s || !out) return -1;
    out[0] = s->last_stats.sweeps;
    out[1] = s->last_stats.ops;
    out[2] = s->last_dense_flops;
    out[3] = s->last_flush_ms;
    return 0;
}

int dgq_recorder_plan_dense(Qureg q, int k, int cap_ops, int *n_ops, int *n_blocks, int *op_qubits, double *mats) {
    QuregState *s = q.state;
    if (!s || !s->recorder || !n_ops) return -1;
    DensePlan D;
    std::string err;
    if (!plan_dense(s->pending, s->n, k, D, err)) return -1;
    *n_ops = (int)D.ops.size();
    if (n_blocks) *n_blocks = (int)D.blocks.size();
    if ((int)D.ops.size() > cap_ops) return 1;
    if (op_qubits) memcpy(op_qubits, D.op_qubits.data(), D.op_qubits.size() * sizeof(int));
    if (mats) memcpy(mats, D.mats.data(), D.mats.size() * sizeof(double));
    return 0;
}

void *dgq_state_pointer(Qureg q, int *prec) {
    QuregState *s = q.state;
    if (!s || s->recorder || s->paged) return nullptr;
    materialize_mapping(s);
    ENG(dq_stream_sync(s->stream));
    // the caller may write through the pointer: no cached P(0) survives it
    s->probs_dirty = true;
    s->det_base = false;
    if (prec) *prec = s->prec;
    return s->d_state;
}

int dgq_paged_info(Qureg q, int *piece_qubits, int *pieces_log2, double stats[10]) {
    QuregState *s = q.state;
    if (!s) return -1;
    if (!s->paged) return 1;
    const PagedStore &P = *s->paged;
    if (piece_qubits) *piece_qubits = P.nl;
    if (pieces_log2) *pieces_log2 = P.K;
    if (stats) {
        const PagedStats &t = P.st;
        const double v[10] = {(double)t.sections, (double)t.remaps, (double)t.pieces_run, (double)t.pieces_skipped,
                              t.copy_ms, t.hidden_ms, t.compute_ms, t.wall_ms, (double)t.bytes_in, (double)t.bytes_out};
        for (int i = 0; i < 10; i++) stats[i] = v[i];
    }
    return 0;
}

}  // extern "C"
