/* dgq_serve.h -- a request server for many registers on one GPU.
 *
 * The reference runs one register at a time through the QuEST API; it has no
 * serving loop.  This interface serves circuits on 10-30+ qubits as requests:
 * each returns sum_q <Z_q> (from P(q = 0) of every qubit, the reduction of
 * src/cpp/QuEST_gpu_kernel.cu:769-807) and a number of seeded samples of the
 * final state, computed in FP32 (or FP64) with an FP64 recheck on a sample.
 *
 * Continuous batching: requests are planned on host threads as they arrive
 * and admitted as device memory (a fixed HBM budget) and stream slots free up.
 * Registers up to batch_max_qubits of one size run together, one launch per
 * fused sweep for the whole batch; larger registers run on streams of their
 * own; work of different streams overlaps on the device.  Every result is
 * the same whichever batch or stream carried it.
 *
 * Gates arrive as the recorded gate list of the QuEST front end
 * (dgq_gate_rec, src/cpp/QuEST_backend.h:160-178 encoding), e.g. exported
 * from a recording register (dgq_recorder_export in dgquest_ext.h).
 *
 * All functions return 0 on success and a negative DQ_ERR_* code otherwise;
 * dgs_last_error() gives the message.  A server object may be used from
 * several threads.
 */
#ifndef DGQ_SERVE_H
#define DGQ_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "dgquest_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dgs_server dgs_server;

typedef struct dgs_config {
    int device;
    int prec;                 /* DQ_F32 (default) or DQ_F64: precision of the served state */
    uint64_t hbm_budget;      /* device bytes for states, workspaces and results (one arena) */
    int planner_threads;      /* host threads planning arriving requests (default 4) */
    int batch_lanes;          /* streams running batches of registers <= batch_max_qubits (default 1) */
    int large_lanes;          /* streams running larger registers, one per launch (default 3) */
    int lane_depth;           /* launches in flight per stream (default 2) */
    int batch_max_qubits;     /* registers up to this size are batched (default 22) */
    int batch_max;            /* registers per batch (default 256) */
    double recheck_fraction;  /* share of requests also run in FP64 and compared (default 0.01) */
    int k, R, W, max_ops;     /* fusion shape for every size; k = 0: per size from the tuning
                                 table (dgq_tune_lookup), max_ops = 0: the table's / default */
} dgs_config;

typedef struct dgs_request {
    int n_qubits;
    uint32_t n_gates;
    const dgq_gate_rec *gates; /* physical qubit positions, as recorded */
    const int *l2p;            /* logical -> physical qubit after the gates (NULL: identity) */
    int init_kind;             /* DQ_INIT_CLASSICAL or DQ_INIT_PLUS */
    uint64_t init_index;
    double init_value;
    uint32_t shots;            /* samples of the final state (0..2^20) */
    uint32_t seed;             /* sample s uses the s-th genrand_res53 of MT19937 seeded
                                  by init_by_array({seed}, 1) */
    uint32_t flags;            /* DGS_RECHECK / DGS_NO_RECHECK force the FP64 recheck on / off */
} dgs_request;

#define DGS_RECHECK 1u
#define DGS_NO_RECHECK 2u

/* Probabilities are those of the normalised final state: P(q = 0) is the
 * reduction's sum over amplitudes with bit q = 0 divided by the state's total
 * probability (which rounding in FP32 moves off 1 by up to ~1e-6). */
typedef struct dgs_result {
    uint64_t id;
    int status;                /* 0, or a DQ_ERR_* code */
    int n_qubits;
    double sum_z;              /* sum over qubits of 2 P(q = 0) - 1 */
    double total_prob;         /* sum_i |a_i|^2 of the served state, before normalising */
    double t_arrive_ms;        /* times since the server started */
    double t_start_ms;         /* launched on the device */
    double t_done_ms;          /* results back on the host */
    int rechecked;
    int batch_size;            /* registers in the launch that carried the request */
    double fidelity;           /* when rechecked: |<a|b>|^2 / (<a|a> <b|b>), a = the FP64 run's
                                  state, b = the served state (1 = same state); else 0 */
    double max_abs_err;        /* max_i |a_i(fp64) - a_i| when rechecked */
    double sum_z_fp64;         /* sum_z of the FP64 run when rechecked */
    uint32_t shot_mismatches;  /* samples that differ between the two runs */
    uint32_t pad;
    double fidelity_raw;       /* when rechecked: |<a|b>|^2 without the norms */
    double norm_fp64;          /* when rechecked: <a|a> of the FP64 run */
} dgs_result;

void dgs_config_default(dgs_config *cfg);
const char *dgs_last_error(void);

/* Starts the planner and stream threads and reserves the arena (hbm_budget
 * bytes of device memory).  NULL on failure. */
dgs_server *dgs_create(const dgs_config *cfg);
/* Finishes all submitted work, then releases everything. */
void dgs_destroy(dgs_server *s);

/* Queue a request (everything it points to is copied).  *id identifies it in
 * dgs_wait.  A request larger than the whole budget completes at once with
 * status DQ_ERR_NOMEM; one that does not fit now waits for memory. */
int dgs_submit(dgs_server *s, const dgs_request *req, uint64_t *id);

/* Wait for request `id` and take its result.  p0 (n + 1 doubles, logical
 * qubit order, total last) and samples (shots logical basis-state indices)
 * may be NULL.  The request is forgotten afterwards. */
int dgs_wait(dgs_server *s, uint64_t id, dgs_result *res, double *p0, uint64_t *samples);

/* 1 while request `id` is still running, 0 once dgs_wait would not block. */
int dgs_poll(dgs_server *s, uint64_t id);

/* Submit reqs[i] at arrive_s[i] seconds after the call (non-decreasing; NULL:
 * all at once), wait for all, and fill res[i], samples (the shots of each
 * request, packed in request order) and p0 (n + 1 doubles per request, packed
 * likewise); samples and p0 may be NULL.  The timed loop of a serving
 * benchmark, native end to end. */
int dgs_replay(dgs_server *s, const dgs_request *reqs, uint32_t count, const double *arrive_s, dgs_result *res,
               uint64_t *samples, double *p0);

/* The plain loop the server is measured against: one request at a time on
 * one stream, in arrival order, each planned, swept and read out with a wait
 * for the result before the next starts.  Device memory for the largest
 * request is allocated once before the loop's clock starts and reused (a
 * loop that keeps its register memory, not a createQureg / destroyQureg per
 * request).  Same arguments as dgs_replay; prec is the state precision (the
 * plain method is DQ_F64). */
int dgs_run_plain(const dgs_config *cfg, const dgs_request *reqs, uint32_t count, const double *arrive_s,
                  dgs_result *res, uint64_t *samples, double *p0);

/* Counters since creation: out[0] requests done, [1] batches (launch groups)
 * issued, [2] registers in batches, [3] peak arena bytes in use, [4] arena
 * bytes, [5] times a lane found the next request waiting for memory,
 * [6] requests rechecked, [7] host planning ms summed over requests. */
int dgs_stats(dgs_server *s, double out[8]);

/* The sample uniforms of a request with this seed: out[s] = the s-th
 * genrand_res53 of MT19937 seeded by init_by_array({seed}, 1). */
int dgs_uniforms(uint32_t seed, uint32_t count, double *out);

/* Host-only check of the arena allocator against a brute-force model over
 * `iters` random allocations and releases; 0 when they agree. */
int dgs_arena_selftest(uint32_t seed, int iters);

#ifdef __cplusplus
}
#endif

#endif /* DGQ_SERVE_H */
