/* dgquest_ext.h -- engine controls beside the QuEST API.
 *
 * None of these exist in the reference; they expose what its
 * time_predictor / QuEST.conf pair decided implicitly
 * (src/cpp/time_predictor.cpp:232-395, src/cpp/QuEST_distributed.cpp:42-56):
 * the fused-block shape, the plan statistics of the last flush, explicit
 * flush / sync points for timing, and a planning-only call that needs no GPU.
 * All return 0 on success and a negative value on a bad argument.
 */
#ifndef DGQUEST_EXT_H
#define DGQUEST_EXT_H

#include "QuEST.h"
#include "dgq_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fused-block shape for this register: k = max target qubits above the six
 * coalesced ones per sweep, R/W = register / wave bits of the tile kernel
 * (tile = 6 + R + W qubits), max_ops = gate cap per sweep.  Values are
 * clamped to what the register size and the built kernels allow. */
int dgq_set_tuning(Qureg qureg, int k, int R, int W, int max_ops);

/* Lane rotation on (1, default) or off (0) for this register's flushes:
 * with it, each fused sweep's store hands the six coalesced qubit positions
 * to the qubits the next sweep needs, which takes fewer sweeps per circuit
 * (each doing more gates); the qubit map absorbs the moves.  Off, those six
 * positions keep their qubits, as in the reference's blocks. */
int dgq_set_lane_rotation(Qureg qureg, int on);

/* The shape searchQuregTaskPartitions chose for this register's recorded
 * circuit: shape = k, R, W, lane rotation (1/0); *predicted_ms = the
 * predicted device time of its plan.  Returns 0 after a search, 1 when the
 * register's shape did not come from one (shape then holds its current
 * setting). */
int dgq_last_search(Qureg qureg, int shape[4], double *predicted_ms);

/* Diagonal fusion on (1, default) or off (0): with it, uncontrolled
 * single-qubit gates run as real 2x2 rotations and their diagonal parts are
 * deferred and merged (fewer floating-point operations per gate). */
int dgq_set_fusion(Qureg qureg, int on);

/* stats[0..7] = sweeps, phases, LDS exchanges, device ops, recorded gates,
 * k, R, W of the last flush; *flush_ms = host time of that flush (planning
 * and launch, not waiting for the GPU). */
int dgq_last_plan_stats(Qureg qureg, int stats[8], double *flush_ms);

/* Run the recorded gates now (enqueue; returns without waiting). */
int dgq_flush(Qureg qureg);

/* Wait until all work of this register has finished on the GPU. */
int dgq_sync(Qureg qureg);

/* Flush, then copy amplitudes start..start+count-1 (logical order) into
 * out[2*count] as {re, im} doubles. */
int dgq_read_state(Qureg qureg, unsigned long long start, unsigned long long count, double *out);

/* The HIP stream (hipStream_t) this register's kernels run on, for event
 * timing around its sweeps. */
void *dgq_stream(Qureg qureg);

/* Algorithmic HBM bytes of the last flush: every sweep reads and writes the
 * state once, except a sweep that generates its input (init fused) which
 * only writes; *launches = number of sweep launches. */
int dgq_last_flush_bytes(Qureg qureg, unsigned long long *bytes, int *launches);

/* Planning only, no GPU needed: plan `ngates` gate records on an n-qubit
 * state with the given shape and report stats[0..7] as above. */
typedef struct dgq_gate_rec {
    int type;                /* numbering of src/cpp/QuEST_backend.h:89-114 */
    int pad;
    unsigned long long tmask;
    unsigned long long cmask;
    double op[8];
} dgq_gate_rec;
int dgq_plan_only(int n, int prec, const dgq_gate_rec *gates, int ngates, int k, int R, int W, int max_ops,
                  int stats[8]);

/* Registers larger than the device budget (DGQ_HBM_BUDGET bytes when set,
 * else the device's free memory less a margin) live in pinned host memory as
 * 2^K pieces of 2^nl amplitudes and are paged through the device.  Returns 1
 * for a register resident on the device, 0 for a paged one, filling nl, K and
 * the statistics of its last flush: sections, high <-> piece qubit moves,
 * pieces run, zero pieces skipped, copy ms, copy ms hidden under compute,
 * compute ms, wall ms, bytes copied in, bytes copied out. */
int dgq_paged_info(Qureg q, int *piece_qubits, int *pieces_log2, double stats[10]);

/* A register that only records and plans, with no device state (no GPU
 * needed): the gate calls of include/QuEST.h validate and record on it as on
 * any register, so the front end and the planner can be checked on any host.
 * Anything that would execute or read the state (flush, probabilities,
 * amplitudes, initStateFromAmps, measurement) ends the program with an
 * error on such a register. */
Qureg dgq_create_recorder(int numQubits);
void dgq_destroy_recorder(Qureg q);

/* Plan what q has recorded (pending initialisation included) with the given
 * shape and diagonal fusion on/off, and copy the plan into the caller's
 * arrays (C ABI of include/dgq_engine.h).  counts[0..2] receive the numbers of
 * blocks, phases and ops; l2p (n ints, may be NULL) the logical -> physical
 * qubit map the plan leaves.  Returns 0, 1 when a capacity is too small
 * (counts then give the sizes; nothing else is written), -1 on bad
 * arguments.  The recording is left as it is. */
int dgq_recorder_plan(Qureg q, int k, int R, int W, int max_ops, int fuse, dq_block *blocks, int cap_blocks,
                      dq_phase *phases, int cap_phases, dq_op *ops, int cap_ops, int counts[3], int *l2p);

/* The recording of q as gate records (physical qubit positions, the encoding
 * of src/cpp/QuEST_backend.h:160-178) for another executor such as the
 * request server of dgq_serve.h: *n_gates receives the count; gates (cap
 * entries) the records; l2p (n ints, may be NULL) the logical -> physical
 * qubit map after them; init_* the pending initialisation (DQ_INIT_*).
 * Returns 0, 1 when cap is too small (only the counts and map are written),
 * -1 on bad arguments. */
int dgq_recorder_export(Qureg q, dgq_gate_rec *gates, int cap, int *n_gates, int *l2p, int *init_kind,
                        unsigned long long *init_index, double *init_value);

/* `shots` samples of q's state without collapsing it: sample s is the first
 * basis state (logical index) whose cumulative probability exceeds u[s]
 * times the total, u[s] in [0, 1).  Flushes recorded gates first.  Returns 0,
 * or -1 on bad arguments.  Not for paged registers. */
int dgq_sample(Qureg q, int shots, const double *u, unsigned long long *out);

/* The fused-sweep shape a new register of n qubits and precision prec
 * (DQ_F32 / DQ_F64) gets: out = {k, R, W, max_ops}.  It comes from
 * DGQ_TUNE=k,R,W,max_ops when set, else from the tuning table measured by
 * dgquest_amd/autotune.py (DGQ_TUNE_TABLE, else dgquest_amd/tune_table.txt;
 * the entry with the nearest n), else the planner defaults.  This replaces
 * the reference's time predictor (src/cpp/time_predictor.cpp:127-395). */
int dgq_tune_lookup(int prec, int n, int out[4]);

/* Reload the tuning table from `path` (NULL: the default location).
 * Returns the number of entries, -1 if the file cannot be read. */
int dgq_tune_reload(const char *path);

/* Precision form: from now on, flushes of this FP32 register run as dense
 * sweeps of width k (5, 6 or 7): the recorded gates are grouped into dense
 * 2^k x 2^k complex matrices, applied on FP32 MFMA with FP32 accumulation
 * inside 13-qubit tiles (dq_apply_dense).  k = 0 returns to fused
 * butterflies.  Gates recorded before the call run first, in the current
 * form.  -1 for a FP64, paged or too small register (k + 5 <= min(n, 13)). */
int dgq_set_dense(Qureg q, int k);

/* Of the last flush: out = {sweeps, dense ops, device GEMM flops (8 * 2^k *
 * 2^n per op), host ms of the flush}. */
int dgq_last_dense_stats(Qureg q, double out[4]);

/* The dense plan of a recording register with width k, for checking on any
 * host: *n_ops ops in order, op i acting on qubits op_qubits[k*i .. k*i+k-1]
 * (matrix-index bit j = qubit op_qubits[k*i+j]) with matrix mats[i]
 * (2^k x 2^k row-major {re, im} doubles).  Returns 0, 1 if cap_ops is too
 * small (only counts written), -1 on bad arguments. */
int dgq_recorder_plan_dense(Qureg q, int k, int cap_ops, int *n_ops, int *n_blocks, int *op_qubits, double *mats);

/* Flush q, move its qubits to their logical positions, wait, and return the
 * device address of its state (2^n interleaved complex amplitudes of the
 * precision written to *prec, DQ_F32 / DQ_F64), for engine calls such as
 * dq_compare_states across registers of the two precisions.  NULL for a
 * recording or paged register.  Valid until the next call on q. */
void *dgq_state_pointer(Qureg q, int *prec);

#ifdef __cplusplus
}
#endif

#endif /* DGQUEST_EXT_H */
