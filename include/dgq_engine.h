/* dgq_engine.h -- C ABI of the gfx950 state-vector engine.
 *
 * This is the seam between the C++ host (gate recorder, DAG planner,
 * executor) and the hand-written HIP kernels.  It replaces the reference's
 * internal C++ launch layer:
 *   launchKernel / calculateSectionGpu      src/cpp/QuEST_backend.h:433-436,
 *                                           src/cpp/QuEST_gpu_kernel.cu:529-587
 *   statevec_fusedKernel + __constant__     src/cpp/QuEST_gpu_kernel.cu:172-359
 *   reorderDataGpu / reorder_data_kernel    src/cpp/QuEST_gpu_kernel.cu:589-654
 *   statevec_calculateProbabilityOfZeroLocal src/cpp/QuEST_backend.h:478,
 *                                           src/cpp/QuEST_gpu_kernel.cu:656-840
 *   statevec_getAmp page read               src/cpp/QuEST_backend_distributed.cpp:649-663
 *   page H2D/D2H copies, pinned pool        src/cpp/QuEST_backend_distributed.cpp:60-73,332-356
 * Every function returns DQ_OK (0) or a negative DQ_ERR_* code and never
 * exits; dq_last_error() gives a message.  Signatures carry only plain
 * pointers, sizes and the POD descriptors below.  Streams are opaque
 * (hipStream_t) pointers; NULL means the default stream.
 *
 * Descriptors travel per call in a caller-owned device workspace (no global
 * __constant__ symbol), so calls on different streams never race.
 *
 * State layout: 2^n interleaved complex amplitudes {re, im} of float
 * (DQ_F32) or double (DQ_F64); qubit q is bit q of the amplitude index.
 */
#ifndef DGQ_ENGINE_H
#define DGQ_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQ_OK 0
#define DQ_ERR_ARG (-1)
#define DQ_ERR_HIP (-2)
#define DQ_ERR_NOMEM (-3)
#define DQ_ERR_UNSUPPORTED (-4)

#define DQ_F32 1
#define DQ_F64 2

/* ---- descriptors of one fused sweep ("block") ------------------------------
 * A block applies an ordered op list to every tile of a state.  A tile is
 * T = 6 + R + W qubits (tile bits); qubits 0..5 are always tile bits 0..5
 * so every wavefront reads and writes 64 consecutive amplitudes.  The other
 * n - T qubits enumerate tiles (one workgroup per tile).
 * Inside a workgroup of 64 << W threads, each thread holds 2^R amplitudes in
 * registers.  A layout says which tile bit each lane bit, register bit and
 * wave bit stands for.  A phase = a layout + the ops run in it; between
 * phases the tile is re-distributed through LDS.  A 1-target op may only
 * target a register bit of its phase; diagonal ops may touch any bit. */

#define DQ_MAX_TILE_BITS 14
#define DQ_MAX_REG_BITS 5
#define DQ_MAX_WAVE_BITS 3

/* A phase: a layout plus the ops run in it.  The layout is given as one-hot
 * tile-bit masks: register bit r of a thread's register index stands for
 * tile bit log2(reg_t[r]); lane bit l for log2(lane_t[l]); wave bit w for
 * log2(wave_t[w]).  Together they must cover the T tile bits exactly once. */
typedef struct dq_phase {
    uint32_t reg_t[DQ_MAX_REG_BITS];
    uint32_t wave_t[DQ_MAX_WAVE_BITS];
    uint32_t lane_t[6];
    uint32_t op_begin; /* index into the op array of the call */
    uint32_t op_count;
} dq_phase;

/* op kinds */
#define DQ_OP_MAT 1   /* 2x2 complex matrix on register bit tbit: m = u00,u01,u10,u11 (re,im) */
#define DQ_OP_NEG 2   /* amplitude *= -1 */
#define DQ_OP_PHASE 3 /* amplitude *= (m[0] + i m[1]) */
#define DQ_OP_SWAP 5  /* X on register bit tbit (exchange the pair) */
#define DQ_OP_RMAT 6  /* real 2x2 matrix [[m0, m1], [m2, m3]] on register bit tbit */
#define DQ_OP_DIAGR 7 /* amplitude of register index j *= table[j]; the table (2^R complex)
                       * follows in the next tbit ops, 4 complex (re, im) per op in m[] */
#define DQ_OP_SHEAR 8 /* two real shears on register bit tbit, in place: x += m0 * y, then
                       * y += m1 * x (x = amplitude with the bit 0, y = with the bit 1);
                       * 4 FMAs per pair, the real part of a factored 1-qubit gate */

/* Conditions: an op acts on an amplitude only if
 *   (reg_index & c_reg) == v_reg, (threadIdx.x & c_thr) == v_thr and
 *   (tile_base & c_wg) == v_wg,
 * where reg_index is the amplitude's register index in its thread and
 * tile_base holds the non-tile qubits of the workgroup.  Controls are bits
 * required to be 1; a condition on a bit being 0 (projections) sets the
 * mask bit and leaves the value bit 0.  For DQ_OP_MAT / DQ_OP_SWAP the
 * condition is tested on the pair's lower index and must not involve tbit.
 * m holds 8 values of the state precision (double for DQ_F64, float for
 * DQ_F32, packed from the start of the array). */
typedef struct dq_op {
    uint32_t kind;
    uint32_t tbit;
    uint32_t c_reg;
    uint32_t v_reg;
    uint32_t c_thr;
    uint32_t v_thr;
    uint64_t c_wg;
    uint64_t v_wg;
    double m[8];
} dq_op;

/* how a block obtains its input amplitudes */
#define DQ_INIT_LOAD 0      /* read from the state */
#define DQ_INIT_CLASSICAL 1 /* |init_index>, nothing read */
#define DQ_INIT_PLUS 2      /* every amplitude = init_value, nothing read */

typedef struct dq_block {
    uint32_t n_qubits;
    uint32_t tile_bits; /* T = 6 + R + W */
    uint32_t reg_bits;  /* R; 0 selects the small-state form (tile = whole state) */
    uint32_t wave_bits; /* W */
    uint32_t init_kind;
    uint32_t phase_begin; /* index into the phase array of the call */
    uint32_t phase_count; /* >= 1 */
    uint32_t pad0;
    uint64_t nontile_mask; /* qubits enumerated by the workgroup index */
    uint64_t init_index;
    double init_value;
    uint64_t pad1;
    /* One-hot qubit masks of register and wave bits when reading (layout of
     * phase 0) and when writing (layout of the last phase).  Lane bit l is
     * always qubit l at both ends, so every wavefront moves 64 consecutive
     * amplitudes.  Reading and writing different qubit masks permutes the
     * tile's qubits (the in-HBM reorder). */
    uint64_t ld_reg[DQ_MAX_REG_BITS];
    uint64_t ld_wave[DQ_MAX_WAVE_BITS];
    uint64_t st_reg[DQ_MAX_REG_BITS];
    uint64_t st_wave[DQ_MAX_WAVE_BITS];
} dq_block;

/* A job = one state and one block applied to it.  A launch runs one or
 * more jobs that share (n, R, W): one state with many tiles, or a batch of
 * small states with one tile each. */
typedef struct dq_job {
    void *state;
    uint32_t block; /* index into the block array of the call */
    uint32_t pad;
    /* Index bits of this state piece inside a larger register (bits at and
     * above the block's n_qubits; 0 for a whole state).  They take part in
     * the ops' workgroup conditions (c_wg / v_wg), never in addresses: a
     * register paged through the device in pieces runs one plan on every
     * piece, controls and diagonals on the piece-index qubits resolved per
     * piece. */
    uint64_t base_hi;
} dq_job;

/* ---- runtime plumbing ------------------------------------------------------ */
int dq_device_init(int device);
const char *dq_last_error(void);
int dq_device_info(char *name, size_t name_len, int *num_cus, size_t *total_mem, size_t *free_mem);
int dq_malloc(void **ptr, size_t bytes);
int dq_free(void *ptr);
int dq_host_alloc(void **ptr, size_t bytes); /* pinned host memory */
int dq_host_free(void *ptr);
int dq_stream_create(void **stream);
int dq_stream_destroy(void *stream);
int dq_stream_sync(void *stream);
int dq_device_sync(void);
#define DQ_H2D 1
#define DQ_D2H 2
#define DQ_D2D 3
int dq_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
int dq_event_create(void **ev);
int dq_event_destroy(void *ev);
int dq_event_record(void *ev, void *stream);
int dq_event_elapsed_ms(void *start, void *stop, float *ms);
int dq_stream_wait_event(void *stream, void *ev);

/* ---- kernels ---------------------------------------------------------------- */

/* Fill a state with |index> (kind DQ_INIT_CLASSICAL) or with `value` in every
 * amplitude (DQ_INIT_PLUS).  Replaces the init branch of
 * src/cpp/QuEST_gpu_kernel.cu:209-226 when no gate follows the init. */
int dq_init_state(void *state, int n_qubits, int prec, int kind, uint64_t index, double value, void *stream);

/* Run `n_launch` launches; launch i runs jobs [job_begin[i], job_begin[i+1])
 * (job_begin has n_launch + 1 entries).  Descriptors are copied into the
 * device workspace `d_ws` (capacity ws_bytes) with one async copy; the host
 * arrays may be reused as soon as the call returns.  Replaces
 * launchKernel/statevec_fused/statevec_fusedKernel
 * (src/cpp/QuEST_gpu_kernel.cu:502-587). */
int dq_apply_blocks(int prec,
                    const dq_block *blocks, uint32_t n_blocks,
                    const dq_phase *phases, uint32_t n_phases,
                    const dq_op *ops, uint32_t n_ops,
                    const dq_job *jobs, const uint32_t *job_begin, uint32_t n_launch,
                    void *d_ws, size_t ws_bytes, void *stream);

/* Bytes of device workspace dq_apply_blocks needs for these counts. */
size_t dq_apply_ws_bytes(uint32_t n_blocks, uint32_t n_phases, uint32_t n_ops, uint32_t n_jobs);

/* The device program dq_apply_blocks builds for one phase's ops (register
 * bits reg_bits), on the host with no GPU: *len gets its size; with out, up
 * to cap bytes are written.  The format is the op interpreter's (the op-code
 * table in the engine's generated interpreter source).  For checking the
 * encoding on any host; it has no counterpart in the reference (the
 * reference's fused kernel reads its gate list directly,
 * src/cpp/QuEST_gpu_kernel.cu:175-359). */
int dq_build_program(int prec, int reg_bits, const dq_op *ops, uint32_t n_ops, unsigned char *out, size_t cap,
                     size_t *len);

/* P(qubit q = 0) for every qubit q < n, and the total probability in
 * out[n] (n + 1 doubles, host memory).  One read of the state plus a small
 * deterministic second stage.  Synchronises `stream`.  Replaces
 * statevec_calculateProbabilityOfZeroLocal (src/cpp/QuEST_gpu_kernel.cu:656-840). */
int dq_prob_zero_all(const void *state, int n_qubits, int prec, double *out,
                     void *d_ws, size_t ws_bytes, void *stream);
size_t dq_prob_ws_bytes(int n_qubits);

/* Same reduction, result left in device memory d_out (n + 1 doubles); no sync. */
int dq_prob_zero_all_async(const void *state, int n_qubits, int prec, double *d_out,
                           void *d_ws, size_t ws_bytes, void *stream);

/* Same reduction with no workspace argument and no allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops).  The partial sums
 * live in a device array of the library (2.2 MiB, one copy on every device
 * it runs on, held until the process ends); dq_device_init must have been
 * called for the current device.  n_qubits <= 36 (larger is refused with
 * DQ_ERR_ARG).  Up to 30 qubits it runs exactly the launches of
 * dq_prob_zero_all_async (bit-identical results); above, the state is
 * reduced in chunks of 2^30 amplitudes whose results a last kernel sums.
 * Asynchronous on `stream`.  Every call on a device shares that array, so
 * calls that overlap on one device (other streams, other threads) race. */
int dq_prob_zero_all_noalloc(const void *state, int n_qubits, int prec, double *d_out, void *stream);

/* ---- dense sweeps: FP32 state, dense 2^k x 2^k blocks on MFMA --------------
 * A dense block loads a tile of T qubits (qubits 0..5 always tile bits
 * 0..5; tile bit b = qubit tile_q[b]) into LDS, applies its dense ops in
 * order, and stores the tile back.  A dense op multiplies, for every
 * assignment of the other tile bits, the 2^k-vector of amplitudes indexed by
 * its k tile bits (matrix-index bit j = tile bit bits[j]) by a 2^k x 2^k
 * complex matrix, computed as one real GEMM per tile on FP32 MFMA with FP32
 * accumulation.  k is 5, 6 or 7, the same for all ops of a call, and
 * k + 5 <= T <= 13. */
typedef struct dq_dense_block {
    uint32_t n_qubits;
    uint32_t tile_bits;   /* T */
    uint32_t op_begin;    /* index into the op array of the call */
    uint32_t op_count;
    uint64_t nontile_mask;
    uint8_t tile_q[16];
} dq_dense_block;

typedef struct dq_dense_op {
    uint32_t bits[8];     /* tile bits of matrix-index bits 0..k-1, ascending */
    uint64_t mat;         /* index of the op's matrix in the call's matrix array */
} dq_dense_op;

/* Apply `n_blocks` dense blocks in order to an FP32 state of n qubits.
 * mats: n_mats complex matrices, 2^k x 2^k each, row-major {re, im} doubles
 * (host memory; rounded to float and laid out for MFMA by this call). */
int dq_apply_dense(int k, const dq_dense_block *blocks, uint32_t n_blocks, const dq_dense_op *ops, uint32_t n_ops,
                   const double *mats, uint32_t n_mats, void *state, void *d_ws, size_t ws_bytes, void *stream);
size_t dq_dense_ws_bytes(int k, uint32_t n_blocks, uint32_t n_ops, uint32_t n_mats);

/* Read-out of `count` states of n qubits each (device pointers in the host
 * array states[]): for every state, P(qubit q = 0) for q < n and the total
 * probability (n + 1 doubles at d_p0 + i * (n + 1); skipped when d_p0 is
 * NULL), and `shots` samples (d_shots + i * shots): sample s is the first
 * amplitude index whose cumulative probability exceeds u[i * shots + s]
 * times the state's total, u in [0, 1) (host array).  One read of each state
 * builds P0 and a tree of probability sums (groups of 2^10 amplitudes, then
 * sums of 2^10 entries per level); each sample then reads at most 2^10
 * values per level.  Asynchronous on `stream`; the host arrays may be reused
 * when the call returns.  The per-state P0 reduction replaces
 * statevec_findProbabilityOfZeroKernel (src/cpp/QuEST_gpu_kernel.cu:769-807)
 * for a batch of registers; sampling has no counterpart in the reference
 * (measure, src/QuEST_common.c:97-115, draws one qubit at a time). */
int dq_readout_many(int prec, int n_qubits, const void *const *states, uint32_t count, const double *u,
                    uint32_t shots, double *d_p0, uint64_t *d_shots, void *d_ws, size_t ws_bytes, void *stream);
size_t dq_readout_ws_bytes(int n_qubits, uint32_t count, uint32_t shots);

/* Agreement of two n-qubit states a, b of any precisions (e.g. an FP64 and
 * an FP32 run of one circuit): d_out (device, 5 doubles) = re <a|b>,
 * im <a|b>, <a|a>, <b|b>, max_i |a_i - b_i|.  Asynchronous on `stream`. */
int dq_compare_states(const void *a, int prec_a, const void *b, int prec_b, int n_qubits, double *d_out,
                      void *d_ws, size_t ws_bytes, void *stream);
size_t dq_compare_ws_bytes(int n_qubits);

/* 0 if all work recorded before `ev` has completed, 1 if not yet (no wait). */
int dq_event_query(void *ev);
/* Wait for the work recorded before `ev`. */
int dq_event_sync(void *ev);

/* Read `count` amplitudes of logical indices start .. start+count-1 as
 * {re, im} doubles into host memory.  l2p[q] is the physical bit that holds
 * logical qubit q (identity if NULL).  Synchronises `stream`.  Replaces
 * statevec_getAmp (src/cpp/QuEST_backend_distributed.cpp:649-663). */
int dq_read_amps(const void *state, int n_qubits, int prec, const int *l2p,
                 uint64_t start, uint64_t count, double *out,
                 void *d_ws, size_t ws_bytes, void *stream);

/* Write `count` amplitudes starting at physical index start from SoA host
 * arrays (converted to the state precision).  Replaces the page fill of
 * statevec_initStateFromAmps (src/cpp/QuEST_backend_distributed.cpp:215-241). */
int dq_write_amps(void *state, int n_qubits, int prec, uint64_t start, uint64_t count,
                  const double *re, const double *im, void *d_ws, size_t ws_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DGQ_ENGINE_H */
