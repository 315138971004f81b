// engine.hip -- gfx950 kernels and the C ABI declared in include/dgq_engine.h.
//
// Kernels:
//   sweep_kernel<T,R,W>  one fused block over every tile of a state (or over a
//                        batch of small states, one tile each).  Replaces the
//                        reference's statevec_fusedKernel
//                        (src/cpp/QuEST_gpu_kernel.cu:175-359) and, through the
//                        store permutation, its reorder_data_kernel (:589-636).
//   small_kernel<T>      whole state in LDS, ops in global-index form; for
//                        registers too small for a 64-lane tile (n < 8).
//   prob_stage1/2<T>     P(q=0) for all qubits + total probability
//                        (replaces :656-840 with wave64 reductions).
//   read/write/init      amplitude gather/scatter and state fill.
//
// Layout: 2^n interleaved complex amplitudes; qubit q = bit q of the index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <type_traits>
#include <vector>

#include "dgq_engine.h"

// ----------------------------------------------------------------------------
// error handling
// ----------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPC(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(DQ_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

extern "C" const char *dq_last_error(void) { return g_err; }

// ----------------------------------------------------------------------------
// device helpers
// ----------------------------------------------------------------------------
template <typename T> struct V2;
template <> struct V2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct V2<float> { typedef float type __attribute__((ext_vector_type(2))); };

// The register tile of one thread: 2^R amplitudes, real and imaginary parts
// as separate scalars.  Every op below updates these registers in place, so
// the tile keeps one register assignment through the op loop.
template <typename T, int R>
struct Regs {
    T re[1 << R];
    T im[1 << R];
};

// Per-phase and per-block records the engine's host code forms for the sweep
// kernels (u32 arrays in the workspace):
//   phase_aux, DQ_PAUX words per phase: [0] byte offset of the phase's device
//     program, [1] its op count; for the exchange from the previous phase into
//     this one, LDS byte offsets under that exchange's swizzle:
//     [DQ_XA + j] / [DQ_XB + j] the uniform part of register index j in the
//     previous / this phase's layout, [DQ_TA + i] / [DQ_TB + i] the part of
//     thread bit i (lane bits 0..5, then wave bits).  Split exchanges (mode
//     1, owned slots) use the same words for their two computed sides, [2] = 1
//     when the exchange needs its middle barrier, and [DQ_TF + i] /
//     [DQ_FJ + j] for the thread's own slots.
//   block_aux, DQ_BAUX words per block: the non-tile qubits as runs of
//     consecutive qubits, [i] = first bit in the tile index | run length << 16,
//     [16 + i] = the run's lowest qubit; runs past the last have length 0.
#define DQ_PAUX 144
#define DQ_XA 8
#define DQ_XB 40
#define DQ_TA 72
#define DQ_TB 81
#define DQ_TF 96
#define DQ_FJ 105
#define DQ_BAUX 32
#define DQ_RUNS 10

// Byte offset of a thread's part of an exchange address: xor of the parts of
// its set thread bits.  The wave bits are wave-uniform (wv is a scalar), so
// their part is formed with scalar selects; each lane bit costs a bit extract
// (0 or all ones), an and and an xor, with no branch.
template <int W>
__device__ __forceinline__ uint32_t thread_part(const uint32_t *__restrict__ t, uint32_t lane, uint32_t wv) {
    uint32_t tv[6 + W];
#pragma unroll
    for (int i = 0; i < 6 + W; i++) tv[i] = t[i];
    uint32_t x = 0;
#pragma unroll
    for (int w = 0; w < W; w++) x ^= ((wv >> w) & 1u) ? tv[6 + w] : 0u;
#pragma unroll
    for (int i = 0; i < 6; i++) x ^= (uint32_t)(-(int32_t)((lane >> i) & 1u)) & tv[i];
    return x;
}

// An LDS element at byte offset `off` of the workgroup's LDS.  The sweep
// kernels declare no static LDS (the engine checks their static group
// segment is 0 when it initialises a device), so their dynamic LDS starts at
// address 0 and an offset is the address itself: no base add per access.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) T *lds_at(uint32_t off) {
    return (__attribute__((address_space(3))) T *)(size_t)off;
}

// Re-distribute the tile from the layout of the previous phase to that of
// this one through LDS (each thread writes its 2^R amplitudes, then reads
// 2^R).  Exchange modes (XM):
//  0  whole amplitudes.  An element's LDS slot is its tile index under a
//     linear swizzle the host picks per exchange (xor of the low bits with
//     images of the high bits) so that the lanes served together by the LDS in
//     either layout hit distinct banks; a byte offset is the thread's part xor
//     a uniform part per register index j, from the host's tables (ex = this
//     phase's phase_aux), one vector xor per access.  A barrier leads every
//     exchange after a block's first (other threads may still read the slots).
//  1  real parts, then imaginary parts, through an LDS array of scalars (half
//     the footprint: more workgroups per CU), in owned slots: every position
//     (thread, register index) of a layout has a fixed slot, own[j] for this
//     thread's register j, under one swizzle per block.  Real parts go from
//     the old positions' slots to their new owners (ex's TB / XB parts: the
//     slot of the previous layout's position of each new element), imaginary
//     parts from their old owners to the new positions' slots (TA / XA).  The
//     imaginary read is from the thread's own slots, which the next
//     exchange's real write reuses, so no barrier leads an exchange; the one
//     between the passes (ex[2]) is needed only when the layout change is not
//     a set of pairwise swaps of positions, since only then may a thread
//     write an imaginary part into a slot another thread has yet to read a
//     real part from.
template <int R, int W, int XM, typename T>
__device__ __forceinline__ void exchange(Regs<T, R> &a, const uint32_t *__restrict__ ex, bool lead_barrier,
                                         uint32_t lane, uint32_t wv, const uint32_t *own) {
    constexpr int NR = 1 << R;
    const uint32_t *__restrict__ xa = ex + DQ_XA;
    const uint32_t *__restrict__ xb = ex + DQ_XB;
    const uint32_t sa = thread_part<W>(ex + DQ_TA, lane, wv);
    const uint32_t sb = thread_part<W>(ex + DQ_TB, lane, wv);
    if constexpr (XM == 0) {
        if (lead_barrier) __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) {
            auto *q = lds_at<T>(sa ^ xa[j]);
            q[0] = a.re[j];
            q[1] = a.im[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const auto *q = lds_at<T>(sb ^ xb[j]);
            a.re[j] = q[0];
            a.im[j] = q[1];
        }
    } else {
        (void)lead_barrier;
#pragma unroll
        for (int j = 0; j < NR; j++) *lds_at<T>(own[j]) = a.re[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) a.re[j] = *lds_at<T>(sb ^ xb[j]);
        if (ex[2]) __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) *lds_at<T>(sa ^ xa[j]) = a.im[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) a.im[j] = *lds_at<T>(own[j]);
    }
}

// Non-tile qubits of tile number t (t < 2^31) from the block's runs.
__device__ __forceinline__ uint64_t tile_base_runs(const uint32_t *__restrict__ ba, uint32_t t) {
    uint64_t base = 0;
#pragma unroll
    for (int i = 0; i < DQ_RUNS; i++) {
        const uint32_t c = ba[i];
        base |= (uint64_t)((t >> (c & 31u)) & ((1u << (c >> 16)) - 1u)) << ba[16 + i];
    }
    return base;
}

// ---- op interpreter --------------------------------------------------------
// The op list of a phase runs in gfx950 assembly generated by gen_ops_asm.py
// (format of the device program and register use documented there).
#include "ops_asm.inc"

// Occupancy target (waves per SIMD) of a sweep instantiation: the register
// tile takes 2^R amplitudes x (2 * sizeof(T) / 4) VGPRs; allow 40 more for
// addresses, coefficients and temporaries.
template <typename T, int R, int XM>
constexpr int sweep_waves_per_eu() {
    constexpr int amp = (1 << R) * (int)(2 * sizeof(T) / 4);
    constexpr int budget = amp + 40;
    constexpr int w = 512 / budget;
    return w > 8 ? 8 : (w < 1 ? 1 : w);
}

// Non-tile qubits of tile number `t`: its bits deposited into nontile_mask.
__device__ __forceinline__ uint64_t tile_base(uint64_t msk, uint64_t t) {
    uint64_t base = 0;
    while (msk) {
        const int p = __builtin_ctzll(msk);
        base |= (t & 1ull) << p;
        t >>= 1;
        msk &= msk - 1;
    }
    return base;
}

// Element `lane` of the 64 consecutive amplitudes starting at the
// wave-uniform address p: a 32-bit lane byte offset on a scalar base, so the
// access is one global load/store with a scalar address and no 64-bit vector
// address arithmetic.
template <typename P>
__device__ __forceinline__ P lane_ptr(P p, uint32_t lane) {
    return (P)((const __attribute__((address_space(1))) char *)p + lane * (uint32_t)sizeof(*p));
}

// Load a tile into registers: lane l = qubit l, register / wave bits =
// ld_reg / ld_wave.  The wave index wv is wave-uniform (a scalar), so every
// address is "scalar base + lane * sizeof(V)" with no vector math.
template <typename T, int R, int W>
__device__ __forceinline__ void load_tile(Regs<T, R> &a, const dq_block *__restrict__ blk,
                                          const __attribute__((address_space(1))) typename V2<T>::type *st,
                                          uint64_t base, uint32_t wv, uint32_t lane) {
    using V = typename V2<T>::type;
    constexpr int NR = 1 << R;
    uint64_t wb = 0;
#pragma unroll
    for (int w = 0; w < W; w++) wb |= ((wv >> w) & 1u) ? blk->ld_wave[w] : 0ull;
    uint64_t qo[R > 0 ? R : 1];
#pragma unroll
    for (int r = 0; r < R; r++) qo[r] = blk->ld_reg[r];
    const uint32_t ik = blk->init_kind;
    if (ik == DQ_INIT_LOAD) {
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t u = base | wb;
#pragma unroll
            for (int r = 0; r < R; r++)
                if ((j >> r) & 1) u |= qo[r];
            const V v = __builtin_nontemporal_load(lane_ptr(st + u, lane));
            a.re[j] = v.x;
            a.im[j] = v.y;
        }
    } else if (ik == DQ_INIT_CLASSICAL) {
        const uint64_t want = blk->init_index;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            uint64_t u = base | wb;
#pragma unroll
            for (int r = 0; r < R; r++)
                if ((j >> r) & 1) u |= qo[r];
            a.re[j] = ((u | lane) == want) ? T(1) : T(0);
            a.im[j] = T(0);
        }
    } else {
        const T v = (T)blk->init_value;
#pragma unroll
        for (int j = 0; j < NR; j++) { a.re[j] = v; a.im[j] = T(0); }
    }
}

// Store a tile: lane l = qubit l, register / wave bits = st_reg / st_wave.
// (A macro expanded in the kernel body: as a function taking the tile by
// reference, the FP32 instantiations kept part of the tile in scratch.)
#define DQ_STORE_TILE(A, BASE)                                                            \
    do {                                                                                  \
        uint64_t wb_ = 0;                                                                 \
        _Pragma("unroll") for (int w = 0; w < W; w++) wb_ |= ((wv >> w) & 1u) ? blk->st_wave[w] : 0ull; \
        uint64_t qo_[R > 0 ? R : 1];                                                      \
        _Pragma("unroll") for (int r = 0; r < R; r++) qo_[r] = blk->st_reg[r];            \
        _Pragma("unroll") for (int j = 0; j < (1 << R); j++) {                            \
            uint64_t u = (BASE) | wb_;                                                    \
            _Pragma("unroll") for (int r = 0; r < R; r++) if ((j >> r) & 1) u |= qo_[r];  \
            typename V2<T>::type v_;                                                      \
            v_.x = (A).re[j];                                                             \
            v_.y = (A).im[j];                                                             \
            __builtin_nontemporal_store(v_, lane_ptr(st + u, lane));                      \
        }                                                                                 \
    } while (0)

// A one-phase block that moves a qubit between wave bits (ld_wave != st_wave)
// stores part of each wave's amplitudes where another wave of the workgroup
// loaded: every load of the tile must be done before the first store.  With
// two or more phases the first LDS exchange already orders them (each wave
// writes its loaded tile to LDS before the exchange barrier).
template <int W>
__device__ __forceinline__ bool store_needs_barrier(const dq_block *__restrict__ blk) {
    if (blk->phase_count != 1) return false;
    bool x = false;
#pragma unroll
    for (int w = 0; w < W; w++) x |= blk->ld_wave[w] != blk->st_wave[w];
    return x;
}

// The phases of a block on one tile: exchanges through LDS between phases,
// the op program of each phase in the assembly interpreter.  lead: a barrier
// before the first exchange (LDS may still be read for a previous tile).
template <typename T, int R, int W, int XM>
__device__ __forceinline__ void process_tile(Regs<T, R> &a, const dq_phase *__restrict__ ph, uint32_t nph,
                                             const uint32_t *__restrict__ pa, const unsigned char *__restrict__ prog,
                                             uint64_t cbase, uint32_t tid, uint32_t wv, bool lead) {
    // owned-slot exchanges: the thread's own slots are the same in every
    // exchange of a block, formed once
    uint32_t own[XM == 1 ? (1 << R) : 1];
    if constexpr (XM == 1) {
        if (nph > 1) {
            const uint32_t *__restrict__ e1 = pa + DQ_PAUX;
            const uint32_t f = thread_part<W>(e1 + DQ_TF, tid & 63u, wv);
#pragma unroll
            for (int j = 0; j < (1 << R); j++) own[j] = f ^ e1[DQ_FJ + j];
        }
    }
    for (uint32_t p = 0; p < nph; p++) {
        if (p > 0) exchange<R, W, XM, T>(a, pa + p * DQ_PAUX, p > 1 || lead, tid & 63u, wv, own);
        AsmOps<T, R>::run(a, prog + pa[p * DQ_PAUX], cbase, tid);
    }
}

template <typename T, int R, int W, int XM>
__global__ void __launch_bounds__(64 << W, (sweep_waves_per_eu<T, R, XM>()))
sweep_kernel(const dq_job *__restrict__ jobs, const dq_block *__restrict__ blocks,
             const dq_phase *__restrict__ phases, const unsigned char *__restrict__ prog,
             const uint32_t *__restrict__ phase_aux, const uint32_t *__restrict__ block_aux) {
    using GV = __attribute__((address_space(1))) typename V2<T>::type;
    const dq_job job = jobs[blockIdx.y];
    const dq_block *__restrict__ blk = blocks + job.block;
    GV *__restrict__ st = (GV *)(job.state);
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lane = tid & 63u;
    const uint64_t base = tile_base_runs(block_aux + job.block * DQ_BAUX, blockIdx.x);
    Regs<T, R> a;
    load_tile<T, R, W>(a, blk, st, base, wv, lane);
    process_tile<T, R, W, XM>(a, phases + blk->phase_begin, blk->phase_count, phase_aux + blk->phase_begin * DQ_PAUX,
                                 prog, base | job.base_hi, tid, wv, false);
    if (store_needs_barrier<W>(blk)) __syncthreads();
    DQ_STORE_TILE(a, base);
}

// Small registers (n <= 13): whole state in LDS, one workgroup per job.  Ops
// are in global-index form: tbit = target qubit, (c_wg, v_wg) = condition
// over the full index.
template <typename T>
__global__ void __launch_bounds__(256)
small_kernel(const dq_job *__restrict__ jobs, const dq_block *__restrict__ blocks,
             const dq_phase *__restrict__ phases, const dq_op *__restrict__ ops) {
    using V = typename V2<T>::type;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *s = reinterpret_cast<V *>(lds_raw);
    const dq_job job = jobs[blockIdx.y];
    const dq_block *__restrict__ blk = blocks + job.block;
    V *__restrict__ st = reinterpret_cast<V *>(job.state);
    const uint32_t n = blk->n_qubits;
    const uint32_t N = 1u << n;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < N; i += blockDim.x) {
        if (blk->init_kind == DQ_INIT_LOAD) s[i] = st[i];
        else if (blk->init_kind == DQ_INIT_CLASSICAL) { s[i].x = (i == blk->init_index) ? T(1) : T(0); s[i].y = T(0); }
        else { s[i].x = (T)blk->init_value; s[i].y = T(0); }
    }
    const dq_phase *__restrict__ ph = phases + blk->phase_begin;
    for (uint32_t p = 0; p < blk->phase_count; p++) {
        for (uint32_t k = 0; k < ph[p].op_count; k++) {
            const dq_op *__restrict__ op = ops + ph[p].op_begin + k;
            const T *__restrict__ m = reinterpret_cast<const T *>(op->m);
            const uint64_t c = op->c_wg, v = op->v_wg, hi = job.base_hi;
            __syncthreads();
            if (op->kind == DQ_OP_MAT || op->kind == DQ_OP_SWAP || op->kind == DQ_OP_RMAT || op->kind == DQ_OP_SHEAR) {
                const uint32_t t = op->tbit, tm = 1u << t;
                for (uint32_t k2 = tid; k2 < N / 2; k2 += blockDim.x) {
                    const uint32_t i0 = ((k2 >> t) << (t + 1)) | (k2 & (tm - 1));
                    const uint32_t i1 = i0 | tm;
                    if (((i0 | hi) & c) != v) continue;
                    const V x = s[i0], y = s[i1];
                    if (op->kind == DQ_OP_SWAP) { s[i0] = y; s[i1] = x; continue; }
                    if (op->kind == DQ_OP_SHEAR) {
                        V nx, ny;
                        nx.x = x.x + m[0] * y.x; nx.y = x.y + m[0] * y.y;
                        ny.x = y.x + m[1] * nx.x; ny.y = y.y + m[1] * nx.y;
                        s[i0] = nx; s[i1] = ny;
                        continue;
                    }
                    V nx, ny;
                    if (op->kind == DQ_OP_RMAT) {
                        nx.x = m[0] * x.x + m[1] * y.x; nx.y = m[0] * x.y + m[1] * y.y;
                        ny.x = m[2] * x.x + m[3] * y.x; ny.y = m[2] * x.y + m[3] * y.y;
                        s[i0] = nx; s[i1] = ny;
                        continue;
                    }
                    nx.x = m[0] * x.x - m[1] * x.y + m[2] * y.x - m[3] * y.y;
                    nx.y = m[0] * x.y + m[1] * x.x + m[2] * y.y + m[3] * y.x;
                    ny.x = m[4] * x.x - m[5] * x.y + m[6] * y.x - m[7] * y.y;
                    ny.y = m[4] * x.y + m[5] * x.x + m[6] * y.y + m[7] * y.x;
                    s[i0] = nx;
                    s[i1] = ny;
                }
            } else {
                for (uint32_t i = tid; i < N; i += blockDim.x) {
                    if (((i | hi) & c) != v) continue;
                    V x = s[i];
                    if (op->kind == DQ_OP_NEG) { x.x = -x.x; x.y = -x.y; s[i] = x; continue; }
                    V y;
                    y.x = m[0] * x.x - m[1] * x.y;
                    y.y = m[0] * x.y + m[1] * x.x;
                    s[i] = y;
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < N; i += blockDim.x) st[i] = s[i];
}

// ----------------------------------------------------------------------------
// probabilities
// ----------------------------------------------------------------------------
// Stage 1: workgroup g reads amplitudes [g * 2^CB, (g+1) * 2^CB) with
// CB = min(n, 16), 256 threads, thread t owning indices with low 8 bits = t.
// It writes 17 partial sums: P0 partials for qubits 0..15 and the total.
#define PROB_WG_BITS 16
#define PROB_SLOTS 17

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__global__ void __launch_bounds__(256) prob_stage1(const typename V2<T>::type *__restrict__ st, int n, double *__restrict__ part) {
    __shared__ double red[4][PROB_SLOTS];
    const uint32_t tid = threadIdx.x;
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    const uint64_t g0 = (uint64_t)blockIdx.x << cb;
    double tot = 0.0;
    double z[8];  // qubits 8..15 (bits of the per-thread loop index)
#pragma unroll
    for (int q = 0; q < 8; q++) z[q] = 0.0;
    if (cb >= 8) {
        const uint32_t npt = 1u << (cb - 8);
        for (uint32_t i = 0; i < npt; i++) {
            const typename V2<T>::type v = __builtin_nontemporal_load(st + g0 + ((uint64_t)i << 8) + tid);
            const double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
            tot += p;
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (!((i >> q) & 1)) z[q] += p;
        }
    } else if (tid < (1u << cb)) {
        const typename V2<T>::type v = st[g0 + tid];
        tot = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    double vals[PROB_SLOTS];
#pragma unroll
    for (int q = 0; q < 8; q++) vals[q] = ((tid >> q) & 1) ? 0.0 : tot;
#pragma unroll
    for (int q = 0; q < 8; q++) vals[8 + q] = z[q];
    vals[16] = tot;
    const uint32_t wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int s = 0; s < PROB_SLOTS; s++) {
        const double r = wave_sum(vals[s]);
        if (lane == 0) red[wave][s] = r;
    }
    __syncthreads();
    if (tid < PROB_SLOTS) {
        const double r = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        part[(uint64_t)blockIdx.x * PROB_SLOTS + tid] = r;
    }
}

// Stage 2: one workgroup per output q (0..n).  q < 16: sum of the partials
// of slot q; q >= 16: sum of workgroup totals whose workgroup-index bit
// (q - 16) is 0; q == n: sum of all totals.
__global__ void __launch_bounds__(256) prob_stage2(const double *__restrict__ part, uint32_t ngroups, int n, double *__restrict__ out) {
    __shared__ double red[4];
    const int q = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    double acc = 0.0;
    for (uint32_t g = tid; g < ngroups; g += 256) {
        double v;
        if (q == n) v = part[(uint64_t)g * PROB_SLOTS + 16];
        else if (q < PROB_WG_BITS) v = part[(uint64_t)g * PROB_SLOTS + q];
        else v = ((g >> (q - PROB_WG_BITS)) & 1) ? 0.0 : part[(uint64_t)g * PROB_SLOTS + 16];
        acc += v;
    }
    acc = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) out[q] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Device memory of the allocation-free P(0) entry point
// (dq_prob_zero_all_noalloc): per-workgroup partial sums of one chunk of at
// most 2^NA_CHUNK_BITS amplitudes, then NA_CHUNK_BITS + 1 results per chunk
// for up to NA_MAX_QUBITS qubits; 2.2 MiB on every device the library runs on.
#define NA_CHUNK_BITS 30
#define NA_MAX_QUBITS 36
#define NA_PART_DOUBLES ((1u << (NA_CHUNK_BITS - PROB_WG_BITS)) * PROB_SLOTS)
#define NA_CHUNK_DOUBLES ((1u << (NA_MAX_QUBITS - NA_CHUNK_BITS)) * (NA_CHUNK_BITS + 1))
static_assert((size_t)NA_PART_DOUBLES * sizeof(double) % 256 == 0, "chunk results stay 256-byte aligned");
static __device__ __align__(256) double scratch[NA_PART_DOUBLES + NA_CHUNK_DOUBLES];

// ----------------------------------------------------------------------------
// read-out of many states of one size: P(0) per qubit, total probability and
// samples drawn by inverse transform over a tree of probability sums
// ----------------------------------------------------------------------------
// Level 1 of the tree holds the sums of |a|^2 over groups of 2^RO_GB
// amplitudes; every higher level sums RO_FAN entries of the level below,
// until a level has at most RO_FAN entries.  A sample walks down from the
// top, reading at most RO_FAN values per level.
#define RO_GB 10
#define RO_FAN_BITS 10
#define RO_FAN (1u << RO_FAN_BITS)
#define RO_GPW 16        // level-1 groups per wave in stage 1
#define RO_MAX_LEVELS 4  // n <= 40

struct RoLevels {
    double *base[RO_MAX_LEVELS];    // level L+1 of state s at base[L] + s * stride[L]
    uint64_t stride[RO_MAX_LEVELS];
    uint64_t count[RO_MAX_LEVELS];  // entries per state
    int nlev;
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Stage 1, grid (ceil(N1 / 64), count), 256 threads.  Wave w of workgroup x
// owns level-1 groups g = 64x + 16w + k (k < 16); lane l reads amplitudes
// g * 2^gb + 64j + l.  Index bits: 0..5 lane, 6..9 j, 10..13 k, 14..15 w,
// 16.. workgroup -- so the 17 partial slots of a workgroup are P0 partials of
// qubits 0..15 and its total, as in prob_stage1, and stage 2 is shared.
template <typename T>
__global__ void __launch_bounds__(256)
readout_stage1(const void *const *__restrict__ states, int n, double *__restrict__ part, uint64_t part_stride,
               double *__restrict__ lvl1, uint64_t lvl1_stride) {
    using V = typename V2<T>::type;
    __shared__ double red[4][PROB_SLOTS];
    const V *__restrict__ st = reinterpret_cast<const V *>(states[blockIdx.y]);
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int gb = n < RO_GB ? n : RO_GB;
    const uint64_t n1 = 1ull << (n - gb);
    const uint32_t gsize = 1u << gb;
    double tot = 0.0;
    double zj[4] = {0.0, 0.0, 0.0, 0.0}, zk[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < RO_GPW; k++) {
        const uint64_t g = (uint64_t)blockIdx.x * (4 * RO_GPW) + w * RO_GPW + k;
        if (g >= n1) break;  // uniform over the wave
        const V *__restrict__ src = st + (g << gb);
        double gs = 0.0;
        if (gsize == (1u << RO_GB)) {
            V v[16];
#pragma unroll
            for (uint32_t j = 0; j < 16; j++) v[j] = __builtin_nontemporal_load(src + (j << 6) + lane);
#pragma unroll
            for (uint32_t j = 0; j < 16; j++) {
                const double p = (double)v[j].x * (double)v[j].x + (double)v[j].y * (double)v[j].y;
                gs += p;
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if (!((j >> b) & 1)) zj[b] += p;
            }
        } else {
            const uint32_t J = gsize >= 64 ? gsize >> 6 : 1;
            for (uint32_t j = 0; j < J; j++) {
                const uint32_t i = (j << 6) + lane;
                if (i >= gsize) continue;
                const V v = src[i];
                const double p = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                gs += p;
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if (!((j >> b) & 1)) zj[b] += p;
            }
        }
        tot += gs;
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (!((k >> b) & 1)) zk[b] += gs;
        const double s = wave_sum(gs);
        if (lane == 0) lvl1[blockIdx.y * lvl1_stride + g] = s;
    }
    double vals[PROB_SLOTS];
#pragma unroll
    for (int q = 0; q < 6; q++) vals[q] = ((lane >> q) & 1) ? 0.0 : tot;
#pragma unroll
    for (int b = 0; b < 4; b++) vals[6 + b] = zj[b];
#pragma unroll
    for (int b = 0; b < 4; b++) vals[10 + b] = zk[b];
#pragma unroll
    for (int b = 0; b < 2; b++) vals[14 + b] = ((w >> b) & 1) ? 0.0 : tot;
    vals[16] = tot;
#pragma unroll
    for (int s = 0; s < PROB_SLOTS; s++) {
        const double r = wave_sum(vals[s]);
        if (lane == 0) red[w][s] = r;
    }
    __syncthreads();
    if (tid < PROB_SLOTS)
        part[blockIdx.y * part_stride + (uint64_t)blockIdx.x * PROB_SLOTS + tid] =
            (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

// Stage 2, grid (n + 1, count): prob_stage2 per state.
__global__ void __launch_bounds__(256) readout_stage2(const double *__restrict__ part, uint64_t part_stride,
                                                      uint32_t ngroups, int n, double *__restrict__ out) {
    __shared__ double red[4];
    const int q = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const double *__restrict__ pp = part + blockIdx.y * part_stride;
    double acc = 0.0;
    for (uint32_t g = tid; g < ngroups; g += 256) {
        double v;
        if (q == n) v = pp[(uint64_t)g * PROB_SLOTS + 16];
        else if (q < PROB_WG_BITS) v = pp[(uint64_t)g * PROB_SLOTS + q];
        else v = ((g >> (q - PROB_WG_BITS)) & 1) ? 0.0 : pp[(uint64_t)g * PROB_SLOTS + 16];
        acc += v;
    }
    acc = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) out[(uint64_t)blockIdx.y * (n + 1) + q] = (red[0] + red[1]) + (red[2] + red[3]);
}

// One level of the tree: entry e = sum of entries RO_FAN*e .. RO_FAN*e + RO_FAN-1
// of the level below.  Grid (ceil(out_count / 4), count), one wave per entry.
__global__ void __launch_bounds__(256) readout_level(const double *__restrict__ in, uint64_t in_count, uint64_t in_stride,
                                                     double *__restrict__ out, uint64_t out_count, uint64_t out_stride) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t e = (uint64_t)blockIdx.x * 4 + w;
    if (e >= out_count) return;  // uniform over the wave; no barrier follows
    const double *__restrict__ src = in + blockIdx.y * in_stride + (e << RO_FAN_BITS);
    const uint64_t lim = in_count - (e << RO_FAN_BITS);
    double s = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < RO_FAN / 64; j++) {
        const uint64_t i = (j << 6) + lane;
        if (i < lim) s += src[i];
    }
    s = wave_sum(s);
    if (lane == 0) out[blockIdx.y * out_stride + e] = s;
}

// Wave-wide pick inside a chunk of len <= 1024 values (lane l holds values
// 16l .. 16l+15): the first index whose running sum exceeds t.  On return t
// holds the remainder inside the picked value.  A target that rounding puts
// at or past the chunk's total picks the last nonzero value.
template <typename F>
__device__ __forceinline__ uint32_t wave_pick(F value, uint32_t len, double &t, uint32_t lane) {
    double v[16];
    double ls = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < 16; j++) {
        const uint32_t i = lane * 16 + j;
        v[j] = i < len ? value(i) : 0.0;
        ls += v[j];
    }
    double inc = ls;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += y;
    }
    double exc = __shfl_up(inc, 1, 64);
    if (lane == 0) exc = 0.0;
    // lane-local choice: the first j past t, else the last nonzero value
    int jsel = -1, jlast = -1;
    double bsel = 0.0, run = exc;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double nr = run + v[j];
        if (jsel < 0 && nr > t) { jsel = j; bsel = run; }
        if (v[j] > 0.0) jlast = j;
        run = nr;
    }
    const uint64_t hit = __ballot(inc > t);
    uint32_t L;
    bool clamp;
    if (hit) {
        L = (uint32_t)__builtin_ctzll(hit);
        clamp = __shfl(jsel, L, 64) < 0;
    } else {
        const uint64_t nz = __ballot(ls > 0.0);
        L = nz ? 63u - (uint32_t)__builtin_clzll(nz) : 0u;
        clamp = true;
    }
    const int j = clamp ? __shfl(jlast, L, 64) : __shfl(jsel, L, 64);
    const int jj = j < 0 ? 0 : j;
    double vsel = v[0];
#pragma unroll
    for (int k = 1; k < 16; k++)
        if (k == jj) vsel = v[k];
    vsel = __shfl(vsel, L, 64);
    const double before = __shfl(bsel, L, 64);
    t = clamp ? 0.5 * vsel : t - before;
    return L * 16 + (uint32_t)jj;
}

// One wave per sample, grid (ceil(shots / 4), count).  u[s] in [0, 1); the
// sample is the first amplitude index whose cumulative probability exceeds
// u[s] times the total.
template <typename T>
__global__ void __launch_bounds__(256)
readout_sample(const void *const *__restrict__ states, int n, RoLevels lv, const double *__restrict__ u, uint32_t shots,
               uint64_t *__restrict__ out) {
    using V = typename V2<T>::type;
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t s = blockIdx.x * 4 + w;
    if (s >= shots) return;  // uniform over the wave; no barrier follows
    const uint32_t sid = blockIdx.y;
    const int gb = n < RO_GB ? n : RO_GB;
    const int top = lv.nlev - 1;
    const double *__restrict__ A = lv.base[top] + sid * lv.stride[top];
    const uint32_t tlen = (uint32_t)lv.count[top];
    double tl = 0.0;
#pragma unroll
    for (uint32_t j = 0; j < 16; j++) {
        const uint32_t i = lane * 16 + j;
        if (i < tlen) tl += A[i];
    }
    const double total = wave_sum(tl);
    double t = u[(uint64_t)sid * shots + s] * total;
    uint64_t e0 = 0;  // first entry of the chunk at the current level
    for (int L = top; L >= 0; L--) {
        const double *__restrict__ B = lv.base[L] + sid * lv.stride[L] + e0;
        const uint64_t rem = lv.count[L] - e0;
        const uint32_t len = rem < RO_FAN ? (uint32_t)rem : RO_FAN;
        const uint32_t e = wave_pick([&](uint32_t i) { return B[i]; }, len, t, lane);
        e0 = (e0 + e) << (L > 0 ? RO_FAN_BITS : gb);
    }
    const V *__restrict__ st = reinterpret_cast<const V *>(states[sid]) + e0;
    const uint32_t e = wave_pick(
        [&](uint32_t i) {
            const V a = st[i];
            return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
        },
        1u << gb, t, lane);
    if (lane == 0) out[(uint64_t)sid * shots + s] = e0 + e;
}

// ----------------------------------------------------------------------------
// comparison of two states of one size (any precisions): <a|b>, <a|a>, <b|b>
// and max |a_i - b_i|
// ----------------------------------------------------------------------------
#define CMP_SLOTS 5

template <typename TA, typename TB>
__global__ void __launch_bounds__(256)
compare_stage1(const typename V2<TA>::type *__restrict__ a, const typename V2<TB>::type *__restrict__ b, int n,
               double *__restrict__ part) {
    __shared__ double red[4][CMP_SLOTS];
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    const uint64_t g0 = (uint64_t)blockIdx.x << cb;
    const uint32_t cnt = 1u << cb;
    double re = 0.0, im = 0.0, aa = 0.0, bb = 0.0, mx = 0.0;
    for (uint32_t i = tid; i < cnt; i += 256) {
        const typename V2<TA>::type x = a[g0 + i];
        const typename V2<TB>::type y = b[g0 + i];
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        re += xr * yr + xi * yi;
        im += xr * yi - xi * yr;
        aa += xr * xr + xi * xi;
        bb += yr * yr + yi * yi;
        const double dr = xr - yr, di = xi - yi;
        mx = fmax(mx, dr * dr + di * di);
    }
    double v[CMP_SLOTS] = {wave_sum(re), wave_sum(im), wave_sum(aa), wave_sum(bb), wave_max(mx)};
    if (lane == 0)
        for (int k = 0; k < CMP_SLOTS; k++) red[w][k] = v[k];
    __syncthreads();
    if (tid < CMP_SLOTS) {
        const double r = tid == 4 ? fmax(fmax(red[0][4], red[1][4]), fmax(red[2][4], red[3][4]))
                                  : (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        part[(uint64_t)blockIdx.x * CMP_SLOTS + tid] = r;
    }
}

// ----------------------------------------------------------------------------
// dense sweeps: FP32 tile in LDS, dense 2^k x 2^k complex blocks on MFMA
// ----------------------------------------------------------------------------
// A dense op on k tile bits is one real GEMM per tile:
//   [Re y; Im y] = [[Ur, -Ui], [Ui, Ur]] [Re x; Im x],
// rows and the contraction index both of size M = 2^(k+1), one column per
// assignment of the other T - k tile bits (NC = 2^(T-k) columns).
// v_mfma_f32_32x32x2f32: lane l supplies A[l % 32][l / 32] and
// B[l / 32][l % 32]; result register v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].  The host lays every real
// matrix out so that a lane's A operands of four consecutive MFMAs (p = 4q ..
// 4q + 3 over the contraction pairs) are 16 contiguous bytes:
//   mat[((rb * M/8 + q) * 64 + l) * 4 + i] = A[32 rb + l % 32][2 (4q + i) + l / 32].
typedef float v16f __attribute__((ext_vector_type(16)));

#define DENSE_T_MAX 13
#define DENSE_THREADS 512
#define DENSE_COLS 256  // column-table entries per op: 2^(T - k) <= 2^(13 - 5)
#define DENSE_WAVES (DENSE_THREADS / 64)

// LDS slot of tile element e: the low 5 bits are xored with bits 5..9 and
// 10..14, spreading power-of-two strides over the 32 banks (a bijection:
// only the low bits change, by a function of the others).
__device__ __forceinline__ uint32_t dswz(uint32_t e) { return e ^ ((e >> 5) & 31u) ^ ((e >> 10) & 31u); }

// One dense op on the tile in LDS.  Work item = CB chunks of 32 columns x RB
// row blocks of 32 rows (RB x CB = 2 result tiles); one item per wave at
// T = 13.  k = 5 takes RB = 2, CB = 1: an item holds all 2^(k+1) rows of its
// columns.  Wider ops take RB = 1, CB = 2 where the tile has two chunks: each
// A operand (the matrix, from L2) then feeds two MFMAs, which halves the
// matrix traffic per MFMA (for k = 7 every workgroup reads 256 KiB of matrix
// per op).  coloff[c] is column c's tile offset (c's bits placed on the tile
// bits outside the op), a table the host forms per op.  The B rows of the
// op's bits 1..k-1 step with the submask increment x' = ((x | ~S) + 1) & S
// (op bits ascending); the LDS swizzle is linear, so a B address is the
// lane's column part xor the swizzle of x, a scalar.
template <int KL, int CB>
__device__ __forceinline__ void dense_op(float *__restrict__ re, float *__restrict__ im, uint32_t T,
                                         const dq_dense_op &op, const float *__restrict__ mat,
                                         const uint32_t *__restrict__ coloff, uint32_t tid) {
    constexpr int K = 1 << KL, M = 2 * K, MB = M / 32;
    constexpr int RB = CB == 2 ? 1 : (MB < 2 ? MB : 2);  // row blocks per work item
    constexpr int NA = RB * CB;                           // result tiles per work item
    constexpr int NQ = K / 4;                             // groups of four contraction pairs (both halves)
    const uint32_t NC = 1u << (T - KL);
    const uint32_t NCG = NC / (32 * CB);                  // column groups
    const uint32_t items = NCG * (MB / RB);
    const uint32_t w = tid >> 6, lane = tid & 63;
    uint32_t used = 0;
    for (int j = 0; j < KL; j++) used |= 1u << op.bits[j];
    const uint32_t b0 = 1u << op.bits[0];
    const uint32_t S = used & ~b0;  // tile bits of matrix-index bits 1..k-1
    __syncthreads();  // the previous op's writes (or the tile load) are done
    v16f acc[NA];
#pragma unroll
    for (int a = 0; a < NA; a++)
        for (int v = 0; v < 16; v++) acc[a][v] = 0.0f;
    const uint32_t item = w;
    const bool busy = item < items;  // uniform over the wave
    const uint32_t cg = item % NCG, rg = item / NCG;
    if (busy) {
        uint32_t lz[CB];
#pragma unroll
        for (int c = 0; c < CB; c++)
            lz[c] = dswz(coloff[(cg * CB + c) * 32 + (lane & 31u)] | ((lane >> 5) ? b0 : 0u));
        const float4 *__restrict__ ma = reinterpret_cast<const float4 *>(mat) + (size_t)(rg * RB) * NQ * 64 + lane;
        // A operands load two groups ahead of their MFMAs into a four-slot
        // ring, B operands (LDS) one group ahead into a two-slot ring: with the
        // loop unrolled by 4 every slot is a fixed register set (no copies
        // that would wait for the loads), and a scheduling barrier keeps each
        // group's loads ahead of its MFMAs.
        float4 ab[4][RB];
        float bb[2][CB][4];
        uint32_t x = 0;
        // B rows of group q for every chunk: four steps of x over the real
        // half (q < NQ/2), then again from 0 over the imaginary half
        auto read_b = [&](int q, float (*b)[4]) {
            if (q == NQ / 2) x = 0;
            const float *__restrict__ src = q < NQ / 2 ? re : im;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t sx = dswz(x);
#pragma unroll
                for (int c = 0; c < CB; c++) b[c][i] = src[sx ^ lz[c]];
                x = ((x | ~S) + 1u) & S;
            }
        };
#pragma unroll
        for (int r = 0; r < RB; r++) {
            ab[0][r] = ma[(size_t)r * NQ * 64];
            ab[1][r] = ma[((size_t)r * NQ + 1) * 64];  // NQ >= 8
        }
        read_b(0, bb[0]);
#pragma unroll 4
        for (int q = 0; q < NQ; q++) {
            if (q + 2 < NQ) {
#pragma unroll
                for (int r = 0; r < RB; r++) ab[(q + 2) & 3][r] = ma[((size_t)r * NQ + q + 2) * 64];
            }
            if (q + 1 < NQ) read_b(q + 1, bb[(q + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int r = 0; r < RB; r++)
#pragma unroll
                    for (int c = 0; c < CB; c++)
                        acc[r * CB + c] =
                            __builtin_amdgcn_mfma_f32_32x32x2f32(ab[q & 3][r][i], bb[q & 1][c][i], acc[r * CB + c], 0, 0, 0);
        }
    }
    // Write in place once every wave has read what it overwrites.  With
    // RB = MB (k = 5) a work item holds all rows of its columns, so a wave
    // reads exactly the tile elements it writes and program order suffices;
    // otherwise a column's rows are split over several waves.
    if constexpr (MB > RB) __syncthreads();
    if (busy) {
        // Result register v of this lane is row 32 rb + 8 (v / 4) + 4 (lane / 32)
        // + v % 4 (rb = rg RB + r); rows below K are real parts.  Its tile element
        // is the row's bits deposited on the op's bits, or the column's bits;
        // deposit and swizzle are linear over GF(2), so the LDS slot is the xor of
        // sd[b] = dswz(1 << bits[b]) over the row's set bits b, xor a lane part.
        uint32_t sd[KL];
#pragma unroll
        for (int b = 0; b < KL; b++) sd[b] = dswz(1u << op.bits[b]);
#pragma unroll
        for (int c = 0; c < CB; c++) {
            const uint32_t lp = dswz(coloff[(cg * CB + c) * 32 + (lane & 31u)]) ^ ((lane >> 5) ? sd[2] : 0u);
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const uint32_t r0 = (rg * RB + r) * 32;  // the row block's first row (bits 5 and up)
                uint32_t ur = 0;
#pragma unroll
                for (int b = 5; b < KL; b++) ur ^= ((r0 >> b) & 1u) ? sd[b] : 0u;
                float *__restrict__ dst = (r0 >> KL) ? im : re;
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const uint32_t u = ur ^ ((v & 1) ? sd[0] : 0u) ^ ((v & 2) ? sd[1] : 0u) ^ ((v & 4) ? sd[3] : 0u) ^
                                       ((v & 8) ? sd[4] : 0u);
                    dst[u ^ lp] = acc[r * CB + c][v];
                }
            }
        }
    }
}

// One dense block over every tile of an FP32 state: grid = tiles, 512 threads;
// at most 128 VGPRs so that two workgroups (64 KiB of LDS each) share a CU.
template <int KL>
__global__ void __launch_bounds__(DENSE_THREADS, 4)
dense_kernel(const dq_dense_block *__restrict__ blk, const dq_dense_op *__restrict__ ops, const float *__restrict__ mats,
             const uint32_t *__restrict__ ctab, typename V2<float>::type *__restrict__ st) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t T = blk->tile_bits;
    const uint32_t NT = 1u << T;
    float *re = reinterpret_cast<float *>(lds_raw);
    float *im = re + NT;
    const uint32_t tid = threadIdx.x;
    const uint64_t base = tile_base(blk->nontile_mask, blockIdx.x);
    // tile_q read as four words (scalar loads; byte loads would put these
    // uniform values in vector registers)
    const uint32_t *__restrict__ tq32 = reinterpret_cast<const uint32_t *>(blk->tile_q);
    uint32_t tq[4];
#pragma unroll
    for (int k = 0; k < 4; k++) tq[k] = tq32[k];
    uint64_t hq[DENSE_T_MAX];
#pragma unroll
    for (int b = 0; b < DENSE_T_MAX; b++) hq[b] = 1ull << ((tq[b >> 2] >> (8 * (b & 3))) & 0xffu);
    // Element e = tid + 512 i of the tile holds the amplitude whose qubits are
    // e's bits placed on tile_q (bits 0..5 are qubits 0..5).  The placement and
    // the LDS swizzle are linear, so the thread's part (bits 0..8) is formed
    // once and the part of i (bits 9 and up) is a scalar.
    // Bits 6..8 of e are the wave index, a scalar: the thread's part of an
    // address is its lane alone (qubits 0..5), a 32-bit offset on a scalar base.
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lane = tid & 63u;
    uint64_t tg = base;
#pragma unroll
    for (int b = 6; b < 9; b++)
        if ((wv >> (b - 6)) & 1u) tg |= hq[b];
    const uint32_t tz = dswz(tid);
    // i < NI = 2^(T - 9) (2..16) runs over the tile bits 9 and up, whose qubits
    // are scalars.  Every thread issues 16 loads before using the first;
    // below T = 13 index i stands for element i mod NI, so the extra loads,
    // LDS writes and stores repeat an element with the same value (no branch).
    const uint32_t NI = NT / DENSE_THREADS;
    const uint64_t h9 = T > 9 ? hq[9] : 0, h10 = T > 10 ? hq[10] : 0, h11 = T > 11 ? hq[11] : 0,
                   h12 = T > 12 ? hq[12] : 0;
    auto gaddr = [&](uint32_t i) {
        return tg | ((i & 1u) ? h9 : 0ull) | ((i & 2u) ? h10 : 0ull) | ((i & 4u) ? h11 : 0ull) | ((i & 8u) ? h12 : 0ull);
    };
    {
        typename V2<float>::type v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = __builtin_nontemporal_load(lane_ptr(st + gaddr((uint32_t)i & (NI - 1)), lane));
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t z = tz ^ dswz(((uint32_t)i & (NI - 1)) * DENSE_THREADS);
            re[z] = v[i].x;
            im[z] = v[i].y;
        }
    }
    const size_t msz = (size_t)(2u << KL) * (2u << KL);
    // ops wider than 5 bits pair two column chunks per work item when the
    // tile has them (T >= k + 6; T is uniform over the launch)
    const bool pair = KL >= 6 && T >= (uint32_t)KL + 6;
    for (uint32_t o = 0; o < blk->op_count; o++) {
        const dq_dense_op op = ops[blk->op_begin + o];
        const float *m = mats + op.mat * msz;
        const uint32_t *ct = ctab + (size_t)(blk->op_begin + o) * DENSE_COLS;
        if (pair) dense_op<KL, 2>(re, im, T, op, m, ct, tid);
        else dense_op<KL, 1>(re, im, T, op, m, ct, tid);
    }
    __syncthreads();
    {
        typename V2<float>::type v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t z = tz ^ dswz(((uint32_t)i & (NI - 1)) * DENSE_THREADS);
            v[i].x = re[z];
            v[i].y = im[z];
        }
        // an opaque copy of the lane: the store addresses are formed again here
        // instead of the load addresses being kept live through every op
        uint32_t lane2 = lane;
        asm volatile("" : "+v"(lane2));
#pragma unroll
        for (int i = 0; i < 16; i++) __builtin_nontemporal_store(v[i], lane_ptr(st + gaddr((uint32_t)i & (NI - 1)), lane2));
    }
}

// one workgroup: out = {re <a|b>, im <a|b>, <a|a>, <b|b>, max |a_i - b_i|}
__global__ void __launch_bounds__(256) compare_stage2(const double *__restrict__ part, uint32_t ngroups, double *__restrict__ out) {
    __shared__ double red[4][CMP_SLOTS];
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    double v[CMP_SLOTS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t g = tid; g < ngroups; g += 256) {
        for (int k = 0; k < 4; k++) v[k] += part[(uint64_t)g * CMP_SLOTS + k];
        v[4] = fmax(v[4], part[(uint64_t)g * CMP_SLOTS + 4]);
    }
    for (int k = 0; k < 4; k++) v[k] = wave_sum(v[k]);
    v[4] = wave_max(v[4]);
    if (lane == 0)
        for (int k = 0; k < CMP_SLOTS; k++) red[w][k] = v[k];
    __syncthreads();
    if (tid < CMP_SLOTS) {
        const double r = tid == 4 ? sqrt(fmax(fmax(red[0][4], red[1][4]), fmax(red[2][4], red[3][4])))
                                  : (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
        out[tid] = r;
    }
}

// ----------------------------------------------------------------------------
// amplitude gather / scatter / fill
// ----------------------------------------------------------------------------
template <typename T>
__global__ void read_kernel(const typename V2<T>::type *__restrict__ st, int n, const int *__restrict__ l2p,
                            uint64_t start, uint64_t count, double *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint64_t li = start + k;
    uint64_t pi = 0;
    for (int q = 0; q < n; q++) pi |= ((li >> q) & 1ull) << l2p[q];
    const typename V2<T>::type v = st[pi];
    out[2 * k] = (double)v.x;
    out[2 * k + 1] = (double)v.y;
}

template <typename T>
__global__ void write_kernel(typename V2<T>::type *__restrict__ st, uint64_t start, uint64_t count,
                             const double *__restrict__ re, const double *__restrict__ im) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    typename V2<T>::type v;
    v.x = (T)re[k];
    v.y = (T)im[k];
    st[start + k] = v;
}

// grid-stride: a launch may not exceed 2^32 threads (n >= 32 states)
template <typename T>
__global__ void fill_kernel(typename V2<T>::type *__restrict__ st, uint64_t N, int kind, uint64_t index, double value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        typename V2<T>::type v;
        v.x = (kind == DQ_INIT_CLASSICAL) ? (i == index ? T(1) : T(0)) : (T)value;
        v.y = T(0);
        __builtin_nontemporal_store(v, st + i);
    }
}

// ----------------------------------------------------------------------------
// kernel table
// ----------------------------------------------------------------------------
typedef void (*sweep_fn)(const dq_job *, const dq_block *, const dq_phase *, const unsigned char *, const uint32_t *,
                         const uint32_t *);

template <typename T>
struct SweepTable {
    sweep_fn fn[2][DQ_MAX_REG_BITS + 1][DQ_MAX_WAVE_BITS + 1];
    SweepTable() {
        memset(fn, 0, sizeof fn);
#define DQ_INST(R_, W_)                                                                  \
    fn[0][R_][W_] = sweep_kernel<T, R_, W_, 0>;                                          \
    fn[1][R_][W_] = sweep_kernel<T, R_, W_, 1>;
        DQ_INST(1, 0) DQ_INST(2, 0) DQ_INST(3, 0) DQ_INST(4, 0) DQ_INST(5, 0)
        DQ_INST(2, 1) DQ_INST(3, 1) DQ_INST(4, 1) DQ_INST(5, 1)
        DQ_INST(2, 2) DQ_INST(3, 2) DQ_INST(4, 2) DQ_INST(5, 2)
        DQ_INST(3, 3) DQ_INST(4, 3)
        // 14-bit tiles (k = 8) for FP32 only: 32 float amplitudes per thread
        // and a 64 KiB split exchange; the FP64 tile would not fit the registers
        if constexpr (std::is_same<T, float>::value) { DQ_INST(5, 3) }
#undef DQ_INST
    }
};

static SweepTable<double> g_tab64;
static SweepTable<float> g_tab32;

// Exchange mode of a tile shape (see exchange): split exchanges (1) for every
// FP32 tile and for FP64 tiles of 12+ bits; half the LDS per workgroup lets
// more FP32 workgroups share a CU (measured 18% faster on 30-qubit FP32
// circuits).  DGQ_SPLIT_EXCHANGE: 0 whole amplitudes everywhere, 1 split
// everywhere, unset the default.
static int g_split_mode = -1;
static int exchange_mode(int prec, int R, int W, int T) {
    (void)R;
    (void)W;
    if (g_split_mode < 0) {
        const char *e = getenv("DGQ_SPLIT_EXCHANGE");
        g_split_mode = e ? atoi(e) : 2;
    }
    if (g_split_mode == 0) return 0;
    if (g_split_mode == 1) return 1;
    return (prec == DQ_F32 || T >= 12) ? 1 : 0;
}

static sweep_fn lookup(int prec, int R, int W, int xm) {
    if (R < 0 || R > DQ_MAX_REG_BITS || W < 0 || W > DQ_MAX_WAVE_BITS || xm < 0 || xm > 1) return nullptr;
    return prec == DQ_F64 ? g_tab64.fn[xm][R][W] : g_tab32.fn[xm][R][W];
}

static size_t elem_bytes(int prec) { return prec == DQ_F64 ? 16 : 8; }

// bytes per LDS element of an exchange in mode xm
static uint32_t exchange_elem_bytes(int prec, int xm) {
    return xm == 0 ? (uint32_t)elem_bytes(prec) : (uint32_t)elem_bytes(prec) / 2;
}

// ----------------------------------------------------------------------------
// staging of descriptors: a pool of pinned host buffers guarded by events.
// An upload takes a buffer whose last copy has completed, so a stream with a
// long queue never holds up uploads on other streams; the pool grows to
// kStageMax buffers before a caller waits.
// ----------------------------------------------------------------------------
namespace {
struct StageSlot {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
};
constexpr unsigned kStageMax = 64;
std::mutex g_stage_mu;
StageSlot g_stage[kStageMax];
unsigned g_stage_used = 0;
unsigned g_stage_next = 0;
int g_device = -1;
int g_num_cus = 256;
// Address of the allocation-free read-out's `scratch` array on each device
// (hipGetSymbolAddress at dq_device_init; null until then).
constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};
}  // namespace

static int stage_upload(void *d_dst, const std::vector<unsigned char> &src, void *stream) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    StageSlot *slot = nullptr;
    for (unsigned k = 0; k < g_stage_used && !slot; k++) {
        StageSlot &c = g_stage[(g_stage_next + k) % g_stage_used];
        const hipError_t e = hipEventQuery(c.ev);
        if (e == hipSuccess) slot = &c;
        else if (e == hipErrorNotReady) (void)hipGetLastError();
        else HIPC(e);
    }
    if (!slot && g_stage_used < kStageMax) {
        slot = &g_stage[g_stage_used++];
        HIPC(hipEventCreateWithFlags(&slot->ev, hipEventDisableTiming));
    }
    if (!slot) {
        slot = &g_stage[g_stage_next % g_stage_used];
        HIPC(hipEventSynchronize(slot->ev));
    }
    g_stage_next = (unsigned)(slot - g_stage) + 1;
    StageSlot &s = *slot;
    if (s.cap < src.size()) {
        if (s.host) HIPC(hipHostFree(s.host));
        s.host = nullptr;
        s.cap = 0;
        size_t cap = 64 * 1024;
        while (cap < src.size()) cap *= 2;
        HIPC(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
        s.cap = cap;
    }
    memcpy(s.host, src.data(), src.size());
    HIPC(hipMemcpyAsync(d_dst, s.host, src.size(), hipMemcpyHostToDevice, (hipStream_t)stream));
    HIPC(hipEventRecord(s.ev, (hipStream_t)stream));
    return DQ_OK;
}

// ----------------------------------------------------------------------------
// C ABI: plumbing
// ----------------------------------------------------------------------------
extern "C" int dq_device_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(DQ_ERR_ARG, "dq_device_init: device %d out of range", device);
    HIPC(hipSetDevice(device));
    if (!g_scratch_dev[device]) {
        // each device holds its own copy of `scratch`; host code uses this
        // address, never the array's name (a host-side stand-in)
        void *p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = (double *)p;
    }
    if (g_device != device) {
        // allow the large LDS tiles (up to 128 KiB) on every sweep instantiation
        for (int sp = 0; sp < 2; sp++)
            for (int R = 0; R <= DQ_MAX_REG_BITS; R++)
                for (int W = 0; W <= DQ_MAX_WAVE_BITS; W++) {
                    for (const void *fn : {(const void *)g_tab64.fn[sp][R][W], (const void *)g_tab32.fn[sp][R][W]}) {
                        if (!fn) continue;
                        HIPC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
                        // the exchange addresses LDS by offset from 0 (lds_at): no static LDS allowed
                        hipFuncAttributes fa;
                        HIPC(hipFuncGetAttributes(&fa, fn));
                        if (fa.sharedSizeBytes != 0)
                            return fail(DQ_ERR_UNSUPPORTED, "sweep kernel (R=%d, W=%d) has %zu bytes of static LDS", R, W,
                                        (size_t)fa.sharedSizeBytes);
                    }
                }
        {
            hipDeviceProp_t prop;
            HIPC(hipGetDeviceProperties(&prop, device));
            g_num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        }
        HIPC(hipFuncSetAttribute((const void *)small_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        HIPC(hipFuncSetAttribute((const void *)small_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        HIPC(hipFuncSetAttribute((const void *)dense_kernel<5>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        HIPC(hipFuncSetAttribute((const void *)dense_kernel<6>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        HIPC(hipFuncSetAttribute((const void *)dense_kernel<7>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        g_device = device;
    }
    return DQ_OK;
}

extern "C" int dq_device_info(char *name, size_t name_len, int *num_cus, size_t *total_mem, size_t *free_mem) {
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPC(hipGetDeviceProperties(&p, dev));
    if (name && name_len) snprintf(name, name_len, "%s", p.gcnArchName);
    if (num_cus) *num_cus = p.multiProcessorCount;
    size_t fr = 0, tot = 0;
    HIPC(hipMemGetInfo(&fr, &tot));
    if (total_mem) *total_mem = tot;
    if (free_mem) *free_mem = fr;
    return DQ_OK;
}

extern "C" int dq_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(DQ_ERR_ARG, "dq_malloc: null out pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();  // a caller may go on (smaller request, other memory): do not leave the error pending
        return fail(DQ_ERR_NOMEM, "dq_malloc: out of device memory (%zu bytes)", bytes);
    }
    HIPC(e);
    return DQ_OK;
}
extern "C" int dq_free(void *ptr) { HIPC(hipFree(ptr)); return DQ_OK; }
extern "C" int dq_host_alloc(void **ptr, size_t bytes) {
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();  // the paging executor falls back to pageable memory: clear the pending error
        return fail(DQ_ERR_NOMEM, "dq_host_alloc: out of pinned memory (%zu bytes)", bytes);
    }
    HIPC(e);
    return DQ_OK;
}
extern "C" int dq_host_free(void *ptr) { HIPC(hipHostFree(ptr)); return DQ_OK; }
extern "C" int dq_stream_create(void **stream) {
    hipStream_t s;
    HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return DQ_OK;
}
extern "C" int dq_stream_destroy(void *stream) { HIPC(hipStreamDestroy((hipStream_t)stream)); return DQ_OK; }
extern "C" int dq_stream_sync(void *stream) { HIPC(hipStreamSynchronize((hipStream_t)stream)); return DQ_OK; }
extern "C" int dq_device_sync(void) { HIPC(hipDeviceSynchronize()); return DQ_OK; }
extern "C" int dq_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    hipMemcpyKind k = kind == DQ_H2D ? hipMemcpyHostToDevice : kind == DQ_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPC(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
    return DQ_OK;
}
extern "C" int dq_event_create(void **ev) {
    hipEvent_t e;
    HIPC(hipEventCreate(&e));
    *ev = (void *)e;
    return DQ_OK;
}
extern "C" int dq_event_destroy(void *ev) { HIPC(hipEventDestroy((hipEvent_t)ev)); return DQ_OK; }
extern "C" int dq_event_record(void *ev, void *stream) { HIPC(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream)); return DQ_OK; }
extern "C" int dq_event_elapsed_ms(void *a, void *b, float *ms) {
    HIPC(hipEventSynchronize((hipEvent_t)b));
    HIPC(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return DQ_OK;
}
extern "C" int dq_stream_wait_event(void *stream, void *ev) {
    HIPC(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
    return DQ_OK;
}

// ----------------------------------------------------------------------------
// C ABI: kernels
// ----------------------------------------------------------------------------
static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Upper bound of the device program built from n_ops descriptors in n_phases
// phases: at most 320 bytes per descriptor.  A complex matrix in FP64 behind
// a workgroup-condition op takes 128.  A register table of 2^R entries spans
// 1 + 2^R / 4 descriptors; fused with shears and emitted once per assignment
// of at most two folded workgroup bits, it takes at most 4 x (32 + 16 +
// 16 * 2^R + 16 * R) bytes (2560 at R = 5, against 9 x 320).  Plus an END op
// per phase, and padding so the interpreter's fixed-size header loads stay in
// the buffer.
static size_t prog_bound(uint32_t n_phases, uint32_t n_ops) { return (size_t)n_ops * 320 + (size_t)n_phases * 16 + 256; }

extern "C" size_t dq_apply_ws_bytes(uint32_t n_blocks, uint32_t n_phases, uint32_t n_ops, uint32_t n_jobs) {
    return align_up(n_blocks * sizeof(dq_block), 256) + align_up(n_phases * sizeof(dq_phase), 256) +
           align_up(n_ops * sizeof(dq_op), 256) + align_up(n_jobs * sizeof(dq_job), 256) +
           align_up((size_t)n_phases * DQ_PAUX * sizeof(uint32_t), 256) +
           align_up((size_t)n_blocks * DQ_BAUX * sizeof(uint32_t), 256) + align_up(prog_bound(n_phases, n_ops), 256);
}

// GF(2) rank of a set of bit vectors.
static int gf2_rank(const uint32_t *v, int n) {
    uint32_t b[32];
    int r = 0;
    for (int i = 0; i < n; i++) {
        uint32_t x = v[i];
        for (int k = 0; k < r; k++) x = std::min(x, x ^ b[k]);
        if (x) b[r++] = x;
    }
    return r;
}

// LDS swizzle of one exchange: tile index x -> x ^ (xor of g[b] over the set
// bits b >= L of x), g[b] < 2^L.  The LDS serves 128 bytes a clock, i.e. the
// 2^L lanes whose lane bits 0..L-1 vary (2^L = 128 / element bytes); they hit
// distinct banks iff the images of their L tile bits (bit b itself for b < L,
// else g[b]) are linearly independent.  Picks g so that this holds for both
// the writing layout A and the reading layout B (search over the at most 2L
// high bits involved); false leaves g = 0 (correct, with bank conflicts).
static bool pick_swizzle(const dq_phase &A, const dq_phase &B, int L, uint32_t *g) {
    memset(g, 0, 32 * sizeof(uint32_t));
    int sa[6], sb[6];
    for (int l = 0; l < L; l++) {
        sa[l] = __builtin_ctz(A.lane_t[l]);
        sb[l] = __builtin_ctz(B.lane_t[l]);
    }
    std::vector<int> vars;
    for (int l = 0; l < L; l++)
        for (int b : {sa[l], sb[l]})
            if (b >= L && std::find(vars.begin(), vars.end(), b) == vars.end()) vars.push_back(b);
    std::vector<bool> set(32, false);
    auto ok = [&](const int *bits) {  // assigned images of one group independent
        uint32_t v[6];
        int n = 0;
        for (int l = 0; l < L; l++) {
            const int b = bits[l];
            if (b < L) v[n++] = 1u << b;
            else if (set[b]) v[n++] = g[b];
        }
        return gf2_rank(v, n) == n;
    };
    std::vector<uint32_t> cand;
    for (int i = 0; i < L; i++) cand.push_back(1u << i);
    for (uint32_t c = 1; c < (1u << L); c++)
        if (c & (c - 1)) cand.push_back(c);
    long budget = 200000;
    std::function<bool(size_t)> dfs = [&](size_t i) -> bool {
        if (i == vars.size()) return true;
        for (uint32_t c : cand) {
            if (--budget < 0) return false;
            g[vars[i]] = c;
            set[vars[i]] = true;
            if (ok(sa) && ok(sb) && dfs(i + 1)) return true;
            set[vars[i]] = false;
        }
        g[vars[i]] = 0;
        return false;
    };
    if (dfs(0)) return true;
    memset(g, 0, 32 * sizeof(uint32_t));
    return false;
}

// Owned-slot exchanges (split mode 1).  A position of a layout is (thread,
// register index) as the bit vector lane bits 0..5, wave bits, register bits;
// every position has a fixed LDS slot, the position under a swizzle that xors
// images g[b] < 2^L of its high bits b >= L into its low L bits.  The bit of a
// tile bit's position in a phase's layout:
static int pos_bit(const dq_phase &P, uint32_t x, int R, int W) {
    for (int l = 0; l < 6; l++)
        if (P.lane_t[l] == x) return l;
    for (int w = 0; w < W; w++)
        if (P.wave_t[w] == x) return 6 + w;
    for (int r = 0; r < R; r++)
        if (P.reg_t[r] == x) return 6 + W + r;
    return -1;
}

// The tile bit at position bit b of a layout.
static uint32_t tile_bit_at(const dq_phase &P, int b, int W) {
    if (b < 6) return P.lane_t[b];
    if (b < 6 + W) return P.wave_t[b - 6];
    return P.reg_t[b - 6 - W];
}

// One swizzle for all exchanges of a block (consecutive exchanges share the
// slots of the layout between them).  A thread's own slots differ only in its
// lane bits among the 2^L lanes the LDS serves together, so they never
// conflict; the computed sides are the real read (the previous layout's
// positions of the lanes' new elements) and the imaginary write (the next
// layout's positions of their old ones).  Picks g so that on both sides of
// every exchange the images of the lanes' low L bits are independent (search
// as in pick_swizzle); false leaves g = 0 (correct, with bank conflicts).
static bool pick_owned_swizzle(const dq_phase *ph, uint32_t nph, int R, int W, int L, uint32_t *g) {
    memset(g, 0, 32 * sizeof(uint32_t));
    std::vector<std::vector<int>> groups;
    for (uint32_t p = 1; p < nph; p++) {
        const dq_phase &A = ph[p - 1], &B = ph[p];
        std::vector<int> g1, g2;
        for (int l = 0; l < L; l++) {
            g1.push_back(pos_bit(A, B.lane_t[l], R, W));
            g2.push_back(pos_bit(B, A.lane_t[l], R, W));
        }
        groups.push_back(g1);
        groups.push_back(g2);
    }
    std::vector<int> vars;
    for (auto &gr : groups)
        for (int b : gr)
            if (b >= L && std::find(vars.begin(), vars.end(), b) == vars.end()) vars.push_back(b);
    std::vector<bool> set(32, false);
    auto ok = [&](const std::vector<int> &bits) {
        uint32_t v[6];
        int n = 0;
        for (int b : bits) {
            if (b < 0) continue;
            if (b < L) v[n++] = 1u << b;
            else if (set[b]) v[n++] = g[b];
        }
        return gf2_rank(v, n) == n;
    };
    std::vector<uint32_t> cand;
    for (int i = 0; i < L; i++) cand.push_back(1u << i);
    for (uint32_t c = 1; c < (1u << L); c++)
        if (c & (c - 1)) cand.push_back(c);
    long budget = 200000;
    std::function<bool(size_t)> dfs = [&](size_t i) -> bool {
        if (i == vars.size()) return true;
        for (uint32_t c : cand) {
            if (--budget < 0) return false;
            g[vars[i]] = c;
            set[vars[i]] = true;
            bool good = true;
            for (auto &gr : groups) good = good && ok(gr);
            if (good && dfs(i + 1)) return true;
            set[vars[i]] = false;
        }
        g[vars[i]] = 0;
        return false;
    };
    if (dfs(0)) return true;
    memset(g, 0, 32 * sizeof(uint32_t));
    return false;
}

// Tables of the owned-slot exchange from layout A to layout B (phase_aux of
// B): [DQ_TF] / [DQ_FJ] the thread's own slots, [DQ_TB] / [DQ_XB] the real
// read (slot of the A-position of B's element), [DQ_TA] / [DQ_XA] the
// imaginary write (slot of the B-position of A's element), [2] = 1 unless the
// layout change is a set of pairwise swaps of position bits.
static bool owned_tables(const dq_phase &A, const dq_phase &B, int R, int W, int L, const uint32_t *g, uint32_t es,
                         uint32_t *pa) {
    const int T = 6 + W + R;
    auto sig = [&](int b) { return ((1u << b) ^ (b >= L ? g[b] : 0u)) * es; };
    int to_b[32], to_a[32];  // position bit of A -> of B, and back
    for (int b = 0; b < T; b++) {
        to_b[b] = pos_bit(B, tile_bit_at(A, b, W), R, W);
        to_a[b] = pos_bit(A, tile_bit_at(B, b, W), R, W);
        if (to_b[b] < 0 || to_a[b] < 0) return false;
    }
    bool swaps = true;
    for (int b = 0; b < T; b++) swaps &= to_b[to_b[b]] == b;
    pa[2] = swaps ? 0u : 1u;
    for (int l = 0; l < 6; l++) {
        pa[DQ_TF + l] = sig(l);
        pa[DQ_TB + l] = sig(to_a[l]);
        pa[DQ_TA + l] = sig(to_b[l]);
    }
    for (int w = 0; w < W; w++) {
        pa[DQ_TF + 6 + w] = sig(6 + w);
        pa[DQ_TB + 6 + w] = sig(to_a[6 + w]);
        pa[DQ_TA + 6 + w] = sig(to_b[6 + w]);
    }
    for (uint32_t j = 0; j < (1u << R); j++) {
        uint32_t f = 0, xb = 0, xa = 0;
        for (int r = 0; r < R; r++)
            if ((j >> r) & 1) {
                f ^= sig(6 + W + r);
                xb ^= sig(to_a[6 + W + r]);
                xa ^= sig(to_b[6 + W + r]);
            }
        pa[DQ_FJ + j] = f;
        pa[DQ_XB + j] = xb;
        pa[DQ_XA + j] = xa;
    }
    return true;
}

// Swizzled tile index of x under images g (bits b >= L).
static uint32_t swizzle_host(uint32_t x, const uint32_t *g, int L) {
    uint32_t y = x;
    for (uint32_t m = x >> L; m; m &= m - 1) y ^= g[L + __builtin_ctz(m)];
    return y;
}

// block_aux record of a block: its non-tile qubits as runs (see block_aux).
static bool block_runs(const dq_block &b, uint32_t *out) {
    memset(out, 0, DQ_BAUX * sizeof(uint32_t));
    uint64_t m = b.nontile_mask;
    uint32_t in = 0;
    int i = 0;
    while (m) {
        if (i == DQ_RUNS) return false;
        const int pos = __builtin_ctzll(m);
        int len = 0;
        while (pos + len < 64 && ((m >> (pos + len)) & 1)) len++;
        out[i] = in | ((uint32_t)len << 16);
        out[16 + i] = (uint32_t)pos;
        in += (uint32_t)len;
        m &= ~(((len == 64) ? ~0ull : ((1ull << len) - 1)) << pos);
        i++;
    }
    return in <= 31;
}

// Device program of one phase (format and op codes in gen_ops_asm.py /
// ops_asm.inc): register conditions resolved into op variants or masks,
// consecutive unconditioned shears on distinct register bits merged into one
// op, unconditioned diagonal ops folded into the table of their run and that
// table fused with the shears after it, a WG op in front of each op with a
// workgroup condition, END last.
namespace {
struct ProgBuilder {
    int prec, R, NR, ncodes;
    size_t ts;
    const DevOpCode *tab;
    std::vector<unsigned char> &out;
    bool bad = false;  // an op with no code (never sent to the device)
    ProgBuilder(int prec_, int R_, std::vector<unsigned char> &o) : prec(prec_), R(R_), NR(1 << R_), out(o) {
        ts = prec == DQ_F64 ? 8 : 4;
        tab = dev_ops(R, ncodes);
    }
    int code(const char *name, int p0 = -1, int p1 = -1, int p2 = -1, int p3 = -1) const {
        for (int c = 0; c < ncodes; c++)
            if (!strcmp(tab[c].name, name) && tab[c].p[0] == p0 && tab[c].p[1] == p1 && tab[c].p[2] == p2 &&
                tab[c].p[3] == p3)
                return c;
        return -1;
    }
    size_t size(int c) const { return (size_t)(prec == DQ_F64 ? tab[c].size64 : tab[c].size32); }
    // one op into `v`: header then coefficient bytes, zero padded to its size
    void encode(std::vector<unsigned char> &v, int c, uint32_t mask, uint32_t cthr, uint32_t vthr, const void *coef,
                size_t nbytes) {
        if (c < 0 || c >= ncodes || nbytes + 16 > size(c)) {
            bad = true;
            return;
        }
        const size_t at = v.size();
        v.resize(at + size(c), 0);
        const uint32_t hdr[4] = {(uint32_t)c, mask, cthr, vthr};
        memcpy(v.data() + at, hdr, 16);
        if (nbytes) memcpy(v.data() + at + 16, coef, nbytes);
    }
    void emit(const std::vector<unsigned char> &op, uint64_t c_wg, uint64_t v_wg) {
        if (c_wg) {
            const uint64_t cond[2] = {c_wg, v_wg};
            encode(out, code("WG"), 0, 0, (uint32_t)op.size(), cond, 16);
        }
        out.insert(out.end(), op.begin(), op.end());
    }
    // register indices (pair ops: lower indices) satisfying (j & c) == v
    uint32_t jmask(uint32_t c, uint32_t v, int pair_bit) const {
        uint32_t m = 0;
        for (int j = 0; j < NR; j++) {
            if (pair_bit >= 0 && ((j >> pair_bit) & 1)) continue;
            if (((uint32_t)j & c) == v) m |= 1u << j;
        }
        return m;
    }
    // NEG / PHASE variant for the register condition (c, v)
    int cond_variant(const char *all, const char *b, const char *bb, const char *gen, uint32_t c, uint32_t v) const {
        const int pc = __builtin_popcount(c);
        if (pc == 0) return code(all);
        if (pc == 1) {
            const int r = __builtin_ctz(c);
            return code(b, r, (int)((v >> r) & 1));
        }
        if (pc == 2) {
            const int r1 = __builtin_ctz(c), r2 = 31 - __builtin_clz(c);
            return code(bb, r1, r2, (int)((v >> r1) & 1), (int)((v >> r2) & 1));
        }
        return code(gen);
    }
    // Value i of op o's coefficient array, in double.
    double val(const dq_op &o, int i) const {
        if (prec == DQ_F64) {
            double v;
            memcpy(&v, reinterpret_cast<const unsigned char *>(o.m) + 8 * i, 8);
            return v;
        }
        float v;
        memcpy(&v, reinterpret_cast<const unsigned char *>(o.m) + 4 * i, 4);
        return (double)v;
    }
    void put(std::vector<unsigned char> &v, double x) const {
        const size_t at = v.size();
        v.resize(at + ts);
        if (prec == DQ_F64) memcpy(v.data() + at, &x, 8);
        else {
            const float f = (float)x;
            memcpy(v.data() + at, &f, 4);
        }
    }
    // A table op (or, with shear slots, a table+shears op) for the diagonal
    // (tr, ti) over the register index: HALF when every factor that is not 1
    // has one register bit set, FULL when at least 3/4 of them are not 1, GEN
    // (a mask of the factors to apply) otherwise.  False if every factor is 1.
    bool encode_table(std::vector<unsigned char> &op, const std::vector<double> &tr, const std::vector<double> &ti,
                      const std::vector<unsigned char> *slots, uint32_t shear_mask) {
        std::vector<unsigned char> table;
        for (int e = 0; e < NR; e++) {
            put(table, tr[e]);
            put(table, ti[e]);
        }
        uint32_t ne1 = 0;  // factors that are not exactly 1 in the state precision
        for (int e = 0; e < NR; e++) {
            double r, i;
            if (prec == DQ_F64) {
                memcpy(&r, table.data() + e * 16, 8);
                memcpy(&i, table.data() + e * 16 + 8, 8);
            } else {
                float fr, fi;
                memcpy(&fr, table.data() + e * 8, 4);
                memcpy(&fi, table.data() + e * 8 + 4, 4);
                r = fr;
                i = fi;
            }
            if (!(r == 1.0 && i == 0.0)) ne1 |= 1u << e;
        }
        if (!ne1) return false;
        int half = -1;
        for (int r = 0; r < R && half < 0; r++)
            if ((ne1 & ~jmask(1u << r, 1u << r, -1)) == 0) half = r;
        std::vector<unsigned char> body;
        int c;
        uint32_t gen = 0;
        if (half >= 0) {
            for (int e = 0; e < NR; e++)
                if ((e >> half) & 1)
                    body.insert(body.end(), table.begin() + (size_t)e * 2 * ts, table.begin() + (size_t)(e + 1) * 2 * ts);
            c = slots ? code("TSH_HALF", half) : code("TABLE_HALF", half);
        } else if (4 * __builtin_popcount(ne1) >= 3 * NR) {
            body = table;
            c = slots ? code("TSH_FULL") : code("TABLE_FULL");
        } else {
            body = table;
            gen = ne1;
            c = slots ? code("TSH_GEN") : code("TABLE_GEN");
        }
        if (slots) {
            body.insert(body.end(), slots->begin(), slots->end());
            encode(op, c, shear_mask, gen, 0, body.data(), body.size());
        } else {
            encode(op, c, gen, 0, 0, body.data(), body.size());
        }
        return true;
    }
    // Diagonal ops commute, so every diagonal op without a lane or workgroup
    // condition in a run of diagonal ops that holds a table is multiplied into
    // that run's table; the table is emitted before the next non-diagonal op,
    // fused with it when that op is a run of unconditioned shears.
    // NEG / PHASE ops of such a run whose only condition is one workgroup bit
    // (controlled-Z or controlled phases with a qubit outside the tile) are
    // folded too, for up to kWgFoldBits distinct bits: the table is emitted
    // once per assignment of those bits, each copy guarded by its assignment
    // and holding the factors that apply there.  A workgroup runs exactly one
    // copy and skips the others, so the conditioned factors cost no vector
    // instruction.
    static constexpr int kWgFoldBits = 2;  // the program size bound (prog_bound) assumes at most 2
    struct WgDiag {
        uint64_t c_wg, v_wg;
        uint32_t c_reg, v_reg;
        double fr, fi;
    };
    void build(const dq_op *ops, uint32_t count, uint32_t &n_out) {
        n_out = 0;
        std::vector<unsigned char> op;
        std::vector<double> tr(NR, 1.0), ti(NR, 0.0);
        bool pending = false;  // tr/ti hold a table not yet emitted
        std::vector<WgDiag> wgd;  // workgroup-conditioned factors of the pending table
        uint64_t wg_bits = 0;     // their workgroup bits
        auto is_diag = [](const dq_op &o) {
            return o.kind == DQ_OP_NEG || o.kind == DQ_OP_PHASE || o.kind == DQ_OP_DIAGR;
        };
        auto mul_into = [&](std::vector<double> &xr, std::vector<double> &xi, uint32_t c, uint32_t v, double fr,
                            double fi) {
            for (int j = 0; j < NR; j++)
                if (((uint32_t)j & c) == v) {
                    const double r = xr[j] * fr - xi[j] * fi, i = xr[j] * fi + xi[j] * fr;
                    xr[j] = r;
                    xi[j] = i;
                }
        };
        auto mul = [&](uint32_t c, uint32_t v, double fr, double fi) { mul_into(tr, ti, c, v, fr, fi); };
        auto flush = [&](const std::vector<unsigned char> *slots, uint32_t shear_mask) {
            if (pending && wg_bits) {
                const int nb = __builtin_popcountll(wg_bits);
                for (uint32_t a = 0; a < (1u << nb); a++) {
                    uint64_t val = 0, m = wg_bits;  // assignment a deposited on the bits
                    for (int i = 0; m; i++, m &= m - 1)
                        if ((a >> i) & 1) val |= m & (~m + 1);
                    std::vector<double> vr = tr, vi = ti;
                    for (const WgDiag &d : wgd)
                        if ((val & d.c_wg) == d.v_wg) mul_into(vr, vi, d.c_reg, d.v_reg, d.fr, d.fi);
                    op.clear();
                    if (!encode_table(op, vr, vi, slots, shear_mask)) {
                        if (!slots) continue;  // nothing to do for these workgroups
                        op.clear();
                        encode(op, code("SHEARS"), shear_mask, 0, 0, slots->data(), slots->size());
                    }
                    emit(op, wg_bits, val);
                    n_out++;
                }
            } else {
                op.clear();
                const bool any = pending && encode_table(op, tr, ti, slots, shear_mask);
                if (any || slots) {
                    if (!any) {
                        op.clear();
                        encode(op, code("SHEARS"), shear_mask, 0, 0, slots->data(), slots->size());
                    }
                    emit(op, 0, 0);
                    n_out++;
                }
            }
            std::fill(tr.begin(), tr.end(), 1.0);
            std::fill(ti.begin(), ti.end(), 0.0);
            pending = false;
            wgd.clear();
            wg_bits = 0;
        };
        // a NEG / PHASE with one workgroup-bit condition, folded into the pending table
        auto fold_wg = [&](const dq_op &o, double fr, double fi) {
            if (o.c_thr || __builtin_popcountll(o.c_wg) != 1) return false;
            if (!(wg_bits & o.c_wg) && __builtin_popcountll(wg_bits) >= kWgFoldBits) return false;
            wgd.push_back({o.c_wg, o.v_wg, o.c_reg, o.v_reg, fr, fi});
            wg_bits |= o.c_wg;
            pending = true;
            return true;
        };
        bool fold_run = false;  // the current run of diagonal ops holds an unconditioned table
        bool in_run = false;
        for (uint32_t k = 0; k < count; k++) {
            const dq_op &o = ops[k];
            const unsigned char *m = reinterpret_cast<const unsigned char *>(o.m);
            if (is_diag(o) && !in_run) {
                fold_run = false;
                for (uint32_t k2 = k; k2 < count && is_diag(ops[k2]); k2 += ops[k2].kind == DQ_OP_DIAGR ? 1 + ops[k2].tbit : 1)
                    if (ops[k2].kind == DQ_OP_DIAGR && !ops[k2].c_wg) fold_run = true;
            }
            in_run = is_diag(o);
            op.clear();
            switch (o.kind) {
                case DQ_OP_SHEAR: {
                    // merge the run of unconditioned shears on distinct bits
                    std::vector<unsigned char> slots(2 * (size_t)R * ts, 0);
                    uint32_t mask = 0;
                    uint32_t k2 = k;
                    while (k2 < count && ops[k2].kind == DQ_OP_SHEAR && !ops[k2].c_wg && !((mask >> ops[k2].tbit) & 1)) {
                        mask |= 1u << ops[k2].tbit;
                        memcpy(slots.data() + 2 * ops[k2].tbit * ts, ops[k2].m, 2 * ts);
                        k2++;
                    }
                    if (k2 == k) {  // a shear with a workgroup condition: on its own
                        flush(nullptr, 0);
                        op.clear();
                        mask = 1u << o.tbit;
                        memcpy(slots.data() + 2 * o.tbit * ts, o.m, 2 * ts);
                        encode(op, code("SHEARS"), mask, 0, 0, slots.data(), slots.size());
                        emit(op, o.c_wg, o.v_wg);
                        n_out++;
                        continue;
                    }
                    flush(&slots, mask);
                    k = k2 - 1;
                    continue;
                }
                case DQ_OP_DIAGR: {
                    if (!o.c_wg) {
                        for (int e = 0; e < NR; e++) {
                            const dq_op &d = ops[k + 1 + e / 4];
                            mul((uint32_t)(NR - 1), (uint32_t)e, val(d, 2 * (e % 4)), val(d, 2 * (e % 4) + 1));
                        }
                        pending = true;
                    } else {  // a table under a workgroup condition: its own op (it commutes with the pending one)
                        std::vector<double> wr(NR), wi(NR);
                        for (int e = 0; e < NR; e++) {
                            const dq_op &d = ops[k + 1 + e / 4];
                            wr[e] = val(d, 2 * (e % 4));
                            wi[e] = val(d, 2 * (e % 4) + 1);
                        }
                        if (encode_table(op, wr, wi, nullptr, 0)) {
                            emit(op, o.c_wg, o.v_wg);
                            n_out++;
                        }
                    }
                    k += o.tbit;
                    continue;
                }
                case DQ_OP_NEG:
                    if (fold_run && !o.c_thr && !o.c_wg) {
                        mul(o.c_reg, o.v_reg, -1.0, 0.0);
                        pending = true;
                        continue;
                    }
                    if (fold_run && o.c_wg && fold_wg(o, -1.0, 0.0)) continue;
                    encode(op, cond_variant("NEG_ALL", "NEG_B", "NEG_BB", "NEG_GEN", o.c_reg, o.v_reg),
                           jmask(o.c_reg, o.v_reg, -1), o.c_thr, o.v_thr, nullptr, 0);
                    break;
                case DQ_OP_PHASE:
                    if (fold_run && !o.c_thr && !o.c_wg) {
                        mul(o.c_reg, o.v_reg, val(o, 0), val(o, 1));
                        pending = true;
                        continue;
                    }
                    if (fold_run && o.c_wg && fold_wg(o, val(o, 0), val(o, 1))) continue;
                    encode(op, cond_variant("PHASE_ALL", "PHASE_B", "PHASE_BB", "PHASE_GEN", o.c_reg, o.v_reg),
                           jmask(o.c_reg, o.v_reg, -1), o.c_thr, o.v_thr, m, 2 * ts);
                    break;
                case DQ_OP_MAT:
                    flush(nullptr, 0);
                    op.clear();
                    encode(op, code("MAT", (int)o.tbit), jmask(o.c_reg, o.v_reg, (int)o.tbit), o.c_thr, o.v_thr, m, 8 * ts);
                    break;
                case DQ_OP_SWAP:
                    flush(nullptr, 0);
                    op.clear();
                    encode(op, code("SWAP", (int)o.tbit), jmask(o.c_reg, o.v_reg, (int)o.tbit), o.c_thr, o.v_thr, nullptr, 0);
                    break;
                case DQ_OP_RMAT:
                    flush(nullptr, 0);
                    op.clear();
                    encode(op, code("RMAT", (int)o.tbit), 0, 0, 0, m, 4 * ts);
                    break;
                default:
                    continue;
            }
            emit(op, o.c_wg, o.v_wg);
            n_out++;
        }
        flush(nullptr, 0);
        op.clear();
        encode(op, code("END"), 0, 0, 0, nullptr, 0);
        out.insert(out.end(), op.begin(), op.end());
    }
    // Walk a built program: every code in range, sizes chaining to END at the
    // end of the buffer (the interpreter jumps through a table on the code).
    bool check(size_t begin) const {
        size_t at = begin;
        while (at + 16 <= out.size()) {
            uint32_t c;
            memcpy(&c, out.data() + at, 4);
            if (c >= (uint32_t)ncodes) return false;
            if (!strcmp(tab[c].name, "END")) return at + size(c) == out.size();
            at += size(c);
        }
        return false;
    }
};
}  // namespace

// The device program of one phase's ops, built on the host as
// dq_apply_blocks builds it (no GPU needed): lets the program encoding be
// checked against the op semantics on any host.
extern "C" int dq_build_program(int prec, int reg_bits, const dq_op *ops, uint32_t n_ops, unsigned char *out,
                                size_t cap, size_t *len) {
    if ((prec != DQ_F32 && prec != DQ_F64) || reg_bits < 1 || reg_bits > DQ_MAX_REG_BITS || !len || (n_ops && !ops))
        return fail(DQ_ERR_ARG, "dq_build_program: bad arguments");
    std::vector<unsigned char> prog;
    ProgBuilder pb(prec, reg_bits, prog);
    uint32_t n_out = 0;
    pb.build(ops, n_ops, n_out);
    if (pb.bad || !pb.check(0)) return fail(DQ_ERR_ARG, "dq_build_program: op not expressible as a device op");
    *len = prog.size();
    if (out) {
        if (cap < prog.size()) return fail(DQ_ERR_ARG, "dq_build_program: %zu bytes needed", prog.size());
        memcpy(out, prog.data(), prog.size());
    }
    return DQ_OK;
}

extern "C" int dq_init_state(void *state, int n, int prec, int kind, uint64_t index, double value, void *stream) {
    if (!state || n < 1 || n > 40 || (prec != DQ_F32 && prec != DQ_F64)) return fail(DQ_ERR_ARG, "dq_init_state: bad arguments");
    if (kind != DQ_INIT_CLASSICAL && kind != DQ_INIT_PLUS) return fail(DQ_ERR_ARG, "dq_init_state: bad kind %d", kind);
    const uint64_t N = 1ull << n;
    const uint64_t want = (N + 255) / 256;
    const unsigned blocks = (unsigned)(want < (1ull << 20) ? want : (1ull << 20));
    if (prec == DQ_F64)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (V2<double>::type *)state, N, kind, index, value);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (V2<float>::type *)state, N, kind, index, value);
    HIPC(hipGetLastError());
    return DQ_OK;
}

static bool onehot64(uint64_t x) { return x && !(x & (x - 1)); }

// Host-side check of every descriptor against the grid and buffers it
// implies: each workgroup reads and writes exactly the 2^T amplitudes of its
// own tile, every LDS index is below 2^T, every op index is in range.
static int validate_block(const dq_block &b, const dq_phase *phases, uint32_t n_phases, const dq_op *ops, uint32_t n_ops) {
    const int n = (int)b.n_qubits;
    if (n < 1 || n > 40) return fail(DQ_ERR_ARG, "block: n_qubits %d out of range", n);
    if (b.phase_count < 1 || b.phase_begin >= n_phases || b.phase_count > n_phases - b.phase_begin)
        return fail(DQ_ERR_ARG, "block: phase range out of bounds");
    for (uint32_t p = 0; p < b.phase_count; p++) {
        const dq_phase &ph = phases[b.phase_begin + p];
        if (ph.op_begin > n_ops || ph.op_count > n_ops - ph.op_begin) return fail(DQ_ERR_ARG, "phase: op range out of bounds");
        for (uint32_t k = 0; k < ph.op_count; k++) {
            const dq_op &o = ops[ph.op_begin + k];
            if (o.c_thr > 0xFFFFu || o.v_thr > 0xFFFFu || (o.v_thr & ~o.c_thr) || (o.v_reg & ~o.c_reg) ||
                (o.v_wg & ~o.c_wg))
                return fail(DQ_ERR_ARG, "op: condition values outside their masks");
            if ((o.kind == DQ_OP_SHEAR || o.kind == DQ_OP_RMAT || o.kind == DQ_OP_DIAGR) && (o.c_reg || o.c_thr))
                return fail(DQ_ERR_ARG, "op: kind %u takes no register or lane condition", o.kind);
            if (o.kind == DQ_OP_DIAGR) {
                if (b.reg_bits == 0 || o.tbit != ((1u << b.reg_bits) + 3) / 4 || k + o.tbit >= ph.op_count)
                    return fail(DQ_ERR_ARG, "phase: register-diagonal table out of bounds");
                k += o.tbit;
            } else if (o.kind == DQ_OP_MAT || o.kind == DQ_OP_SWAP || o.kind == DQ_OP_RMAT || o.kind == DQ_OP_SHEAR) {
                if (b.reg_bits ? o.tbit >= b.reg_bits : o.tbit >= b.n_qubits)
                    return fail(DQ_ERR_ARG, "op: target bit %u out of range", o.tbit);
            } else if (o.kind != DQ_OP_NEG && o.kind != DQ_OP_PHASE) {
                return fail(DQ_ERR_ARG, "op: unknown kind %u", o.kind);
            }
        }
    }
    if (b.reg_bits == 0) {  // small-state form
        if ((int)b.tile_bits != n || n > 13) return fail(DQ_ERR_ARG, "small block: tile_bits must equal n <= 13");
        return DQ_OK;
    }
    const int T = (int)b.tile_bits, R = (int)b.reg_bits, W = (int)b.wave_bits;
    if (R > DQ_MAX_REG_BITS || W > DQ_MAX_WAVE_BITS || T != 6 + R + W || T > n || T > DQ_MAX_TILE_BITS)
        return fail(DQ_ERR_ARG, "block: tile shape T=%d R=%d W=%d n=%d", T, R, W, n);
    const uint64_t full = (1ull << n) - 1;
    for (int end = 0; end < 2; end++) {
        const uint64_t *rg = end ? b.st_reg : b.ld_reg;
        const uint64_t *wv = end ? b.st_wave : b.ld_wave;
        uint64_t m = 0x3Full;
        int cnt = 6;
        for (int r = 0; r < R; r++) { if (!onehot64(rg[r])) return fail(DQ_ERR_ARG, "block: register mask not one-hot"); m |= rg[r]; cnt++; }
        for (int w = 0; w < W; w++) { if (!onehot64(wv[w])) return fail(DQ_ERR_ARG, "block: wave mask not one-hot"); m |= wv[w]; cnt++; }
        if (__builtin_popcountll(m) != cnt || (m & b.nontile_mask) || ((m | b.nontile_mask) != full))
            return fail(DQ_ERR_ARG, "block: tile and non-tile qubit masks inconsistent");
    }
    for (uint32_t p = 0; p < b.phase_count; p++) {
        const dq_phase &ph = phases[b.phase_begin + p];
        uint32_t m = 0;
        int cnt = 0;
        auto add = [&](uint32_t x) { m |= x; cnt++; return x && !(x & (x - 1)) && x < (1u << T); };
        bool ok = true;
        for (int r = 0; r < R; r++) ok &= add(ph.reg_t[r]);
        for (int w = 0; w < W; w++) ok &= add(ph.wave_t[w]);
        for (int l = 0; l < 6; l++) ok &= add(ph.lane_t[l]);
        if (!ok || m != (1u << T) - 1 || __builtin_popcount(m) != cnt)
            return fail(DQ_ERR_ARG, "phase %u: layout is not a permutation of the tile bits", p);
    }
    return DQ_OK;
}

extern "C" int dq_apply_blocks(int prec, const dq_block *blocks, uint32_t n_blocks, const dq_phase *phases,
                               uint32_t n_phases, const dq_op *ops, uint32_t n_ops, const dq_job *jobs,
                               const uint32_t *job_begin, uint32_t n_launch, void *d_ws, size_t ws_bytes, void *stream) {
    if (prec != DQ_F32 && prec != DQ_F64) return fail(DQ_ERR_ARG, "dq_apply_blocks: bad precision %d", prec);
    if (n_launch == 0) return DQ_OK;
    if (!blocks || !phases || !jobs || !job_begin || !d_ws) return fail(DQ_ERR_ARG, "dq_apply_blocks: null argument");
    const uint32_t n_jobs = job_begin[n_launch];
    const size_t need = dq_apply_ws_bytes(n_blocks, n_phases, n_ops, n_jobs);
    if (need > ws_bytes) return fail(DQ_ERR_ARG, "dq_apply_blocks: workspace %zu < %zu bytes", ws_bytes, need);
    for (uint32_t i = 0; i < n_blocks; i++) {
        int rc = validate_block(blocks[i], phases, n_phases, ops, n_ops);
        if (rc) return rc;
    }
    for (uint32_t j = 0; j < n_jobs; j++)
        if (jobs[j].block >= n_blocks || !jobs[j].state) return fail(DQ_ERR_ARG, "job %u: bad block or state", j);
    for (uint32_t j = 0; j < n_jobs; j++)
        if (jobs[j].base_hi & ((1ull << blocks[jobs[j].block].n_qubits) - 1))
            return fail(DQ_ERR_ARG, "job %u: base_hi overlaps the piece's own qubits", j);
    for (uint32_t l = 0; l < n_launch; l++) {
        if (job_begin[l + 1] <= job_begin[l]) return fail(DQ_ERR_ARG, "launch %u: empty job range", l);
        const dq_block &b0 = blocks[jobs[job_begin[l]].block];
        for (uint32_t j = job_begin[l]; j < job_begin[l + 1]; j++) {
            const dq_block &b = blocks[jobs[j].block];
            if (b.n_qubits != b0.n_qubits || b.reg_bits != b0.reg_bits || b.wave_bits != b0.wave_bits || b.tile_bits != b0.tile_bits)
                return fail(DQ_ERR_ARG, "launch %u: jobs differ in shape", l);
        }
        if (b0.reg_bits && !lookup(prec, b0.reg_bits, b0.wave_bits,
                                   exchange_mode(prec, (int)b0.reg_bits, (int)b0.wave_bits, (int)b0.tile_bits)))
            return fail(DQ_ERR_UNSUPPORTED, "no sweep kernel for R=%d W=%d", b0.reg_bits, b0.wave_bits);
    }

    // per phase of a sweep block: its device program and exchange offsets
    // (phase_aux); per block: the runs of its non-tile qubits (block_aux)
    std::vector<uint32_t> paux((size_t)n_phases * DQ_PAUX, 0), baux((size_t)n_blocks * DQ_BAUX, 0);
    std::vector<int> prog_r(n_phases, -1);
    std::vector<uint32_t> prog_es(n_phases, 0);
    std::vector<unsigned char> prog;
    for (uint32_t i = 0; i < n_blocks; i++) {
        const dq_block &b = blocks[i];
        if (b.reg_bits == 0) continue;
        if (!block_runs(b, baux.data() + (size_t)i * DQ_BAUX))
            return fail(DQ_ERR_UNSUPPORTED, "block %u: non-tile qubits in more than %d runs", i, DQ_RUNS);
        const int bxm = exchange_mode(prec, (int)b.reg_bits, (int)b.wave_bits, (int)b.tile_bits);
        const uint32_t es = exchange_elem_bytes(prec, bxm);
        uint32_t gown[32];
        if (bxm == 1) {
            if (b.tile_bits != 6 + b.wave_bits + b.reg_bits)
                return fail(DQ_ERR_ARG, "block %u: tile of %u bits is not 6 + %u + %u", i, b.tile_bits, b.wave_bits, b.reg_bits);
            const int Lo = es == 4 ? 5 : (es == 8 ? 4 : 3);
            pick_owned_swizzle(phases + b.phase_begin, b.phase_count, (int)b.reg_bits, (int)b.wave_bits, Lo, gown);
        }
        for (uint32_t p = b.phase_begin; p < b.phase_begin + b.phase_count; p++) {
            uint32_t *pa = paux.data() + (size_t)p * DQ_PAUX;
            if (prog_r[p] != (int)b.reg_bits || prog_es[p] != es) {
                if (prog_r[p] >= 0) return fail(DQ_ERR_ARG, "phase %u shared by blocks of different tile shapes", p);
                prog_r[p] = (int)b.reg_bits;
                prog_es[p] = es;
                pa[0] = (uint32_t)prog.size();
                ProgBuilder pb(prec, (int)b.reg_bits, prog);
                const size_t start = prog.size();
                pb.build(ops + phases[p].op_begin, phases[p].op_count, pa[1]);
                if (pb.bad || !pb.check(start)) return fail(DQ_ERR_ARG, "phase %u: op not expressible as a device op", p);
            }
            if (p == b.phase_begin) continue;  // no exchange into a block's first phase
            // the exchange from phase p - 1: its swizzle and offset tables
            const dq_phase &A = phases[p - 1], &B = phases[p];
            const int L = es == 4 ? 5 : (es == 8 ? 4 : 3);
            if (bxm == 1) {
                uint32_t t[DQ_PAUX];
                memcpy(t, pa, sizeof t);
                if (!owned_tables(A, B, (int)b.reg_bits, (int)b.wave_bits, L, gown, es, t))
                    return fail(DQ_ERR_ARG, "phase %u: layouts of the exchange do not place the same tile bits", p);
                // a phase shared with an earlier block must get the same tables
                bool filled = false;
                for (int k = 2; k < DQ_PAUX; k++) filled |= pa[k] != 0;
                if (filled && memcmp(t + 2, pa + 2, (DQ_PAUX - 2) * sizeof(uint32_t)))
                    return fail(DQ_ERR_ARG, "phase %u: shared by blocks with different exchange slots", p);
                memcpy(pa, t, sizeof t);
                continue;
            }
            uint32_t g[32];
            pick_swizzle(A, B, L, g);
            for (uint32_t j = 0; j < (1u << b.reg_bits); j++) {
                uint32_t xa = 0, xb = 0;
                for (uint32_t r = 0; r < b.reg_bits; r++)
                    if ((j >> r) & 1) {
                        xa ^= swizzle_host(A.reg_t[r], g, L) * es;
                        xb ^= swizzle_host(B.reg_t[r], g, L) * es;
                    }
                pa[DQ_XA + j] = xa;
                pa[DQ_XB + j] = xb;
            }
            for (int l = 0; l < 6; l++) {
                pa[DQ_TA + l] = swizzle_host(A.lane_t[l], g, L) * es;
                pa[DQ_TB + l] = swizzle_host(B.lane_t[l], g, L) * es;
            }
            for (uint32_t w = 0; w < b.wave_bits; w++) {
                pa[DQ_TA + 6 + w] = swizzle_host(A.wave_t[w], g, L) * es;
                pa[DQ_TB + 6 + w] = swizzle_host(B.wave_t[w], g, L) * es;
            }
        }
    }
    if (prog.size() + 256 > prog_bound(n_phases, n_ops)) return fail(DQ_ERR_ARG, "dq_apply_blocks: device program overflow");

    // pack and upload
    std::vector<unsigned char> buf(need, 0);
    size_t off_b = 0, off_p = align_up(n_blocks * sizeof(dq_block), 256);
    size_t off_o = off_p + align_up(n_phases * sizeof(dq_phase), 256);
    size_t off_j = off_o + align_up(n_ops * sizeof(dq_op), 256);
    size_t off_pa = off_j + align_up(n_jobs * sizeof(dq_job), 256);
    size_t off_ba = off_pa + align_up((size_t)n_phases * DQ_PAUX * sizeof(uint32_t), 256);
    size_t off_pg = off_ba + align_up((size_t)n_blocks * DQ_BAUX * sizeof(uint32_t), 256);
    memcpy(buf.data() + off_b, blocks, n_blocks * sizeof(dq_block));
    memcpy(buf.data() + off_p, phases, n_phases * sizeof(dq_phase));
    if (n_ops) memcpy(buf.data() + off_o, ops, n_ops * sizeof(dq_op));
    memcpy(buf.data() + off_j, jobs, n_jobs * sizeof(dq_job));
    if (n_phases) memcpy(buf.data() + off_pa, paux.data(), paux.size() * sizeof(uint32_t));
    memcpy(buf.data() + off_ba, baux.data(), baux.size() * sizeof(uint32_t));
    if (!prog.empty()) memcpy(buf.data() + off_pg, prog.data(), prog.size());
    buf.resize(off_pg + prog.size() + 256, 0);  // upload only what is used (+ padding)
    int rc = stage_upload(d_ws, buf, stream);
    if (rc) return rc;
    unsigned char *d = (unsigned char *)d_ws;
    const dq_block *d_blocks = (const dq_block *)(d + off_b);
    const dq_phase *d_phases = (const dq_phase *)(d + off_p);
    const dq_op *d_ops = (const dq_op *)(d + off_o);
    const dq_job *d_jobs = (const dq_job *)(d + off_j);
    const uint32_t *d_paux = (const uint32_t *)(d + off_pa);
    const uint32_t *d_baux = (const uint32_t *)(d + off_ba);
    const unsigned char *d_prog = d + off_pg;

    for (uint32_t l = 0; l < n_launch; l++) {
        const uint32_t j0 = job_begin[l], nj = job_begin[l + 1] - j0;
        const dq_block &b0 = blocks[jobs[j0].block];
        const int n = b0.n_qubits;
        if (b0.reg_bits == 0) {
            const size_t lds = ((size_t)1 << n) * elem_bytes(prec);
            if (prec == DQ_F64)
                hipLaunchKernelGGL(small_kernel<double>, dim3(1, nj), dim3(256), lds, (hipStream_t)stream, d_jobs + j0, d_blocks, d_phases, d_ops);
            else
                hipLaunchKernelGGL(small_kernel<float>, dim3(1, nj), dim3(256), lds, (hipStream_t)stream, d_jobs + j0, d_blocks, d_phases, d_ops);
        } else {
            bool multi = false;
            for (uint32_t j = j0; j < j0 + nj; j++) multi |= blocks[jobs[j].block].phase_count > 1;
            const int xm = exchange_mode(prec, (int)b0.reg_bits, (int)b0.wave_bits, (int)b0.tile_bits);
            const size_t lds = multi ? ((size_t)1 << b0.tile_bits) * exchange_elem_bytes(prec, xm) : 0;
            if (lds > 131072) return fail(DQ_ERR_UNSUPPORTED, "tile of %d bits needs %zu bytes of LDS", b0.tile_bits, lds);
            const uint64_t tiles = 1ull << (n - b0.tile_bits);
            if (tiles > 0x7fffffffull || nj > 65535 || ((tiles << (6 + b0.wave_bits)) >> 32))
                return fail(DQ_ERR_UNSUPPORTED, "grid too large (at most 2^32 threads per launch)");
            sweep_fn fn = lookup(prec, b0.reg_bits, b0.wave_bits, xm);
            if (!fn) return fail(DQ_ERR_UNSUPPORTED, "no sweep kernel for R=%u W=%u exchange mode %d", b0.reg_bits, b0.wave_bits, xm);
            hipLaunchKernelGGL(fn, dim3((unsigned)tiles, nj), dim3(64u << b0.wave_bits), lds, (hipStream_t)stream,
                               d_jobs + j0, d_blocks, d_phases, d_prog, d_paux, d_baux);
        }
        HIPC(hipGetLastError());
    }
    return DQ_OK;
}

// bytes of per-workgroup partial sums of prob_stage1
static size_t prob_part_bytes(int n) {
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    const uint64_t groups = 1ull << (n - cb);
    return align_up(groups * PROB_SLOTS * sizeof(double), 256);
}

extern "C" size_t dq_prob_ws_bytes(int n) { return prob_part_bytes(n) + 64 * sizeof(double); }

extern "C" int dq_prob_zero_all_async(const void *state, int n, int prec, double *d_out, void *d_ws, size_t ws_bytes, void *stream) {
    if (!state || !d_out || !d_ws || n < 1 || n > 40 || (prec != DQ_F32 && prec != DQ_F64))
        return fail(DQ_ERR_ARG, "dq_prob_zero_all: bad arguments");
    if (ws_bytes < prob_part_bytes(n)) return fail(DQ_ERR_ARG, "dq_prob_zero_all_async: workspace too small");
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    const uint64_t groups = 1ull << (n - cb);
    double *part = (double *)d_ws;
    if (prec == DQ_F64)
        hipLaunchKernelGGL(prob_stage1<double>, dim3((unsigned)groups), dim3(256), 0, (hipStream_t)stream,
                           (const V2<double>::type *)state, n, part);
    else
        hipLaunchKernelGGL(prob_stage1<float>, dim3((unsigned)groups), dim3(256), 0, (hipStream_t)stream,
                           (const V2<float>::type *)state, n, part);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(prob_stage2, dim3(n + 1), dim3(256), 0, (hipStream_t)stream, part, (uint32_t)groups, n, d_out);
    HIPC(hipGetLastError());
    return DQ_OK;
}

extern "C" int dq_prob_zero_all(const void *state, int n, int prec, double *out, void *d_ws, size_t ws_bytes, void *stream) {
    if (!out) return fail(DQ_ERR_ARG, "dq_prob_zero_all: null output");
    if (ws_bytes < dq_prob_ws_bytes(n)) return fail(DQ_ERR_ARG, "dq_prob_zero_all: workspace too small");
    const size_t part_bytes = prob_part_bytes(n);
    double *d_out = (double *)((unsigned char *)d_ws + part_bytes);
    int rc = dq_prob_zero_all_async(state, n, prec, d_out, d_ws, part_bytes, stream);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(out, d_out, (n + 1) * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPC(hipStreamSynchronize((hipStream_t)stream));
    return DQ_OK;
}

// ---- P(0) of every qubit without allocation or workspace --------------------
// For callers that may not allocate device memory (code under stream capture,
// serving loops): the partial sums live in a file-scope device array.  A
// state is reduced in chunks of at most 2^NA_CHUNK_BITS amplitudes, each by
// prob_stage1/prob_stage2 (for n <= NA_CHUNK_BITS exactly the launches of
// dq_prob_zero_all_async); prob_combine sums the chunks' results.  The
// array `scratch` (defined with the P(0) kernels) holds the partial sums of
// one chunk, then the results of every chunk.

// out[q] for q < n and the total out[n] from nchunks chunk results of
// NA_CHUNK_BITS + 1 doubles each (chunk c = amplitudes with index bits
// NA_CHUNK_BITS.. equal to c); one thread per output, chunks summed in order.
__global__ void __launch_bounds__(64) prob_combine(const double *__restrict__ chunks, uint32_t nchunks, int n,
                                                   double *__restrict__ out) {
    const int q = threadIdx.x;
    if (q > n) return;
    constexpr int S = NA_CHUNK_BITS + 1;
    double acc = 0.0;
    for (uint32_t c = 0; c < nchunks; c++) {
        if (q < NA_CHUNK_BITS) acc += chunks[(uint64_t)c * S + q];
        else if (q == n || !((c >> (q - NA_CHUNK_BITS)) & 1u)) acc += chunks[(uint64_t)c * S + NA_CHUNK_BITS];
    }
    out[q] = acc;
}

extern "C" int dq_prob_zero_all_noalloc(const void *state, int n, int prec, double *d_out, void *stream) {
    if (!state || !d_out || n < 1 || (prec != DQ_F32 && prec != DQ_F64))
        return fail(DQ_ERR_ARG, "dq_prob_zero_all_noalloc: bad arguments");
    if (n > NA_MAX_QUBITS)
        return fail(DQ_ERR_ARG, "dq_prob_zero_all_noalloc: %d qubits, above its limit of %d", n, NA_MAX_QUBITS);
    int dev = -1;
    HIPC(hipGetDevice(&dev));
    double *scr = (dev >= 0 && dev < kMaxDevices) ? g_scratch_dev[dev] : nullptr;
    if (!scr) return fail(DQ_ERR_ARG, "dq_prob_zero_all_noalloc: dq_device_init(%d) has not been called", dev);
    if (n <= NA_CHUNK_BITS)
        return dq_prob_zero_all_async(state, n, prec, d_out, scr, (size_t)NA_PART_DOUBLES * sizeof(double), stream);
    double *chunk_res = scr + NA_PART_DOUBLES;
    const uint32_t nchunks = 1u << (n - NA_CHUNK_BITS);
    const size_t amp_bytes = prec == DQ_F64 ? 16 : 8;
    for (uint32_t c = 0; c < nchunks; c++) {
        const void *piece = (const unsigned char *)state + ((size_t)c << NA_CHUNK_BITS) * amp_bytes;
        const int rc = dq_prob_zero_all_async(piece, NA_CHUNK_BITS, prec, chunk_res + (size_t)c * (NA_CHUNK_BITS + 1), scr,
                                              (size_t)NA_PART_DOUBLES * sizeof(double), stream);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(prob_combine, dim3(1), dim3(64), 0, (hipStream_t)stream, chunk_res, nchunks, n, d_out);
    HIPC(hipGetLastError());
    return DQ_OK;
}

// ---- read-out of many states ------------------------------------------------
namespace {
struct RoPlan {
    int gb = 0;
    uint64_t wgs = 0;  // stage-1 workgroups per state
    RoLevels lv;
    size_t off_ptr = 0, off_u = 0, off_part = 0, off_lvl[RO_MAX_LEVELS] = {0, 0, 0, 0}, total = 0;
};

RoPlan ro_plan(int n, uint32_t count, uint32_t shots) {
    RoPlan p;
    memset(&p.lv, 0, sizeof p.lv);
    p.gb = n < RO_GB ? n : RO_GB;
    const uint64_t n1 = 1ull << (n - p.gb);
    p.wgs = (n1 + 4 * RO_GPW - 1) / (4 * RO_GPW);
    size_t off = 0;
    p.off_ptr = off;
    off += align_up((size_t)count * sizeof(void *), 256);
    p.off_u = off;
    off += align_up((size_t)count * shots * sizeof(double), 256);
    p.off_part = off;
    off += align_up((size_t)count * p.wgs * PROB_SLOTS * sizeof(double), 256);
    uint64_t c = n1;
    int L = 0;
    for (;;) {
        p.lv.count[L] = c;
        p.lv.stride[L] = c;
        p.off_lvl[L] = off;
        off += align_up((size_t)count * c * sizeof(double), 256);
        L++;
        if (c <= RO_FAN) break;
        c = (c + RO_FAN - 1) >> RO_FAN_BITS;
    }
    p.lv.nlev = L;
    p.total = off;
    return p;
}
}  // namespace

extern "C" size_t dq_readout_ws_bytes(int n, uint32_t count, uint32_t shots) {
    if (n < 1 || n > 40 || count == 0) return 0;
    return ro_plan(n, count, shots).total;
}

extern "C" int dq_readout_many(int prec, int n, const void *const *states, uint32_t count, const double *u,
                               uint32_t shots, double *d_p0, uint64_t *d_shots, void *d_ws, size_t ws_bytes,
                               void *stream) {
    if ((prec != DQ_F32 && prec != DQ_F64) || n < 1 || n > 40 || !states || count == 0 || count > 65535)
        return fail(DQ_ERR_ARG, "dq_readout_many: bad arguments");
    if (shots > (1u << 24)) return fail(DQ_ERR_ARG, "dq_readout_many: at most 2^24 samples per state");
    if (shots && (!u || !d_shots)) return fail(DQ_ERR_ARG, "dq_readout_many: samples need u and d_shots");
    for (uint32_t i = 0; i < count; i++)
        if (!states[i]) return fail(DQ_ERR_ARG, "dq_readout_many: state %u is null", i);
    for (uint64_t i = 0; i < (uint64_t)count * shots; i++)
        if (!(u[i] >= 0.0 && u[i] < 1.0)) return fail(DQ_ERR_ARG, "dq_readout_many: u[%llu] not in [0, 1)", (unsigned long long)i);
    const RoPlan p = ro_plan(n, count, shots);
    if (!d_ws || ws_bytes < p.total) return fail(DQ_ERR_ARG, "dq_readout_many: workspace %zu < %zu bytes", ws_bytes, p.total);
    if (p.wgs > 0x7fffffffull) return fail(DQ_ERR_UNSUPPORTED, "dq_readout_many: grid too large");
    // state pointers and uniforms travel in one staged copy
    std::vector<unsigned char> buf(p.off_part - p.off_ptr, 0);
    memcpy(buf.data(), states, (size_t)count * sizeof(void *));
    if (shots) memcpy(buf.data() + (p.off_u - p.off_ptr), u, (size_t)count * shots * sizeof(double));
    unsigned char *d = (unsigned char *)d_ws;
    int rc = stage_upload(d + p.off_ptr, buf, stream);
    if (rc) return rc;
    const void *const *d_states = (const void *const *)(d + p.off_ptr);
    double *part = (double *)(d + p.off_part);
    RoLevels lv = p.lv;
    for (int L = 0; L < lv.nlev; L++) lv.base[L] = (double *)(d + p.off_lvl[L]);
    const uint64_t part_stride = p.wgs * PROB_SLOTS;
    hipStream_t s = (hipStream_t)stream;
    if (prec == DQ_F64)
        hipLaunchKernelGGL(readout_stage1<double>, dim3((unsigned)p.wgs, count), dim3(256), 0, s, d_states, n, part,
                           part_stride, lv.base[0], lv.stride[0]);
    else
        hipLaunchKernelGGL(readout_stage1<float>, dim3((unsigned)p.wgs, count), dim3(256), 0, s, d_states, n, part,
                           part_stride, lv.base[0], lv.stride[0]);
    HIPC(hipGetLastError());
    if (d_p0) {
        hipLaunchKernelGGL(readout_stage2, dim3(n + 1, count), dim3(256), 0, s, part, part_stride, (uint32_t)p.wgs, n, d_p0);
        HIPC(hipGetLastError());
    }
    if (!shots) return DQ_OK;
    for (int L = 1; L < lv.nlev; L++) {
        hipLaunchKernelGGL(readout_level, dim3((unsigned)((lv.count[L] + 3) / 4), count), dim3(256), 0, s,
                           lv.base[L - 1], lv.count[L - 1], lv.stride[L - 1], lv.base[L], lv.count[L], lv.stride[L]);
        HIPC(hipGetLastError());
    }
    const double *d_u = (const double *)(d + p.off_u);
    if (prec == DQ_F64)
        hipLaunchKernelGGL(readout_sample<double>, dim3((shots + 3) / 4, count), dim3(256), 0, s, d_states, n, lv, d_u,
                           shots, d_shots);
    else
        hipLaunchKernelGGL(readout_sample<float>, dim3((shots + 3) / 4, count), dim3(256), 0, s, d_states, n, lv, d_u,
                           shots, d_shots);
    HIPC(hipGetLastError());
    return DQ_OK;
}

// ---- comparison of two states -------------------------------------------------
extern "C" size_t dq_compare_ws_bytes(int n) {
    if (n < 1 || n > 40) return 0;
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    return align_up(((size_t)1 << (n - cb)) * CMP_SLOTS * sizeof(double), 256);
}

template <typename TA, typename TB>
static void launch_compare(const void *a, const void *b, int n, unsigned groups, double *part, hipStream_t s) {
    hipLaunchKernelGGL((compare_stage1<TA, TB>), dim3(groups), dim3(256), 0, s, (const typename V2<TA>::type *)a,
                       (const typename V2<TB>::type *)b, n, part);
}

extern "C" int dq_compare_states(const void *a, int prec_a, const void *b, int prec_b, int n, double *d_out,
                                 void *d_ws, size_t ws_bytes, void *stream) {
    const bool pa = prec_a == DQ_F32 || prec_a == DQ_F64, pb = prec_b == DQ_F32 || prec_b == DQ_F64;
    if (!a || !b || !d_out || !pa || !pb || n < 1 || n > 40) return fail(DQ_ERR_ARG, "dq_compare_states: bad arguments");
    if (!d_ws || ws_bytes < dq_compare_ws_bytes(n)) return fail(DQ_ERR_ARG, "dq_compare_states: workspace too small");
    const int cb = n < PROB_WG_BITS ? n : PROB_WG_BITS;
    const uint64_t groups = 1ull << (n - cb);
    double *part = (double *)d_ws;
    hipStream_t s = (hipStream_t)stream;
    if (prec_a == DQ_F64 && prec_b == DQ_F64) launch_compare<double, double>(a, b, n, (unsigned)groups, part, s);
    else if (prec_a == DQ_F64) launch_compare<double, float>(a, b, n, (unsigned)groups, part, s);
    else if (prec_b == DQ_F64) launch_compare<float, double>(a, b, n, (unsigned)groups, part, s);
    else launch_compare<float, float>(a, b, n, (unsigned)groups, part, s);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(compare_stage2, dim3(1), dim3(256), 0, s, part, (uint32_t)groups, d_out);
    HIPC(hipGetLastError());
    return DQ_OK;
}

// ---- dense sweeps -------------------------------------------------------------
namespace {
size_t dense_mat_floats(int k) { return (size_t)(2u << k) * (2u << k); }
size_t dense_lds_bytes(uint32_t T) { return 2 * ((size_t)4 << T); }
}  // namespace

extern "C" size_t dq_dense_ws_bytes(int k, uint32_t n_blocks, uint32_t n_ops, uint32_t n_mats) {
    if (k < 5 || k > 7) return 0;
    return align_up(n_blocks * sizeof(dq_dense_block), 256) + align_up(n_ops * sizeof(dq_dense_op), 256) +
           align_up((size_t)n_mats * dense_mat_floats(k) * sizeof(float), 256) +
           align_up((size_t)n_ops * DENSE_COLS * sizeof(uint32_t), 256);
}

extern "C" int dq_apply_dense(int k, const dq_dense_block *blocks, uint32_t n_blocks, const dq_dense_op *ops,
                              uint32_t n_ops, const double *mats, uint32_t n_mats, void *state, void *d_ws,
                              size_t ws_bytes, void *stream) {
    if (k < 5 || k > 7) return fail(DQ_ERR_ARG, "dq_apply_dense: k must be 5, 6 or 7");
    if (!blocks || !state || !d_ws || (n_ops && !ops) || (n_mats && !mats))
        return fail(DQ_ERR_ARG, "dq_apply_dense: null argument");
    const size_t need = dq_dense_ws_bytes(k, n_blocks, n_ops, n_mats);
    if (ws_bytes < need) return fail(DQ_ERR_ARG, "dq_apply_dense: workspace %zu < %zu bytes", ws_bytes, need);
    for (uint32_t i = 0; i < n_blocks; i++) {
        const dq_dense_block &b = blocks[i];
        const uint32_t n = b.n_qubits, T = b.tile_bits;
        if (n < 6 || n > 40 || T > DENSE_T_MAX || T > n || (int)T < k + 5)
            return fail(DQ_ERR_ARG, "dense block %u: need k + 5 <= T <= min(n, 13) (n=%u T=%u k=%d)", i, n, T, k);
        uint64_t m = 0;
        for (uint32_t t = 0; t < T; t++) {
            if (t < 6 && b.tile_q[t] != t) return fail(DQ_ERR_ARG, "dense block %u: tile bits 0..5 must be qubits 0..5", i);
            if (b.tile_q[t] >= n || ((m >> b.tile_q[t]) & 1)) return fail(DQ_ERR_ARG, "dense block %u: bad tile qubits", i);
            m |= 1ull << b.tile_q[t];
        }
        if ((m | b.nontile_mask) != ((n == 64) ? ~0ull : ((1ull << n) - 1)) || (m & b.nontile_mask))
            return fail(DQ_ERR_ARG, "dense block %u: tile and non-tile masks inconsistent", i);
        if ((uint64_t)b.op_begin + b.op_count > n_ops) return fail(DQ_ERR_ARG, "dense block %u: op range", i);
        for (uint32_t o = b.op_begin; o < b.op_begin + b.op_count; o++) {
            for (int j = 0; j < k; j++)
                if (ops[o].bits[j] >= T || (j && ops[o].bits[j] <= ops[o].bits[j - 1]))
                    return fail(DQ_ERR_ARG, "dense op %u: bits must be ascending tile bits", o);
            if (ops[o].mat >= n_mats) return fail(DQ_ERR_ARG, "dense op %u: matrix index", o);
        }
        if (n - T > 31) return fail(DQ_ERR_UNSUPPORTED, "dense block %u: grid too large", i);
    }
    // upload: blocks | ops | real matrices in MFMA operand order
    const size_t ob = align_up(n_blocks * sizeof(dq_dense_block), 256);
    const size_t mo = ob + align_up(n_ops * sizeof(dq_dense_op), 256);
    const int K = 1 << k, M = 2 * K;
    const size_t co = mo + align_up((size_t)n_mats * dense_mat_floats(k) * sizeof(float), 256);
    std::vector<unsigned char> buf(co + (size_t)n_ops * DENSE_COLS * sizeof(uint32_t), 0);
    memcpy(buf.data(), blocks, n_blocks * sizeof(dq_dense_block));
    if (n_ops) memcpy(buf.data() + ob, ops, n_ops * sizeof(dq_dense_op));
    float *dst = reinterpret_cast<float *>(buf.data() + mo);
    for (uint32_t mi = 0; mi < n_mats; mi++) {
        const double *U = mats + (size_t)mi * K * K * 2;
        float *D = dst + (size_t)mi * dense_mat_floats(k);
        for (int rb = 0; rb < M / 32; rb++)
            for (int p = 0; p < M / 2; p++)
                for (int l = 0; l < 64; l++) {
                    const int r = rb * 32 + (l & 31), c = 2 * p + (l >> 5);
                    const int ur = r & (K - 1), uc = c & (K - 1);
                    const double Ure = U[2 * ((size_t)ur * K + uc)], Uim = U[2 * ((size_t)ur * K + uc) + 1];
                    double a;
                    if (r < K) a = c < K ? Ure : -Uim;
                    else a = c < K ? Uim : Ure;
                    D[(((size_t)rb * (M / 8) + p / 4) * 64 + l) * 4 + p % 4] = (float)a;
                }
    }
    // per op: column c's tile offset, c's bits placed on the block's tile bits
    // outside the op (in ascending order)
    uint32_t *ct = reinterpret_cast<uint32_t *>(buf.data() + co);
    std::vector<int> op_T(n_ops, -1);
    for (uint32_t i = 0; i < n_blocks; i++) {
        const uint32_t T = blocks[i].tile_bits;
        for (uint32_t o = blocks[i].op_begin; o < blocks[i].op_begin + blocks[i].op_count; o++) {
            if (op_T[o] >= 0 && op_T[o] != (int)T) return fail(DQ_ERR_ARG, "dense op %u: in blocks of different tiles", o);
            op_T[o] = (int)T;
            uint32_t used = 0, cb[DENSE_T_MAX];
            for (int j = 0; j < k; j++) used |= 1u << ops[o].bits[j];
            int nc = 0;
            for (uint32_t b = 0; b < T; b++)
                if (!((used >> b) & 1u)) cb[nc++] = b;
            for (uint32_t c = 0; c < (1u << nc); c++) {
                uint32_t v = 0;
                for (int b = 0; b < nc; b++) v |= ((c >> b) & 1u) << cb[b];
                ct[(size_t)o * DENSE_COLS + c] = v;
            }
        }
    }
    int rc = stage_upload(d_ws, buf, stream);
    if (rc) return rc;
    unsigned char *d = (unsigned char *)d_ws;
    const dq_dense_block *d_blocks = (const dq_dense_block *)d;
    const dq_dense_op *d_ops = (const dq_dense_op *)(d + ob);
    const float *d_mats = (const float *)(d + mo);
    const uint32_t *d_ct = (const uint32_t *)(d + co);
    hipStream_t s = (hipStream_t)stream;
    for (uint32_t i = 0; i < n_blocks; i++) {
        const dq_dense_block &b = blocks[i];
        const unsigned grid = 1u << (b.n_qubits - b.tile_bits);
        const size_t lds = dense_lds_bytes(b.tile_bits);
        auto *st = (typename V2<float>::type *)state;
        if (k == 5) hipLaunchKernelGGL(dense_kernel<5>, dim3(grid), dim3(DENSE_THREADS), lds, s, d_blocks + i, d_ops, d_mats, d_ct, st);
        else if (k == 6) hipLaunchKernelGGL(dense_kernel<6>, dim3(grid), dim3(DENSE_THREADS), lds, s, d_blocks + i, d_ops, d_mats, d_ct, st);
        else hipLaunchKernelGGL(dense_kernel<7>, dim3(grid), dim3(DENSE_THREADS), lds, s, d_blocks + i, d_ops, d_mats, d_ct, st);
        HIPC(hipGetLastError());
    }
    return DQ_OK;
}

extern "C" int dq_event_query(void *ev) {
    if (!ev) return fail(DQ_ERR_ARG, "dq_event_query: null event");
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: keep it out of the next launch check
        return 1;
    }
    return fail(DQ_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(e));
}

extern "C" int dq_event_sync(void *ev) {
    if (!ev) return fail(DQ_ERR_ARG, "dq_event_sync: null event");
    HIPC(hipEventSynchronize((hipEvent_t)ev));
    return DQ_OK;
}

extern "C" int dq_read_amps(const void *state, int n, int prec, const int *l2p, uint64_t start, uint64_t count,
                            double *out, void *d_ws, size_t ws_bytes, void *stream) {
    if (!state || !out || n < 1 || n > 40 || (prec != DQ_F32 && prec != DQ_F64)) return fail(DQ_ERR_ARG, "dq_read_amps: bad arguments");
    const uint64_t N = 1ull << n;
    if (start > N || count > N - start) return fail(DQ_ERR_ARG, "dq_read_amps: range out of bounds");
    if (count == 0) return DQ_OK;
    int map[64];
    for (int q = 0; q < n; q++) {
        map[q] = l2p ? l2p[q] : q;
        if (map[q] < 0 || map[q] >= n) return fail(DQ_ERR_ARG, "dq_read_amps: bad qubit map");
    }
    const size_t need = align_up(64 * sizeof(int), 256) + count * 2 * sizeof(double);
    if (!d_ws || ws_bytes < need) return fail(DQ_ERR_ARG, "dq_read_amps: workspace %zu < %zu bytes", ws_bytes, need);
    int *d_map = (int *)d_ws;
    double *d_out = (double *)((unsigned char *)d_ws + align_up(64 * sizeof(int), 256));
    std::vector<unsigned char> m((unsigned char *)map, (unsigned char *)map + 64 * sizeof(int));
    int rc = stage_upload(d_map, m, stream);
    if (rc) return rc;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    if (prec == DQ_F64)
        hipLaunchKernelGGL(read_kernel<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const V2<double>::type *)state, n, d_map, start, count, d_out);
    else
        hipLaunchKernelGGL(read_kernel<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const V2<float>::type *)state, n, d_map, start, count, d_out);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, d_out, count * 2 * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPC(hipStreamSynchronize((hipStream_t)stream));
    return DQ_OK;
}

extern "C" int dq_write_amps(void *state, int n, int prec, uint64_t start, uint64_t count, const double *re,
                             const double *im, void *d_ws, size_t ws_bytes, void *stream) {
    if (!state || !re || !im || n < 1 || n > 40 || (prec != DQ_F32 && prec != DQ_F64)) return fail(DQ_ERR_ARG, "dq_write_amps: bad arguments");
    const uint64_t N = 1ull << n;
    if (start > N || count > N - start) return fail(DQ_ERR_ARG, "dq_write_amps: range out of bounds");
    if (count == 0) return DQ_OK;
    const size_t half = align_up(count * sizeof(double), 256);
    if (!d_ws || ws_bytes < 2 * half) return fail(DQ_ERR_ARG, "dq_write_amps: workspace too small");
    double *d_re = (double *)d_ws;
    double *d_im = (double *)((unsigned char *)d_ws + half);
    HIPC(hipMemcpyAsync(d_re, re, count * sizeof(double), hipMemcpyHostToDevice, (hipStream_t)stream));
    HIPC(hipMemcpyAsync(d_im, im, count * sizeof(double), hipMemcpyHostToDevice, (hipStream_t)stream));
    const unsigned blocks = (unsigned)((count + 255) / 256);
    if (prec == DQ_F64)
        hipLaunchKernelGGL(write_kernel<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (V2<double>::type *)state, start, count, d_re, d_im);
    else
        hipLaunchKernelGGL(write_kernel<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (V2<float>::type *)state, start, count, d_re, d_im);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize((hipStream_t)stream));
    return DQ_OK;
}
