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

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// LDS address swizzle: xor the low 3 bits of an element index with the xor of
// its higher 3-bit groups.  A bijection, and linear over GF(2), so the swizzle
// of a sum of disjoint bits is the xor of their swizzles.
__device__ __forceinline__ uint32_t swz(uint32_t x) {
    return x ^ (((x >> 3) ^ (x >> 6) ^ (x >> 9) ^ (x >> 12)) & 7u);
}

template <int W>
__device__ __forceinline__ uint32_t thread_tile_addr(const dq_phase *__restrict__ ph, uint32_t tid) {
    uint32_t t = 0;
#pragma unroll
    for (int l = 0; l < 6; l++) t |= ((tid >> l) & 1u) ? ph->lane_t[l] : 0u;
#pragma unroll
    for (int w = 0; w < W; w++) t |= ((tid >> (6 + w)) & 1u) ? ph->wave_t[w] : 0u;
    return t;
}

// The register tile of one thread: 2^R amplitudes, real and imaginary parts
// as separate scalars.  Every op below updates these registers in place, so
// the tile keeps one register assignment through the op loop.
template <typename T, int R>
struct Regs {
    T re[1 << R];
    T im[1 << R];
};

// Re-distribute the tile from the layout of phase A to that of phase B
// through LDS (each thread writes its 2^R amplitudes, then reads 2^R).
// SPLIT moves the real parts, then the imaginary parts, through an LDS array
// of scalars: half the LDS footprint (more workgroups per CU), two more
// barriers.  Offsets are in bytes: the thread's part xor a uniform part per
// register (the swizzle is linear), one vector xor per access.
template <int R, int W, bool SPLIT, typename T>
__device__ __forceinline__ void exchange(Regs<T, R> &a, void *lds_raw, const dq_phase *__restrict__ A,
                                         const dq_phase *__restrict__ B, bool lead_barrier, uint32_t tid) {
    constexpr int NR = 1 << R;
    constexpr uint32_t ES = SPLIT ? sizeof(T) : 2 * sizeof(T);
    const uint32_t sa = swz(thread_tile_addr<W>(A, tid)) * ES;
    const uint32_t sb = swz(thread_tile_addr<W>(B, tid)) * ES;
    uint32_t ra[R > 0 ? R : 1], rb[R > 0 ? R : 1];
#pragma unroll
    for (int r = 0; r < R; r++) { ra[r] = swz(A->reg_t[r]) * ES; rb[r] = swz(B->reg_t[r]) * ES; }
    // register j's offset = thread part xor a uniform (scalar) part, formed at
    // each access: one vector xor, no offset arrays held in VGPRs
    auto offa = [&](int j) {
        uint32_t x = 0;
#pragma unroll
        for (int r = 0; r < R; r++)
            if ((j >> r) & 1) x ^= ra[r];
        return sa ^ x;
    };
    auto offb = [&](int j) {
        uint32_t x = 0;
#pragma unroll
        for (int r = 0; r < R; r++)
            if ((j >> r) & 1) x ^= rb[r];
        return sb ^ x;
    };
    unsigned char *lds = reinterpret_cast<unsigned char *>(lds_raw);
    if (!SPLIT) {
        if (lead_barrier) __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) {
            *reinterpret_cast<T *>(lds + offa(j)) = a.re[j];
            *reinterpret_cast<T *>(lds + offa(j) + sizeof(T)) = a.im[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) {
            a.re[j] = *reinterpret_cast<const T *>(lds + offb(j));
            a.im[j] = *reinterpret_cast<const T *>(lds + offb(j) + sizeof(T));
        }
    } else {
        if (lead_barrier) __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) *reinterpret_cast<T *>(lds + offa(j)) = a.re[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) a.re[j] = *reinterpret_cast<const T *>(lds + offb(j));
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) *reinterpret_cast<T *>(lds + offa(j)) = a.im[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NR; j++) a.im[j] = *reinterpret_cast<const T *>(lds + offb(j));
    }
}

// ---- op interpreter --------------------------------------------------------
// The op list of a phase runs in gfx950 assembly generated by gen_ops_asm.py
// (format of the device program and register use documented there).
#include "ops_asm.inc"

// Occupancy target (waves per SIMD) of a sweep instantiation: the register
// tile takes 2^R amplitudes x (2 * sizeof(T) / 4) VGPRs; allow 40 more for
// addresses, coefficients and temporaries.
template <typename T, int R>
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
            const V v = __builtin_nontemporal_load(st + u + lane);
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
            GV *p_ = st + u;                                                              \
            typename V2<T>::type v_;                                                      \
            v_.x = (A).re[j];                                                             \
            v_.y = (A).im[j];                                                             \
            __builtin_nontemporal_store(v_, p_ + lane);                                   \
        }                                                                                 \
    } while (0)

// The phases of a block on one tile: exchanges through LDS between phases,
// the op program of each phase in the assembly interpreter.  lead: a barrier
// before the first exchange (LDS may still be read for a previous tile).
template <typename T, int R, int W, bool SPLIT>
__device__ __forceinline__ void process_tile(Regs<T, R> &a, void *lds, const dq_phase *__restrict__ ph, uint32_t nph,
                                             const uint32_t *__restrict__ pp, const unsigned char *__restrict__ prog,
                                             uint64_t cbase, uint32_t tid, bool lead) {
    for (uint32_t p = 0; p < nph; p++) {
        if (p > 0) exchange<R, W, SPLIT, T>(a, lds, ph + (p - 1), ph + p, p > 1 || lead, tid);
        AsmOps<T, R>::run(a, prog + pp[2 * p], pp[2 * p + 1], cbase, tid);
    }
}

template <typename T, int R, int W, bool SPLIT>
__global__ void __launch_bounds__(64 << W, (sweep_waves_per_eu<T, R>()))
sweep_kernel(const dq_job *__restrict__ jobs, const dq_block *__restrict__ blocks,
             const dq_phase *__restrict__ phases, const unsigned char *__restrict__ prog,
             const uint32_t *__restrict__ pprog) {
    using GV = __attribute__((address_space(1))) typename V2<T>::type;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const dq_job job = jobs[blockIdx.y];
    const dq_block *__restrict__ blk = blocks + job.block;
    GV *__restrict__ st = (GV *)(job.state);
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lane = tid & 63u;
    const uint64_t base = tile_base(blk->nontile_mask, blockIdx.x);
    Regs<T, R> a;
    load_tile<T, R, W>(a, blk, st, base, wv, lane);
    process_tile<T, R, W, SPLIT>(a, lds_raw, phases + blk->phase_begin, blk->phase_count,
                                 pprog + 2 * blk->phase_begin, prog, base | job.base_hi, tid, false);
    DQ_STORE_TILE(a, base);
}

// Occupancy target of the pipelined form: two register tiles live at once.
template <typename T, int R>
constexpr int pipe_waves_per_eu() {
    constexpr int amp = (1 << R) * (int)(2 * sizeof(T) / 4);
    constexpr int budget = 2 * amp + 40;
    constexpr int w = 512 / budget;
    return w > 8 ? 8 : (w < 1 ? 1 : w);
}

// Persistent, software-pipelined sweep of one state: workgroup g runs tiles
// g, g + G, g + 2G, ... (G = gridDim.x) and issues the loads of its next tile
// into a second register tile before it computes the current one, so each
// wave keeps reads in flight while its VALU, LDS and scalar work runs.  The
// interpreter's asm statement is a compiler barrier for memory operations,
// which keeps the prefetch loads where they are issued.
template <typename T, int R, int W, bool SPLIT>
__global__ void __launch_bounds__(64 << W, (pipe_waves_per_eu<T, R>()))
sweep_pipe_kernel(const dq_job *__restrict__ jobs, const dq_block *__restrict__ blocks,
                  const dq_phase *__restrict__ phases, const unsigned char *__restrict__ prog,
                  const uint32_t *__restrict__ pprog) {
    using GV = __attribute__((address_space(1))) typename V2<T>::type;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const dq_job job = jobs[blockIdx.y];
    const dq_block *__restrict__ blk = blocks + job.block;
    GV *__restrict__ st = (GV *)(job.state);
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lane = tid & 63u;
    const uint64_t msk = blk->nontile_mask;
    const uint64_t ntiles = 1ull << __builtin_popcountll(msk);
    const uint64_t G = gridDim.x;
    const dq_phase *__restrict__ ph = phases + blk->phase_begin;
    const uint32_t nph = blk->phase_count;
    const uint32_t *__restrict__ pp = pprog + 2 * blk->phase_begin;
    const uint64_t hi = job.base_hi;
    uint64_t t = blockIdx.x;
    if (t >= ntiles) return;
    Regs<T, R> a, b;
    uint64_t ba = tile_base(msk, t);
    load_tile<T, R, W>(a, blk, st, ba, wv, lane);
    bool lead = false;
    // The prefetch loads are unconditional (a tile past the end is clamped to
    // the last tile and its data left unused): a conditional load would make
    // the compiler merge the two register tiles with copies, and a copy waits
    // for the data, which turns the prefetch into a plain load.
    while (true) {
        const uint64_t t1 = t + G;
        const uint64_t bb = tile_base(msk, t1 < ntiles ? t1 : ntiles - 1);
        load_tile<T, R, W>(b, blk, st, bb, wv, lane);
        process_tile<T, R, W, SPLIT>(a, lds_raw, ph, nph, pp, prog, ba | hi, tid, lead);
        DQ_STORE_TILE(a, ba);
        lead = true;
        if (t1 >= ntiles) break;
        const uint64_t t2 = t1 + G;
        ba = tile_base(msk, t2 < ntiles ? t2 : ntiles - 1);
        load_tile<T, R, W>(a, blk, st, ba, wv, lane);
        process_tile<T, R, W, SPLIT>(b, lds_raw, ph, nph, pp, prog, bb | hi, tid, lead);
        DQ_STORE_TILE(b, bb);
        if (t2 >= ntiles) break;
        t = t2;
    }
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

template <typename T>
__global__ void fill_kernel(typename V2<T>::type *__restrict__ st, uint64_t N, int kind, uint64_t index, double value) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    typename V2<T>::type v;
    v.x = (kind == DQ_INIT_CLASSICAL) ? (i == index ? T(1) : T(0)) : (T)value;
    v.y = T(0);
    __builtin_nontemporal_store(v, st + i);
}

// ----------------------------------------------------------------------------
// kernel table
// ----------------------------------------------------------------------------
typedef void (*sweep_fn)(const dq_job *, const dq_block *, const dq_phase *, const unsigned char *, const uint32_t *);

template <typename T>
struct SweepTable {
    sweep_fn fn[2][DQ_MAX_REG_BITS + 1][DQ_MAX_WAVE_BITS + 1];
    sweep_fn pipe[2][DQ_MAX_REG_BITS + 1][DQ_MAX_WAVE_BITS + 1];
    SweepTable() {
        memset(fn, 0, sizeof fn);
        memset(pipe, 0, sizeof pipe);
#define DQ_INST(R_, W_)                                                                  \
    fn[0][R_][W_] = sweep_kernel<T, R_, W_, false>;                                      \
    fn[1][R_][W_] = sweep_kernel<T, R_, W_, true>;                                       \
    pipe[0][R_][W_] = sweep_pipe_kernel<T, R_, W_, false>;                               \
    pipe[1][R_][W_] = sweep_pipe_kernel<T, R_, W_, true>;
        DQ_INST(1, 0) DQ_INST(2, 0) DQ_INST(3, 0) DQ_INST(4, 0) DQ_INST(5, 0)
        DQ_INST(2, 1) DQ_INST(3, 1) DQ_INST(4, 1) DQ_INST(5, 1)
        DQ_INST(2, 2) DQ_INST(3, 2) DQ_INST(4, 2) DQ_INST(5, 2)
        DQ_INST(3, 3) DQ_INST(4, 3)
#undef DQ_INST
    }
};

static SweepTable<double> g_tab64;
static SweepTable<float> g_tab32;

// split exchanges: on for FP64 tiles of 12+ bits unless DGQ_SPLIT_EXCHANGE=0
static int g_split_mode = -1;
static bool use_split(int prec, int T) {
    if (g_split_mode < 0) {
        const char *e = getenv("DGQ_SPLIT_EXCHANGE");
        g_split_mode = e ? atoi(e) : 2;
    }
    if (g_split_mode == 0) return false;
    if (g_split_mode == 1) return true;
    return prec == DQ_F64 && T >= 12;
}

// persistent pipelined sweeps for single-state launches (DGQ_PIPE=1); off by
// default: with two register tiles live the allocator copies prefetched
// values, and each copy waits for its load (measured slower on MI355X)
static int g_pipe_mode = -1;
static bool use_pipe() {
    if (g_pipe_mode < 0) {
        const char *e = getenv("DGQ_PIPE");
        g_pipe_mode = e ? atoi(e) : 0;
    }
    return g_pipe_mode != 0;
}

static sweep_fn lookup_pipe(int prec, int R, int W, bool split) {
    if (R < 0 || R > DQ_MAX_REG_BITS || W < 0 || W > DQ_MAX_WAVE_BITS) return nullptr;
    return prec == DQ_F64 ? g_tab64.pipe[split][R][W] : g_tab32.pipe[split][R][W];
}

static sweep_fn lookup(int prec, int R, int W, bool split = false) {
    if (R < 0 || R > DQ_MAX_REG_BITS || W < 0 || W > DQ_MAX_WAVE_BITS) return nullptr;
    return prec == DQ_F64 ? g_tab64.fn[split][R][W] : g_tab32.fn[split][R][W];
}

static size_t elem_bytes(int prec) { return prec == DQ_F64 ? 16 : 8; }

// ----------------------------------------------------------------------------
// staging of descriptors: a ring of pinned host buffers guarded by events
// ----------------------------------------------------------------------------
namespace {
struct StageSlot {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
};
std::mutex g_stage_mu;
StageSlot g_stage[16];
unsigned g_stage_next = 0;
int g_device = -1;
int g_num_cus = 256;
}  // namespace

static int stage_upload(void *d_dst, const std::vector<unsigned char> &src, void *stream) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    StageSlot &s = g_stage[g_stage_next++ % 16];
    if (!s.ev) HIPC(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    HIPC(hipEventSynchronize(s.ev));
    if (s.cap < src.size()) {
        if (s.host) HIPC(hipHostFree(s.host));
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
    HIPC(hipSetDevice(device));
    if (g_device != device) {
        // allow the large LDS tiles (up to 128 KiB) on every sweep instantiation
        for (int sp = 0; sp < 2; sp++)
            for (int R = 0; R <= DQ_MAX_REG_BITS; R++)
                for (int W = 0; W <= DQ_MAX_WAVE_BITS; W++) {
                    if (g_tab64.fn[sp][R][W])
                        HIPC(hipFuncSetAttribute((const void *)g_tab64.fn[sp][R][W], hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
                    if (g_tab32.fn[sp][R][W])
                        HIPC(hipFuncSetAttribute((const void *)g_tab32.fn[sp][R][W], hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
                    if (g_tab64.pipe[sp][R][W])
                        HIPC(hipFuncSetAttribute((const void *)g_tab64.pipe[sp][R][W], hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
                    if (g_tab32.pipe[sp][R][W])
                        HIPC(hipFuncSetAttribute((const void *)g_tab32.pipe[sp][R][W], hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
                }
        {
            hipDeviceProp_t prop;
            HIPC(hipGetDeviceProperties(&prop, device));
            g_num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        }
        HIPC(hipFuncSetAttribute((const void *)small_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
        HIPC(hipFuncSetAttribute((const void *)small_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
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
    if (e == hipErrorOutOfMemory) return fail(DQ_ERR_NOMEM, "dq_malloc: out of device memory (%zu bytes)", bytes);
    HIPC(e);
    return DQ_OK;
}
extern "C" int dq_free(void *ptr) { HIPC(hipFree(ptr)); return DQ_OK; }
extern "C" int dq_host_alloc(void **ptr, size_t bytes) {
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault);
    if (e == hipErrorOutOfMemory) return fail(DQ_ERR_NOMEM, "dq_host_alloc: out of pinned memory (%zu bytes)", bytes);
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

// Upper bound of the device program built from n_ops descriptors: at most
// 96 bytes per descriptor (a complex matrix in FP64; a register table of 2^R
// entries spans 1 + 2^R / 4 descriptors and 32 + 16 * 2^R bytes), plus
// padding so the interpreter's fixed-size header loads stay in the buffer.
static size_t prog_bound(uint32_t n_ops) { return (size_t)n_ops * 96 + 256; }

extern "C" size_t dq_apply_ws_bytes(uint32_t n_blocks, uint32_t n_phases, uint32_t n_ops, uint32_t n_jobs) {
    return align_up(n_blocks * sizeof(dq_block), 256) + align_up(n_phases * sizeof(dq_phase), 256) +
           align_up(n_ops * sizeof(dq_op), 256) + align_up(n_jobs * sizeof(dq_job), 256) +
           align_up(n_phases * 2 * sizeof(uint32_t), 256) + align_up(prog_bound(n_ops), 256);
}

// Device program of one phase (format documented in gen_ops_asm.py): one
// entry per op, a 32-byte header and the coefficients in the state precision,
// register conditions turned into a mask over register indices.
static void build_prog(int prec, int R, const dq_op *ops, uint32_t count, std::vector<unsigned char> &out,
                       uint32_t &n_out) {
    const size_t ts = prec == DQ_F64 ? 8 : 4;
    const uint32_t NR = 1u << R;
    n_out = 0;
    for (uint32_t k = 0; k < count; k++) {
        const dq_op &o = ops[k];
        auto reg_mask = [&](bool pairs) {
            uint32_t m = 0;
            for (uint32_t j = 0; j < NR; j++) {
                if (pairs && ((j >> o.tbit) & 1)) continue;
                if ((j & o.c_reg) == o.v_reg) m |= 1u << j;
            }
            return m;
        };
        uint32_t kind = 0, bit = 0, jmask = 0;
        size_t body = 0;
        const unsigned char *src = reinterpret_cast<const unsigned char *>(o.m);
        std::vector<unsigned char> table;
        switch (o.kind) {
            case DQ_OP_SHEAR: kind = 1; bit = o.tbit; body = 2 * ts; break;
            case DQ_OP_DIAGR: {  // jmask: the entries that are not exactly 1
                kind = 2;
                table.resize(NR * 2 * ts);
                for (uint32_t e = 0; e < NR; e++) {
                    const unsigned char *ent = reinterpret_cast<const unsigned char *>(ops[k + 1 + e / 4].m) + (e % 4) * 2 * ts;
                    memcpy(table.data() + e * 2 * ts, ent, 2 * ts);
                    bool one;
                    if (prec == DQ_F64) {
                        double v[2];
                        memcpy(v, ent, sizeof v);
                        one = v[0] == 1.0 && v[1] == 0.0;
                    } else {
                        float v[2];
                        memcpy(v, ent, sizeof v);
                        one = v[0] == 1.0f && v[1] == 0.0f;
                    }
                    if (!one) jmask |= 1u << e;
                }
                src = table.data();
                body = table.size();
                k += o.tbit;
                break;
            }
            case DQ_OP_NEG: kind = 3; jmask = reg_mask(false); break;
            case DQ_OP_PHASE: kind = 4; jmask = reg_mask(false); body = 2 * ts; break;
            case DQ_OP_MAT: kind = 5; bit = o.tbit; jmask = reg_mask(true); body = 8 * ts; break;
            case DQ_OP_SWAP: kind = 6; bit = o.tbit; jmask = reg_mask(true); break;
            case DQ_OP_RMAT: kind = 7; bit = o.tbit; body = 4 * ts; break;
            default: break;
        }
        const size_t size = align_up(32 + body, 32);
        const size_t at = out.size();
        out.resize(at + size, 0);
        uint32_t hdr[4] = {kind * 8 + bit, jmask, o.c_thr | (o.v_thr << 16), (uint32_t)size};
        memcpy(out.data() + at, hdr, 16);
        memcpy(out.data() + at + 16, &o.c_wg, 8);
        memcpy(out.data() + at + 24, &o.v_wg, 8);
        if (body) memcpy(out.data() + at + 32, src, body);
        n_out++;
    }
}

extern "C" int dq_init_state(void *state, int n, int prec, int kind, uint64_t index, double value, void *stream) {
    if (!state || n < 1 || n > 40 || (prec != DQ_F32 && prec != DQ_F64)) return fail(DQ_ERR_ARG, "dq_init_state: bad arguments");
    if (kind != DQ_INIT_CLASSICAL && kind != DQ_INIT_PLUS) return fail(DQ_ERR_ARG, "dq_init_state: bad kind %d", kind);
    const uint64_t N = 1ull << n;
    const unsigned blocks = (unsigned)((N + 255) / 256);
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
        if (b0.reg_bits && !lookup(prec, b0.reg_bits, b0.wave_bits))
            return fail(DQ_ERR_UNSUPPORTED, "no sweep kernel for R=%d W=%d", b0.reg_bits, b0.wave_bits);
    }

    // device programs of the phases of sweep blocks: (byte offset, op count) per phase
    std::vector<uint32_t> pprog(2 * (size_t)n_phases, 0);
    std::vector<int> prog_r(n_phases, -1);
    std::vector<unsigned char> prog;
    for (uint32_t i = 0; i < n_blocks; i++) {
        const dq_block &b = blocks[i];
        if (b.reg_bits == 0) continue;
        for (uint32_t p = b.phase_begin; p < b.phase_begin + b.phase_count; p++) {
            if (prog_r[p] == (int)b.reg_bits) continue;
            if (prog_r[p] >= 0) return fail(DQ_ERR_ARG, "phase %u shared by blocks of different register bits", p);
            prog_r[p] = (int)b.reg_bits;
            pprog[2 * p] = (uint32_t)prog.size();
            build_prog(prec, (int)b.reg_bits, ops + phases[p].op_begin, phases[p].op_count, prog, pprog[2 * p + 1]);
        }
    }
    if (prog.size() + 256 > prog_bound(n_ops)) return fail(DQ_ERR_ARG, "dq_apply_blocks: device program overflow");

    // pack and upload
    std::vector<unsigned char> buf(need, 0);
    size_t off_b = 0, off_p = align_up(n_blocks * sizeof(dq_block), 256);
    size_t off_o = off_p + align_up(n_phases * sizeof(dq_phase), 256);
    size_t off_j = off_o + align_up(n_ops * sizeof(dq_op), 256);
    size_t off_pp = off_j + align_up(n_jobs * sizeof(dq_job), 256);
    size_t off_pg = off_pp + align_up(n_phases * 2 * sizeof(uint32_t), 256);
    memcpy(buf.data() + off_b, blocks, n_blocks * sizeof(dq_block));
    memcpy(buf.data() + off_p, phases, n_phases * sizeof(dq_phase));
    if (n_ops) memcpy(buf.data() + off_o, ops, n_ops * sizeof(dq_op));
    memcpy(buf.data() + off_j, jobs, n_jobs * sizeof(dq_job));
    if (n_phases) memcpy(buf.data() + off_pp, pprog.data(), pprog.size() * sizeof(uint32_t));
    if (!prog.empty()) memcpy(buf.data() + off_pg, prog.data(), prog.size());
    buf.resize(off_pg + prog.size() + 256, 0);  // upload only what is used (+ padding)
    int rc = stage_upload(d_ws, buf, stream);
    if (rc) return rc;
    unsigned char *d = (unsigned char *)d_ws;
    const dq_block *d_blocks = (const dq_block *)(d + off_b);
    const dq_phase *d_phases = (const dq_phase *)(d + off_p);
    const dq_op *d_ops = (const dq_op *)(d + off_o);
    const dq_job *d_jobs = (const dq_job *)(d + off_j);
    const uint32_t *d_pprog = (const uint32_t *)(d + off_pp);
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
            const bool split = use_split(prec, b0.tile_bits);
            const size_t lds = multi ? ((size_t)1 << b0.tile_bits) * (split ? elem_bytes(prec) / 2 : elem_bytes(prec)) : 0;
            if (lds > 131072) return fail(DQ_ERR_UNSUPPORTED, "tile of %d bits needs %zu bytes of LDS", b0.tile_bits, lds);
            const uint64_t tiles = 1ull << (n - b0.tile_bits);
            if (tiles > 0x7fffffffull || nj > 65535) return fail(DQ_ERR_UNSUPPORTED, "grid too large");
            // (R = 5 keeps the one-pass form: two 32-amplitude tiles do not fit the registers)
            sweep_fn pf = (nj == 1 && b0.reg_bits <= 4 && use_pipe()) ? lookup_pipe(prec, b0.reg_bits, b0.wave_bits, split)
                                                                     : nullptr;
            int per_cu = 0;
            if (pf) HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)pf, 64 << b0.wave_bits, lds));
            const uint64_t resident = (uint64_t)(per_cu > 0 ? per_cu : 0) * (uint64_t)g_num_cus;
            if (pf && resident > 0 && tiles >= 2 * resident) {
                // persistent: every resident workgroup walks its share of the tiles
                hipLaunchKernelGGL(pf, dim3((unsigned)resident, 1), dim3(64u << b0.wave_bits), lds, (hipStream_t)stream,
                                   d_jobs + j0, d_blocks, d_phases, d_prog, d_pprog);
            } else {
                sweep_fn fn = lookup(prec, b0.reg_bits, b0.wave_bits, split);
                hipLaunchKernelGGL(fn, dim3((unsigned)tiles, nj), dim3(64u << b0.wave_bits), lds, (hipStream_t)stream,
                                   d_jobs + j0, d_blocks, d_phases, d_prog, d_pprog);
            }
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
