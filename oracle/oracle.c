/* oracle.c -- CPU oracle for the fused-gate state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the product library
 * never links or calls it.
 *
 * It is an eager, gate-by-gate FP64 restatement of the reference's QuEST
 * semantics, with OpenMP over amplitudes:
 *   - gate encoding (type, targetMask, controlMask, operand; rotations store
 *     half angles; diagonal gates keep their qubit in controlMask):
 *       src/cpp/QuEST_distributed.cpp:118-364, src/QuEST.c:93-423,
 *       src/QuEST_common.c:56-63,154-183
 *   - per-gate arithmetic, operation by operation:
 *       src/cpp/QuEST_gpu_kernel.cu:17-170 (the *_func_reg bodies) and the
 *       control test of :253-352 (MATCH_BIT on the amplitude index)
 *   - host-side sin/cos of the stored angle: src/cpp/QuEST_gpu_kernel.cu:450-455
 *   - probabilities: P(0) by direct summation, P(1) = 1 - P(0)
 *       (src/cpp/QuEST_backend_distributed.cpp:735-737)
 *   - measurement outcome rule: src/QuEST_common.c:97-115
 *   - MT19937 (Matsumoto & Nishimura, mt19937ar): src/mt19937ar.c:60-180
 *
 * Amplitudes are AoS interleaved {re, im} doubles, qubit q = bit q of the
 * index, as in the reference (include/QuEST_precision.h:41).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } oc_amp;

/* Gate types: same numbering as src/cpp/QuEST_backend.h:89-114.  Values 0..7
 * are the phase gates e^{i k pi/4} applied when every qubit of controlMask
 * is 1.  OC_COLLAPSE (this oracle's own) projects onto one outcome of the
 * target qubit and rescales. */
enum {
    OC_ONE = 0, OC_T = 1, OC_S = 2, OC_T_CONJ_NEG = 3, OC_Z = 4, OC_T_NEG = 5,
    OC_S_CONJ = 6, OC_T_CONJ = 7, OC_PHASE_SHIFT = 8, OC_ROTATE_Z_FAKE = 9,
    OC_ROTATE_Z = 10, OC_X = 11, OC_Y = 12, OC_Y_CONJ = 13, OC_H = 14,
    OC_ROTATE_X = 15, OC_ROTATE_Y = 16, OC_COMPACT = 17, OC_UNITARY = 18,
    OC_COLLAPSE = 19
};

#define RSQRT2 0.7071067811865475244008443621 /* src/cpp/QuEST_backend.h:40 */

static int match(uint64_t idx, uint64_t mask) { return (idx & mask) == mask; }

/* Apply one recorded gate.  op[] holds the GateOperand union as 8 doubles:
 * angle at op[0]; compact alpha=(op0,op1) beta=(op2,op3); unitary
 * r0c0,r0c1,r1c0,r1c1 = (op0,op1),(op2,op3),(op4,op5),(op6,op7);
 * collapse: outcome at op[0], scale at op[1]. */
void oc_apply(oc_amp *s, int n, int type, uint64_t tmask, uint64_t cmask, const double *op) {
    const int64_t N = (int64_t)1 << n;
    if (type < OC_PHASE_SHIFT) {
        /* diagonal phase gates, QuEST_gpu_kernel.cu:17-64 */
        if (type == OC_ONE) return;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if (!match((uint64_t)i, cmask)) continue;
            double r = s[i].re, m = s[i].im;
            switch (type) {
                case OC_Z: s[i].re = -r; s[i].im = -m; break;
                case OC_S: s[i].re = -m; s[i].im = +r; break;
                case OC_S_CONJ: s[i].re = +m; s[i].im = -r; break;
                case OC_T: s[i].re = RSQRT2 * (r - m); s[i].im = RSQRT2 * (m + r); break;
                case OC_T_CONJ: s[i].re = RSQRT2 * (r + m); s[i].im = RSQRT2 * (m - r); break;
                case OC_T_NEG: s[i].re = -RSQRT2 * (r - m); s[i].im = -RSQRT2 * (m + r); break;
                case OC_T_CONJ_NEG: s[i].re = -RSQRT2 * (r + m); s[i].im = -RSQRT2 * (m - r); break;
            }
        }
        return;
    }
    if (type == OC_PHASE_SHIFT) {
        /* QuEST_gpu_kernel.cu:66-71, angle -> (cos, sin) on the host */
        const double c = cos(op[0]), sn = sin(op[0]);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            if (!match((uint64_t)i, cmask)) continue;
            double r = s[i].re, m = s[i].im;
            s[i].re = c * r - sn * m;
            s[i].im = c * m + sn * r;
        }
        return;
    }
    if (type == OC_COLLAPSE) {
        const int outcome = (int)op[0];
        const double scale = op[1];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            int bit = (i & (int64_t)tmask) != 0;
            if (bit == outcome) { s[i].re *= scale; s[i].im *= scale; }
            else { s[i].re = 0.0; s[i].im = 0.0; }
        }
        return;
    }
    /* one-target gates on pairs (i0, i0 | tmask) */
    int t = 0;
    while (((uint64_t)1 << t) != tmask) t++;
    const int64_t half = N >> 1;
    const int64_t low = ((int64_t)1 << t) - 1;
    double c = 0, sn = 0;
    if (type == OC_ROTATE_Z || type == OC_ROTATE_X || type == OC_ROTATE_Y) { c = cos(op[0]); sn = sin(op[0]); }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < half; k++) {
        int64_t i0 = ((k & ~low) << 1) | (k & low);
        int64_t i1 = i0 | (int64_t)tmask;
        if (!match((uint64_t)i0, cmask)) continue;
        double r0 = s[i0].re, m0 = s[i0].im, r1 = s[i1].re, m1 = s[i1].im;
        double R0, M0, R1, M1;
        switch (type) {
            case OC_ROTATE_Z: /* :73-83 */
                R0 = c * r0 + sn * m0; M0 = c * m0 - sn * r0;
                R1 = c * r1 - sn * m1; M1 = c * m1 + sn * r1; break;
            case OC_X: /* :85-95 */
                R0 = r1; M0 = m1; R1 = r0; M1 = m0; break;
            case OC_Y: /* :97-107 */
                R0 = +m1; M0 = -r1; R1 = -m0; M1 = +r0; break;
            case OC_Y_CONJ: /* :109-119 */
                R0 = -m1; M0 = +r1; R1 = +m0; M1 = -r0; break;
            case OC_H: /* :121-131 */
                R0 = RSQRT2 * (r0 + r1); M0 = RSQRT2 * (m0 + m1);
                R1 = RSQRT2 * (r0 - r1); M1 = RSQRT2 * (m0 - m1); break;
            case OC_ROTATE_X: /* :133-143 */
                R0 = c * r0 + sn * m1; M0 = c * m0 - sn * r1;
                R1 = c * r1 + sn * m0; M1 = c * m1 - sn * r0; break;
            case OC_ROTATE_Y: /* :145-155 */
                R0 = c * r0 - sn * r1; M0 = c * m0 - sn * m1;
                R1 = sn * r0 + c * r1; M1 = sn * m0 + c * m1; break;
            case OC_COMPACT: { /* :157-166 */
                double ar = op[0], ai = op[1], br = op[2], bi = op[3];
                R0 = ar * r0 - ai * m0 - br * r1 - bi * m1;
                M0 = ar * m0 + ai * r0 - br * m1 + bi * r1;
                R1 = br * r0 - bi * m0 + ar * r1 + ai * m1;
                M1 = br * m0 + bi * r0 + ar * m1 - ai * r1; break;
            }
            case OC_UNITARY: { /* :168-177 */
                R0 = op[0] * r0 - op[1] * m0 + op[2] * r1 - op[3] * m1;
                M0 = op[0] * m0 + op[1] * r0 + op[2] * m1 + op[3] * r1;
                R1 = op[4] * r0 - op[5] * m0 + op[6] * r1 - op[7] * m1;
                M1 = op[4] * m0 + op[5] * r0 + op[6] * m1 + op[7] * r1; break;
            }
            default: R0 = r0; M0 = m0; R1 = r1; M1 = m1; break;
        }
        s[i0].re = R0; s[i0].im = M0; s[i1].re = R1; s[i1].im = M1;
    }
}

/* ---- state initialisation (src/cpp/QuEST_backend.cpp:21-47) ----------- */
void oc_init_classical(oc_amp *s, int n, int64_t k) {
    const int64_t N = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) { s[i].re = (i == k) ? 1.0 : 0.0; s[i].im = 0.0; }
}

void oc_init_plus(oc_amp *s, int n) {
    const int64_t N = (int64_t)1 << n;
    const double f = 1.0 / sqrt((double)N);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) { s[i].re = f; s[i].im = 0.0; }
}

/* ---- probabilities ------------------------------------------------------ */
double oc_prob_zero(const oc_amp *s, int n, int q) {
    const int64_t N = (int64_t)1 << n;
    double p = 0.0;
#pragma omp parallel for reduction(+ : p) schedule(static)
    for (int64_t i = 0; i < N; i++)
        if (!((i >> q) & 1)) p += s[i].re * s[i].re + s[i].im * s[i].im;
    return p;
}

double oc_prob_of_outcome(const oc_amp *s, int n, int q, int outcome) {
    double p0 = oc_prob_zero(s, n, q);
    return outcome == 1 ? 1.0 - p0 : p0;
}

double oc_total_prob(const oc_amp *s, int n) {
    const int64_t N = (int64_t)1 << n;
    double p = 0.0;
#pragma omp parallel for reduction(+ : p) schedule(static)
    for (int64_t i = 0; i < N; i++) p += s[i].re * s[i].re + s[i].im * s[i].im;
    return p;
}

/* ---- MT19937, restated from the published mt19937ar algorithm ---------- */
#define MT_N 624
#define MT_M 397
static uint32_t mt[MT_N];
static int mti = MT_N + 1;

void oc_mt_init_genrand(uint32_t seed) {
    mt[0] = seed;
    for (mti = 1; mti < MT_N; mti++)
        mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
}

void oc_mt_init_by_array(const uint64_t *key, int len) {
    oc_mt_init_genrand(19650218u);
    int i = 1, j = 0;
    for (int k = (MT_N > len ? MT_N : len); k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + (uint32_t)key[j] + (uint32_t)j;
        i++; j++;
        if (i >= MT_N) { mt[0] = mt[MT_N - 1]; i = 1; }
        if (j >= len) j = 0;
    }
    for (int k = MT_N - 1; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= MT_N) { mt[0] = mt[MT_N - 1]; i = 1; }
    }
    mt[0] = 0x80000000u;
}

uint32_t oc_mt_genrand_int32(void) {
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    uint32_t y;
    if (mti >= MT_N) {
        if (mti == MT_N + 1) oc_mt_init_genrand(5489u);
        int kk;
        for (kk = 0; kk < MT_N - MT_M; kk++) {
            y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
            mt[kk] = mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < MT_N - 1; kk++) {
            y = (mt[kk] & 0x80000000u) | (mt[kk + 1] & 0x7fffffffu);
            mt[kk] = mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (mt[MT_N - 1] & 0x80000000u) | (mt[0] & 0x7fffffffu);
        mt[MT_N - 1] = mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1u];
        mti = 0;
    }
    y = mt[mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

double oc_mt_genrand_real1(void) { return oc_mt_genrand_int32() * (1.0 / 4294967295.0); }

double oc_mt_genrand_res53(void) {
    uint32_t a = oc_mt_genrand_int32() >> 5, b = oc_mt_genrand_int32() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

/* ---- measurement (src/QuEST_common.c:97-115) ---------------------------- */
#define OC_REAL_EPS 1e-13

int oc_generate_outcome(double zeroProb, double *outcomeProb) {
    int outcome;
    if (zeroProb < OC_REAL_EPS) outcome = 1;
    else if (1 - zeroProb < OC_REAL_EPS) outcome = 0;
    else outcome = (oc_mt_genrand_real1() > zeroProb);
    *outcomeProb = (outcome == 0) ? zeroProb : 1 - zeroProb;
    return outcome;
}

/* collapseToOutcome: zero the other branch, rescale by 1/sqrt(prob). */
double oc_collapse(oc_amp *s, int n, int q, int outcome) {
    double prob = oc_prob_of_outcome(s, n, q, outcome);
    double op[8] = {(double)outcome, 1.0 / sqrt(prob), 0, 0, 0, 0, 0, 0};
    oc_apply(s, n, OC_COLLAPSE, (uint64_t)1 << q, 0, op);
    return prob;
}

int oc_measure(oc_amp *s, int n, int q, double *outcomeProb) {
    double p0 = oc_prob_of_outcome(s, n, q, 0);
    double prob;
    int outcome = oc_generate_outcome(p0, &prob);
    double op[8] = {(double)outcome, 1.0 / sqrt(prob), 0, 0, 0, 0, 0, 0};
    oc_apply(s, n, OC_COLLAPSE, (uint64_t)1 << q, 0, op);
    if (outcomeProb) *outcomeProb = prob;
    return outcome;
}

/* ---- API-level gate encoding --------------------------------------------
 * oc_gate(name, qubits, params) restates how each QuEST API call is encoded
 * (src/cpp/QuEST_distributed.cpp:118-364, src/QuEST.c:93-423) and applies it
 * at once.  Returns 0 on success, -1 for an unknown name. */
static void apply1(oc_amp *s, int n, int type, int t, uint64_t cm, const double *op) {
    oc_apply(s, n, type, (uint64_t)1 << t, cm, op);
}
static void diag(oc_amp *s, int n, int type, uint64_t cm, const double *op) {
    oc_apply(s, n, type, 0, cm, op);
}
static void unitary_m(oc_amp *s, int n, int t, uint64_t cm, const double m[8]) {
    apply1(s, n, OC_UNITARY, t, cm, m);
}

int oc_gate(oc_amp *s, int n, const char *name, const int *q, int nq, const double *p) {
    double op[8] = {0};
    uint64_t cm = 0;
#define IS(x) (strcmp(name, x) == 0)
    if (IS("hadamard")) apply1(s, n, OC_H, q[0], 0, op);
    else if (IS("pauliX")) apply1(s, n, OC_X, q[0], 0, op);
    else if (IS("pauliY")) apply1(s, n, OC_Y, q[0], 0, op);
    else if (IS("pauliZ")) diag(s, n, OC_Z, 1ull << q[0], op);
    else if (IS("sGate")) diag(s, n, OC_S, 1ull << q[0], op);
    else if (IS("tGate")) diag(s, n, OC_T, 1ull << q[0], op);
    else if (IS("phaseShift")) { op[0] = p[0]; diag(s, n, OC_PHASE_SHIFT, 1ull << q[0], op); }
    else if (IS("controlledPhaseShift")) { op[0] = p[0]; diag(s, n, OC_PHASE_SHIFT, (1ull << q[0]) | (1ull << q[1]), op); }
    else if (IS("controlledPhaseFlip")) diag(s, n, OC_Z, (1ull << q[0]) | (1ull << q[1]), op);
    else if (IS("multiControlledPhaseFlip")) { for (int i = 0; i < nq; i++) cm |= 1ull << q[i]; diag(s, n, OC_Z, cm, op); }
    else if (IS("multiControlledPhaseShift")) { for (int i = 0; i < nq; i++) cm |= 1ull << q[i]; op[0] = p[0]; diag(s, n, OC_PHASE_SHIFT, cm, op); }
    else if (IS("rotateX")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_X, q[0], 0, op); }
    else if (IS("rotateY")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_Y, q[0], 0, op); }
    else if (IS("rotateZ")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_Z, q[0], 0, op); }
    else if (IS("controlledRotateX")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_X, q[1], 1ull << q[0], op); }
    else if (IS("controlledRotateY")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_Y, q[1], 1ull << q[0], op); }
    else if (IS("controlledRotateZ")) { op[0] = p[0] * 0.5; apply1(s, n, OC_ROTATE_Z, q[1], 1ull << q[0], op); }
    else if (IS("controlledNot")) apply1(s, n, OC_X, q[1], 1ull << q[0], op);
    else if (IS("controlledPauliY")) apply1(s, n, OC_Y, q[1], 1ull << q[0], op);
    else if (IS("compactUnitary") || IS("controlledCompactUnitary")) {
        int c = IS("controlledCompactUnitary");
        op[0] = p[0]; op[1] = p[1]; op[2] = p[2]; op[3] = p[3];
        apply1(s, n, OC_COMPACT, q[c], c ? 1ull << q[0] : 0, op);
    } else if (IS("unitary") || IS("controlledUnitary")) {
        int c = IS("controlledUnitary");
        unitary_m(s, n, q[c], c ? 1ull << q[0] : 0, p);
    } else if (IS("multiControlledUnitary")) {
        /* qubits = controls..., target last */
        for (int i = 0; i < nq - 1; i++) cm |= 1ull << q[i];
        unitary_m(s, n, q[nq - 1], cm, p);
    } else if (IS("rotateAroundAxis") || IS("controlledRotateAroundAxis")) {
        /* src/QuEST_common.c:56-63: angle, axis(x,y,z) -> compact(alpha, beta) */
        int c = IS("controlledRotateAroundAxis");
        double mag = sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
        double ux = p[1] / mag, uy = p[2] / mag, uz = p[3] / mag;
        op[0] = cos(p[0] / 2.0);
        op[1] = -sin(p[0] / 2.0) * uz;
        op[2] = sin(p[0] / 2.0) * uy;
        op[3] = -sin(p[0] / 2.0) * ux;
        apply1(s, n, OC_COMPACT, q[c], c ? 1ull << q[0] : 0, op);
    } else if (IS("u1Gate")) { /* src/QuEST.c:320-333 */
        double m[8] = {1, 0, 0, 0, 0, 0, cos(p[0]), sin(p[0])};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("u2Gate")) { /* src/QuEST.c:335-351 */
        double is2 = 1 / sqrt(2);
        double m[8] = {is2, 0, -cos(p[1]) * is2, -sin(p[1]) * is2,
                       cos(p[0]) * is2, sin(p[0]) * is2, cos(p[0] + p[1]) * is2, sin(p[0] + p[1]) * is2};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("u3Gate")) { /* src/QuEST.c:353-370 */
        double ct = cos(p[0] / 2), st = sin(p[0] / 2);
        double m[8] = {ct, 0, -cos(p[2]) * st, -sin(p[2]) * st,
                       cos(p[1]) * st, sin(p[1]) * st, cos(p[1] + p[2]) * ct, sin(p[1] + p[2]) * ct};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("SqX")) { /* src/QuEST.c:372-385 */
        double is2 = 1 / sqrt(2);
        double m[8] = {is2, 0, 0, -is2, 0, -is2, is2, 0};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("SqY")) { /* src/QuEST.c:387-400 */
        double is2 = 1 / sqrt(2);
        double m[8] = {is2, 0, -is2, 0, is2, 0, is2, 0};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("SqW")) { /* src/QuEST.c:402-415 */
        double is2 = 1 / sqrt(2);
        double m[8] = {is2, 0, -0.5, -0.5, 0.5, -0.5, is2, 0};
        unitary_m(s, n, q[0], 0, m);
    } else if (IS("fSim")) { /* src/QuEST.c:418-423 */
        int a = q[0], b = q[1];
        double ps[1] = {-p[1]};
        int qa[2] = {a, b}, qb[2] = {b, a};
        double rx[1] = {p[0] * 2};
        oc_gate(s, n, "controlledPhaseShift", qa, 2, ps);
        oc_gate(s, n, "controlledNot", qa, 2, NULL);
        oc_gate(s, n, "controlledRotateX", qb, 2, rx);
        oc_gate(s, n, "controlledNot", qa, 2, NULL);
    } else if (IS("swapGate")) {
        /* textbook SWAP as three CNOTs (the reference declares swapGate,
         * include/QuEST.h:1541, but never defines it) */
        int ab[2] = {q[0], q[1]}, ba[2] = {q[1], q[0]};
        oc_gate(s, n, "controlledNot", ab, 2, NULL);
        oc_gate(s, n, "controlledNot", ba, 2, NULL);
        oc_gate(s, n, "controlledNot", ab, 2, NULL);
    } else return -1;
#undef IS
    return 0;
}

int oc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
