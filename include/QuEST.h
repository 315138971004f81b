/* QuEST.h -- the QuEST C API served by the gfx950 state-vector engine.
 *
 * Drop-in for the reference's include/QuEST.h: the same type names, field
 * names and function signatures (include/QuEST.h:58-104 for the types,
 * :125-1549 for the functions), so the reference's user programs
 * (the testcases directory) compile unchanged against this header.
 *
 * Conventions kept from the reference:
 *   - Qureg and QuESTEnv are passed by value and carry an opaque backend
 *     pointer (include/QuEST.h:79-104).
 *   - Gate calls only record a gate; the circuit runs when a result is read
 *     (getAmp, calcProbOfOutcome, ...)  (src/cpp/QuEST_distributed.cpp:106-116).
 *   - Invalid arguments print a QuEST error and exit() with the error code
 *     (src/QuEST_validation.c:82-88).
 *
 * Functions the reference declares but never defines (measure, swapGate,
 * calcTotalProb, getRealAmp, ...) are implemented here.  The density-matrix,
 * noise, QASM and report families are neither declared nor implemented.
 */
#ifndef QUEST_H
#define QUEST_H

#include "QuEST_precision.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long long u64;

/** One complex number. */
typedef struct Complex {
    qreal real;
    qreal imag;
} Complex;

/** A 2x2 complex matrix. */
typedef struct ComplexMatrix2 {
    Complex r0c0, r0c1;
    Complex r1c0, r1c1;
} ComplexMatrix2;

/** A real 3-vector. */
typedef struct Vector {
    qreal x, y, z;
} Vector;

/** Execution environment (one process, one GPU). */
typedef struct {
    struct QuESTEnvValue *value;
    int rank;
} QuESTEnv;

/** A register of qubits; qubits are zero-based, qubit q is bit q of an index. */
typedef struct Qureg {
    int isDensityMatrix;
    u64 numQubitsRepresented;
    u64 numQubitsInStateVec;
    u64 numAmpsTotal;
    QuESTEnv env;
    struct QuregState *state;
} Qureg;

/* ---- environment ------------------------------------------------------ */
QuESTEnv createQuESTEnv(void);
void destroyQuESTEnv(QuESTEnv env);
void seedQuESTDefault(void);
void seedQuEST(unsigned long int *seedArray, int numSeeds);
int getQuEST_PREC(void);

/* ---- registers -------------------------------------------------------- */
Qureg createQureg(int numQubits, QuESTEnv env);
void destroyQureg(Qureg qureg, QuESTEnv env);
void searchQuregTaskPartitions(Qureg qureg, QuESTEnv env);
int getNumQubits(Qureg qureg);
int getNumAmps(Qureg qureg);

/* ---- state initialisation --------------------------------------------- */
void initZeroState(Qureg qureg);
void initPlusState(Qureg qureg);
void initClassicalState(Qureg qureg, long long int stateInd);
void initStateFromAmps(Qureg qureg, qreal *reals, qreal *imags);
void setAmps(Qureg qureg, long long int startInd, qreal *reals, qreal *imags, long long int numAmps);

/* ---- unitary gates ---------------------------------------------------- */
void phaseShift(Qureg qureg, const int targetQubit, qreal angle);
void controlledPhaseShift(Qureg qureg, const int idQubit1, const int idQubit2, qreal angle);
void multiControlledPhaseShift(Qureg qureg, int *controlQubits, int numControlQubits, qreal angle);
void controlledPhaseFlip(Qureg qureg, const int idQubit1, const int idQubit2);
void multiControlledPhaseFlip(Qureg qureg, int *controlQubits, int numControlQubits);
void sGate(Qureg qureg, const int targetQubit);
void tGate(Qureg qureg, const int targetQubit);
void compactUnitary(Qureg qureg, const int targetQubit, Complex alpha, Complex beta);
void unitary(Qureg qureg, const int targetQubit, ComplexMatrix2 u);
void rotateX(Qureg qureg, const int rotQubit, qreal angle);
void rotateY(Qureg qureg, const int rotQubit, qreal angle);
void rotateZ(Qureg qureg, const int rotQubit, qreal angle);
void rotateAroundAxis(Qureg qureg, const int rotQubit, qreal angle, Vector axis);
void controlledRotateX(Qureg qureg, const int controlQubit, const int targetQubit, qreal angle);
void controlledRotateY(Qureg qureg, const int controlQubit, const int targetQubit, qreal angle);
void controlledRotateZ(Qureg qureg, const int controlQubit, const int targetQubit, qreal angle);
void controlledRotateAroundAxis(Qureg qureg, const int controlQubit, const int targetQubit, qreal angle, Vector axis);
void controlledCompactUnitary(Qureg qureg, const int controlQubit, const int targetQubit, Complex alpha, Complex beta);
void controlledUnitary(Qureg qureg, const int controlQubit, const int targetQubit, ComplexMatrix2 u);
void multiControlledUnitary(Qureg qureg, int *controlQubits, const int numControlQubits, const int targetQubit, ComplexMatrix2 u);
void pauliX(Qureg qureg, const int targetQubit);
void pauliY(Qureg qureg, const int targetQubit);
void pauliZ(Qureg qureg, const int targetQubit);
void hadamard(Qureg qureg, const int targetQubit);
void controlledNot(Qureg qureg, const int controlQubit, const int targetQubit);
void controlledPauliY(Qureg qureg, const int controlQubit, const int targetQubit);
void swapGate(Qureg qureg, int qb1, int qb2);
void u1Gate(Qureg qureg, int targetQubit, qreal lambda);
void u2Gate(Qureg qureg, int targetQubit, qreal phi, qreal lambda);
void u3Gate(Qureg qureg, int targetQubit, qreal theta, qreal phi, qreal lambda);
void SqX(Qureg qureg, int targetQubit);
void SqY(Qureg qureg, int targetQubit);
void SqW(Qureg qureg, int targetQubit);
void fSim(Qureg qureg, int targetQubit1, int targetQubit2, qreal theta, qreal phi);

/* ---- reading results (each flushes the recorded circuit) --------------- */
Complex getAmp(Qureg qureg, long long int index);
qreal getRealAmp(Qureg qureg, long long int index);
qreal getImagAmp(Qureg qureg, long long int index);
qreal getProbAmp(Qureg qureg, long long int index);
qreal calcProbOfOutcome(Qureg qureg, const int measureQubit, int outcome);
qreal calcTotalProb(Qureg qureg);
Complex calcInnerProduct(Qureg bra, Qureg ket);       /* <bra|ket> */
qreal calcFidelity(Qureg qureg, Qureg pureState);     /* |<pureState|qureg>|^2 */

/* ---- measurement (seeded by seedQuEST / seedQuESTDefault) -------------- */
qreal collapseToOutcome(Qureg qureg, const int measureQubit, int outcome);
int measure(Qureg qureg, int measureQubit);
int measureWithStats(Qureg qureg, int measureQubit, qreal *outcomeProb);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_H */
