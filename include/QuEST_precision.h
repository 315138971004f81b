/* QuEST_precision.h -- compile-time choice of the amplitude precision.
 *
 * Same contract as the reference's include/QuEST_precision.h:18-64:
 *   QuEST_PREC == 1  ->  qreal is float,  REAL_EPS 1e-5
 *   QuEST_PREC == 2  ->  qreal is double, REAL_EPS 1e-13   (default)
 * Unlike the reference (which defines qComplex only for PREC 2, through
 * cuComplex.h), both precisions compile here and no GPU vendor header is
 * pulled in: the complex amplitude type is a plain struct of two qreals.
 * Quad precision (PREC 4) has no device arithmetic on gfx950 and is refused.
 */
#ifndef QUEST_PRECISION_H
#define QUEST_PRECISION_H

#include <math.h>

#ifndef QuEST_PREC
#define QuEST_PREC 2
#endif

#if QuEST_PREC == 1
#define qreal float
#define REAL_STRING_FORMAT "%.8f"
#define REAL_QASM_FORMAT "%.8g"
#define REAL_EPS 1e-5
#define absReal(X) fabs(X)
#elif QuEST_PREC == 2
#define qreal double
#define REAL_STRING_FORMAT "%.14f"
#define REAL_QASM_FORMAT "%.14g"
#define REAL_EPS 1e-13
#define absReal(X) fabs(X)
#else
#error "QuEST_PREC must be 1 (float) or 2 (double) for the gfx950 engine"
#endif

/* Interleaved complex amplitude as stored in device memory: {re, im}. */
typedef struct qcomplex {
    qreal re;
    qreal im;
} qcomplex;

#endif /* QUEST_PRECISION_H */
