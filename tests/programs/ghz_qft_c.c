/* A plain C user program of the QuEST API: GHZ + textbook QFT on n qubits,
 * then P(q = 1) for every qubit, the total probability and a few amplitudes
 * printed one per line for the test to check against the closed form. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "QuEST.h"

int main(int argc, char **argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 20;
    QuESTEnv env = createQuESTEnv();
    Qureg q = createQureg(n, env);
    hadamard(q, 0);
    for (int i = 0; i + 1 < n; i++) controlledNot(q, i, i + 1);
    for (int j = 0; j < n; j++) {
        hadamard(q, j);
        for (int m = 1; j + m < n; m++) controlledPhaseShift(q, j, j + m, M_PI / (double)(1 << m));
    }
    for (int i = 0; i < n / 2; i++) swapGate(q, i, n - 1 - i);
    for (int t = 0; t < n; t++) printf("p1 %d %.17g\n", t, calcProbOfOutcome(q, t, 1));
    printf("total %.17g\n", calcTotalProb(q));
    const long long idx[4] = {0, 1, 5, (1LL << n) - 1};
    for (int k = 0; k < 4; k++) {
        Complex a = getAmp(q, idx[k]);
        printf("amp %lld %.17g %.17g\n", idx[k], a.real, a.imag);
    }
    destroyQureg(q, env);
    destroyQuESTEnv(env);
    return 0;
}
