/*
 * q1_solve -- a C program on the public interfaces only: the Q1 stiffness and
 * mass matrices of the E^3 mesh assembled on the device, the solver set up by
 * GCGE_HIP_Solver_Init (the device counterpart of the reference's
 * GCGE_CSR_Solver_Init), solved, and the eigenvalues printed one per line as
 * "eval <i> <value>".
 *
 *   q1_solve E nev [-gcge_* flags]
 */
#include <stdio.h>
#include <stdlib.h>

#include "gcge_hip_solver.h"

int main(int argc, char *argv[])
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s E nev [-gcge_* flags]\n", argv[0]);
        return 2;
    }
    const int E = atoi(argv[1]), nev = atoi(argv[2]);
    if (GCGE_HIP_Init(0) != 0) {
        fprintf(stderr, "no usable GPU: %s\n", GCGE_HIP_LastError());
        return 1;
    }
    void *K = GCGE_HIP_MatCreateQ1(E, NULL, 0, NULL);
    void *M = GCGE_HIP_MatCreateQ1(E, NULL, 1, NULL);
    if (!K || !M) {
        fprintf(stderr, "assembly failed: %s\n", GCGE_HIP_LastError());
        return 1;
    }
    /* the flags after E and nev, behind a program name */
    GCGE_SOLVER *solver = GCGE_HIP_Solver_Init(K, M, nev, argc - 2, argv + 2);
    GCGE_SOLVER_Solve(solver);
    for (int i = 0; i < nev; ++i) printf("eval %d %.17g\n", i, solver->eval[i]);
    printf("num_iter %d num_conv %d\n", (int)solver->para->num_iter,
           (int)solver->para->num_conv_continuous);
    GCGE_SOLVER_Free_All(&solver);
    GCGE_HIP_MatDestroy(K);
    GCGE_HIP_MatDestroy(M);
    GCGE_HIP_Finalize();
    return 0;
}
