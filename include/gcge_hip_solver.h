/*
 * gcge_hip_solver.h -- solver set-up on the MI355X backend, the counterparts
 * of the reference CSR application's GCGE_SOLVER_SetCSROps and
 * GCGE_CSR_Solver_Init (reference app/csr/gcge_app_csr.c:86-124,
 * app/csr/gcge_app_csr.h:82-84).
 *
 * Header-only: the functions call the solver facade (gcge.h, reference
 * src/gcge_solver.h) and GCGE_HIP_SetOps, so a program links the solver core
 * it uses (this repository's libgcge_amd.so, or the reference's own core) and
 * libgcge_hip.so, and neither library depends on the other.
 */
#ifndef GCGE_HIP_SOLVER_H_
#define GCGE_HIP_SOLVER_H_

#include <stdlib.h>

#include "gcge.h"
#include "gcge_hip.h"

/* GCGE_SOLVER_SetCSROps (reference app/csr/gcge_app_csr.c:86-89) */
static inline void GCGE_SOLVER_SetHIPOps(GCGE_SOLVER *solver) { GCGE_HIP_SetOps(solver->ops); }

/* GCGE_CSR_Solver_Init (reference app/csr/gcge_app_csr.c:92-124): a solver for
 * A x = lambda B x (B may be NULL) on device matrices (GCGE_HIP_MatCreateCSR,
 * GCGE_HIP_MatCreateQ1), nev from num_eigenvalues unless it is -1, the other
 * parameters from the -gcge_* command-line flags, eigenvalues in a calloc'ed
 * host array and eigenvectors in one device multivec of nev columns made by
 * the table's MultiVecCreateByMat; set up and ready for GCGE_SOLVER_Solve.
 * GCGE_SOLVER_Free_All releases both (the matrices stay the caller's).
 * The backend must be initialised (GCGE_HIP_Init). */
static inline GCGE_SOLVER *GCGE_HIP_Solver_Init(void *A, void *B, int num_eigenvalues, int argc,
                                                char *argv[])
{
    GCGE_SOLVER *solver = NULL;
    GCGE_SOLVER_Create(&solver);
    if (num_eigenvalues != -1) solver->para->nev = num_eigenvalues;
    GCGE_PARA_SetFromCommandLine(solver->para, argc, argv);
    const int nev = solver->para->nev;
    GCGE_DOUBLE *eval = (GCGE_DOUBLE *)calloc((size_t)nev, sizeof(GCGE_DOUBLE));
    GCGE_SOLVER_SetMatA(solver, A);
    if (B != NULL) GCGE_SOLVER_SetMatB(solver, B);
    GCGE_SOLVER_SetHIPOps(solver);
    void **evec = NULL;
    solver->ops->MultiVecCreateByMat(&evec, nev, A, solver->ops);
    GCGE_SOLVER_SetEigenvalues(solver, eval);
    GCGE_SOLVER_SetEigenvectors(solver, evec);
    GCGE_SOLVER_Setup(solver);
    return solver;
}

#endif /* GCGE_HIP_SOLVER_H_ */
