/*
 * Workspace sizing and allocation, and the solver facade.
 *
 * Sizes follow reference src/gcge_workspace.c:61-139:
 *   max_dim_x  = nev + 20
 *   V          = max_dim_x + 2*block_size columns   ([X, P, W])
 *   V_tmp      = max_dim_x - nev + block_size columns
 *   CG_p       = block_size columns
 *   dense      = (max_dim_x + 2*block_size)^2 for the projected matrix and its
 *                eigenvectors, 2*m^2 + 40*m doubles and 100*m ints of scratch
 * The dense buffers keep the reference sizes and are reused with a leading
 * dimension that changes from one iteration to the next, as the reference
 * does; the Rayleigh-Ritz and P steps rely on that layout.
 * Facade: reference src/gcge_solver.c:25-128.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

void GCGE_WORKSPACE_Create(GCGE_WORKSPACE **workspace)
{
    *workspace = (GCGE_WORKSPACE *)calloc(1, sizeof(GCGE_WORKSPACE));
}

void GCGE_WORKSPACE_Setup(GCGE_WORKSPACE *ws, GCGE_PARA *para, GCGE_OPS *ops, void *A)
{
    GCGE_INT nev = para->nev, bs = para->block_size;
    GCGE_INT max_dim_x = nev + 20;
    ws->V_size = max_dim_x + 2 * bs;
    ws->V_tmp_size = max_dim_x - nev + bs;
    ws->CG_p_size = bs;
    ops->MultiVecCreateByMat(&ws->V, ws->V_size, A, ops);
    ops->MultiVecCreateByMat(&ws->V_tmp, ws->V_tmp_size, A, ops);
    ops->MultiVecCreateByMat(&ws->CG_p, ws->CG_p_size, A, ops);

    ws->max_dim_x = max_dim_x;
    ws->dim_x = nev;
    ws->dim_xp = 0;
    ws->dim_xpw = nev;
    ws->conv_bs = 0;
    ws->num_soft_locked_in_last_iter = 0;
    ws->unconv_bs = nev < bs ? nev : bs;

    size_t m = (size_t)(max_dim_x + 2 * bs);
    ws->eval = (GCGE_DOUBLE *)calloc(m, sizeof(GCGE_DOUBLE));
    ws->subspace_dtmp = (GCGE_DOUBLE *)calloc(2 * m * m + 40 * m, sizeof(GCGE_DOUBLE));
    ws->subspace_itmp = (GCGE_INT *)calloc(100 * m, sizeof(GCGE_INT));
    ws->subspace_evec = (GCGE_DOUBLE *)calloc(m * m, sizeof(GCGE_DOUBLE));
    ws->subspace_matrix = (GCGE_DOUBLE *)calloc(m * m, sizeof(GCGE_DOUBLE));
    ws->unlock = (GCGE_INT *)calloc((size_t)max_dim_x, sizeof(GCGE_INT));
    for (GCGE_INT i = 0; i < max_dim_x - para->num_conv; ++i) ws->unlock[i] = i + para->num_conv;
}

void GCGE_WORKSPACE_Free(GCGE_WORKSPACE **workspace, GCGE_PARA *para, GCGE_OPS *ops)
{
    (void)para;
    GCGE_WORKSPACE *ws = *workspace;
    if (!ws) return;
    free(ws->unlock);
    free(ws->subspace_matrix);
    free(ws->subspace_evec);
    free(ws->eval);
    free(ws->subspace_dtmp);
    free(ws->subspace_itmp);
    if (ws->V) {
        ops->MultiVecDestroy(&ws->V, ws->V_size, ops);
        ops->MultiVecDestroy(&ws->V_tmp, ws->V_tmp_size, ops);
        ops->MultiVecDestroy(&ws->CG_p, ws->CG_p_size, ops);
    }
    free(ws);
    *workspace = NULL;
}

/* ---- solver facade ---------------------------------------------------- */

void GCGE_SOLVER_Create(GCGE_SOLVER **solver)
{
    GCGE_SOLVER *s = (GCGE_SOLVER *)calloc(1, sizeof(GCGE_SOLVER));
    GCGE_PARA_Create(&s->para);
    GCGE_OPS_Create(&s->ops);
    GCGE_WORKSPACE_Create(&s->workspace);
    *solver = s;
}

GCGE_SOLVER *GCGE_SOLVER_CreateByOps(void *A, void *B, GCGE_INT nev, GCGE_DOUBLE *eval,
                                     void **evec, GCGE_OPS *ops)
{
    GCGE_SOLVER *s = (GCGE_SOLVER *)calloc(1, sizeof(GCGE_SOLVER));
    GCGE_PARA_Create(&s->para);
    s->para->nev = nev;
    s->ops = ops;
    GCGE_WORKSPACE_Create(&s->workspace);
    s->A = A;
    s->B = B;
    s->eval = eval;
    s->evec = evec;
    GCGE_SOLVER_Setup(s);
    return s;
}

void GCGE_SOLVER_Free(GCGE_SOLVER **solver)
{
    GCGE_WORKSPACE_Free(&(*solver)->workspace, (*solver)->para, (*solver)->ops);
    GCGE_PARA_Free(&(*solver)->para);
    GCGE_OPS_Free(&(*solver)->ops);
    free(*solver);
    *solver = NULL;
}

void GCGE_SOLVER_Free_Some(GCGE_SOLVER **solver)
{
    GCGE_WORKSPACE_Free(&(*solver)->workspace, (*solver)->para, (*solver)->ops);
    GCGE_PARA_Free(&(*solver)->para);
    free(*solver);
    *solver = NULL;
}

void GCGE_SOLVER_Free_All(GCGE_SOLVER **solver)
{
    GCGE_INT nev = (*solver)->para->nev;
    if ((*solver)->evec) (*solver)->ops->MultiVecDestroy(&(*solver)->evec, nev, (*solver)->ops);
    free((*solver)->eval);
    (*solver)->eval = NULL;
    GCGE_SOLVER_Free(solver);
}

void GCGE_SOLVER_Setup(GCGE_SOLVER *solver)
{
    GCGE_PARA_Setup(solver->para);
    GCGE_OPS_Setup(solver->ops);
    GCGE_WORKSPACE_Setup(solver->workspace, solver->para, solver->ops, solver->A);
}

void GCGE_SOLVER_SetMatA(GCGE_SOLVER *solver, void *A) { solver->A = A; }
void GCGE_SOLVER_SetMatB(GCGE_SOLVER *solver, void *B) { solver->B = B; }
void GCGE_SOLVER_SetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE *eval) { solver->eval = eval; }
void GCGE_SOLVER_SetEigenvectors(GCGE_SOLVER *solver, void **evec) { solver->evec = evec; }
void GCGE_SOLVER_SetNumEigen(GCGE_SOLVER *solver, GCGE_INT nev) { GCGE_PARA_SetNumEigen(solver->para, nev); }
void GCGE_SOLVER_SetOpsLinearSolverWorkspace(GCGE_SOLVER *solver, void *ws)
{
    GCGE_OPS_SetLinearSolverWorkspace(solver->ops, ws);
}
void GCGE_SOLVER_GetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE **eval) { *eval = solver->eval; }
void GCGE_SOLVER_GetEigenvectors(GCGE_SOLVER *solver, void ***evec) { *evec = solver->evec; }

void GCGE_SOLVER_Solve(GCGE_SOLVER *solver)
{
    GCGE_Solve(solver->A, solver->B, solver->eval, solver->evec, solver->para, solver->ops,
               solver->workspace);
}

/* ---- one-call driver ---------------------------------------------------- */
int GCGE_SolveWithBackend(const GCGE_BACKEND *backend, void *A, void *B, GCGE_INT n,
                          GCGE_INT nev, int argc, char **argv, const GCGE_DOUBLE *x0,
                          GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_DOUBLE *res,
                          GCGE_DOUBLE *info, GCGE_DOUBLE *history, GCGE_INT history_cap)
{
    GCGE_SOLVER *s;
    GCGE_SOLVER_Create(&s);
    s->para->nev = nev;
    s->para->print_para = 0;
    s->para->print_conv = 0;
    s->para->print_result = 0;
    s->para->print_final_part_time = 0;
    GCGE_PARA_SetFromCommandLine(s->para, argc, argv);
    if (s->para->nev != nev) {
        /* eval/evec/res/x0 are the caller's arrays of nev columns */
        fprintf(stderr, "GCGE_SolveWithBackend: -gcge_nev %d differs from nev %d\n",
                (int)s->para->nev, (int)nev);
        GCGE_SOLVER_Free(&s);
        return 2;
    }
    backend->SetOps(s->ops);
    if (GCGE_OPS_Setup(s->ops) != 0) {
        GCGE_SOLVER_Free(&s);
        return 1;
    }
    s->A = A;
    s->B = B;
    s->eval = (GCGE_DOUBLE *)calloc((size_t)nev, sizeof(GCGE_DOUBLE));
    s->ops->MultiVecCreateByMat(&s->evec, nev, A, s->ops);
    if (x0) {
        backend->MultiVecSetHost(s->evec, 0, nev, x0, n);
        s->para->num_init_evec = nev;
    }
    s->para->history = history;
    s->para->history_cap = history ? history_cap : 0;
    GCGE_SOLVER_Setup(s);
    GCGE_SOLVER_Solve(s);
    if (eval) memcpy(eval, s->eval, (size_t)nev * sizeof(GCGE_DOUBLE));
    if (evec) backend->MultiVecGetHost(s->evec, 0, nev, evec, n);
    if (res) memcpy(res, s->para->res, (size_t)nev * sizeof(GCGE_DOUBLE));
    if (info) {
        info[0] = s->para->num_iter;
        info[1] = s->para->num_conv_continuous;
        info[2] = s->para->solve_time;
        info[3] = s->para->block_size;
    }
    GCGE_SOLVER_Free_All(&s);
    return 0;
}
