/*
 * gcge.h -- the GCG eigensolver core: parameters, workspace, solver object
 * and the algorithm entry points.
 *
 * This is the host side of the MI355X backend.  It mirrors the reference
 * solver API so that a program written against the reference keeps working:
 *   GCGE_SOLVER_*       reference src/gcge_solver.h:42-60
 *   GCGE_PARA           reference src/gcge_para.h:26-116 (same field names)
 *   GCGE_WORKSPACE      reference src/gcge_workspace.h:23-61
 *   GCGE_Solve          reference src/gcge_eigsol.c:65
 * The core reaches vector data only through the GCGE_OPS table
 * (gcge_ops.h).  A backend may also publish GCGE_OPS_EXT through
 * ops->ops; the core then uses the block operations in it where the
 * reference loops over single vectors with the same arithmetic meaning.
 */
#ifndef GCGE_H_
#define GCGE_H_

#include <stddef.h>
#include "gcge_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- optional block extensions a backend can publish in ops->ops ---- */
#define GCGE_OPS_EXT_MAGIC 0x47434745u /* "GCGE" */

typedef struct GCGE_OPS_EXT_ {
    unsigned int magic; /* GCGE_OPS_EXT_MAGIC */
    /* For columns j = start .. end-1 of X:
     *   rnorm[j-start]  = || A x_j - eval[j] B x_j ||_2   (B == NULL: identity)
     *   xnorm[j-start]  = || x_j ||_2
     *   bxnorm[j-start] = || B x_j ||_2                  (may be NULL)
     * Same quantities as the per-vector loop of the convergence check
     * (reference src/gcge_eigsol.c:689-710). */
    void (*MultiVecResidualNorms)(void *A, void *B, void **X, GCGE_INT start, GCGE_INT end,
                                  const GCGE_DOUBLE *eval, GCGE_DOUBLE *rnorm,
                                  GCGE_DOUBLE *xnorm, GCGE_DOUBLE *bxnorm, GCGE_OPS *ops);
    /* Block CG with the argument list and the arithmetic of the reference
     * GCGE_BCG (src/gcge_cg.c:302-501). */
    void (*BlockCG)(void *A, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B,
                    void **RHS, void **V, GCGE_INT *start, GCGE_INT *end,
                    GCGE_INT max_it, GCGE_DOUBLE rate, GCGE_OPS *ops,
                    void **CG_P, void **CG_W, GCGE_DOUBLE *dtmp, GCGE_INT *itmp,
                    GCGE_INT by_column);
    /* Wait for all queued device work (timers call it before reading the clock). */
    void (*Synchronize)(GCGE_OPS *ops);
    /* Y(:, ycols[c]) = M X(:, xcols[c]) for c < k: the column loop of
     * MatDotVec calls over listed columns, as one call. */
    void (*MatDotMultiVecCols)(void *mat, void **X, const GCGE_INT *xcols, void **Y,
                               const GCGE_INT *ycols, GCGE_INT k, GCGE_OPS *ops);
    /* Y(:, ycols[c]) = a[c] X(:, xcols[c]) + b[c] Y(:, ycols[c]) for c < k,
     * each column with the VecAxpby rules (a[c] == 0: X not read; b[c] == 0:
     * Y not read). */
    void (*MultiVecAxpbyCols)(const GCGE_DOUBLE *a, void **X, const GCGE_INT *xcols,
                              const GCGE_DOUBLE *b, void **Y, const GCGE_INT *ycols, GCGE_INT k,
                              GCGE_OPS *ops);
    /* The table's DenseMatEigenSolver contract (reference src/gcge_ops.c:401-432)
     * on the backend's own solver; returns 0 when it solved the problem,
     * nonzero to decline (the core then calls ops->DenseMatEigenSolver). */
    int (*DenseMatEigenSolver)(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT nrows, GCGE_DOUBLE *eval,
                               GCGE_DOUBLE *evec, GCGE_INT lde, GCGE_INT il, GCGE_INT iu);
    void *backend; /* backend private state */
    /* BlockCG whose solution columns leave as the correction to the values
     * they start from (w - w0 instead of w; the same iteration).  May be
     * NULL. */
    void (*BlockCGCorrection)(void *A, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B,
                              void **RHS, void **V, GCGE_INT *start, GCGE_INT *end,
                              GCGE_INT max_it, GCGE_DOUBLE rate, GCGE_OPS *ops,
                              void **CG_P, void **CG_W, GCGE_DOUBLE *dtmp, GCGE_INT *itmp,
                              GCGE_INT by_column);
    /* R(:, rcols[c]) = s1[c] B x - s2[c] A x for x = X(:, xcols[c]) (B x = x
     * when B is NULL), from the products of the last MultiVecResidualNorms
     * call on X, whose columns the caller has left unchanged since; returns
     * nonzero, doing nothing, when that call did not cover these columns of
     * this X with these A and B.  May be NULL. */
    int (*RhsFromResidual)(void *A, void *B, void **X, const GCGE_INT *xcols,
                           const GCGE_DOUBLE *s1, const GCGE_DOUBLE *s2, void **R,
                           const GCGE_INT *rcols, GCGE_INT k, GCGE_OPS *ops);
    /* BlockCG from the initial residual: RHS holds b - A x0 on entry and the
     * solution columns receive the correction x - x0 (x0 is not read).  May
     * be NULL. */
    void (*BlockCGFromResidual)(void *A, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B,
                                void **RHS, void **V, GCGE_INT *start, GCGE_INT *end,
                                GCGE_INT max_it, GCGE_DOUBLE rate, GCGE_OPS *ops,
                                void **CG_P, void **CG_W, GCGE_DOUBLE *dtmp, GCGE_INT *itmp,
                                GCGE_INT by_column);
    /* V(:, out:out+k) = V(:, in:in+m) C within one multivector, C m x k
     * (leading dimension ldc); the output columns may overlap the input
     * ones.  Returns nonzero, doing nothing, for shapes the backend cannot
     * combine in place.  May be NULL. */
    int (*MultiVecLinearCombInPlace)(void **V, GCGE_INT in, GCGE_INT m, const GCGE_DOUBLE *C,
                                     GCGE_INT ldc, GCGE_INT out, GCGE_INT k, GCGE_OPS *ops);
} GCGE_OPS_EXT;

/* ---- statistics (reference src/gcge_statistics.h:23-42) ---- */
typedef struct GCGE_PART_TIME_ {
    GCGE_DOUBLE x_axpy_time;
    GCGE_DOUBLE x_orth_time;
    GCGE_DOUBLE p_axpy_time;
    GCGE_DOUBLE p_orth_time;
    GCGE_DOUBLE w_line_time; /* linear solves */
    GCGE_DOUBLE w_orth_time;
    GCGE_DOUBLE rr_mat_time; /* Rayleigh-Ritz projection */
    GCGE_DOUBLE rr_eigen_time;
    GCGE_DOUBLE conv_time;
} GCGE_PART_TIME;

typedef struct GCGE_OPS_TIME_COUNT_ {
    GCGE_DOUBLE mv_time;
    GCGE_INT mv_count;
    GCGE_DOUBLE ip_time;
    GCGE_INT ip_count;
    GCGE_DOUBLE axpy_time;
    GCGE_INT axpy_count;
} GCGE_OPS_TIME_COUNT;

typedef struct GCGE_STATISTIC_PARA_ {
    GCGE_PART_TIME *part_time_total;
    GCGE_PART_TIME *part_time_one_iter;
    GCGE_OPS_TIME_COUNT *ops_time_count_total;
    GCGE_OPS_TIME_COUNT *ops_time_count_linear_solver;
} GCGE_STATISTIC_PARA;

/* ---- parameters (reference src/gcge_para.h:26-116) ---- */
typedef struct GCGE_ORTH_PARA_ {
    GCGE_DOUBLE orth_zero_tol;
    GCGE_DOUBLE reorth_tol;
    GCGE_DOUBLE scbgs_reorth_tol;
    GCGE_DOUBLE criterion_tol;
    GCGE_INT max_reorth_time;
    GCGE_INT scbgs_wself_max_reorth_time;
    GCGE_INT print_orth_zero;
    GCGE_INT x_orth_block_size;
    GCGE_INT w_orth_block_size;
    GCGE_INT max_direct_orth_length;
} GCGE_ORTH_PARA;

typedef struct GCGE_PARA_ {
    GCGE_INT nev;
    GCGE_INT ev_max_it;
    GCGE_INT block_size;
    GCGE_INT if_lobgcg;
    GCGE_INT num_init_evec;
    GCGE_DOUBLE ev_tol;
    /* "R": ||r||/(||x|| max(1,|lambda|)); "O": omega norm; "A" (anything
     * else): ||r||/||x||; "RB" (added): ||r||/(|lambda| ||Bx||) */
    char *conv_type;
    GCGE_DOUBLE conv_omega_norm;
    GCGE_INT num_unlock;
    GCGE_INT num_conv;
    GCGE_INT num_conv_continuous;
    char *eval_type;
    char *orth_type;
    char *w_orth_type;
    char *p_orth_type;
    char *x_orth_type;
    GCGE_ORTH_PARA *orth_para;
    GCGE_DOUBLE multi_tol;
    GCGE_DOUBLE multi_tol_for_lock;
    GCGE_INT if_use_cg;
    GCGE_INT cg_max_it;
    GCGE_DOUBLE cg_rate;
    GCGE_INT cg_type;
    GCGE_INT if_use_bcg;
    GCGE_INT use_bcg_continuous;
    GCGE_INT num_iter;
    GCGE_DOUBLE max_res;
    GCGE_DOUBLE max_ind;
    GCGE_DOUBLE min_res;
    GCGE_DOUBLE min_ind;
    GCGE_DOUBLE sum_res;
    GCGE_DOUBLE *res;
    GCGE_INT dirichlet_boundary;
    GCGE_INT use_mpi_bcast;
    GCGE_INT use_shift;
    GCGE_INT if_shift;
    GCGE_DOUBLE shift;
    char *partition_type;
    GCGE_INT print_eval;
    GCGE_INT print_cg_error;
    GCGE_INT print_orthzero;
    GCGE_INT print_para;
    GCGE_INT print_result;
    GCGE_INT print_matlab;
    GCGE_INT print_part_time;
    GCGE_INT print_final_part_time;
    GCGE_INT print_level;
    GCGE_INT print_conv;
    GCGE_INT print_split;
    GCGE_STATISTIC_PARA *stat_para;
    GCGE_INT opt_rr_eig_partly;
    GCGE_INT opt_bcast;
    GCGE_INT opt_allgatherv;
    GCGE_INT min_nev_each_proc;
    /* added: wall-clock seconds of the last GCGE_Solve, and the block
     * extensions switch (1: use GCGE_OPS_EXT when the backend has it) */
    GCGE_DOUBLE solve_time;
    GCGE_INT use_ops_ext;
    /* added: 1 keeps the converged prefix in the Rayleigh-Ritz step with its
     * computed coupling to the new directions (they still get no new W
     * directions); 0 (default) zeroes that coupling and solves only the
     * trailing block, as src/gcge_rayleighritz.c:507-805 does */
    GCGE_INT full_rr;
    /* added: 1 keeps the block CG scalars of every column under that column
     * (a plain CG per column); 0 (default) stores them by position in the
     * list of unconverged columns and reads beta by column, as
     * src/gcge_cg.c:405,467-476 does */
    GCGE_INT bcg_by_column;
    /* added: optional per-check record; row r (r < history_cap) holds
     * num_iter, num_conv_continuous, then eval[0..nev-1] after the r-th
     * convergence check; history_len counts the rows written */
    GCGE_DOUBLE *history;
    GCGE_INT history_cap;
    GCGE_INT history_len;
    /* added, w_orth_type "svqb" only: the second SVQB pass is skipped when the
     * first kept every column and the scaled Gram's condition number
     * lambda_max / lambda_min is below this (B-orthonormality error of one
     * pass ~ u times that number); 0 always runs two passes */
    GCGE_DOUBLE svqb_one_pass_cond;
    /* added, w_orth_type "svqb" only: eta > 0 returns W as the correction to
     * its starting vector (w - x or w - x / lambda, the same span once W is
     * projected against [X, P]) and runs a second projection pass only when
     * a column keeps less than eta of its B-norm in the first (Daniel,
     * Gragg, Kaufman, Stewart 1976); 0 forms W as the reference does and
     * runs up to max_reorth_time passes */
    GCGE_DOUBLE svqb_dgks_eta;
} GCGE_PARA;

/* ---- workspace (reference src/gcge_workspace.h:23-61) ---- */
typedef struct GCGE_WORKSPACE_ {
    GCGE_INT V_size;
    GCGE_INT V_tmp_size;
    GCGE_INT CG_p_size;
    GCGE_INT CG_r_size;
    GCGE_INT RitzVec_size;
    void **V;        /* [X, P, W] basis */
    void **V_tmp;    /* scratch columns */
    void **RitzVec;
    void **evec;     /* caller's eigenvector block (nev columns) */
    void **CG_p;     /* CG search directions */
    void **CG_r;
    GCGE_INT *unlock;
    GCGE_INT max_dim_x;
    GCGE_INT dim_x;
    GCGE_INT dim_xp;
    GCGE_INT dim_xpw;
    GCGE_INT last_dim_xpw;
    GCGE_INT last_dim_x;
    GCGE_INT conv_bs;
    GCGE_INT unconv_bs;
    GCGE_INT num_soft_locked_in_last_iter;
    void *dense_matrix;
    GCGE_DOUBLE *subspace_matrix;
    GCGE_DOUBLE *subspace_evec;
    GCGE_DOUBLE *subspace_dtmp;
    GCGE_INT *subspace_itmp;
    GCGE_DOUBLE *eval;
    GCGE_INT *orth_ind;
    /* added: 1 from a convergence check that formed the residual products
     * of evec on the backend until the next inner solve reads them */
    GCGE_INT residual_products_fresh;
} GCGE_WORKSPACE;

typedef struct GCGE_SOLVER_ {
    void *A;
    void *B;
    void **evec;
    GCGE_DOUBLE *eval;
    GCGE_PARA *para;
    GCGE_OPS *ops;
    GCGE_WORKSPACE *workspace;
} GCGE_SOLVER;

/* ---- solver facade (reference src/gcge_solver.c:25-128) ---- */
void GCGE_SOLVER_Create(GCGE_SOLVER **solver);
GCGE_SOLVER *GCGE_SOLVER_CreateByOps(void *A, void *B, GCGE_INT nev, GCGE_DOUBLE *eval,
                                     void **evec, GCGE_OPS *ops);
void GCGE_SOLVER_Free(GCGE_SOLVER **solver);
void GCGE_SOLVER_Free_Some(GCGE_SOLVER **solver);
void GCGE_SOLVER_Free_All(GCGE_SOLVER **solver);
void GCGE_SOLVER_SetMatA(GCGE_SOLVER *solver, void *A);
void GCGE_SOLVER_SetMatB(GCGE_SOLVER *solver, void *B);
void GCGE_SOLVER_SetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE *eval);
void GCGE_SOLVER_SetEigenvectors(GCGE_SOLVER *solver, void **evec);
void GCGE_SOLVER_SetNumEigen(GCGE_SOLVER *solver, GCGE_INT nev);
void GCGE_SOLVER_Setup(GCGE_SOLVER *solver);
void GCGE_SOLVER_SetOpsLinearSolverWorkspace(GCGE_SOLVER *solver, void *linear_solver_workspace);
void GCGE_SOLVER_GetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE **eval);
void GCGE_SOLVER_GetEigenvectors(GCGE_SOLVER *solver, void ***evec);
void GCGE_SOLVER_Solve(GCGE_SOLVER *solver);

/* ---- parameters (reference src/gcge_para.c) ---- */
void GCGE_PARA_Create(GCGE_PARA **para);
void GCGE_PARA_Free(GCGE_PARA **para);
GCGE_INT GCGE_PARA_SetFromCommandLine(GCGE_PARA *para, GCGE_INT argc, char **argv);
void GCGE_PARA_SetNumEigen(GCGE_PARA *para, GCGE_INT nev);
void GCGE_PARA_Setup(GCGE_PARA *para);
void GCGE_PrintParaInfo(GCGE_PARA *para);
void GCGE_PrintIterationInfo(GCGE_DOUBLE *eval, GCGE_PARA *para);
void GCGE_PrintFinalInfo(GCGE_DOUBLE *eval, GCGE_PARA *para);
void GCGE_STATISTIC_PARA_Create(GCGE_STATISTIC_PARA **stat_para);
void GCGE_STATISTIC_PARA_Free(GCGE_STATISTIC_PARA **stat_para);

/* ---- ops table (reference src/gcge_ops.c:773-985) ---- */
void GCGE_OPS_Create(GCGE_OPS **ops);
GCGE_INT GCGE_OPS_Setup(GCGE_OPS *ops);
void GCGE_OPS_Free(GCGE_OPS **ops);
void GCGE_OPS_SetLinearSolverWorkspace(GCGE_OPS *ops, void *linear_solver_workspace);
/* the block extensions published by the backend, or NULL */
GCGE_OPS_EXT *GCGE_OPS_GetExt(GCGE_OPS *ops);

/* ---- workspace (reference src/gcge_workspace.c:27-195) ---- */
void GCGE_WORKSPACE_Create(GCGE_WORKSPACE **workspace);
void GCGE_WORKSPACE_Setup(GCGE_WORKSPACE *workspace, GCGE_PARA *para, GCGE_OPS *ops, void *A);
void GCGE_WORKSPACE_Free(GCGE_WORKSPACE **workspace, GCGE_PARA *para, GCGE_OPS *ops);

/* ---- algorithm ---- */
void GCGE_Solve(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_PARA *para,
                GCGE_OPS *ops, GCGE_WORKSPACE *workspace);
void GCGE_CheckConvergence(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_WORKSPACE *workspace);
void GCGE_ComputeX(void **V, GCGE_DOUBLE *subspace_evec, void **X, GCGE_INT size_X,
                   GCGE_INT size_V, GCGE_OPS *ops, GCGE_WORKSPACE *workspace);
void GCGE_ComputeW(void *A, void *B, void **V, GCGE_DOUBLE *eval, GCGE_OPS *ops,
                   GCGE_PARA *para, GCGE_WORKSPACE *workspace);
void GCGE_ComputeP(GCGE_DOUBLE *subspace_evec, void **V, GCGE_OPS *ops, GCGE_PARA *para,
                   GCGE_WORKSPACE *workspace);
void GCGE_CG(void *Matrix, void *b, void *x, GCGE_OPS *ops, GCGE_PARA *para,
             void **V_tmp1, void **V_tmp2);
void GCGE_BCG(void *Matrix, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B, void **RHS,
              void **V, GCGE_INT *start, GCGE_INT *end, GCGE_INT max_it, GCGE_DOUBLE rate,
              GCGE_OPS *ops, void *V_tmp1, void **V_tmp2, void **V_tmp3,
              GCGE_DOUBLE *subspace_dtmp, GCGE_INT *subspace_itmp);
/* GCGE_BCG with the scalar bookkeeping selectable (by_column: see
 * GCGE_PARA.bcg_by_column) */
void GCGE_BlockCG(void *Matrix, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B, void **RHS,
                  void **V, GCGE_INT *start, GCGE_INT *end, GCGE_INT max_it, GCGE_DOUBLE rate,
                  GCGE_OPS *ops, void *V_tmp1, void **V_tmp2, void **V_tmp3,
                  GCGE_DOUBLE *subspace_dtmp, GCGE_INT *subspace_itmp, GCGE_INT by_column);
void GCGE_Orthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                             GCGE_PARA *para, void **V_tmp, GCGE_DOUBLE *d_tmp);
void GCGE_StableMultiOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                        GCGE_OPS *ops, GCGE_PARA *para,
                                        GCGE_WORKSPACE *workspace);
void GCGE_StableMultiOrthonormalizationPartly(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                              GCGE_OPS *ops, GCGE_PARA *para,
                                              GCGE_WORKSPACE *workspace);
void GCGE_SubOrthonormalization(void **V, GCGE_INT *start, GCGE_INT *end, void *B,
                                void **V_tmp, GCGE_DOUBLE *subspace_dtmp, GCGE_OPS *ops);
void GCGE_SVQBOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                                 GCGE_PARA *para, GCGE_WORKSPACE *ws);
void GCGE_MultiOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                  GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *workspace);
void GCGE_SubOrthonormalizationSelfBGS(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                       GCGE_PARA *para, GCGE_OPS *ops,
                                       GCGE_WORKSPACE *workspace);
void GCGE_OrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                       GCGE_INT start, GCGE_INT *end, GCGE_DOUBLE *B,
                                       GCGE_INT ldB, GCGE_ORTH_PARA *orth_para, GCGE_OPS *ops);
void GCGE_MultiOrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                            GCGE_INT start, GCGE_INT *end, void *B,
                                            GCGE_INT ldB, GCGE_ORTH_PARA *orth_para,
                                            GCGE_WORKSPACE *workspace, GCGE_OPS *ops);
void GCGE_StableMultiOrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                                  GCGE_INT start, GCGE_INT *end, void *B,
                                                  GCGE_INT ldB, GCGE_ORTH_PARA *orth_para,
                                                  GCGE_WORKSPACE *workspace, GCGE_OPS *ops);
void GCGE_ComputeSubspaceMatrixVTAW(void **V, void *A, GCGE_INT start_W, GCGE_INT start_V,
                                    GCGE_INT end_V, GCGE_DOUBLE *subspace_mat, GCGE_OPS *ops,
                                    void **workspace);
void GCGE_ComputeSubspaceMatrix(void *A, void **V, GCGE_OPS *ops, GCGE_PARA *para,
                                GCGE_WORKSPACE *workspace);
void GCGE_ComputeSubspaceEigenpairs(GCGE_DOUBLE *subspace_matrix, GCGE_DOUBLE *eval,
                                    GCGE_DOUBLE *subspace_evec, GCGE_OPS *ops, GCGE_PARA *para,
                                    GCGE_WORKSPACE *workspace);
void GCGE_SortEigenpairs(GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_INT nev, GCGE_INT ldv,
                         GCGE_DOUBLE *work);

/* ---- dense host routines used as the table defaults ---- */
/* all eigenpairs of the symmetric matrix in the upper triangle of a
 * (column-major, lda); ascending eval, vector j in evec[j*lde ..], lde >= n;
 * work holds n*n + 5n doubles */
void GCGE_Dense_SymEig(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT n, GCGE_DOUBLE *eval,
                       GCGE_DOUBLE *evec, GCGE_INT lde, GCGE_DOUBLE *work);

/* ---- one-call driver ------------------------------------------------- */
/* A backend as seen by the driver: its table filler and host copies of
 * multivec columns (column-major host arrays with leading dimension ld). */
typedef struct GCGE_BACKEND_ {
    void (*SetOps)(GCGE_OPS *ops);
    void (*MultiVecSetHost)(void **mv, GCGE_INT start, GCGE_INT ncols, const GCGE_DOUBLE *host,
                            GCGE_INT ld);
    void (*MultiVecGetHost)(void **mv, GCGE_INT start, GCGE_INT ncols, GCGE_DOUBLE *host,
                            GCGE_INT ld);
} GCGE_BACKEND;

/* Solve A x = lambda B x (B may be NULL) for nev pairs with the backend's
 * matrices A and B.  argv carries -gcge_* flags (parsed like the reference
 * command line).  x0 (n x nev, may be NULL) gives the initial vectors, in
 * which case num_init_evec = nev.  Outputs (each may be NULL): eval[nev],
 * evec[n x nev], res[nev] (last residuals), info[4] = {num_iter,
 * num_conv_continuous, solve seconds, block_size}, history[cap x (nev+2)].
 * The output arrays are sized by nev, so a -gcge_nev flag must repeat the
 * nev argument.  Returns 0, 1 when the operations table is incomplete, or 2
 * when a -gcge_nev flag differs from nev (nothing is solved). */
int GCGE_SolveWithBackend(const GCGE_BACKEND *backend, void *A, void *B, GCGE_INT n,
                          GCGE_INT nev, int argc, char **argv, const GCGE_DOUBLE *x0,
                          GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_DOUBLE *res,
                          GCGE_DOUBLE *info, GCGE_DOUBLE *history, GCGE_INT history_cap);

/* ---- utilities (reference src/gcge_utilities.c) ---- */
GCGE_DOUBLE GCGE_GetTime(void);
void GCGE_Printf(const char *fmt, ...);

#ifdef __cplusplus
}
#endif

#endif /* GCGE_H_ */
