/*
 * gcge_ops.h -- the GCGE operations table.
 *
 * The GCG eigensolver core never touches matrix or vector memory.  Every
 * large-scale operation goes through this table of function pointers, and a
 * backend fills it.  The member order, the member types and the start/end
 * range conventions are kept identical to the reference table
 * (src/gcge_ops.h:29-150) so that a driver compiled against the reference
 * header can be handed a table filled by GCGE_HIP_SetOps (gcge_hip.h).
 *
 * Conventions (reference src/gcge_ops.c:114-309):
 *   - a "multivec" is an opaque void** handle; column j is obtained with
 *     GetVecFromMultiVec and must be given back with RestoreVecForMultiVec;
 *   - start[2]/end[2] are half-open column ranges, [0] for the first
 *     multivec argument and [1] for the second;
 *   - dense arrays (inner-product results, linear-combination coefficients,
 *     Rayleigh-Ritz matrices) are host memory, column-major, with an explicit
 *     leading dimension;
 *   - every operation is synchronous from the caller's point of view: host
 *     outputs are valid when the call returns.
 */
#ifndef GCGE_OPS_H_
#define GCGE_OPS_H_

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GCGE_DOUBLE
#define GCGE_DOUBLE double
#endif
#ifndef GCGE_INT
#define GCGE_INT int
#endif

struct GCGE_OPS_;
typedef struct GCGE_OPS_ GCGE_OPS;

struct GCGE_OPS_ {
    /* backend context; the core never reads it */
    void *ops;

    /* single vectors ------------------------------------------------ */
    void (*VecSetRandomValue)(void *vec, GCGE_OPS *ops);
    /* r = mat * x */
    void (*MatDotVec)(void *mat, void *x, void *r, GCGE_OPS *ops);
    /* r = mat^T * x */
    void (*MatTransposeDotVec)(void *mat, void *x, void *r, GCGE_OPS *ops);
    /* y = a x + b y */
    void (*VecAxpby)(GCGE_DOUBLE a, void *x, GCGE_DOUBLE b, void *y, GCGE_OPS *ops);
    /* *value_ip = x^T y (global) */
    void (*VecInnerProd)(void *x, void *y, GCGE_DOUBLE *value_ip, GCGE_OPS *ops);
    /* *value_ip = x^T y (this process's rows only) */
    void (*VecLocalInnerProd)(void *x, void *y, GCGE_DOUBLE *value_ip, GCGE_OPS *ops);
    void (*VecCreateByVec)(void **des_vec, void *src_vec, GCGE_OPS *ops);
    void (*VecCreateByMat)(void **vec, void *mat, GCGE_OPS *ops);
    void (*VecDestroy)(void **vec, GCGE_OPS *ops);

    /* column views of a multivec */
    void (*GetVecFromMultiVec)(void **V, GCGE_INT j, void **x, GCGE_OPS *ops);
    void (*RestoreVecForMultiVec)(void **V, GCGE_INT j, void **x, GCGE_OPS *ops);

    /* multigrid hooks (used by layers above the solver only) */
    void (*MultiGridCreate)(void ***A_array, void ***B_array, void ***P_array,
                            GCGE_INT *num_levels, void *A, void *B, GCGE_OPS *ops);
    void (*MultiGridDestroy)(void ***A_array, void ***B_array, void ***P_array,
                             GCGE_INT *num_levels, GCGE_OPS *ops);
    void (*VecFromItoJ)(void **P_array, GCGE_INT level_i, GCGE_INT level_j,
                        void *vec_i, void *vec_j, void **vec_tmp, GCGE_OPS *ops);
    void (*MultiVecFromItoJ)(void **P_array, GCGE_INT level_i, GCGE_INT level_j,
                             void **multi_vec_i, void **multi_vec_j, void ***multi_vec_tmp,
                             GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops);

    /* user linear solvers for the W equations */
    void (*LinearSolver)(void *Matrix, void *b, void *x, GCGE_OPS *ops);
    void *linear_solver_workspace;
    /* x(:, start[1]:end[1]) = Matrix^{-1} b(:, start[0]:end[0]); five-element ranges */
    void (*MultiLinearSolver)(void *Matrix, void **b, void **x, int *start, int *end, GCGE_OPS *ops);

    void (*Orthonormalization)(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                               GCGE_DOUBLE orth_zero_tol, void **work);
    void (*DenseMatCreate)(void **densemat, GCGE_INT nrows, GCGE_INT ncols);
    void (*DenseMatDestroy)(void **mat);

    /* multivecs ------------------------------------------------------ */
    void (*MultiVecCreateByVec)(void ***multi_vec, GCGE_INT n_vec, void *vec, GCGE_OPS *ops);
    void (*MultiVecCreateByMat)(void ***multi_vec, GCGE_INT n_vec, void *mat, GCGE_OPS *ops);
    void (*MultiVecCreateByMultiVec)(void ***multi_vec, GCGE_INT n_vec, void **init_vec, GCGE_OPS *ops);
    void (*MultiVecDestroy)(void ***multi_vec, GCGE_INT n_vec, GCGE_OPS *ops);
    /* random values in columns start .. start+n_vec-1 */
    void (*MultiVecSetRandomValue)(void **multi_vec, GCGE_INT start, GCGE_INT n_vec, GCGE_OPS *ops);
    /* y(:, start[1]:end[1]) = mat * x(:, start[0]:end[0]) */
    void (*MatDotMultiVec)(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops);
    void (*MatTransposeDotMultiVec)(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops);
    /* y(:, start[1]:end[1]) = a x(:, start[0]:end[0]) + b y(:, start[1]:end[1]) */
    void (*MultiVecAxpby)(GCGE_DOUBLE a, void **x, GCGE_DOUBLE b, void **y,
                          GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops);
    /* y(:, col_y) = a x(:, col_x) + b y(:, col_y) */
    void (*MultiVecAxpbyColumn)(GCGE_DOUBLE a, void **x, GCGE_INT col_x, GCGE_DOUBLE b,
                                void **y, GCGE_INT col_y, GCGE_OPS *ops);
    /* y(:, j) = beta y(:, j) + sum_i x(:, i) a[(j-start[1])*lda + (i-start[0])];
     * alpha is accepted and ignored, as in the reference default */
    void (*MultiVecLinearComb)(void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                               GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT if_Vec,
                               GCGE_DOUBLE alpha, GCGE_DOUBLE beta, GCGE_OPS *ops);
    /* a[(j-start[1])*lda + (i-start[0])] = V(:, i)^T W(:, j); is_sym "sym" takes the
     * upper triangle (i <= j) of an equal-range product and mirrors it */
    void (*MultiVecInnerProd)(void **V, void **W, GCGE_DOUBLE *a, char *is_sym,
                              GCGE_INT *start, GCGE_INT *end, GCGE_INT lda,
                              GCGE_INT if_Vec, GCGE_OPS *ops);
    void (*MultiVecInnerProdLocal)(void **V, void **W, GCGE_DOUBLE *a, char *is_sym,
                                   GCGE_INT *start, GCGE_INT *end, GCGE_INT lda,
                                   GCGE_INT if_Vec, GCGE_OPS *ops);
    /* exchange V_1(:, start[0]:end[0]) and V_2(:, start[1]:end[1]) */
    void (*MultiVecSwap)(void **V_1, void **V_2, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops);
    void (*MultiVecPrint)(void **x, GCGE_INT n, GCGE_OPS *ops);

    void (*SetDirichletBoundary)(void **Vecs, GCGE_INT nev, void *A, void *B);

    /* dense (host) helpers ------------------------------------------- */
    /* eigenpairs il..iu (1-based, ascending) of the upper triangle of a */
    void (*DenseMatEigenSolver)(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT nrows,
                                GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_INT lde,
                                GCGE_INT il, GCGE_INT iu, GCGE_INT *iwork, GCGE_DOUBLE *dwork);
    /* dgemm-shaped: c = alpha op(a) op(b) + beta c */
    void (*DenseMatDotDenseMat)(char *transa, char *transb, GCGE_INT *nrows, GCGE_INT *ncols,
                                GCGE_INT *mid, GCGE_DOUBLE *alpha, GCGE_DOUBLE *a,
                                GCGE_INT *lda, GCGE_DOUBLE *b, GCGE_INT *ldb,
                                GCGE_DOUBLE *beta, GCGE_DOUBLE *c, GCGE_INT *ldc);
    /* dsymm-shaped: c = alpha a b + beta c with a symmetric */
    void (*DenseSymMatDotDenseMat)(char *side, char *uplo, GCGE_INT *nrows, GCGE_INT *ncols,
                                   GCGE_DOUBLE *alpha, GCGE_DOUBLE *a, GCGE_INT *lda,
                                   GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta,
                                   GCGE_DOUBLE *c, GCGE_INT *ldc);

    GCGE_DOUBLE (*ArrayDotArray)(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length);
    GCGE_DOUBLE (*ArrayNorm)(GCGE_DOUBLE *x, GCGE_INT length);
    void (*ArrayAXPBY)(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_DOUBLE b, GCGE_DOUBLE *y, GCGE_INT length);
    void (*ArrayCopy)(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length);
    void (*ArrayScale)(GCGE_DOUBLE alpha, GCGE_DOUBLE *a, GCGE_INT length);
};

#ifdef __cplusplus
}
#endif

#endif /* GCGE_OPS_H_ */
