/*
 * The operations table: creation, the mandatory-operation check, and the
 * column-by-column defaults a backend inherits for every multivec operation
 * it does not provide.
 *
 * Behaviour follows reference src/gcge_ops.c:51-398 and :773-975:
 *   - MatDotVec, VecAxpby, VecInnerProd, VecLocalInnerProd, VecCreateByVec,
 *     VecCreateByMat, VecDestroy and VecSetRandomValue are mandatory;
 *     GCGE_OPS_Setup reports the first missing one and returns 1;
 *   - default multivec operations loop over columns with the single-vector
 *     operations; MultiVecLinearComb ignores alpha and applies beta with the
 *     first column only; the "sym" inner product takes the upper triangle
 *     and mirrors it; MultiVecSwap exchanges the column handles.
 * The device backend overrides all multivec operations with block kernels,
 * so these defaults serve third-party single-vector backends.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

/* dense defaults (gcge_dense.c) */
void GCGE_Default_DenseMatDotDenseMat(char *, char *, GCGE_INT *, GCGE_INT *, GCGE_INT *,
                                      GCGE_DOUBLE *, GCGE_DOUBLE *, GCGE_INT *, GCGE_DOUBLE *,
                                      GCGE_INT *, GCGE_DOUBLE *, GCGE_DOUBLE *, GCGE_INT *);
void GCGE_Default_DenseSymMatDotDenseMat(char *, char *, GCGE_INT *, GCGE_INT *, GCGE_DOUBLE *,
                                         GCGE_DOUBLE *, GCGE_INT *, GCGE_DOUBLE *, GCGE_INT *,
                                         GCGE_DOUBLE *, GCGE_DOUBLE *, GCGE_INT *);
void GCGE_Default_DenseMatEigenSolver(GCGE_DOUBLE *, GCGE_INT, GCGE_INT, GCGE_DOUBLE *,
                                      GCGE_DOUBLE *, GCGE_INT, GCGE_INT, GCGE_INT, GCGE_INT *,
                                      GCGE_DOUBLE *);
GCGE_DOUBLE GCGE_Default_ArrayDotArray(GCGE_DOUBLE *, GCGE_DOUBLE *, GCGE_INT);
GCGE_DOUBLE GCGE_Default_ArrayNorm(GCGE_DOUBLE *, GCGE_INT);
void GCGE_Default_ArrayAXPBY(GCGE_DOUBLE, GCGE_DOUBLE *, GCGE_DOUBLE, GCGE_DOUBLE *, GCGE_INT);
void GCGE_Default_ArrayCopy(GCGE_DOUBLE *, GCGE_DOUBLE *, GCGE_INT);
void GCGE_Default_ArrayScale(GCGE_DOUBLE, GCGE_DOUBLE *, GCGE_INT);

static void def_get(void **V, GCGE_INT j, void **x, GCGE_OPS *ops)
{
    (void)ops;
    *x = V[j];
}

static void def_restore(void **V, GCGE_INT j, void **x, GCGE_OPS *ops)
{
    (void)ops;
    V[j] = *x;
}

static void def_mv_create_by_vec(void ***mv, GCGE_INT n, void *vec, GCGE_OPS *ops)
{
    *mv = (void **)malloc((size_t)(n > 0 ? n : 1) * sizeof(void *));
    for (GCGE_INT i = 0; i < n; ++i) ops->VecCreateByVec(*mv + i, vec, ops);
}

static void def_mv_create_by_mat(void ***mv, GCGE_INT n, void *mat, GCGE_OPS *ops)
{
    *mv = (void **)malloc((size_t)(n > 0 ? n : 1) * sizeof(void *));
    for (GCGE_INT i = 0; i < n; ++i) ops->VecCreateByMat(*mv + i, mat, ops);
}

static void def_mv_create_by_mv(void ***mv, GCGE_INT n, void **init, GCGE_OPS *ops)
{
    *mv = (void **)malloc((size_t)(n > 0 ? n : 1) * sizeof(void *));
    for (GCGE_INT i = 0; i < n; ++i) ops->VecCreateByVec(*mv + i, init[0], ops);
}

static void def_mv_destroy(void ***mv, GCGE_INT n, GCGE_OPS *ops)
{
    if (!*mv) return;
    for (GCGE_INT i = 0; i < n; ++i) {
        ops->VecDestroy(&(*mv)[i], ops);
        (*mv)[i] = NULL;
    }
    free(*mv);
    *mv = NULL;
}

static void def_mv_random(void **mv, GCGE_INT start, GCGE_INT n, GCGE_OPS *ops)
{
    void *x;
    for (GCGE_INT i = 0; i < n; ++i) {
        ops->GetVecFromMultiVec(mv, start + i, &x, ops);
        ops->VecSetRandomValue(x, ops);
        ops->RestoreVecForMultiVec(mv, start + i, &x, ops);
    }
}

static void def_mat_dot_mv(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                           GCGE_OPS *ops)
{
    void *xs, *ys;
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i) {
        ops->GetVecFromMultiVec(x, start[0] + i, &xs, ops);
        ops->GetVecFromMultiVec(y, start[1] + i, &ys, ops);
        ops->MatDotVec(mat, xs, ys, ops);
        ops->RestoreVecForMultiVec(x, start[0] + i, &xs, ops);
        ops->RestoreVecForMultiVec(y, start[1] + i, &ys, ops);
    }
}

static void def_mat_t_dot_mv(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                             GCGE_OPS *ops)
{
    void *xs, *ys;
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i) {
        ops->GetVecFromMultiVec(x, start[0] + i, &xs, ops);
        ops->GetVecFromMultiVec(y, start[1] + i, &ys, ops);
        ops->MatTransposeDotVec(mat, xs, ys, ops);
        ops->RestoreVecForMultiVec(x, start[0] + i, &xs, ops);
        ops->RestoreVecForMultiVec(y, start[1] + i, &ys, ops);
    }
}

static void def_mv_axpby(GCGE_DOUBLE a, void **x, GCGE_DOUBLE b, void **y, GCGE_INT *start,
                         GCGE_INT *end, GCGE_OPS *ops)
{
    void *xs, *ys;
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i) {
        ops->GetVecFromMultiVec(x, start[0] + i, &xs, ops);
        ops->GetVecFromMultiVec(y, start[1] + i, &ys, ops);
        ops->VecAxpby(a, xs, b, ys, ops);
        ops->RestoreVecForMultiVec(x, start[0] + i, &xs, ops);
        ops->RestoreVecForMultiVec(y, start[1] + i, &ys, ops);
    }
}

static void def_mv_axpby_column(GCGE_DOUBLE a, void **x, GCGE_INT cx, GCGE_DOUBLE b, void **y,
                                GCGE_INT cy, GCGE_OPS *ops)
{
    void *xs, *ys;
    ops->GetVecFromMultiVec(x, cx, &xs, ops);
    if (x == y && cx == cy) {
        ys = xs;
    } else {
        ops->GetVecFromMultiVec(y, cy, &ys, ops);
    }
    ops->VecAxpby(a, xs, b, ys, ops);
    ops->RestoreVecForMultiVec(x, cx, &xs, ops);
    if (!(x == y && cx == cy)) ops->RestoreVecForMultiVec(y, cy, &ys, ops);
}

static void def_mv_lincomb(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a,
                           GCGE_INT lda, GCGE_INT if_Vec, GCGE_DOUBLE alpha, GCGE_DOUBLE beta,
                           GCGE_OPS *ops)
{
    (void)if_Vec;
    (void)alpha;
    void *xs, *ys;
    for (GCGE_INT iy = start[1]; iy < end[1]; ++iy) {
        const GCGE_DOUBLE *col = a + (size_t)(iy - start[1]) * lda;
        ops->GetVecFromMultiVec(y, iy, &ys, ops);
        ops->GetVecFromMultiVec(x, start[0], &xs, ops);
        ops->VecAxpby(col[0], xs, beta, ys, ops);
        ops->RestoreVecForMultiVec(x, start[0], &xs, ops);
        for (GCGE_INT ix = start[0] + 1; ix < end[0]; ++ix) {
            ops->GetVecFromMultiVec(x, ix, &xs, ops);
            ops->VecAxpby(col[ix - start[0]], xs, 1.0, ys, ops);
            ops->RestoreVecForMultiVec(x, ix, &xs, ops);
        }
        ops->RestoreVecForMultiVec(y, iy, &ys, ops);
    }
}

static void def_mv_inner(void **V, void **W, GCGE_DOUBLE *a, char *is_sym, GCGE_INT *start,
                         GCGE_INT *end, GCGE_INT lda, GCGE_INT if_Vec, GCGE_OPS *ops)
{
    (void)if_Vec;
    void *vs, *ws;
    if (strcmp(is_sym, "sym") == 0) {
        GCGE_INT len = end[1] - start[1];
        for (GCGE_INT iw = 0; iw < len; ++iw) {
            ops->GetVecFromMultiVec(W, iw + start[1], &ws, ops);
            for (GCGE_INT iv = 0; iv <= iw; ++iv) {
                ops->GetVecFromMultiVec(V, iv + start[0], &vs, ops);
                ops->VecLocalInnerProd(vs, ws, a + (size_t)iw * lda + iv, ops);
                ops->RestoreVecForMultiVec(V, iv + start[0], &vs, ops);
            }
            ops->RestoreVecForMultiVec(W, iw + start[1], &ws, ops);
        }
        for (GCGE_INT iw = 0; iw < len; ++iw)
            for (GCGE_INT iv = iw + 1; iv < len; ++iv)
                a[(size_t)iw * lda + iv] = a[(size_t)iv * lda + iw];
    } else {
        for (GCGE_INT iw = start[1]; iw < end[1]; ++iw) {
            ops->GetVecFromMultiVec(W, iw, &ws, ops);
            for (GCGE_INT iv = start[0]; iv < end[0]; ++iv) {
                ops->GetVecFromMultiVec(V, iv, &vs, ops);
                ops->VecLocalInnerProd(vs, ws, a + (size_t)(iw - start[1]) * lda + (iv - start[0]),
                                       ops);
                ops->RestoreVecForMultiVec(V, iv, &vs, ops);
            }
            ops->RestoreVecForMultiVec(W, iw, &ws, ops);
        }
    }
}

static void def_mv_swap(void **V1, void **V2, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops)
{
    void *vs, *ws;
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i) {
        ops->GetVecFromMultiVec(V1, start[0] + i, &vs, ops);
        ops->GetVecFromMultiVec(V2, start[1] + i, &ws, ops);
        ops->RestoreVecForMultiVec(V1, start[0] + i, &ws, ops);
        ops->RestoreVecForMultiVec(V2, start[1] + i, &vs, ops);
    }
}

void GCGE_OPS_Create(GCGE_OPS **ops)
{
    /* calloc: every pointer, including the backend context, starts NULL */
    *ops = (GCGE_OPS *)calloc(1, sizeof(GCGE_OPS));
}

#define REQUIRE(field)                                                        \
    do {                                                                      \
        if (ops->field == NULL) {                                             \
            printf("ERROR: Please provide the " #field " operation!\n");      \
            return 1;                                                         \
        }                                                                     \
    } while (0)
#define DEFAULT(field, fn)                  \
    do {                                    \
        if (ops->field == NULL) ops->field = fn; \
    } while (0)

GCGE_INT GCGE_OPS_Setup(GCGE_OPS *ops)
{
    REQUIRE(MatDotVec);
    REQUIRE(VecAxpby);
    REQUIRE(VecInnerProd);
    REQUIRE(VecLocalInnerProd);
    REQUIRE(VecCreateByVec);
    REQUIRE(VecCreateByMat);
    REQUIRE(VecDestroy);
    REQUIRE(VecSetRandomValue);
    DEFAULT(GetVecFromMultiVec, def_get);
    DEFAULT(RestoreVecForMultiVec, def_restore);
    DEFAULT(MultiVecCreateByVec, def_mv_create_by_vec);
    DEFAULT(MultiVecCreateByMat, def_mv_create_by_mat);
    DEFAULT(MultiVecCreateByMultiVec, def_mv_create_by_mv);
    DEFAULT(MultiVecDestroy, def_mv_destroy);
    DEFAULT(MultiVecSetRandomValue, def_mv_random);
    DEFAULT(MatDotMultiVec, def_mat_dot_mv);
    DEFAULT(MatTransposeDotMultiVec, def_mat_t_dot_mv);
    DEFAULT(MultiVecAxpby, def_mv_axpby);
    DEFAULT(MultiVecAxpbyColumn, def_mv_axpby_column);
    DEFAULT(MultiVecLinearComb, def_mv_lincomb);
    DEFAULT(MultiVecInnerProd, def_mv_inner);
    DEFAULT(MultiVecInnerProdLocal, def_mv_inner);
    DEFAULT(MultiVecSwap, def_mv_swap);
    DEFAULT(DenseMatEigenSolver, GCGE_Default_DenseMatEigenSolver);
    DEFAULT(DenseMatDotDenseMat, GCGE_Default_DenseMatDotDenseMat);
    DEFAULT(DenseSymMatDotDenseMat, GCGE_Default_DenseSymMatDotDenseMat);
    DEFAULT(ArrayDotArray, GCGE_Default_ArrayDotArray);
    DEFAULT(ArrayNorm, GCGE_Default_ArrayNorm);
    DEFAULT(ArrayAXPBY, GCGE_Default_ArrayAXPBY);
    DEFAULT(ArrayCopy, GCGE_Default_ArrayCopy);
    DEFAULT(ArrayScale, GCGE_Default_ArrayScale);
    return 0;
}

void GCGE_OPS_Free(GCGE_OPS **ops)
{
    free(*ops);
    *ops = NULL;
}

void GCGE_OPS_SetLinearSolverWorkspace(GCGE_OPS *ops, void *ws) { ops->linear_solver_workspace = ws; }

GCGE_OPS_EXT *GCGE_OPS_GetExt(GCGE_OPS *ops)
{
    GCGE_OPS_EXT *ext = (GCGE_OPS_EXT *)ops->ops;
    if (ext && ext->magic == GCGE_OPS_EXT_MAGIC) return ext;
    return NULL;
}
