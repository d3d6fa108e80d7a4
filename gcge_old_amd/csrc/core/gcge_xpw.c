/*
 * Ritz vectors X, inner-solve directions W and conjugate directions P.
 *
 *   GCGE_ComputeX  reference src/gcge_xpw.c:47-76
 *       X(:, nsl:nev) = V(:, nsl:dim_v) * C(nsl:dim_v, nsl:nev)
 *   GCGE_ComputeW  reference src/gcge_xpw.c:118-260
 *       rhs_i = (lambda_i + shift) B x_i with initial guess x_i when
 *       |lambda_i| <= 1, rhs_i = B x_i with guess x_i / (lambda_i + shift)
 *       otherwise; then MultiLinearSolver / LinearSolver / block CG / CG.
 *   GCGE_ComputeP  reference src/gcge_xpw.c:329-478
 *       P coefficients from the unconverged Ritz coefficients with their X
 *       rows zeroed, orthonormalised in the small space, then one linear
 *       combination of V into V_tmp and a swap into V.
 * When the backend publishes BlockCG it replaces the column-by-column
 * GCGE_BCG with the same argument list and arithmetic.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

void GCGE_ComputeX(void **V, GCGE_DOUBLE *subspace_evec, void **X, GCGE_INT size_X,
                   GCGE_INT size_V, GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    GCGE_INT sx = ws->num_soft_locked_in_last_iter;
    GCGE_INT start[2] = {sx, sx};
    GCGE_INT end[2] = {size_V, size_X};
    ws->residual_products_fresh = 0;  /* X changes */
    ops->MultiVecLinearComb(V, X, start, end, subspace_evec + (size_t)sx * size_V + sx, size_V, 0,
                            1.0, 0.0, ops);
}

static int compute_w_solve(void *A, void *B, void **V, GCGE_DOUBLE *eval, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_WORKSPACE *ws, int want_correction);

/* With the svqb W orthonormalisation and svqb_dgks_eta > 0, W leaves as the
 * correction to its starting vector: w - a x, a = 1 where |lambda| <= 1 and
 * 1 / (lambda + shift) elsewhere (the starting vector set below).  x is a
 * column of V, so [X, P, W] spans the same space and the projected W is the
 * same as with the reference's W; without the dominant x component (W is
 * close to a x once the pair converges) the projection loses no digits to
 * cancellation, which is what lets a single projection pass suffice. */
void GCGE_ComputeW(void *A, void *B, void **V, GCGE_DOUBLE *eval, GCGE_OPS *ops,
                   GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const int want = para->svqb_dgks_eta > 0.0 && strcmp(para->w_orth_type, "svqb") == 0;
    const int corrected = compute_w_solve(A, B, V, eval, ops, para, ws, want);
    ws->residual_products_fresh = 0;  /* the inner solve uses evec as scratch */
    const GCGE_INT w_start = ws->dim_xp;
    const GCGE_INT w_length = ws->unconv_bs;
    if (!want || corrected || w_length <= 0) return;
    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
    GCGE_DOUBLE *a = (GCGE_DOUBLE *)malloc(2 * (size_t)w_length * sizeof(GCGE_DOUBLE));
    GCGE_INT *cols = (GCGE_INT *)malloc(2 * (size_t)w_length * sizeof(GCGE_INT));
    GCGE_DOUBLE *b = a + w_length;
    for (GCGE_INT idx = 0; idx < w_length; ++idx) {
        const GCGE_INT xi = ws->unlock[idx];
        a[idx] = -(fabs(eval[xi]) <= 1.0 ? 1.0 : 1.0 / (eval[xi] + para->shift));
        b[idx] = 1.0;
        cols[idx] = xi;
        cols[w_length + idx] = w_start + idx;
    }
    if (para->use_ops_ext && ext && ext->MultiVecAxpbyCols) {
        ext->MultiVecAxpbyCols(a, V, cols, b, V, cols + w_length, w_length, ops);
    } else {
        for (GCGE_INT idx = 0; idx < w_length; ++idx) {
            void *xc, *wc;
            ops->GetVecFromMultiVec(V, cols[idx], &xc, ops);
            ops->GetVecFromMultiVec(V, cols[w_length + idx], &wc, ops);
            ops->VecAxpby(a[idx], xc, 1.0, wc, ops);
            ops->RestoreVecForMultiVec(V, cols[idx], &xc, ops);
            ops->RestoreVecForMultiVec(V, cols[w_length + idx], &wc, ops);
        }
    }
    free(a);
    free(cols);
}

/* returns 1 when the block CG already left W as the correction (the
 * backend's BlockCGCorrection, asked for by want_correction) */
static int compute_w_solve(void *A, void *B, void **V, GCGE_DOUBLE *eval, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_WORKSPACE *ws, int want_correction)
{
    const GCGE_INT w_start = ws->dim_xp;
    const GCGE_INT w_length = ws->unconv_bs;
    GCGE_INT *unlock = ws->unlock;
    void **V_tmp = ws->V_tmp;
    void *rhs, *xc, *wc;
    GCGE_INT mv_s[5], mv_e[5];
    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);

    /* ranges: [0] rhs in V_tmp, [1] solution in V, [2] CG_p, [3] scratch, [4] spare */
    mv_s[0] = 0; mv_e[0] = w_length;
    mv_s[1] = w_start; mv_e[1] = w_start + w_length;
    mv_s[2] = 0; mv_e[2] = w_length;
    mv_s[3] = 0; mv_e[3] = w_length;
    mv_s[4] = 0; mv_e[4] = w_length;

    /* the correction straight from the residual check's products: with
     * rhs = s1 B x and w0 = s2 x (as set below), r0 = rhs - A w0 =
     * s1 B x - s2 A x is formed from the B X and A X of the convergence check
     * (X = evec, unchanged since), and the block CG starts from it */
    if (want_correction && ws->residual_products_fresh && para->use_ops_ext && ext &&
        ext->RhsFromResidual &&
        ext->BlockCGFromResidual && w_length > 0 && para->if_use_bcg == 1 &&
        para->if_shift != 1 && !para->use_bcg_continuous && !ops->MultiLinearSolver &&
        !ops->LinearSolver) {
        GCGE_INT *cols = (GCGE_INT *)malloc(2 * (size_t)w_length * sizeof(GCGE_INT));
        GCGE_DOUBLE *s12 = (GCGE_DOUBLE *)malloc(2 * (size_t)w_length * sizeof(GCGE_DOUBLE));
        for (GCGE_INT idx = 0; idx < w_length; ++idx) {
            const GCGE_INT xi = unlock[idx];
            const int small = fabs(eval[xi]) <= 1.0;
            cols[idx] = xi;
            cols[w_length + idx] = idx;
            s12[idx] = small ? eval[xi] + para->shift : 1.0;
            s12[w_length + idx] = small ? 1.0 : 1.0 / (eval[xi] + para->shift);
        }
        const int rc = ext->RhsFromResidual(A, B, ws->evec, cols, s12, s12 + w_length, V_tmp,
                                            cols + w_length, w_length, ops);
        free(cols);
        free(s12);
        if (rc == 0) {
            ext->BlockCGFromResidual(A, para->if_shift, para->shift, B, V_tmp, V, mv_s, mv_e,
                                     para->cg_max_it, para->cg_rate, ops, ws->CG_p, ws->evec,
                                     ws->subspace_dtmp, ws->subspace_itmp, para->bcg_by_column);
            return 1;
        }
    }

    if (para->use_ops_ext && ext && ext->MatDotMultiVecCols && ext->MultiVecAxpbyCols &&
        w_length > 0) {
        /* the same per-column operations as the loop below, one call each:
         * rhs = B x (or x), rhs *= lambda + shift where |lambda| <= 1
         * (y = 1*y elsewhere), w = x or x / (lambda + shift) */
        GCGE_INT *cols = (GCGE_INT *)malloc(3 * (size_t)w_length * sizeof(GCGE_INT));
        GCGE_DOUBLE *coef = (GCGE_DOUBLE *)malloc(4 * (size_t)w_length * sizeof(GCGE_DOUBLE));
        GCGE_INT *xcols = cols, *rcols = cols + w_length, *wcols = cols + 2 * w_length;
        GCGE_DOUBLE *ra = coef, *rb = coef + w_length, *wa = coef + 2 * w_length,
                    *wb = coef + 3 * w_length;
        int any_small = 0;
        for (GCGE_INT idx = 0; idx < w_length; ++idx) {
            const GCGE_INT xi = unlock[idx];
            xcols[idx] = xi;
            rcols[idx] = idx;
            wcols[idx] = w_start + idx;
            const int small = fabs(eval[xi]) <= 1.0;
            any_small |= small;
            ra[idx] = 0.0;
            rb[idx] = small ? eval[xi] + para->shift : 1.0;
            wa[idx] = small ? 1.0 : 1.0 / (eval[xi] + para->shift);
            wb[idx] = 0.0;
        }
        if (B == NULL) {
            /* rhs = x: a = 1, b = 0 (a copy) */
            GCGE_DOUBLE *ones = (GCGE_DOUBLE *)malloc((size_t)w_length * sizeof(GCGE_DOUBLE));
            for (GCGE_INT idx = 0; idx < w_length; ++idx) ones[idx] = 1.0;
            ext->MultiVecAxpbyCols(ones, V, xcols, wb, V_tmp, rcols, w_length, ops);
            free(ones);
        } else {
            ext->MatDotMultiVecCols(B, V, xcols, V_tmp, rcols, w_length, ops);
        }
        /* the scaling pass only where some |lambda| <= 1 (the column loop
         * below scales those columns only; y = 0 x + 1 y elsewhere) */
        if (any_small) ext->MultiVecAxpbyCols(ra, V_tmp, rcols, rb, V_tmp, rcols, w_length, ops);
        ext->MultiVecAxpbyCols(wa, V, xcols, wb, V, wcols, w_length, ops);
        free(cols);
        free(coef);
    } else
    for (GCGE_INT idx = 0; idx < w_length; ++idx) {
        GCGE_INT xi = unlock[idx], wi = w_start + idx;
        ops->GetVecFromMultiVec(V_tmp, idx, &rhs, ops);
        ops->GetVecFromMultiVec(V, xi, &xc, ops);
        ops->GetVecFromMultiVec(V, wi, &wc, ops);
        if (B == NULL)
            ops->VecAxpby(1.0, xc, 0.0, rhs, ops);
        else
            ops->MatDotVec(B, xc, rhs, ops);
        if (fabs(eval[xi]) <= 1.0) {
            ops->VecAxpby(0.0, rhs, eval[xi] + para->shift, rhs, ops);
            ops->VecAxpby(1.0, xc, 0.0, wc, ops);
        } else {
            ops->VecAxpby(1.0 / (eval[xi] + para->shift), xc, 0.0, wc, ops);
        }
        ops->RestoreVecForMultiVec(V, xi, &xc, ops);
        ops->RestoreVecForMultiVec(V, wi, &wc, ops);
        ops->RestoreVecForMultiVec(V_tmp, idx, &rhs, ops);
    }

    if (ops->MultiLinearSolver) {
        ops->MultiLinearSolver(A, V_tmp, V, mv_s, mv_e, ops);
        return 0;
    }
    if (ops->LinearSolver) {
        for (GCGE_INT idx = 0; idx < w_length; ++idx) {
            ops->GetVecFromMultiVec(V_tmp, idx, &rhs, ops);
            ops->GetVecFromMultiVec(V, w_start + idx, &wc, ops);
            ops->LinearSolver(A, rhs, wc, ops);
            ops->RestoreVecForMultiVec(V, w_start + idx, &wc, ops);
            ops->RestoreVecForMultiVec(V_tmp, idx, &rhs, ops);
        }
        return 0;
    }
    if (para->if_use_bcg == 1) {
        if (para->if_shift == 1) para->use_bcg_continuous = 0;
        if (para->use_bcg_continuous) {
            fprintf(stderr, "gcge: use_bcg_continuous is not implemented in this library\n");
            exit(1);
        }
        GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
        if (para->use_ops_ext && ext && want_correction && ext->BlockCGCorrection) {
            ext->BlockCGCorrection(A, para->if_shift, para->shift, B, V_tmp, V, mv_s, mv_e,
                                   para->cg_max_it, para->cg_rate, ops, ws->CG_p, ws->evec,
                                   ws->subspace_dtmp, ws->subspace_itmp, para->bcg_by_column);
            return 1;
        } else if (para->use_ops_ext && ext && ext->BlockCG) {
            ext->BlockCG(A, para->if_shift, para->shift, B, V_tmp, V, mv_s, mv_e, para->cg_max_it,
                         para->cg_rate, ops, ws->CG_p, ws->evec, ws->subspace_dtmp,
                         ws->subspace_itmp, para->bcg_by_column);
        } else {
            GCGE_BlockCG(A, para->if_shift, para->shift, B, V_tmp, V, mv_s, mv_e, para->cg_max_it,
                         para->cg_rate, ops, ws->CG_p, ws->evec, NULL, ws->subspace_dtmp,
                         ws->subspace_itmp, para->bcg_by_column);
        }
    } else {
        for (GCGE_INT idx = 0; idx < w_length; ++idx) {
            ops->GetVecFromMultiVec(V_tmp, idx, &rhs, ops);
            ops->GetVecFromMultiVec(V, w_start + idx, &wc, ops);
            GCGE_CG(A, rhs, wc, ops, para, ws->V_tmp, ws->evec);
            ops->RestoreVecForMultiVec(V, w_start + idx, &wc, ops);
            ops->RestoreVecForMultiVec(V_tmp, idx, &rhs, ops);
        }
    }
    return 0;
}

void GCGE_ComputeP(GCGE_DOUBLE *subspace_evec, void **V, GCGE_OPS *ops, GCGE_PARA *para,
                   GCGE_WORKSPACE *ws)
{
    GCGE_DOUBLE t1 = GCGE_GetTime(), t2;
    GCGE_DOUBLE *coef = subspace_evec;
    const GCGE_INT ldc = ws->dim_xpw;
    GCGE_INT *unlock = ws->unlock;
    GCGE_INT p_start = ws->dim_x;
    GCGE_INT p_ncols = ws->unconv_bs;
    const GCGE_INT xx_nrows = ws->last_dim_x;
    void **V_tmp = ws->V_tmp;
    GCGE_INT mv_s[2], mv_e[2];

    /* coefficient blocks (columns: new X | new P; rows: old X, P, W):
     *   XX  PX        PX := 0
     *   XP  PP        PP, PW := rows below the old X of unconverged X columns
     *   XW  PW                                                             */
    GCGE_DOUBLE *px = coef + (size_t)p_start * ldc;
    GCGE_DOUBLE *pp = px + xx_nrows;
    GCGE_DOUBLE *xp = coef + xx_nrows;
    memset(px, 0, (size_t)p_ncols * ldc * sizeof(GCGE_DOUBLE));
    for (GCGE_INT cp = 0; cp < p_ncols; ++cp)
        memcpy(pp + (size_t)cp * ldc, xp + (size_t)unlock[cp] * ldc,
               (size_t)(ldc - xx_nrows) * sizeof(GCGE_DOUBLE));

    /* orthonormalise the trailing block (from the first unlocked row/column) */
    const GCGE_INT orth_start = ws->num_soft_locked_in_last_iter;
    const GCGE_INT orth_length = ldc - orth_start;
    GCGE_DOUBLE *sub = coef + (size_t)orth_start * ldc + orth_start;
    p_start -= orth_start;
    GCGE_INT p_end = p_start + p_ncols;
    if (strcmp(para->p_orth_type, "multi") == 0) {
        GCGE_StableMultiOrthonormalizationInSubspace(sub, ldc, orth_length, p_start, &p_end, NULL,
                                                     -1, para->orth_para, ws, ops);
    } else if (strcmp(para->p_orth_type, "scbgs") == 0 || strcmp(para->p_orth_type, "bgs") == 0) {
        fprintf(stderr, "gcge: p_orth_type \"%s\" is not implemented in this library\n",
                para->p_orth_type);
        exit(1);
    } else {
        GCGE_OrthonormalizationInSubspace(sub, ldc, orth_length, p_start, &p_end, NULL, -1,
                                          para->orth_para, ops);
    }
    p_ncols = p_end - p_start;
    p_end += orth_start;
    t2 = GCGE_GetTime();
    para->stat_para->part_time_one_iter->p_orth_time = t2 - t1;
    para->stat_para->part_time_total->p_orth_time += t2 - t1;

    /* V_tmp(:, 0:p_end-nev) = V(:, orth_start:ldc) * coef(orth_start:ldc, nev:p_end),
     * then swap into V(:, nev:p_end) */
    const GCGE_INT nev = para->nev;
    const GCGE_INT p_end_tmp = p_end - nev;
    GCGE_OPS_EXT *ext_ip = GCGE_OPS_GetExt(ops);
    if (strcmp(para->w_orth_type, "svqb") == 0 && para->use_ops_ext && ext_ip &&
        ext_ip->MultiVecLinearCombInPlace &&
        ext_ip->MultiVecLinearCombInPlace(V, orth_start, ldc - orth_start,
                                          coef + (size_t)nev * ldc + orth_start, ldc, nev,
                                          p_end_tmp, ops) == 0) {
        /* V(:, nev:p_end) formed in place: V_tmp is not needed */
    } else if (strcmp(para->w_orth_type, "svqb") == 0) {
        /* V_tmp is written before it is read again (the next right-hand
         * sides, projections and SVQB products), so a copy does what the
         * swap does for V with half the traffic on a device backend */
        mv_s[0] = orth_start; mv_e[0] = ldc;
        mv_s[1] = 0; mv_e[1] = p_end_tmp;
        ops->MultiVecLinearComb(V, V_tmp, mv_s, mv_e, coef + (size_t)nev * ldc + orth_start, ldc, 0,
                                1.0, 0.0, ops);
        mv_s[0] = 0; mv_e[0] = p_end_tmp;
        mv_s[1] = nev; mv_e[1] = p_end;
        ops->MultiVecAxpby(1.0, V_tmp, 0.0, V, mv_s, mv_e, ops);
    } else {
        mv_s[0] = orth_start; mv_e[0] = ldc;
        mv_s[1] = 0; mv_e[1] = p_end_tmp;
        ops->MultiVecLinearComb(V, V_tmp, mv_s, mv_e, coef + (size_t)nev * ldc + orth_start, ldc, 0,
                                1.0, 0.0, ops);
        mv_s[0] = nev; mv_e[0] = p_end;
        mv_s[1] = 0; mv_e[1] = p_end_tmp;
        ops->MultiVecSwap(V, V_tmp, mv_s, mv_e, ops);
    }
    ws->dim_xp = p_end;

    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
    if (ext && ext->Synchronize) ext->Synchronize(ops);
    t1 = GCGE_GetTime();
    para->stat_para->part_time_one_iter->p_axpy_time = t1 - t2;
    para->stat_para->part_time_total->p_axpy_time += t1 - t2;
    (void)p_ncols;
}
