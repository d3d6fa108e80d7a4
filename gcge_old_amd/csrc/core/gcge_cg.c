/*
 * Inner linear solves for W: plain CG and the column-batched CG.
 *
 * GCGE_CG follows reference src/gcge_cg.c:26-113 (cg_type 1: explicit r^T r;
 * cg_type 2: r^T r from the recurrence).
 * GCGE_BCG follows reference src/gcge_cg.c:302-501: one CG per column with
 * its own stopping test (||r|| < rate ||r0|| or ||r|| < 1e-15), the columns
 * still running kept in a compacted list.  The scalar arrays are indexed
 * exactly as the reference indexes them: the initial rho2 by column, the
 * per-iteration p^T w, rho1 and rho2 by position in the compacted list, and
 * beta read by column.  Once a column has left the list these two indexings
 * differ, and the iterates depend on it, so the device block CG reproduces
 * the same indexing (see the BlockCG extension).
 * Ranges: [0] right-hand sides (overwritten with residuals), [1] solutions,
 * [2] search directions, [3] A p.
 */
#include <math.h>
#include <stdio.h>

#include "gcge.h"

void GCGE_CG(void *Matrix, void *b, void *x, GCGE_OPS *ops, GCGE_PARA *para, void **V_tmp1,
             void **V_tmp2)
{
    void *r, *p, *w;
    GCGE_INT niter = 0, max_it = para->cg_max_it, type = para->cg_type;
    GCGE_DOUBLE rate = para->cg_rate;
    GCGE_DOUBLE alpha, beta, rho1 = 0.0, rho2, error, last_error, pTw, wTw;
    ops->GetVecFromMultiVec(V_tmp1, 1, &r, ops);
    ops->GetVecFromMultiVec(V_tmp2, 0, &p, ops);
    ops->GetVecFromMultiVec(V_tmp2, 1, &w, ops);
    ops->MatDotVec(Matrix, x, r, ops);
    ops->VecAxpby(1.0, b, -1.0, r, ops);
    ops->VecInnerProd(r, r, &rho2, ops);
    error = sqrt(rho2);
    niter = 1;
    last_error = error;
    while (last_error / error >= rate && niter < max_it) {
        if (niter == 1) {
            ops->VecAxpby(1.0, r, 0.0, p, ops);
        } else {
            beta = rho2 / rho1;
            ops->VecAxpby(1.0, r, beta, p, ops);
        }
        ops->MatDotVec(Matrix, p, w, ops);
        if (type == 1) {
            ops->VecInnerProd(p, w, &pTw, ops);
            alpha = rho2 / pTw;
            ops->VecAxpby(alpha, p, 1.0, x, ops);
            ops->VecAxpby(-alpha, w, 1.0, r, ops);
            rho1 = rho2;
            ops->VecInnerProd(r, r, &rho2, ops);
        } else {
            ops->VecInnerProd(p, w, &pTw, ops);
            ops->VecInnerProd(w, w, &wTw, ops);
            alpha = rho2 / pTw;
            ops->VecAxpby(alpha, p, 1.0, x, ops);
            ops->VecAxpby(-alpha, w, 1.0, r, ops);
            rho1 = rho2;
            rho2 = rho1 - 2.0 * alpha * pTw + alpha * alpha * wTw;
        }
        last_error = sqrt(rho2);
        niter++;
    }
    ops->RestoreVecForMultiVec(V_tmp1, 1, &r, ops);
    ops->RestoreVecForMultiVec(V_tmp2, 0, &p, ops);
    ops->RestoreVecForMultiVec(V_tmp2, 1, &w, ops);
}

/* w = M v (+ shift * B v, or + shift * v when B is NULL) */
static void shifted_matvec(void *M, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B, void *v,
                           void *w, void *Bv, GCGE_OPS *ops)
{
    ops->MatDotVec(M, v, w, ops);
    if (if_shift == 1) {
        if (B == NULL) {
            ops->VecAxpby(shift, v, 1.0, w, ops);
        } else {
            ops->MatDotVec(B, v, Bv, ops);
            ops->VecAxpby(shift, Bv, 1.0, w, ops);
        }
    }
}

void GCGE_BCG(void *Matrix, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B, void **RHS,
              void **V, GCGE_INT *start, GCGE_INT *end, GCGE_INT max_it, GCGE_DOUBLE rate,
              GCGE_OPS *ops, void *V_tmp1, void **V_tmp2, void **V_tmp3,
              GCGE_DOUBLE *subspace_dtmp, GCGE_INT *subspace_itmp)
{
    GCGE_BlockCG(Matrix, if_shift, shift, B, RHS, V, start, end, max_it, rate, ops, V_tmp1, V_tmp2,
                 V_tmp3, subspace_dtmp, subspace_itmp, 0);
}

void GCGE_BlockCG(void *Matrix, GCGE_INT if_shift, GCGE_DOUBLE shift, void *B, void **RHS,
                  void **V, GCGE_INT *start, GCGE_INT *end, GCGE_INT max_it, GCGE_DOUBLE rate,
                  GCGE_OPS *ops, void *V_tmp1, void **V_tmp2, void **V_tmp3,
                  GCGE_DOUBLE *subspace_dtmp, GCGE_INT *subspace_itmp, GCGE_INT by_column)
{
    void *x, *r, *p, *w, *b, *Bx = NULL;
    void **CG_R = RHS, **CG_P = (void **)V_tmp1, **CG_W = V_tmp2, **CG_X = V;
    const GCGE_INT n = end[1] - start[1];
    if (if_shift == 1 && B != NULL) ops->VecCreateByMat(&Bx, Matrix, ops);
    if (V_tmp3 != NULL) {
        GCGE_INT s[2] = {start[0], start[4]}, e[2] = {end[0], end[4]};
        CG_R = V_tmp3;
        ops->MultiVecAxpby(1.0, RHS, 0.0, V_tmp3, s, e, ops);
    }
    GCGE_DOUBLE *rho1 = subspace_dtmp, *rho2 = rho1 + n, *error = rho2 + n,
                *last_error = error + n, *ptw = last_error + n;
    GCGE_INT *unlock = subspace_itmp, num_unlock = 0, old_num_unlock;

    for (GCGE_INT idx = 0; idx < n; ++idx) {
        ops->GetVecFromMultiVec(CG_X, start[1] + idx, &x, ops);
        ops->GetVecFromMultiVec(CG_W, start[3] + idx, &r, ops);
        shifted_matvec(Matrix, if_shift, shift, B, x, r, Bx, ops);
        ops->RestoreVecForMultiVec(CG_X, start[1] + idx, &x, ops);
        ops->GetVecFromMultiVec(CG_R, start[0] + idx, &b, ops);
        ops->VecAxpby(-1.0, r, 1.0, b, ops);
        ops->VecLocalInnerProd(b, b, rho2 + idx, ops);
        ops->RestoreVecForMultiVec(CG_R, start[0] + idx, &b, ops);
        ops->RestoreVecForMultiVec(CG_W, start[3] + idx, &r, ops);
    }
    for (GCGE_INT idx = 0; idx < n; ++idx) {
        error[idx] = sqrt(rho2[idx]);
        last_error[idx] = error[idx];
        if (error[idx] >= 1e-15) unlock[num_unlock++] = idx;
    }
    GCGE_INT niter = 1;
    while (num_unlock > 0 && niter < max_it) {
        for (GCGE_INT id = 0; id < num_unlock; ++id) {
            GCGE_INT idx = unlock[id];
            ops->GetVecFromMultiVec(CG_P, start[2] + idx, &p, ops);
            ops->GetVecFromMultiVec(CG_R, start[0] + idx, &r, ops);
            if (niter == 1)
                ops->VecAxpby(1.0, r, 0.0, p, ops);
            else
                ops->VecAxpby(1.0, r, rho2[idx] / rho1[idx], p, ops);
            ops->RestoreVecForMultiVec(CG_R, start[0] + idx, &r, ops);
            ops->GetVecFromMultiVec(CG_W, start[3] + idx, &w, ops);
            shifted_matvec(Matrix, if_shift, shift, B, p, w, Bx, ops);
            ops->VecLocalInnerProd(p, w, ptw + id, ops);
            ops->RestoreVecForMultiVec(CG_P, start[2] + idx, &p, ops);
            ops->RestoreVecForMultiVec(CG_W, start[3] + idx, &w, ops);
        }
        for (GCGE_INT id = 0; id < num_unlock; ++id) {
            GCGE_INT idx = unlock[id];
            /* the reference takes alpha's numerator by position */
            const GCGE_INT q = by_column ? idx : id;
            GCGE_DOUBLE alpha = rho2[q] / ptw[id];
            ops->GetVecFromMultiVec(CG_P, start[2] + idx, &p, ops);
            ops->GetVecFromMultiVec(CG_X, start[1] + idx, &x, ops);
            ops->VecAxpby(alpha, p, 1.0, x, ops);
            ops->RestoreVecForMultiVec(CG_P, start[2] + idx, &p, ops);
            ops->RestoreVecForMultiVec(CG_X, start[1] + idx, &x, ops);
            ops->GetVecFromMultiVec(CG_R, start[0] + idx, &r, ops);
            ops->GetVecFromMultiVec(CG_W, start[3] + idx, &w, ops);
            ops->VecAxpby(-alpha, w, 1.0, r, ops);
            rho1[q] = rho2[q];
            ops->VecLocalInnerProd(r, r, rho2 + q, ops);
            ops->RestoreVecForMultiVec(CG_R, start[0] + idx, &r, ops);
            ops->RestoreVecForMultiVec(CG_W, start[3] + idx, &w, ops);
        }
        for (GCGE_INT id = 0; id < num_unlock; ++id)
            last_error[unlock[id]] = sqrt(fabs(rho2[by_column ? unlock[id] : id]));
        old_num_unlock = num_unlock;
        num_unlock = 0;
        for (GCGE_INT id = 0; id < old_num_unlock; ++id) {
            GCGE_INT idx = unlock[id];
            if (last_error[idx] >= rate * error[idx] && last_error[idx] >= 1e-15)
                unlock[num_unlock++] = idx;
        }
        niter++;
    }
    if (Bx) ops->VecDestroy(&Bx, ops);
}
