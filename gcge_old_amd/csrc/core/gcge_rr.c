/*
 * Rayleigh-Ritz: the projected matrix V^T A V and its eigenpairs.
 *
 * GCGE_ComputeSubspaceMatrix follows reference src/gcge_rayleighritz.c:121-190:
 * with V = [X, P, W] and the previous projected matrix H and coefficients C,
 *     X^T A X = diag(eval),  X^T A P = 0,  P^T A P = C_p^T H C_p,
 * and the new columns [X2, P, W]^T A W come from one A W product and two
 * inner-product blocks (GCGE_ComputeSubspaceMatrixVTAW, :40-90); rows of
 * the soft-locked prefix are skipped (the reference's opt_rr_eig_partly
 * test is an assignment, so this is always on).
 * GCGE_ComputeSubspaceEigenpairs follows the single-process branch of
 * :507-805: the locked rows of the W columns are zeroed, eigenpairs
 * unlock[0]+1 .. min(max_dim_x, ldm) of the trailing block are computed,
 * and num_soft_locked_in_last_iter becomes unlock[0].
 *
 * With full_rr the converged prefix stays in both steps: its rows of the
 * new columns are computed, the whole projected matrix is solved and every
 * Ritz vector is recomputed.  Zeroing those rows treats the converged
 * vectors as exact; they are only converged to the tolerance, which caps the
 * attainable residual of a vector at relative eigenvalue distance g from
 * them near tol / g and can stall the iteration when g < 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

void GCGE_ComputeSubspaceMatrixVTAW(void **V, void *A, GCGE_INT start_W, GCGE_INT start_V,
                                    GCGE_INT end_V, GCGE_DOUBLE *subspace_mat, GCGE_OPS *ops,
                                    void **work)
{
    const GCGE_INT w_length = end_V - start_W;
    const GCGE_INT xp_length = end_V - w_length;
    GCGE_INT mv_s[2] = {start_W, 0}, mv_e[2] = {end_V, w_length};
    /* work(:, 0:w_length) = A W */
    ops->MatDotMultiVec(A, V, work, mv_s, mv_e, ops);
    memset(subspace_mat, 0, (size_t)end_V * w_length * sizeof(GCGE_DOUBLE));
    /* [X2, P]^T (A W) */
    mv_s[0] = start_V; mv_e[0] = xp_length;
    mv_s[1] = 0; mv_e[1] = w_length;
    ops->MultiVecInnerProd(V, work, subspace_mat + start_V, "nonsym", mv_s, mv_e, end_V, 0, ops);
    /* W^T (A W), symmetric */
    mv_s[0] = start_W; mv_e[0] = end_V;
    mv_s[1] = 0; mv_e[1] = w_length;
    ops->MultiVecInnerProd(V, work, subspace_mat + xp_length, "sym", mv_s, mv_e, end_V, 0, ops);
}

/* The full projection (full_rr, every row from 0): one inner-product block
 * [X, P, W]^T (A W) instead of [X, P]^T (A W) and the "sym" W^T (A W) (one
 * Gram launch and one transfer fewer, A W read once); the W^T A W part then
 * takes its lower triangle from the upper one, as the "sym" block does. */
static void vtaw_one_block(void **V, void *A, GCGE_INT start_W, GCGE_INT end_V,
                           GCGE_DOUBLE *subspace_mat, GCGE_OPS *ops, void **work)
{
    const GCGE_INT w_length = end_V - start_W;
    GCGE_INT mv_s[2] = {start_W, 0}, mv_e[2] = {end_V, w_length};
    ops->MatDotMultiVec(A, V, work, mv_s, mv_e, ops);
    mv_s[0] = 0; mv_e[0] = end_V;
    mv_s[1] = 0; mv_e[1] = w_length;
    ops->MultiVecInnerProd(V, work, subspace_mat, "nonsym", mv_s, mv_e, end_V, 0, ops);
    for (GCGE_INT j = 0; j < w_length; ++j)
        for (GCGE_INT i = j + 1; i < w_length; ++i)
            subspace_mat[(size_t)j * end_V + start_W + i] =
                subspace_mat[(size_t)i * end_V + start_W + j];
}

void GCGE_ComputeSubspaceMatrix(void *A, void **V, GCGE_OPS *ops, GCGE_PARA *para,
                                GCGE_WORKSPACE *ws)
{
    const GCGE_INT dim_xp = ws->dim_xp, ldm = ws->dim_xpw;
    GCGE_DOUBLE *H = ws->subspace_matrix;
    if (dim_xp != 0) {
        GCGE_INT lde = ws->last_dim_xpw;
        GCGE_INT dim_x = ws->dim_x;
        GCGE_INT dim_p = dim_xp - dim_x;
        GCGE_DOUBLE one = 1.0, zero = 0.0;
        GCGE_DOUBLE *pp = ws->subspace_evec + (size_t)dim_x * lde;
        GCGE_DOUBLE *AApp = ws->subspace_dtmp;
        GCGE_DOUBLE *PTAP = H + (size_t)dim_x * ldm + dim_x;
        /* AApp = H_last C_p, then H := 0, P^T A P = C_p^T AApp */
        ops->DenseSymMatDotDenseMat("L", "U", &lde, &dim_p, &one, H, &lde, pp, &lde, &zero, AApp,
                                    &lde);
        memset(H, 0, (size_t)ldm * ldm * sizeof(GCGE_DOUBLE));
        ops->DenseMatDotDenseMat("T", "N", &dim_p, &dim_p, &lde, &one, pp, &lde, AApp, &lde,
                                 &zero, PTAP, (GCGE_INT *)&ldm);
        for (GCGE_INT i = 0; i < dim_x; ++i) H[(size_t)i * ldm + i] = ws->eval[i];
    }
    para->opt_rr_eig_partly = 1;
    const GCGE_INT start_V = para->full_rr ? 0 : ws->num_soft_locked_in_last_iter;
    if (para->full_rr && dim_xp > 0)
        vtaw_one_block(V, A, dim_xp, ldm, H + (size_t)dim_xp * ldm, ops, ws->evec);
    else
        GCGE_ComputeSubspaceMatrixVTAW(V, A, dim_xp, start_V, ldm, H + (size_t)dim_xp * ldm, ops,
                                       ws->evec);
    if (para->if_shift == 1)
        for (GCGE_INT i = 0; i < ldm; ++i) H[(size_t)i * ldm + i] += para->shift;
}

/* ascending-magnitude quicksort carrying an index array (reference
 * src/gcge_rayleighritz.c:842-874, same partition scheme so that ties end
 * in the same order) */
static void sort_by_magnitude(GCGE_DOUBLE *a, GCGE_INT *idx, GCGE_INT left, GCGE_INT right)
{
    if (left > right) return;
    GCGE_INT i = left, j = right;
    GCGE_INT idx_pivot = idx[left];
    GCGE_DOUBLE pivot = a[left], m_pivot = fabs(a[left]);
    while (i != j) {
        while (fabs(a[j]) >= m_pivot && j > i) j--;
        if (j > i) {
            idx[i] = idx[j];
            a[i++] = a[j];
        }
        while (fabs(a[i]) <= m_pivot && j > i) i++;
        if (j > i) {
            idx[j] = idx[i];
            a[j--] = a[i];
        }
    }
    a[i] = pivot;
    idx[i] = idx_pivot;
    sort_by_magnitude(a, idx, left, i - 1);
    sort_by_magnitude(a, idx, i + 1, right);
}

void GCGE_SortEigenpairs(GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_INT nev, GCGE_INT ldv,
                         GCGE_DOUBLE *work)
{
    for (GCGE_INT head = 0; head < nev; ++head) {
        GCGE_INT tail = ldv - 1 - head;
        if (head >= tail) break;
        memcpy(work, evec + (size_t)head * ldv, (size_t)ldv * sizeof(GCGE_DOUBLE));
        memcpy(evec + (size_t)head * ldv, evec + (size_t)tail * ldv, (size_t)ldv * sizeof(GCGE_DOUBLE));
        memcpy(evec + (size_t)tail * ldv, work, (size_t)ldv * sizeof(GCGE_DOUBLE));
        GCGE_DOUBLE t = eval[head];
        eval[head] = 1.0 / eval[tail];
        eval[tail] = 1.0 / t;
    }
}

void GCGE_ComputeSubspaceEigenpairs(GCGE_DOUBLE *subspace_matrix, GCGE_DOUBLE *eval,
                                    GCGE_DOUBLE *subspace_evec, GCGE_OPS *ops, GCGE_PARA *para,
                                    GCGE_WORKSPACE *ws)
{
    GCGE_DOUBLE t1 = GCGE_GetTime();
    const GCGE_INT ldm = ws->dim_xpw;
    GCGE_DOUBLE *temp = ws->subspace_dtmp;
    GCGE_DOUBLE *dwork = temp + (size_t)ldm * ldm;
    const GCGE_INT max_dim_x = ws->max_dim_x;
    GCGE_INT nsl = ws->unlock[0];
    if (para->opt_rr_eig_partly == 0 || para->full_rr) nsl = 0;
    GCGE_INT il = nsl + 1;
    ws->num_soft_locked_in_last_iter = nsl;
    GCGE_INT iu = max_dim_x < ldm ? max_dim_x : ldm;
    if (strcmp(para->eval_type, "sm") == 0) iu = ldm;
    const GCGE_INT real_nev_in_rr = iu;
    const GCGE_INT m = iu - il + 1;
    GCGE_INT *itmp = ws->subspace_itmp;

    for (GCGE_INT i = 0; i < ws->dim_xpw - ws->dim_xp; ++i)
        memset(subspace_matrix + (size_t)(ws->dim_xp + i) * ldm, 0, (size_t)nsl * sizeof(GCGE_DOUBLE));
    memcpy(temp, subspace_matrix, (size_t)ldm * ldm * sizeof(GCGE_DOUBLE));

    GCGE_INT nrows = ldm - nsl;
    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
    if (!(para->use_ops_ext && ext && ext->DenseMatEigenSolver &&
          ext->DenseMatEigenSolver(temp + (size_t)nsl * ldm + nsl, ldm, nrows, eval + nsl,
                                   subspace_evec + (size_t)nsl * ldm + nsl, ldm, il - nsl,
                                   iu - nsl) == 0))
        ops->DenseMatEigenSolver(temp + (size_t)nsl * ldm + nsl, ldm, nrows, eval + nsl,
                                 subspace_evec + (size_t)nsl * ldm + nsl, ldm, il - nsl, iu - nsl,
                                 itmp, dwork);
    (void)m;

    if (strcmp(para->eval_type, "sm") == 0) {
        /* smallest magnitude: reorder the computed pairs by |lambda| */
        for (GCGE_INT i = 0; i < nrows; ++i) itmp[i] = i;
        sort_by_magnitude(eval + nsl, itmp, 0, nrows - 1);
        GCGE_INT eval_length = (max_dim_x < ldm ? max_dim_x : ldm) - nsl;
        GCGE_DOUBLE *evec_start = subspace_evec + (size_t)nsl * ldm + nsl;
        for (GCGE_INT i = 0; i < eval_length; ++i)
            memcpy(temp + (size_t)i * ldm, evec_start + (size_t)itmp[i] * ldm,
                   (size_t)nrows * sizeof(GCGE_DOUBLE));
        memcpy(evec_start, temp, ((size_t)(eval_length - 1) * ldm + nrows) * sizeof(GCGE_DOUBLE));
    }
    if (para->if_shift == 1) {
        for (GCGE_INT i = nsl; i < real_nev_in_rr; ++i) eval[i] -= para->shift;
        for (GCGE_INT i = 0; i < ldm; ++i) subspace_matrix[(size_t)i * ldm + i] -= para->shift;
    }
    GCGE_DOUBLE t2 = GCGE_GetTime();
    para->stat_para->part_time_one_iter->rr_eigen_time = t2 - t1;
    para->stat_para->part_time_total->rr_eigen_time += t2 - t1;
    if (strcmp(para->orth_type, "A") == 0)
        GCGE_SortEigenpairs(eval, subspace_evec, iu, ldm, ws->subspace_dtmp);
}
