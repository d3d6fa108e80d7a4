/*
 * The GCG outer iteration and the convergence check.
 *
 * GCGE_Solve follows reference src/gcge_eigsol.c:65-646:
 *   init:  random X (until nev independent columns survive), B-orthonormalise,
 *          Rayleigh-Ritz on X, Ritz vectors, convergence check, then W from
 *          the inner solves, W orthonormalisation, Rayleigh-Ritz on [X, W];
 *   loop:  P from the subspace coefficients, X swapped into V, W, W
 *          orthonormalisation, Rayleigh-Ritz on [X, P, W], Ritz vectors,
 *          convergence check, until nev pairs converged in a row or
 *          ev_max_it iterations;
 *   end:   eigenvalues out, soft-locked vectors copied back to evec.
 * GCGE_CheckConvergence follows :663-887 (prefix counting, multiplicity
 * test for unlocking, padding of the unlock list to block_size, the stop
 * after 10 consecutive unconverged pairs).  When the backend publishes
 * MultiVecResidualNorms the residual norms of the candidate columns come
 * from one block call; the bookkeeping that consumes them is unchanged and
 * stores only the values the column-by-column loop would have stored.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

static void sync_ops(GCGE_OPS *ops)
{
    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
    if (ext && ext->Synchronize) ext->Synchronize(ops);
}

static GCGE_DOUBLE tick(GCGE_OPS *ops)
{
    sync_ops(ops);
    return GCGE_GetTime();
}

#define PART(field, t1, t2)                                          \
    do {                                                             \
        stat->part_time_one_iter->field = (t2) - (t1);               \
        stat->part_time_total->field += (t2) - (t1);                 \
    } while (0)

static void unsupported(const char *what, const char *name)
{
    fprintf(stderr, "gcge: %s \"%s\" is not implemented in this library\n", what, name);
    exit(1);
}

/* X orthonormalisation dispatch of the initial phase (reference
 * src/gcge_eigsol.c:194-213 and :221-241) */
static void orth_x(void **X, GCGE_INT start, GCGE_INT *end, void *Orth_mat, GCGE_OPS *ops,
                   GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    if (strcmp(para->x_orth_type, "multi") == 0)
        GCGE_StableMultiOrthonormalization(X, start, end, Orth_mat, ops, para, ws);
    else if (strcmp(para->x_orth_type, "scbgs") == 0 || strcmp(para->x_orth_type, "cbgs") == 0)
        unsupported("x_orth_type", para->x_orth_type);
    else
        GCGE_Orthonormalization(X, start, end, Orth_mat, ops, para, ws->V_tmp, ws->subspace_dtmp);
}

/* W orthonormalisation dispatch; `in_loop` selects the table of the main
 * loop (reference :509-534) or of the initial [X, W] step (:380-399). */
static void orth_w(void **V, GCGE_INT start, GCGE_INT *end, void *Orth_mat, GCGE_OPS *ops,
                   GCGE_PARA *para, GCGE_WORKSPACE *ws, int in_loop)
{
    const char *t = para->w_orth_type;
    if (strcmp(t, "multi") == 0 || strcmp(t, "svqb") == 0)
        GCGE_StableMultiOrthonormalization(V, start, end, Orth_mat, ops, para, ws);
    else if (strcmp(t, "cbgs") == 0 || strcmp(t, "scbgs") == 0 || (in_loop && strcmp(t, "bgs") == 0))
        unsupported("w_orth_type", t);
    else
        GCGE_Orthonormalization(V, start, end, Orth_mat, ops, para, ws->V_tmp, ws->subspace_dtmp);
}

void GCGE_Solve(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_PARA *para,
                GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    GCGE_PrintParaInfo(para);
    GCGE_DOUBLE solve_t0 = tick(ops);
    const GCGE_INT nev = para->nev;
    const GCGE_INT ev_max_it = para->ev_max_it;
    const GCGE_INT max_dim_x = ws->max_dim_x;
    GCGE_INT num_init_evec = para->num_init_evec;
    void **V = ws->V;
    GCGE_DOUBLE *subspace_matrix = ws->subspace_matrix;
    GCGE_DOUBLE *subspace_evec = ws->subspace_evec;
    GCGE_STATISTIC_PARA *stat = para->stat_para;
    GCGE_INT mv_s[2] = {0, 0}, mv_e[2] = {0, 0};
    GCGE_DOUBLE t1, t2;

    ws->evec = evec;
    /* orth_type "B": B inner product, A in the projection (else the roles swap) */
    void *Orth_mat = strcmp(para->orth_type, "B") ? A : B;
    void *RayleighRitz_mat = strcmp(para->orth_type, "B") ? B : A;

    t1 = tick(ops);
    if (num_init_evec < nev) {
        do {
            ops->MultiVecSetRandomValue(evec, num_init_evec, nev - num_init_evec, ops);
            if (strcmp(para->conv_type, "O") == 0 && num_init_evec == 0) {
                /* estimate of the omega norm of A from the random block
                 * (reference :141-178) */
                void *v;
                GCGE_DOUBLE nrm = 0.0, fro_omega = 0.0, fro_a_omega = 0.0;
                for (GCGE_INT i = 0; i < nev; ++i) {
                    ops->GetVecFromMultiVec(evec, i, &v, ops);
                    ops->VecInnerProd(v, v, &nrm, ops);
                    ops->VecAxpby(0.0, v, 1.0 / sqrt(nrm), v, ops);
                    ops->RestoreVecForMultiVec(evec, i, &v, ops);
                }
                for (GCGE_INT i = 0; i < nev; ++i) {
                    ops->GetVecFromMultiVec(evec, i, &v, ops);
                    ops->VecInnerProd(v, v, &nrm, ops);
                    ops->RestoreVecForMultiVec(evec, i, &v, ops);
                    fro_omega += nrm;
                }
                mv_s[0] = 0; mv_e[0] = nev; mv_s[1] = 0; mv_e[1] = nev;
                ops->MatDotMultiVec(A, evec, V, mv_s, mv_e, ops);
                for (GCGE_INT i = 0; i < nev; ++i) {
                    ops->GetVecFromMultiVec(V, i, &v, ops);
                    ops->VecInnerProd(v, v, &nrm, ops);
                    ops->RestoreVecForMultiVec(V, i, &v, ops);
                    fro_a_omega += nrm;
                }
                para->conv_omega_norm = sqrt(fro_a_omega / fro_omega);
            }
            if (para->dirichlet_boundary == 1 && ops->SetDirichletBoundary)
                ops->SetDirichletBoundary(evec, nev, A, B);
            ws->dim_x = nev;
            orth_x(evec, num_init_evec, &ws->dim_x, Orth_mat, ops, para, ws);
            num_init_evec = ws->dim_x;
        } while (num_init_evec < nev);
    } else {
        orth_x(evec, para->num_conv, &ws->dim_x, Orth_mat, ops, para, ws);
    }
    t2 = tick(ops);
    PART(x_orth_time, t1, t2);

    /* V(:, 0:nev) = evec */
    mv_s[0] = 0; mv_e[0] = nev; mv_s[1] = 0; mv_e[1] = nev;
    ops->MultiVecAxpby(1.0, evec, 0.0, V, mv_s, mv_e, ops);

    ws->dim_xpw = ws->dim_x;
    ws->last_dim_x = ws->dim_x;
    ws->dim_xp = 0;

    t1 = tick(ops);
    GCGE_ComputeSubspaceMatrix(RayleighRitz_mat, V, ops, para, ws);
    t2 = tick(ops);
    PART(rr_mat_time, t1, t2);
    GCGE_ComputeSubspaceEigenpairs(subspace_matrix, ws->eval, subspace_evec, ops, para, ws);
    t1 = tick(ops);
    GCGE_ComputeX(V, subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
    t2 = tick(ops);
    PART(x_axpy_time, t1, t2);
    t1 = tick(ops);
    GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
    t2 = tick(ops);
    PART(conv_time, t1, t2);
    GCGE_PrintIterationInfo(ws->eval, para);

    if (para->num_conv_continuous < nev) {
        mv_s[0] = ws->num_soft_locked_in_last_iter; mv_e[0] = nev;
        mv_s[1] = ws->num_soft_locked_in_last_iter; mv_e[1] = nev;
        t1 = tick(ops);
        ops->MultiVecSwap(V, evec, mv_s, mv_e, ops);
        t2 = tick(ops);
        PART(x_axpy_time, t1, t2);
        ws->dim_xp = ws->dim_x;

        t1 = tick(ops);
        GCGE_ComputeW(A, B, V, ws->eval, ops, para, ws);
        t2 = tick(ops);
        PART(w_line_time, t1, t2);
        ws->last_dim_xpw = ws->dim_xpw;
        ws->dim_xpw = ws->dim_xp + ws->unconv_bs;

        t1 = tick(ops);
        orth_w(V, ws->dim_x, &ws->dim_xpw, Orth_mat, ops, para, ws, 0);
        t2 = tick(ops);
        PART(w_orth_time, t1, t2);
        ws->dim_xp = nev;

        t1 = tick(ops);
        GCGE_ComputeSubspaceMatrix(RayleighRitz_mat, V, ops, para, ws);
        t2 = tick(ops);
        PART(rr_mat_time, t1, t2);
        GCGE_ComputeSubspaceEigenpairs(subspace_matrix, ws->eval, subspace_evec, ops, para, ws);
        t1 = tick(ops);
        GCGE_ComputeX(V, subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
        t2 = tick(ops);
        PART(x_axpy_time, t1, t2);
        t1 = tick(ops);
        GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
        t2 = tick(ops);
        PART(conv_time, t1, t2);
        GCGE_PrintIterationInfo(ws->eval, para);
    }

    while (para->num_conv_continuous < nev && para->num_iter < ev_max_it) {
        GCGE_ComputeP(subspace_evec, V, ops, para, ws);

        mv_s[0] = ws->num_soft_locked_in_last_iter; mv_e[0] = nev;
        mv_s[1] = ws->num_soft_locked_in_last_iter; mv_e[1] = nev;
        t1 = tick(ops);
        /* The reference swaps V(:, nsl:nev) and evec(:, nsl:nev) here
         * (src/gcge_eigsol.c:478); only the V side is read afterwards: evec is
         * the block CG's scratch in GCGE_ComputeW and is rewritten by
         * GCGE_ComputeX before it is read, and its locked columns come back
         * from V at the end.  One copy moves half the bytes of the swap. */
        ops->MultiVecAxpby(1.0, evec, 0.0, V, mv_s, mv_e, ops);
        t2 = tick(ops);
        PART(x_axpy_time, t1, t2);

        t1 = tick(ops);
        GCGE_ComputeW(A, B, V, ws->eval, ops, para, ws);
        t2 = tick(ops);
        PART(w_line_time, t1, t2);
        ws->last_dim_xpw = ws->dim_xpw;
        ws->dim_xpw = ws->dim_xp + ws->unconv_bs;

        t1 = tick(ops);
        orth_w(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws, 1);
        t2 = tick(ops);
        PART(w_orth_time, t1, t2);

        t1 = tick(ops);
        GCGE_ComputeSubspaceMatrix(RayleighRitz_mat, V, ops, para, ws);
        t2 = tick(ops);
        PART(rr_mat_time, t1, t2);
        GCGE_ComputeSubspaceEigenpairs(subspace_matrix, ws->eval, subspace_evec, ops, para, ws);

        ws->last_dim_x = ws->dim_x;
        ws->dim_x = max_dim_x;

        t1 = tick(ops);
        GCGE_ComputeX(V, subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
        t2 = tick(ops);
        PART(x_axpy_time, t1, t2);
        t1 = tick(ops);
        GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
        t2 = tick(ops);
        PART(conv_time, t1, t2);
        GCGE_PrintIterationInfo(ws->eval, para);
    }

    memcpy(eval, ws->eval, (size_t)nev * sizeof(GCGE_DOUBLE));
    mv_s[0] = 0; mv_e[0] = ws->num_soft_locked_in_last_iter;
    mv_s[1] = 0; mv_e[1] = ws->num_soft_locked_in_last_iter;
    ops->MultiVecAxpby(1.0, V, 0.0, evec, mv_s, mv_e, ops);
    para->solve_time = tick(ops) - solve_t0;
    GCGE_PrintFinalInfo(eval, para);
}

/* Residual norms for columns start .. stop-1 of evec, one block call when
 * the backend has it; otherwise nothing is precomputed and the caller runs
 * the column loop. */
static int block_residuals(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_INT start, GCGE_INT stop, GCGE_DOUBLE *rnorm,
                           GCGE_DOUBLE *xnorm, GCGE_DOUBLE *bxnorm)
{
    GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
    if (!para->use_ops_ext || !ext || !ext->MultiVecResidualNorms || stop <= start) return 0;
    ext->MultiVecResidualNorms(A, B, evec, start, stop, eval + start, rnorm, xnorm, bxnorm, ops);
    return 1;
}

/* ||A x - lambda B x||, ||x||, ||B x|| of column idx by single-vector calls
 * (the reference's loop, src/gcge_eigsol.c:683-721); tmp1, tmp2 scratch */
static void column_residual(void *A, void *B, GCGE_DOUBLE lambda, void **evec, GCGE_INT idx,
                            void *tmp1, void *tmp2, GCGE_OPS *ops, GCGE_DOUBLE *res_norm,
                            GCGE_DOUBLE *evec_norm, GCGE_DOUBLE *bx_norm)
{
    void *ev;
    ops->GetVecFromMultiVec(evec, idx, &ev, ops);
    ops->MatDotVec(A, ev, tmp1, ops);
    if (B == NULL) {
        ops->VecAxpby(-lambda, ev, 1.0, tmp1, ops);
    } else {
        ops->MatDotVec(B, ev, tmp2, ops);
        ops->VecAxpby(-lambda, tmp2, 1.0, tmp1, ops);
    }
    ops->VecInnerProd(tmp1, tmp1, res_norm, ops);
    *res_norm = sqrt(*res_norm);
    ops->VecInnerProd(ev, ev, evec_norm, ops);
    *evec_norm = sqrt(*evec_norm);
    if (B == NULL) {
        *bx_norm = *evec_norm;
    } else {
        ops->VecInnerProd(tmp2, tmp2, bx_norm, ops);
        *bx_norm = sqrt(*bx_norm);
    }
    ops->RestoreVecForMultiVec(evec, idx, &ev, ops);
}

void GCGE_CheckConvergence(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_INT *unlock = ws->unlock;
    GCGE_INT max_conv_idx = unlock[0] - 1;
    const GCGE_INT start = unlock[0];
    const GCGE_INT dim_x = ws->dim_x;
    GCGE_INT flag = 0, num_unlock = 0;
    GCGE_INT num_conv_continuous = unlock[0];
    GCGE_INT max_ind = 0, min_ind = 0;
    const GCGE_DOUBLE multi_tol = para->multi_tol_for_lock;
    const GCGE_DOUBLE ev_tol = para->ev_tol;
    GCGE_DOUBLE max_res = 0.0, min_res = 0.0, sum_res = 0.0;
    GCGE_DOUBLE *res = para->res;
    const char *conv_type = para->conv_type;
    const GCGE_INT nev = para->nev;

    /* block path: all candidate columns at once */
    GCGE_DOUBLE *rn = NULL, *xn = NULL, *bxn = NULL;
    int have_block = 0;
    ws->residual_products_fresh = 0;
    if (nev > start) {
        rn = (GCGE_DOUBLE *)malloc(3 * (size_t)(nev - start) * sizeof(GCGE_DOUBLE));
        xn = rn + (nev - start);
        bxn = xn + (nev - start);
        have_block = block_residuals(A, B, eval, evec, ops, para, start, nev, rn, xn, bxn);
        ws->residual_products_fresh = have_block && evec == ws->evec;
    }

    void *tmp1 = NULL, *tmp2 = NULL;
    if (!have_block) {
        ops->GetVecFromMultiVec(ws->V_tmp, 0, &tmp1, ops);
        ops->GetVecFromMultiVec(ws->V_tmp, 1, &tmp2, ops);
    }
    unlock[0] = -1;
    for (GCGE_INT idx = start; idx < nev; ++idx) {
        GCGE_DOUBLE res_norm, evec_norm, bx_norm;
        if (have_block) {
            res_norm = rn[idx - start];
            evec_norm = xn[idx - start];
            bx_norm = bxn[idx - start];
        } else {
            column_residual(A, B, eval[idx], evec, idx, tmp1, tmp2, ops, &res_norm, &evec_norm,
                            &bx_norm);
        }
        GCGE_DOUBLE residual = res_norm / evec_norm;
        const GCGE_DOUBLE abs_residual = residual;
        if (strcmp(conv_type, "R") == 0) {
            residual /= (1.0 > fabs(eval[idx]) ? 1.0 : fabs(eval[idx]));
        } else if (strcmp(conv_type, "O") == 0) {
            residual /= (para->conv_omega_norm + fabs(eval[idx]));
        } else if (strcmp(conv_type, "RB") == 0) {
            residual = res_norm / (fabs(eval[idx]) * bx_norm);
        }
        res[idx] = residual;

        if (residual < ev_tol && abs_residual < 1e-2) {
            if (idx > 0) {
                if (res[idx - 1] < ev_tol) {
                    max_conv_idx = idx;
                    if (flag == 0) num_conv_continuous = idx + 1;
                } else {
                    GCGE_DOUBLE me = fabs(eval[idx] - eval[idx - 1]) / fabs(eval[idx - 1]);
                    if (me < multi_tol)
                        unlock[num_unlock++] = idx;
                    else
                        max_conv_idx = idx;
                }
            } else {
                max_conv_idx = 0;
                num_conv_continuous = 1;
            }
        } else {
            flag = 1;
            if (idx > 0 && res[idx - 1] < ev_tol) {
                /* unconverged after a converged one: unlock the whole cluster
                 * of (near) multiple eigenvalues that ends here */
                GCGE_INT first = idx;
                for (GCGE_INT i = idx; i > 0; --i) {
                    GCGE_DOUBLE me = fabs(eval[i] - eval[i - 1]) / fabs(eval[i - 1]);
                    if (me < multi_tol)
                        first = i - 1;
                    else
                        break;
                }
                for (GCGE_INT i = first; i <= idx; ++i) unlock[num_unlock++] = i;
            } else {
                unlock[num_unlock++] = idx;
            }
        }
        sum_res += residual;
        if (idx == start) {
            max_res = residual;
            min_res = residual;
        } else {
            if (residual > max_res) {
                max_res = residual;
                max_ind = idx;
            }
            if (residual < min_res) {
                min_res = residual;
                min_ind = idx;
            }
        }
        if (idx - max_conv_idx > 9) break;
    }
    if (!have_block) {
        ops->RestoreVecForMultiVec(ws->V_tmp, 0, &tmp1, ops);
        ops->RestoreVecForMultiVec(ws->V_tmp, 1, &tmp2, ops);
    }
    free(rn);

    if (para->use_shift == 1) {
        para->if_shift = 1;
        para->shift = (eval[dim_x - 1] - 100.0 * eval[0]) / 99.0;
    }
    /* pad the unlock list with the following indices up to block_size */
    if (num_conv_continuous < nev && num_unlock < para->block_size) {
        if (num_unlock > 0 && unlock[num_unlock - 1] < nev - 1) {
            for (GCGE_INT i = num_unlock; i < para->block_size; ++i) {
                unlock[i] = unlock[i - 1] + 1;
                num_unlock += 1;
                if (unlock[i] > nev - 2) break;
            }
        }
    }
    if (num_conv_continuous < nev && unlock[0] < num_conv_continuous) num_conv_continuous = unlock[0];
    if (num_unlock > para->block_size) num_unlock = para->block_size;

    /* With -gcge_full_rr the converged prefix [0, start) stays in the
     * Rayleigh-Ritz step, so its vectors keep moving after their check: before
     * the solve ends, check them again; the first one above the tolerance
     * and the columns after it go back into the unlock list. */
    if (para->full_rr && num_conv_continuous >= nev && start > 0) {
        GCGE_INT first_bad = -1;
        GCGE_DOUBLE *pr = (GCGE_DOUBLE *)malloc(3 * (size_t)start * sizeof(GCGE_DOUBLE));
        GCGE_DOUBLE *px = pr + start, *pb = px + start;
        if (!block_residuals(A, B, eval, evec, ops, para, 0, start, pr, px, pb)) {
            void *t1 = NULL, *t2 = NULL;
            ops->GetVecFromMultiVec(ws->V_tmp, 0, &t1, ops);
            ops->GetVecFromMultiVec(ws->V_tmp, 1, &t2, ops);
            for (GCGE_INT idx = 0; idx < start; ++idx)
                column_residual(A, B, eval[idx], evec, idx, t1, t2, ops, &pr[idx], &px[idx], &pb[idx]);
            ops->RestoreVecForMultiVec(ws->V_tmp, 0, &t1, ops);
            ops->RestoreVecForMultiVec(ws->V_tmp, 1, &t2, ops);
        }
        {
            for (GCGE_INT idx = 0; idx < start; ++idx) {
                GCGE_DOUBLE r = pr[idx] / px[idx];
                if (strcmp(conv_type, "R") == 0)
                    r /= (1.0 > fabs(eval[idx]) ? 1.0 : fabs(eval[idx]));
                else if (strcmp(conv_type, "O") == 0)
                    r /= (para->conv_omega_norm + fabs(eval[idx]));
                else if (strcmp(conv_type, "RB") == 0)
                    r = pr[idx] / (fabs(eval[idx]) * pb[idx]);
                res[idx] = r;
                if (first_bad < 0 && !(r < ev_tol && pr[idx] / px[idx] < 1e-2)) first_bad = idx;
            }
        }
        free(pr);
        if (first_bad >= 0) {
            num_unlock = 0;
            for (GCGE_INT i = first_bad; i < nev && num_unlock < para->block_size; ++i)
                unlock[num_unlock++] = i;
            num_conv_continuous = first_bad;
        }
    }

    para->num_unlock = num_unlock;
    para->num_conv_continuous = num_conv_continuous;
    ws->unconv_bs = num_unlock;
    para->num_iter += 1;
    para->max_res = max_res;
    para->max_ind = max_ind;
    para->min_res = min_res;
    para->min_ind = min_ind;
    para->sum_res = sum_res;
    if (para->history && para->history_len < para->history_cap) {
        GCGE_DOUBLE *row = para->history + (size_t)para->history_len * (nev + 2);
        row[0] = para->num_iter;
        row[1] = num_conv_continuous;
        memcpy(row + 2, eval, (size_t)nev * sizeof(GCGE_DOUBLE));
        para->history_len++;
    }
}
