/*
 * B-orthonormalisation of large blocks and orthonormalisation of small
 * coefficient blocks.
 *
 *   GCGE_Orthonormalization        reference src/gcge_orthogonal.c:112-308
 *       (default for X: column-by-column Gram-Schmidt with block inner
 *       products, re-orthogonalisation while the norm drops below
 *       reorth_tol, zero columns swapped to the end and dropped)
 *   GCGE_StableMultiOrthonormalization (+Partly)   :1909-2019
 *       (default for W: project out V(:, 0:start) up to max_reorth_time
 *       times while ||coef||_F > scbgs_reorth_tol, then the recursive
 *       self-orthonormalisation; chunked when W exceeds V_tmp)
 *   GCGE_SubOrthonormalization     :2022-2071   V2 -= V1 (V1^T B V2)
 *   GCGE_MultiOrthonormalization   :2087-2178   recursive bisection
 *   GCGE_SubOrthonormalizationSelfBGS :2180-2263 (leaves)
 *   GCGE_SVQBOrthonormalization    no reference counterpart: w_orth_type
 *       "svqb" replaces the recursive self-orthonormalisation of W by two
 *       SVQB passes (one Gram block and one linear combination each)
 *   GCGE_OrthonormalizationInSubspace :584-648
 *   GCGE_MultiOrthonormalizationInSubspace / StableMulti...InSubspace
 *       :2278-2395 (default for the P coefficients)
 * In the two in-subspace routines the update V2 -= V1 D runs over ldV rows
 * although only nrows rows belong to the columns, exactly like the
 * reference's dgemm calls; the extra rows are the leading (soft-locked)
 * rows of the next columns of the coefficient matrix.  This is kept so that
 * the coefficients, and hence the iterates, match the reference.
 * In SelfBGS the elimination of the current column from the later ones is
 * one linear combination with beta = 1 instead of a loop of axpby calls;
 * each column gets y + c x either way.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

void GCGE_Orthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                             GCGE_PARA *para, void **V_tmp, GCGE_DOUBLE *d_tmp)
{
    GCGE_ORTH_PARA *op = para->orth_para;
    GCGE_INT mv_s[2], mv_e[2];
    GCGE_DOUBLE vin = 0.0, vout = 0.0, ratio, norm_value = 0.0, ipv;
    const GCGE_DOUBLE sum_res = para->sum_res;
    void *vc, *vt, *vt2;
    for (GCGE_INT current = start; current < *end; ++current) {
        GCGE_INT reorth_count = 0;
        if (sum_res > 0.0 && sum_res < op->criterion_tol) {
            /* single-vector modified Gram-Schmidt (the reference never
             * increments the count on this path, so it stops on the ratio
             * or on a zero vector only) */
            ops->GetVecFromMultiVec(V, current, &vc, ops);
            if (B == NULL) {
                ops->VecInnerProd(vc, vc, &norm_value, ops);
            } else {
                ops->GetVecFromMultiVec(V_tmp, 1, &vt2, ops);
                ops->MatDotVec(B, vc, vt2, ops);
                ops->VecInnerProd(vc, vt2, &norm_value, ops);
                ops->RestoreVecForMultiVec(V_tmp, 1, &vt2, ops);
            }
            vout = sqrt(norm_value);
            do {
                for (GCGE_INT j = 0; j < current; ++j) {
                    ops->GetVecFromMultiVec(V, j, &vt, ops);
                    if (B == NULL) {
                        ops->VecInnerProd(vc, vt, &ipv, ops);
                    } else {
                        ops->GetVecFromMultiVec(V_tmp, 1, &vt2, ops);
                        ops->MatDotVec(B, vc, vt2, ops);
                        ops->VecInnerProd(vt, vt2, &ipv, ops);
                        ops->RestoreVecForMultiVec(V_tmp, 1, &vt2, ops);
                    }
                    ops->VecAxpby(-ipv, vt, 1.0, vc, ops);
                    ops->RestoreVecForMultiVec(V, j, &vt, ops);
                }
                vin = vout;
                if (B == NULL) {
                    ops->VecInnerProd(vc, vc, &norm_value, ops);
                } else {
                    ops->GetVecFromMultiVec(V_tmp, 1, &vt2, ops);
                    ops->MatDotVec(B, vc, vt2, ops);
                    ops->VecInnerProd(vc, vt2, &norm_value, ops);
                    ops->RestoreVecForMultiVec(V_tmp, 1, &vt2, ops);
                }
                vout = sqrt(norm_value);
                ratio = vout / vin;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time &&
                     vout > op->orth_zero_tol);
            ops->RestoreVecForMultiVec(V, current, &vc, ops);
        } else {
            do {
                /* V_tmp[0] = B v_current; d = V(:, 0:current+1)^T V_tmp[0] */
                mv_s[0] = current; mv_e[0] = current + 1;
                mv_s[1] = 0; mv_e[1] = 1;
                if (B == NULL)
                    ops->MultiVecAxpby(1.0, V, 0.0, V_tmp, mv_s, mv_e, ops);
                else
                    ops->MatDotMultiVec(B, V, V_tmp, mv_s, mv_e, ops);
                mv_s[0] = 0; mv_e[0] = current + 1;
                mv_s[1] = 0; mv_e[1] = 1;
                ops->MultiVecInnerProd(V, V_tmp, d_tmp, "ns", mv_s, mv_e, current + 1, 0, ops);
                if (current != 0) {
                    /* V_tmp[0] = V(:, 0:current) d; v_current -= V_tmp[0] */
                    mv_s[0] = 0; mv_e[0] = current;
                    mv_s[1] = 0; mv_e[1] = 1;
                    ops->MultiVecLinearComb(V, V_tmp, mv_s, mv_e, d_tmp, current, 0, 1.0, 0.0, ops);
                    mv_s[0] = 0; mv_e[0] = 1;
                    mv_s[1] = current; mv_e[1] = current + 1;
                    ops->MultiVecAxpby(-1.0, V_tmp, 1.0, V, mv_s, mv_e, ops);
                }
                vin = sqrt(d_tmp[current]);
                d_tmp[current] = d_tmp[current] - ops->ArrayDotArray(d_tmp, d_tmp, current);
                vout = sqrt(d_tmp[current]);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time &&
                     vout > op->orth_zero_tol);
        }
        if (vout > op->orth_zero_tol) {
            ops->GetVecFromMultiVec(V, current, &vc, ops);
            ops->VecAxpby(0.0, vc, 1.0 / vout, vc, ops);
            ops->RestoreVecForMultiVec(V, current, &vc, ops);
        } else {
            if (current < *end - 1) {
                mv_s[0] = current; mv_e[0] = current + 1;
                mv_s[1] = *end - 1; mv_e[1] = *end;
                ops->MultiVecSwap(V, V, mv_s, mv_e, ops);
            }
            (*end)--;
            current--;
            if (op->print_orth_zero == 1)
                GCGE_Printf("In Orthonormalization, there is a zero vector!, current = %d, start = "
                            "%d, end = %d\n", current, start, *end);
        }
    }
}

void GCGE_StableMultiOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                        GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const GCGE_INT chunk = ws->V_tmp_size;
    if (chunk >= *end - start) {
        GCGE_StableMultiOrthonormalizationPartly(V, start, end, B, ops, para, ws);
        return;
    }
    GCGE_INT cs = start, ce = start + chunk;
    GCGE_INT mv_s[2], mv_e[2];
    while (ce <= *end) {
        GCGE_INT old_end = ce;
        GCGE_StableMultiOrthonormalizationPartly(V, cs, &ce, B, ops, para, ws);
        if (old_end < *end) {
            if (ce < old_end) {
                /* zero columns were dropped inside the chunk: move the same
                 * number of columns from the tail into the gap */
                GCGE_INT n_zero = old_end - ce;
                GCGE_INT copy_start = *end - n_zero;
                if (old_end <= copy_start) {
                    mv_s[0] = ce; mv_e[0] = old_end;
                    mv_s[1] = copy_start; mv_e[1] = *end;
                } else {
                    mv_s[0] = ce; mv_e[0] = copy_start;
                    mv_s[1] = old_end; mv_e[1] = *end;
                }
                ops->MultiVecSwap(V, V, mv_s, mv_e, ops);
                *end = copy_start;
            }
            cs = ce;
            ce = (*end < cs + chunk) ? *end : cs + chunk;
        } else {
            ce = *end + 1;
        }
    }
}

static GCGE_DOUBLE frobenius(const GCGE_DOUBLE *d, GCGE_INT len)
{
    GCGE_DOUBLE s = 0.0;
    for (GCGE_INT i = 0; i < len; ++i) s += d[i] * d[i];
    return sqrt(s);
}

/* One projection pass of the svqb path: GCGE_SubOrthonormalization with the
 * subtraction folded into the linear combination, V2 <- V2 + V1 (-d) with
 * beta = 1 (one pass over V2 instead of a temporary block plus an axpby).
 * subspace_dtmp holds -d afterwards (the same Frobenius norm).  A backend
 * that forms the combination first and adds it once gives the bits of the
 * two-step form; the default column loops accumulate into V2 term by term
 * and round differently. */
static void sub_orth_fused(void **V, GCGE_INT *start, GCGE_INT *end, void *B, void **V_tmp,
                           GCGE_DOUBLE *subspace_dtmp, GCGE_OPS *ops)
{
    const GCGE_INT len1 = end[0] - start[0], len2 = end[1] - start[1];
    GCGE_INT mv_s[2], mv_e[2];
    mv_s[0] = start[1]; mv_e[0] = end[1];
    mv_s[1] = 0; mv_e[1] = len2;
    if (B == NULL)
        ops->MultiVecAxpby(1.0, V, 0.0, V_tmp, mv_s, mv_e, ops);
    else
        ops->MatDotMultiVec(B, V, V_tmp, mv_s, mv_e, ops);
    mv_s[0] = start[0]; mv_e[0] = end[0];
    mv_s[1] = 0; mv_e[1] = len2;
    ops->MultiVecInnerProd(V, V_tmp, subspace_dtmp, "ns", mv_s, mv_e, len1, 0, ops);
    for (GCGE_INT i = 0; i < len1 * len2; ++i) subspace_dtmp[i] = -subspace_dtmp[i];
    mv_s[0] = start[0]; mv_e[0] = end[0];
    mv_s[1] = start[1]; mv_e[1] = end[1];
    ops->MultiVecLinearComb(V, V, mv_s, mv_e, subspace_dtmp, len1, 0, 1.0, 1.0, ops);
}

static int svqb_orth(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                     GCGE_PARA *para, GCGE_WORKSPACE *ws, int bw_ready, const GCGE_DOUBLE *dgks_d,
                     GCGE_INT dgks_ld);

void GCGE_StableMultiOrthonormalizationPartly(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                              GCGE_OPS *ops, GCGE_PARA *para,
                                              GCGE_WORKSPACE *ws)
{
    int bw_ready = 0;
    if (start > 0) {
        GCGE_INT mv_s[2] = {0, start}, mv_e[2] = {start, *end};
        GCGE_DOUBLE value_inner = 1.0;
        GCGE_INT iter = 0;
        const GCGE_DOUBLE tol = para->orth_para->scbgs_reorth_tol;
        const int fused = strcmp(para->w_orth_type, "svqb") == 0;
        if (fused && para->svqb_dgks_eta > 0.0 && para->orth_para->max_reorth_time > 1) {
            /* one pass, then the SVQB on its B-product: it runs to the end
             * when every column kept at least eta of its B-norm in the pass
             * (the norms come from the SVQB Gram and the pass's
             * coefficients), and otherwise returns before changing W, and
             * the remaining passes run as below */
            iter = 1;
            sub_orth_fused(V, mv_s, mv_e, B, ws->V_tmp, ws->subspace_dtmp, ops);
            value_inner = frobenius(ws->subspace_dtmp, start * (*end - start));
            if (svqb_orth(V, start, end, B, ops, para, ws, 1, ws->subspace_dtmp, start) == 0) return;
        }
        while (value_inner > tol && iter < para->orth_para->max_reorth_time) {
            iter++;
            if (fused)
                sub_orth_fused(V, mv_s, mv_e, B, ws->V_tmp, ws->subspace_dtmp, ops);
            else
                GCGE_SubOrthonormalization(V, mv_s, mv_e, B, ws->V_tmp, ws->subspace_dtmp, ops);
            value_inner = frobenius(ws->subspace_dtmp, start * (*end - start));
        }
        /* after a second pass (or a first one below the tolerance) V_tmp holds
         * B W' for the W' before the last correction W = W' - V d, and the
         * SVQB Gram W^T (B W') = W^T B W + (W^T B V) d differs from W^T B W
         * only by the product of two rounding-level factors (W is B-orthogonal
         * to V up to rounding, d at most the size of W) */
        bw_ready = fused && (iter >= 2 || value_inner <= tol);
    }
    if (strcmp(para->w_orth_type, "svqb") == 0)
        svqb_orth(V, start, end, B, ops, para, ws, bw_ready, NULL, 0);
    else
        GCGE_MultiOrthonormalization(V, start, end, B, ops, para, ws);
}

/* B-orthonormalise columns start .. *end-1 among themselves by SVQB
 * passes: G = W^T B W, D = diag(G)^(-1/2) (columns with sqrt(G_jj) <=
 * orth_zero_tol are zero columns, as in the recursive leaves, and dropped),
 * D G D = U diag(l) U^T, W <- W D U diag(l)^(-1/2) over the directions with
 * l > 1e-14 max(l), largest first.  One Gram and one linear combination per
 * pass instead of one Gram per bisection level and leaf; the span of the
 * surviving columns is the same as with the recursive algorithm.  A second
 * pass runs unless the first kept every column with a scaled-Gram condition
 * number below para->svqb_one_pass_cond.  bw_ready: V_tmp(:, 0:k) already
 * holds B W for the first pass (from the projection). */
void GCGE_SVQBOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                                 GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    svqb_orth(V, start, end, B, ops, para, ws, 0, NULL, 0);
}

/* dgks_d (leading dimension dgks_ld): the coefficients d of the projection
 * pass just made, W = W' - V d.  Column j then kept the fraction
 * sqrt(G_jj / (G_jj + |d_j|^2)) of its B-norm; below svqb_dgks_eta for any
 * column the call returns 1 before W is changed (another projection pass
 * is due), else 0. */
static int svqb_orth(void **V, GCGE_INT start, GCGE_INT *end, void *B, GCGE_OPS *ops,
                     GCGE_PARA *para, GCGE_WORKSPACE *ws, int bw_ready, const GCGE_DOUBLE *dgks_d,
                     GCGE_INT dgks_ld)
{
    const GCGE_DOUBLE zero_tol = para->orth_para->orth_zero_tol;
    GCGE_INT k = *end - start;
    int rc = 0;
    if (k <= 0) return 0;
    GCGE_DOUBLE *G = (GCGE_DOUBLE *)malloc((size_t)k * k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *S = (GCGE_DOUBLE *)malloc((size_t)k * k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *U = (GCGE_DOUBLE *)malloc((size_t)k * k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *C = (GCGE_DOUBLE *)malloc((size_t)k * k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *lam = (GCGE_DOUBLE *)malloc((size_t)k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *dsc = (GCGE_DOUBLE *)malloc((size_t)k * sizeof(GCGE_DOUBLE));
    GCGE_DOUBLE *work = (GCGE_DOUBLE *)malloc(((size_t)k * k + 5 * (size_t)k) * sizeof(GCGE_DOUBLE));
    GCGE_INT *keep = (GCGE_INT *)malloc((size_t)k * sizeof(GCGE_INT));
    GCGE_INT mv_s[2], mv_e[2];
    for (int pass = 0; pass < 2 && k > 0; ++pass) {
        /* V_tmp = B W; G = W^T V_tmp */
        mv_s[0] = start; mv_e[0] = *end;
        mv_s[1] = 0; mv_e[1] = k;
        if (pass > 0 || !bw_ready) {
            if (B == NULL)
                ops->MultiVecAxpby(1.0, V, 0.0, ws->V_tmp, mv_s, mv_e, ops);
            else
                ops->MatDotMultiVec(B, V, ws->V_tmp, mv_s, mv_e, ops);
        }
        ops->MultiVecInnerProd(V, ws->V_tmp, G, "ns", mv_s, mv_e, k, 0, ops);
        if (pass == 0 && dgks_d) {
            const GCGE_DOUBLE eta2 = para->svqb_dgks_eta * para->svqb_dgks_eta;
            for (GCGE_INT j = 0; j < k && !rc; ++j) {
                const GCGE_DOUBLE *dj = dgks_d + (size_t)j * dgks_ld;
                GCGE_DOUBLE dn = 0.0;
                for (GCGE_INT i = 0; i < dgks_ld; ++i) dn += dj[i] * dj[i];
                const GCGE_DOUBLE gjj = G[(size_t)j * k + j];
                if (!(gjj >= eta2 * (gjj + dn))) rc = 1;
            }
            if (rc) break;  /* W unchanged: another projection pass first */
        }
        GCGE_INT m = 0;
        for (GCGE_INT j = 0; j < k; ++j) {
            const GCGE_DOUBLE gjj = G[(size_t)j * k + j];
            const GCGE_DOUBLE nrm = gjj > 0.0 ? sqrt(gjj) : 0.0;
            if (nrm > zero_tol) {
                keep[m++] = j;
                dsc[j] = 1.0 / nrm;
            } else {
                dsc[j] = 0.0;
            }
        }
        GCGE_INT r = 0;
        if (m > 0) {
            /* scaled, symmetrised Gram of the kept columns (upper triangle) */
            for (GCGE_INT b = 0; b < m; ++b)
                for (GCGE_INT a = 0; a <= b; ++a) {
                    const GCGE_INT ja = keep[a], jb = keep[b];
                    S[(size_t)b * m + a] = 0.5 * (G[(size_t)jb * k + ja] + G[(size_t)ja * k + jb]) *
                                           dsc[ja] * dsc[jb];
                }
            GCGE_Dense_SymEig(S, m, m, lam, U, m, work);
            const GCGE_DOUBLE lmax = lam[m - 1];
            /* coefficients over all k columns, directions largest first */
            for (GCGE_INT q = m - 1; q >= 0; --q) {
                if (!(lam[q] > 1e-14 * lmax)) break;
                GCGE_DOUBLE *c = C + (size_t)r * k;
                for (GCGE_INT j = 0; j < k; ++j) c[j] = 0.0;
                const GCGE_DOUBLE f = 1.0 / sqrt(lam[q]);
                for (GCGE_INT a = 0; a < m; ++a) c[keep[a]] = U[(size_t)q * m + a] * dsc[keep[a]] * f;
                ++r;
            }
        }
        GCGE_OPS_EXT *ext = GCGE_OPS_GetExt(ops);
        if (r > 0 && !(para->use_ops_ext && ext && ext->MultiVecLinearCombInPlace &&
                       ext->MultiVecLinearCombInPlace(V, start, k, C, k, start, r, ops) == 0)) {
            /* V_tmp(:, 0:r) = W C; W(:, 0:r) = V_tmp(:, 0:r) */
            mv_s[0] = start; mv_e[0] = *end;
            mv_s[1] = 0; mv_e[1] = r;
            ops->MultiVecLinearComb(V, ws->V_tmp, mv_s, mv_e, C, k, 0, 1.0, 0.0, ops);
            mv_s[0] = 0; mv_e[0] = r;
            mv_s[1] = start; mv_e[1] = start + r;
            ops->MultiVecAxpby(1.0, ws->V_tmp, 0.0, V, mv_s, mv_e, ops);
        }
        if (r < k && para->orth_para->print_orth_zero == 1)
            GCGE_Printf("In GCGE_SVQBOrthonormalization, %d of %d columns dropped\n", k - r, k);
        /* one pass is enough when it kept every column and the scaled Gram
         * is well conditioned: the B-orthonormality error of an SVQB pass is
         * about u * lambda_max / lambda_min */
        const int kept_all = (r == k && r == m);
        const GCGE_DOUBLE cond = r > 0 ? lam[m - 1] / lam[m - r] : 0.0;
        *end = start + r;
        k = r;
        if (pass == 0 && kept_all && cond < para->svqb_one_pass_cond) break;
    }
    free(G);
    free(S);
    free(U);
    free(C);
    free(lam);
    free(dsc);
    free(work);
    free(keep);
    return rc;
}

void GCGE_SubOrthonormalization(void **V, GCGE_INT *start, GCGE_INT *end, void *B,
                                void **V_tmp, GCGE_DOUBLE *subspace_dtmp, GCGE_OPS *ops)
{
    const GCGE_INT len1 = end[0] - start[0], len2 = end[1] - start[1];
    GCGE_INT mv_s[2], mv_e[2];
    /* V_tmp = B V2 */
    mv_s[0] = start[1]; mv_e[0] = end[1];
    mv_s[1] = 0; mv_e[1] = len2;
    if (B == NULL)
        ops->MultiVecAxpby(1.0, V, 0.0, V_tmp, mv_s, mv_e, ops);
    else
        ops->MatDotMultiVec(B, V, V_tmp, mv_s, mv_e, ops);
    /* d = V1^T V_tmp */
    mv_s[0] = start[0]; mv_e[0] = end[0];
    mv_s[1] = 0; mv_e[1] = len2;
    ops->MultiVecInnerProd(V, V_tmp, subspace_dtmp, "ns", mv_s, mv_e, len1, 0, ops);
    /* V_tmp = V1 d; V2 -= V_tmp */
    ops->MultiVecLinearComb(V, V_tmp, mv_s, mv_e, subspace_dtmp, len1, 0, 1.0, 0.0, ops);
    mv_s[0] = 0; mv_e[0] = len2;
    mv_s[1] = start[1]; mv_e[1] = end[1];
    ops->MultiVecAxpby(-1.0, V_tmp, 1.0, V, mv_s, mv_e, ops);
}

void GCGE_MultiOrthonormalization(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                  GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const GCGE_INT length = *end - start;
    const GCGE_INT len1 = length / 2, len2 = length - len1;
    GCGE_INT mid = start + len1;
    GCGE_ORTH_PARA *op = para->orth_para;
    GCGE_INT mv_s[2], mv_e[2];

    if (length <= op->max_direct_orth_length) {
        GCGE_SubOrthonormalizationSelfBGS(V, start, end, B, para, ops, ws);
        return;
    }
    GCGE_MultiOrthonormalization(V, start, &mid, B, ops, para, ws);
    if (mid + len2 < *end) {
        /* the left half lost columns: pull the right half up against it */
        GCGE_INT new_end2 = mid + len2;
        GCGE_INT n_zero = *end - new_end2;
        mv_s[0] = mid; mv_e[0] = mid + n_zero;
        mv_s[1] = new_end2; mv_e[1] = *end;
        ops->MultiVecSwap(V, V, mv_s, mv_e, ops);
        *end = new_end2;
    }
    if (mid - start > 0) {
        GCGE_DOUBLE value_inner = 1.0;
        GCGE_INT iter = 0;
        while (value_inner > op->scbgs_reorth_tol) {
            iter++;
            mv_s[0] = start; mv_e[0] = mid;
            mv_s[1] = mid; mv_e[1] = *end;
            GCGE_SubOrthonormalization(V, mv_s, mv_e, B, ws->V_tmp, ws->subspace_dtmp, ops);
            /* the reference measures the first len1*len2 entries, the sizes
             * before any column was dropped */
            value_inner = frobenius(ws->subspace_dtmp, len1 * len2);
            if (iter >= op->max_reorth_time) break;
        }
        if (value_inner > op->scbgs_reorth_tol && op->print_orth_zero == 1)
            GCGE_Printf("in GCGE_SubOrthonormalization, V1: %d-%d, V2: %d-%d, reorth_count: %d, "
                        "value_inner: %e\n", start, mid - 1, mid, *end, op->max_reorth_time,
                        value_inner);
    }
    GCGE_MultiOrthonormalization(V, mid, end, B, ops, para, ws);
}

void GCGE_SubOrthonormalizationSelfBGS(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                                       GCGE_PARA *para, GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    void **V_tmp = ws->V_tmp;
    GCGE_DOUBLE *d = ws->subspace_dtmp;
    const GCGE_DOUBLE zero_tol = para->orth_para->orth_zero_tol;
    GCGE_INT mv_s[2], mv_e[2];
    for (GCGE_INT current = start; current < *end; ++current) {
        /* V_tmp[0] = B v; d = V(:, current:end)^T V_tmp[0] */
        mv_s[0] = current; mv_e[0] = current + 1;
        mv_s[1] = 0; mv_e[1] = 1;
        if (B == NULL)
            ops->MultiVecAxpby(1.0, V, 0.0, V_tmp, mv_s, mv_e, ops);
        else
            ops->MatDotMultiVec(B, V, V_tmp, mv_s, mv_e, ops);
        mv_s[0] = current; mv_e[0] = *end;
        mv_s[1] = 0; mv_e[1] = 1;
        const GCGE_INT len = *end - current;
        ops->MultiVecInnerProd(V, V_tmp, d, "ns", mv_s, mv_e, len, 0, ops);
        const GCGE_DOUBLE norm = sqrt(d[0]);
        if (norm > zero_tol) {
            void *vc;
            ops->GetVecFromMultiVec(V, current, &vc, ops);
            ops->VecAxpby(0.0, vc, 1.0 / norm, vc, ops);
            ops->RestoreVecForMultiVec(V, current, &vc, ops);
            if (len > 1) {
                /* v_j -= (d_j / norm) v_current for the later columns */
                GCGE_DOUBLE *coef = d + len;
                for (GCGE_INT j = 1; j < len; ++j) coef[j - 1] = -d[j] / norm;
                mv_s[0] = current; mv_e[0] = current + 1;
                mv_s[1] = current + 1; mv_e[1] = *end;
                ops->MultiVecLinearComb(V, V, mv_s, mv_e, coef, 1, 0, 1.0, 1.0, ops);
            }
        } else {
            if (current < *end - 1) {
                mv_s[0] = current; mv_e[0] = current + 1;
                mv_s[1] = *end - 1; mv_e[1] = *end;
                ops->MultiVecSwap(V, V, mv_s, mv_e, ops);
            }
            (*end)--;
            current--;
            if (para->orth_para->print_orth_zero == 1)
                GCGE_Printf("In GCGE_SubOrthonormalizationSelfBGS, there is a zero vector!, "
                            "current = %d, start = %d, end = %d\n", current, start, *end);
        }
    }
}

/* ---- small coefficient blocks (host arrays) ---------------------------- */

void GCGE_OrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                       GCGE_INT start, GCGE_INT *end, GCGE_DOUBLE *B,
                                       GCGE_INT ldB, GCGE_ORTH_PARA *op, GCGE_OPS *ops)
{
    (void)B;
    (void)ldB;
    for (GCGE_INT current = start; current < *end; ++current) {
        GCGE_DOUBLE *vc = V + (size_t)current * ldV;
        GCGE_DOUBLE vout = ops->ArrayNorm(vc, nrows), vin, ratio;
        if (current != 0) {
            GCGE_INT reorth_count = 0;
            do {
                vin = vout;
                for (GCGE_INT idx = 0; idx < current; ++idx) {
                    GCGE_DOUBLE *vi = V + (size_t)idx * ldV;
                    GCGE_DOUBLE ip = ops->ArrayDotArray(vi, vc, nrows);
                    ops->ArrayAXPBY(-ip, vi, 1.0, vc, nrows);
                }
                vout = ops->ArrayNorm(vc, nrows);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time &&
                     vout > op->orth_zero_tol);
        }
        if (vout > op->orth_zero_tol) {
            ops->ArrayScale(1.0 / vout, vc, nrows);
        } else {
            if (current != *end - 1) ops->ArrayCopy(V + (size_t)(*end - 1) * ldV, vc, nrows);
            (*end)--;
            current--;
            if (op->print_orth_zero == 1)
                GCGE_Printf("In OrthonormalizationInSubspace, there is a zero vector!, current = "
                            "%d, start = %d, end = %d\n", current, start, *end);
        }
    }
}

void GCGE_MultiOrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                            GCGE_INT start, GCGE_INT *end, void *B, GCGE_INT ldB,
                                            GCGE_ORTH_PARA *op, GCGE_WORKSPACE *ws, GCGE_OPS *ops)
{
    GCGE_INT length = *end - start;
    GCGE_INT len1 = length / 2, len2 = length - len1;
    GCGE_INT mid = start + len1;
    GCGE_DOUBLE *d = ws->subspace_dtmp;
    if (length <= 0) return; /* nothing to do (the reference would recurse forever) */
    if (length == 1) {
        GCGE_DOUBLE nrm = ops->ArrayNorm(V + (size_t)start * ldV, nrows);
        if (nrm > op->orth_zero_tol)
            ops->ArrayScale(1.0 / nrm, V + (size_t)start * ldV, nrows);
        else
            *end = start;
        return;
    }
    GCGE_MultiOrthonormalizationInSubspace(V, ldV, nrows, start, &mid, B, ldB, op, ws, ops);
    if (mid + len2 < *end) {
        GCGE_INT new_end2 = mid + len2;
        GCGE_INT n_zero = *end - new_end2;
        len1 -= n_zero;
        memcpy(V + (size_t)mid * ldV, V + (size_t)new_end2 * ldV,
               (size_t)n_zero * ldV * sizeof(GCGE_DOUBLE));
        *end = new_end2;
    }
    if (len1 > 0) {
        for (GCGE_INT r = 0; r < op->max_reorth_time; ++r) {
            GCGE_DOUBLE one = 1.0, zero = 0.0, mone = -1.0;
            GCGE_INT m = ldV;
            /* d = V1^T V2 over nrows rows; V2 -= V1 d over ldV rows */
            ops->DenseMatDotDenseMat("T", "N", &len1, &len2, &nrows, &one, V + (size_t)start * ldV,
                                     &ldV, V + (size_t)mid * ldV, &ldV, &zero, d, &len1);
            ops->DenseMatDotDenseMat("N", "N", &m, &len2, &len1, &mone, V + (size_t)start * ldV,
                                     &ldV, d, &len1, &one, V + (size_t)mid * ldV, &ldV);
        }
    }
    GCGE_MultiOrthonormalizationInSubspace(V, ldV, nrows, mid, end, B, ldB, op, ws, ops);
}

void GCGE_StableMultiOrthonormalizationInSubspace(GCGE_DOUBLE *V, GCGE_INT ldV, GCGE_INT nrows,
                                                  GCGE_INT start, GCGE_INT *end, void *B,
                                                  GCGE_INT ldB, GCGE_ORTH_PARA *op,
                                                  GCGE_WORKSPACE *ws, GCGE_OPS *ops)
{
    GCGE_DOUBLE *d = ws->subspace_dtmp;
    if (start != 0) {
        for (GCGE_INT r = 0; r < op->max_reorth_time; ++r) {
            GCGE_INT len = *end - start;
            GCGE_DOUBLE one = 1.0, zero = 0.0, mone = -1.0;
            GCGE_INT m = ldV;
            ops->DenseMatDotDenseMat("T", "N", &start, &len, &nrows, &one, V, &ldV,
                                     V + (size_t)start * ldV, &ldV, &zero, d, &start);
            ops->DenseMatDotDenseMat("N", "N", &m, &len, &start, &mone, V, &ldV, d, &start, &one,
                                     V + (size_t)start * ldV, &ldV);
        }
    }
    GCGE_MultiOrthonormalizationInSubspace(V, ldV, nrows, start, end, B, ldB, op, ws, ops);
}
