/*
 * Small dense host kernels used by the GCG core: the gemm/symm-shaped
 * products of the Rayleigh-Ritz step and of the in-subspace
 * orthonormalisation, the Array* helpers, and a symmetric eigensolver.
 *
 * They are the table defaults for DenseMatDotDenseMat,
 * DenseSymMatDotDenseMat, DenseMatEigenSolver and Array*, which the
 * reference routes to BLAS/LAPACK (src/gcge_ops.c:401-646).  The matrices
 * are at most (nev+20+2*block)^2, so plain column-major loops are enough
 * here; the large tall-skinny products live in the device backend.
 *
 * Eigensolver: Householder reduction to tridiagonal form followed by the
 * implicit-shift QL iteration, then an ascending sort.  The QL sweep
 * (tridiag_ql below) follows the statement order of the tqli procedure in
 * Press, Teukolsky, Vetterling and Flannery, Numerical Recipes in C, 2nd ed.,
 * section 11.3 (itself a restatement of the Wilkinson-Reinsch tql2 of the
 * Handbook for Automatic Computation vol. II), including its early exit when
 * a rotation radius underflows to zero; the rotations of a sweep are recorded
 * and applied to the eigenvector columns in one pass.  GCGE_Default_DenseMatEigenSolver keeps the reference dispatch
 * (src/gcge_ops.c:401-432): the whole spectrum when il == 1 and iu == n is
 * computed in place in `a` and then copied to evec as n*n contiguous
 * doubles (which is what the reference's memcpy after dsyev does, also
 * when lda != n); otherwise eigenpairs il..iu are written to evec with
 * leading dimension lde.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

/* The dense products and the eigensolver are compiled twice, for baseline
 * x86-64 and for AVX2 (helpers inlined into each clone), and the clone is
 * picked at load time where the host has AVX2.  Their vector loops are
 * element-wise and their sums keep their order (no reassociation), and the
 * clone targets AVX2 without fused multiply-adds, so both compute the same
 * roundings: the results are bit-identical, the AVX2 one 4 doubles per
 * instruction instead of 2. */
#ifndef GCGE_VECTOR_CLONES
#if defined(__GNUC__) && !defined(__clang__) && defined(__x86_64__)
#define GCGE_VECTOR_CLONES __attribute__((target_clones("avx2", "default")))
#else
#define GCGE_VECTOR_CLONES
#endif
#endif
#define GCGE_INLINE static inline __attribute__((always_inline))

static inline int lsame(const char *c, char u) { return (*c == u) || (*c == (char)(u + 32)); }

/* c = alpha op(a) op(b) + beta c, all column-major (Fortran dgemm contract:
 * with beta == 0, c is not read). */
GCGE_VECTOR_CLONES
void GCGE_Default_DenseMatDotDenseMat(char *transa, char *transb, GCGE_INT *nrows,
                                      GCGE_INT *ncols, GCGE_INT *mid, GCGE_DOUBLE *alpha,
                                      GCGE_DOUBLE *a, GCGE_INT *lda, GCGE_DOUBLE *b,
                                      GCGE_INT *ldb, GCGE_DOUBLE *beta, GCGE_DOUBLE *c,
                                      GCGE_INT *ldc)
{
    const int M = *nrows, N = *ncols, K = *mid;
    const int ta = lsame(transa, 'T') || lsame(transa, 'C');
    const int tb = lsame(transb, 'T') || lsame(transb, 'C');
    const double al = *alpha, be = *beta;
    const int LA = *lda, LB = *ldb, LC = *ldc;
    if (M <= 0 || N <= 0) return;
    for (int j = 0; j < N; ++j) {
        double *cj = c + (size_t)j * LC;
        if (be == 0.0) {
            for (int i = 0; i < M; ++i) cj[i] = 0.0;
        } else if (be != 1.0) {
            for (int i = 0; i < M; ++i) cj[i] *= be;
        }
        if (al == 0.0 || K <= 0) continue;
        if (!ta) {
            /* column sweep: c(:,j) += alpha * a(:,l) * op(b)(l,j) */
            for (int l = 0; l < K; ++l) {
                double blj = tb ? b[j + (size_t)l * LB] : b[l + (size_t)j * LB];
                double t = al * blj;
                if (t == 0.0) continue;
                const double *al_col = a + (size_t)l * LA;
                for (int i = 0; i < M; ++i) cj[i] += t * al_col[i];
            }
        } else {
            /* dot form: c(i,j) += alpha * a(:,i) . op(b)(:,j); four rows i at a
             * time (independent sums, each over l in order) */
            int i = 0;
            if (!tb) {
                const double *bj = b + (size_t)j * LB;
                for (; i + 4 <= M; i += 4) {
                    const double *a0 = a + (size_t)i * LA, *a1 = a0 + LA, *a2 = a1 + LA, *a3 = a2 + LA;
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                    for (int l = 0; l < K; ++l) {
                        const double bl = bj[l];
                        s0 += a0[l] * bl;
                        s1 += a1[l] * bl;
                        s2 += a2[l] * bl;
                        s3 += a3[l] * bl;
                    }
                    cj[i] += al * s0;
                    cj[i + 1] += al * s1;
                    cj[i + 2] += al * s2;
                    cj[i + 3] += al * s3;
                }
            }
            for (; i < M; ++i) {
                const double *ai = a + (size_t)i * LA;
                double s = 0.0;
                if (!tb) {
                    const double *bj = b + (size_t)j * LB;
                    for (int l = 0; l < K; ++l) s += ai[l] * bj[l];
                } else {
                    for (int l = 0; l < K; ++l) s += ai[l] * b[j + (size_t)l * LB];
                }
                cj[i] += al * s;
            }
        }
    }
}

/* c = alpha a b + beta c (side 'L') or alpha b a + beta c (side 'R'), with
 * a symmetric and only its `uplo` triangle referenced.  The triangle is first
 * expanded to a full square in scratch memory; every c(i,j) is then the sum
 * over k in increasing order of a(i,k) b(k,j) (side 'L') or b(i,k) a(k,j)
 * (side 'R'), formed for a whole column of c at a time. */
GCGE_VECTOR_CLONES
void GCGE_Default_DenseSymMatDotDenseMat(char *side, char *uplo, GCGE_INT *nrows, GCGE_INT *ncols,
                                         GCGE_DOUBLE *alpha, GCGE_DOUBLE *a, GCGE_INT *lda,
                                         GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta,
                                         GCGE_DOUBLE *c, GCGE_INT *ldc)
{
    const int M = *nrows, N = *ncols, LA = *lda, LB = *ldb, LC = *ldc;
    const int left = lsame(side, 'L'), upper = lsame(uplo, 'U');
    const double al = *alpha, be = *beta;
    if (M <= 0 || N <= 0) return;
    const int S = left ? M : N; /* order of a */
    double *full = (double *)malloc(((size_t)S * S + (size_t)M) * sizeof(double));
    if (!full) {
        fprintf(stderr, "GCGE_Default_DenseSymMatDotDenseMat: out of memory\n");
        abort();
    }
    double *s = full + (size_t)S * S; /* one column of sums */
    for (int k = 0; k < S; ++k)
        for (int i = 0; i < S; ++i)
            full[i + (size_t)k * S] = (upper ? (i <= k) : (i >= k)) ? a[i + (size_t)k * LA]
                                                                  : a[k + (size_t)i * LA];
    for (int j = 0; j < N; ++j) {
        for (int i = 0; i < M; ++i) s[i] = 0.0;
        if (left) {
            for (int k = 0; k < M; ++k) {
                const double bkj = b[k + (size_t)j * LB];
                const double *ak = full + (size_t)k * S;
                for (int i = 0; i < M; ++i) s[i] += ak[i] * bkj;
            }
        } else {
            for (int k = 0; k < N; ++k) {
                const double akj = full[k + (size_t)j * S];
                const double *bk = b + (size_t)k * LB;
                for (int i = 0; i < M; ++i) s[i] += bk[i] * akj;
            }
        }
        double *cj = c + (size_t)j * LC;
        for (int i = 0; i < M; ++i) cj[i] = (be == 0.0 ? 0.0 : be * cj[i]) + al * s[i];
    }
    free(full);
}

GCGE_DOUBLE GCGE_Default_ArrayDotArray(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length)
{
    double s = 0.0;
    for (int i = 0; i < length; ++i) s += x[i] * y[i];
    return s;
}

/* Euclidean norm with scaling, so that very large or tiny entries neither
 * overflow nor underflow. */
GCGE_DOUBLE GCGE_Default_ArrayNorm(GCGE_DOUBLE *x, GCGE_INT length)
{
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < length; ++i) {
        if (x[i] == 0.0) continue;
        double ax = fabs(x[i]);
        if (scale < ax) {
            double r = scale / ax;
            ssq = 1.0 + ssq * r * r;
            scale = ax;
        } else {
            double r = ax / scale;
            ssq += r * r;
        }
    }
    return scale * sqrt(ssq);
}

/* y = a x + b y: y is first scaled by b (when b != 1), then a x is added,
 * the order of the reference default (src/gcge_ops.c:608-622). */
void GCGE_Default_ArrayAXPBY(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_DOUBLE b, GCGE_DOUBLE *y,
                             GCGE_INT length)
{
    if (b != 1.0)
        for (int i = 0; i < length; ++i) y[i] *= b;
    if (a != 0.0)
        for (int i = 0; i < length; ++i) y[i] += a * x[i];
}

void GCGE_Default_ArrayCopy(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length)
{
    if (length > 0) memmove(y, x, (size_t)length * sizeof(double));
}

void GCGE_Default_ArrayScale(GCGE_DOUBLE alpha, GCGE_DOUBLE *a, GCGE_INT length)
{
    for (int i = 0; i < length; ++i) a[i] *= alpha;
}

/* ---- symmetric eigensolver -------------------------------------------- */
/* Householder tridiagonalisation, explicit accumulation of the orthogonal
 * factor and implicit-shift QL on the tridiagonal, all on column-major
 * storage so that every inner loop runs over one contiguous column (the
 * matrix-vector product, the rank-2 update, the reflector applications and
 * the plane rotations).  Backward stable; eigenvalues agree with LAPACK
 * dsyev to rounding, eigenvectors up to sign and the choice of basis inside
 * multiple eigenvalues. */



GCGE_INLINE double hypot_safe(double a, double b)
{
    double aa = fabs(a), bb = fabs(b);
    if (aa > bb) {
        double r = bb / aa;
        return aa * sqrt(1.0 + r * r);
    }
    if (bb == 0.0) return 0.0;
    double r = aa / bb;
    return bb * sqrt(1.0 + r * r);
}

/* sqrt(a^2 + b^2) with one square root when the squares neither overflow
 * nor underflow (the division of the scaled form sits on the QL
 * recurrence's dependency chain), the scaled form otherwise */
GCGE_INLINE double hypot_fast(double a, double b)
{
    const double s = a * a + b * b;
    if (s > 1e-290 && s < 1e290) return sqrt(s);
    return hypot_safe(a, b);
}

/* 2-norm of x[0..m-1]: sum of squares, scaled form when out of range */
GCGE_INLINE double norm2(const double *x, int m)
{
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += x[i] * x[i];
    if (s > 1e-290 && s < 1e290) return sqrt(s);
    double r = 0.0;
    for (int i = 0; i < m; ++i) r = hypot_safe(r, x[i]);
    return r;
}

/* A (n x n, column-major, full symmetric) -> tridiagonal (d, e) with
 * A = Q T Q^T; e[i] couples i and i+1.  The Householder vectors are left in
 * the columns of A below the subdiagonal (unit leading entry implied),
 * their factors in tau. */
GCGE_INLINE void sym_tridiag(int n, double *A, double *d, double *e, double *tau, double *p)
{
#define AC(i, j) A[(size_t)(j) * n + (i)]
    for (int j = 0; j < n - 2; ++j) {
        const int m = n - j - 1; /* length of the column below the diagonal */
        double *x = &AC(j + 1, j);
        double alpha = x[0];
        const double xnorm = m > 1 ? norm2(x + 1, m - 1) : 0.0;
        d[j] = AC(j, j);
        if (xnorm == 0.0) {
            tau[j] = 0.0;
            e[j] = alpha;
            continue;
        }
        double beta = -copysign(hypot_safe(alpha, xnorm), alpha);
        tau[j] = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (int i = 1; i < m; ++i) x[i] *= scal;
        x[0] = 1.0;
        e[j] = beta;
        /* p = tau A22 v, A22 = A(j+1:, j+1:) */
        for (int i = 0; i < m; ++i) p[i] = 0.0;
        for (int c = 0; c < m; ++c) {
            const double vc = tau[j] * x[c];
            const double *col = &AC(j + 1, j + 1 + c);
            for (int i = 0; i < m; ++i) p[i] += col[i] * vc;
        }
        /* w = p - (tau/2)(p.v) v, stored in p */
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * x[i];
        const double h = 0.5 * tau[j] * pv;
        for (int i = 0; i < m; ++i) p[i] -= h * x[i];
        /* A22 -= v w^T + w v^T */
        for (int c = 0; c < m; ++c) {
            double *col = &AC(j + 1, j + 1 + c);
            const double wc = p[c], vc = x[c];
            for (int i = 0; i < m; ++i) col[i] -= x[i] * wc + p[i] * vc;
        }
    }
    if (n >= 2) {
        d[n - 2] = AC(n - 2, n - 2);
        e[n - 2] = AC(n - 1, n - 2);
        tau[n - 2] = 0.0;
    }
    d[n - 1] = AC(n - 1, n - 1);
    e[n - 1] = 0.0;
#undef AC
}

/* Q = H_0 H_1 ... H_{n-3} into Z (column-major, ldz), reflectors applied
 * from the last to the first so each touches only the trailing block. */
GCGE_INLINE void form_q(int n, const double *A, const double *tau, double *Z, int ldz)
{
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) Z[(size_t)c * ldz + i] = (i == c) ? 1.0 : 0.0;
    for (int j = n - 3; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        const int m = n - j - 1;
        const double *v = A + (size_t)j * n + j + 1;
        for (int c = j + 1; c < n; ++c) {
            double *z = Z + (size_t)c * ldz + j + 1;
            double s = z[0];
            for (int i = 1; i < m; ++i) s += v[i] * z[i];
            s *= tau[j];
            z[0] -= s;
            for (int i = 1; i < m; ++i) z[i] -= s * v[i];
        }
    }
}

/* Apply the plane rotations i = hi-1 .. lo of one QL sweep to columns
 * lo .. hi of Z, as the sweep would one at a time:
 *     (z_i, z_{i+1}) <- (c_i z_i - s_i z_{i+1}, s_i z_i + c_i z_{i+1}).
 * Rows are taken in chunks whose running column stays in registers, so every
 * element is loaded and stored once per sweep. */
GCGE_INLINE void apply_sweep(double *Z, int ldz, int n, int lo, int hi, const double *cs,
                        const double *sn)
{
    enum { CH = 16 };
    for (int k0 = 0; k0 < n; k0 += CH) {
        const int kn = (n - k0 < CH) ? n - k0 : CH;
        double t[CH];
        const double *zh = Z + (size_t)hi * ldz + k0;
        for (int k = 0; k < kn; ++k) t[k] = zh[k];
        for (int i = hi - 1; i >= lo; --i) {
            const double c = cs[i], s = sn[i];
            const double *zi = Z + (size_t)i * ldz + k0;
            double *zn = Z + (size_t)(i + 1) * ldz + k0;
            for (int k = 0; k < kn; ++k) {
                const double ze = zi[k];
                zn[k] = s * ze + c * t[k];
                t[k] = c * ze - s * t[k];
            }
        }
        double *zl = Z + (size_t)lo * ldz + k0;
        for (int k = 0; k < kn; ++k) zl[k] = t[k];
    }
}

/* Implicit-shift QL on the tridiagonal (d, e), rotating columns of Z; the
 * rotations of a sweep are recorded in cs/sn (n entries each) and applied
 * together.  Returns 0 on success, 1 when an eigenvalue needs more than 60
 * sweeps. */
GCGE_INLINE int tridiag_ql(int n, double *d, double *e, double *Z, int ldz, double *cs, double *sn)
{
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1.1102230246251565e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) return 1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot_safe(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    double b = c * e[i];
                    r = hypot_fast(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    const double ri = 1.0 / r;
                    s = f * ri;
                    c = g * ri;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    cs[i] = c;
                    sn[i] = s;
                }
                /* rotations m-1 .. i+1 were taken (all of them when i < l) */
                if (i + 1 < m) apply_sweep(Z, ldz, n, i + 1, m, cs, sn);
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return 0;
}

/* All eigenpairs of the symmetric n x n matrix whose upper triangle is in
 * a (column-major, lda).  Eigenvalues ascending in eval; eigenvector j in
 * evec[j*lde .. j*lde+n-1] (lde >= n).  work needs n*n + 5n doubles. */
GCGE_VECTOR_CLONES
void GCGE_Dense_SymEig(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT n, GCGE_DOUBLE *eval,
                       GCGE_DOUBLE *evec, GCGE_INT lde, GCGE_DOUBLE *work)
{
    if (n <= 0) return;
    double *A = work, *d = work + (size_t)n * n, *e = d + n, *tau = e + n, *p = tau + n,
           *sn = p + n;
    /* full symmetric copy from the upper triangle */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            const double v = a[i + (size_t)j * lda];
            A[(size_t)j * n + i] = v;
            A[(size_t)i * n + j] = v;
        }
    sym_tridiag(n, A, d, e, tau, p);
    form_q(n, A, tau, evec, lde);
    /* tau and p are free again: they hold the rotations of a sweep */
    tridiag_ql(n, d, e, evec, lde, tau, sn);
    /* selection sort of the eigenvalues, swapping vector columns */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double pv = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < pv) {
                k = j;
                pv = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = pv;
            double *zi = evec + (size_t)i * lde, *zk = evec + (size_t)k * lde;
            for (int r = 0; r < n; ++r) {
                const double t = zi[r];
                zi[r] = zk[r];
                zk[r] = t;
            }
        }
    }
    for (int j = 0; j < n; ++j) eval[j] = d[j];
}

void GCGE_Default_DenseMatEigenSolver(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT nrows,
                                      GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_INT lde,
                                      GCGE_INT il, GCGE_INT iu, GCGE_INT *iwork,
                                      GCGE_DOUBLE *dwork)
{
    (void)iwork;
    (void)dwork;
    const int n = nrows;
    if (n <= 0) return;
    double *vals = (double *)malloc((size_t)n * sizeof(double));
    double *vecs = (double *)malloc((size_t)n * n * sizeof(double));
    double *work = (double *)malloc(((size_t)n * n + 5 * (size_t)n) * sizeof(double));
    GCGE_Dense_SymEig(a, lda, n, vals, vecs, n, work);
    if (il == 1 && iu == n) {
        /* whole spectrum: vectors replace a (leading dimension lda), then n*n
         * contiguous doubles are copied from a to evec */
        for (int j = 0; j < n; ++j) {
            eval[j] = vals[j];
            memcpy(a + (size_t)j * lda, vecs + (size_t)j * n, (size_t)n * sizeof(double));
        }
        memcpy(evec, a, (size_t)n * n * sizeof(double));
    } else {
        for (int j = il - 1; j < iu; ++j) {
            eval[j - (il - 1)] = vals[j];
            memcpy(evec + (size_t)(j - (il - 1)) * lde, vecs + (size_t)j * n,
                   (size_t)n * sizeof(double));
        }
    }
    free(vals);
    free(vecs);
    free(work);
}
