/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement of the reference CSR backend and of the reference
 * column-by-column multivec semantics.  It is the checker for the device
 * backend: tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it; the product never does.
 *
 *   CSR matrix/vector ops ........ reference app/csr/gcge_app_csr.c:284-396
 *     (MatDotVec zeroes r, then row-wise r_i += a_ij x_j in stored order;
 *      VecAxpby with the a == 0 / b == 0 / copy special cases;
 *      VecInnerProd as a serial sum)
 *   CSR text file read/write ..... app/csr/gcge_app_csr.c:185-236
 *   Dirichlet zeroing ............ app/csr/gcge_app_csr.c:130-183
 *   multivec defaults ............ src/gcge_ops.c:102-398 (column loops;
 *      LinearComb ignores alpha; "sym" inner product mirrors the upper
 *      triangle; Swap exchanges column handles)
 *   dense eigensolver ............ the contract of LAPACK dsyev/dsyevx as
 *      called from src/gcge_ops.c:401-432, computed here by cyclic Jacobi
 *      rotations (an algorithm independent of the product's QL solver)
 *   random initial vectors ....... uniform(-1, 1) from a counter-based
 *      generator keyed by (seed, draw number, column, row); the product's
 *      device generator uses the same key and mixer (SplitMix64 finaliser,
 *      Steele, Lea and Flood 2014), so both backends start from identical
 *      vectors.  (The reference draws uniform[0,1) from rand(), which no
 *      device can reproduce; see app/csr/gcge_app_csr.c:284-294.)
 *
 * Every function here is pinned by a check in tests/: brute-force dense
 * products and the reference's fixtures (SpMM, Gram, linear combination),
 * closed-form spectra and numpy (eigensolver), statistical properties
 * (generator), the reference's own block-CG runs (block CG).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge_ops.h"
#include "oracle_csr.h"

/* ---- CSR matrix and vectors ------------------------------------------- */

ORACLE_CSR *ORACLE_CSR_Create(int nrows, int ncols, int nnz, const int *rowptr, const int *colidx,
                              const double *vals)
{
    ORACLE_CSR *m = (ORACLE_CSR *)calloc(1, sizeof(ORACLE_CSR));
    m->nrows = nrows;
    m->ncols = ncols;
    m->nnz = nnz;
    m->rowptr = (int *)malloc((size_t)(nrows + 1) * sizeof(int));
    m->colidx = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    m->vals = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    memcpy(m->rowptr, rowptr, (size_t)(nrows + 1) * sizeof(int));
    if (nnz > 0) {
        memcpy(m->colidx, colidx, (size_t)nnz * sizeof(int));
        memcpy(m->vals, vals, (size_t)nnz * sizeof(double));
    }
    return m;
}

void ORACLE_CSR_Free(ORACLE_CSR *m)
{
    if (!m) return;
    free(m->rowptr);
    free(m->colidx);
    free(m->vals);
    free(m);
}

int ORACLE_CSR_Nrows(const ORACLE_CSR *m) { return m->nrows; }
int ORACLE_CSR_Nnz(const ORACLE_CSR *m) { return m->nnz; }

void ORACLE_CSR_Export(const ORACLE_CSR *m, int *rowptr, int *colidx, double *vals)
{
    memcpy(rowptr, m->rowptr, (size_t)(m->nrows + 1) * sizeof(int));
    memcpy(colidx, m->colidx, (size_t)m->nnz * sizeof(int));
    memcpy(vals, m->vals, (size_t)m->nnz * sizeof(double));
}

/* The reference text format: nrow, ncol, nnz, then ia[nrow+1], ja[nnz],
 * aa[nnz], one number per line, 0-based.  Returns NULL when the file cannot
 * be opened (the reference exits instead). */
ORACLE_CSR *ORACLE_CSR_ReadFile(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    int nrow = 0, ncol = 0, nnz = 0;
    if (fscanf(f, "%d %d %d", &nrow, &ncol, &nnz) != 3 || nrow < 0 || nnz < 0) {
        fclose(f);
        return NULL;
    }
    int *ia = (int *)malloc((size_t)(nrow + 1) * sizeof(int));
    int *ja = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    double *aa = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    int ok = 1;
    for (int i = 0; i < nrow + 1 && ok; ++i) ok = fscanf(f, "%d", ia + i) == 1;
    for (int i = 0; i < nnz && ok; ++i) ok = fscanf(f, "%d", ja + i) == 1;
    for (int i = 0; i < nnz && ok; ++i) ok = fscanf(f, "%lf", aa + i) == 1;
    fclose(f);
    ORACLE_CSR *m = ok ? ORACLE_CSR_Create(nrow, ncol, nnz, ia, ja, aa) : NULL;
    free(ia);
    free(ja);
    free(aa);
    return m;
}

int ORACLE_CSR_WriteFile(const ORACLE_CSR *m, const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "%d\n%d\n%d\n", m->nrows, m->ncols, m->nnz);
    for (int i = 0; i <= m->nrows; ++i) fprintf(f, "%d\n", m->rowptr[i]);
    for (int i = 0; i < m->nnz; ++i) fprintf(f, "%d\n", m->colidx[i]);
    for (int i = 0; i < m->nnz; ++i) fprintf(f, "%.17g\n", m->vals[i]);
    fclose(f);
    return 0;
}

typedef struct {
    int size;
    double *data;
} ORACLE_VEC;

static ORACLE_VEC *vec_new(int size)
{
    ORACLE_VEC *v = (ORACLE_VEC *)malloc(sizeof(ORACLE_VEC));
    v->size = size;
    v->data = (double *)calloc((size_t)(size > 0 ? size : 1), sizeof(double));
    return v;
}

/* ---- counter-based uniform(-1, 1) generator ---------------------------- */

static uint64_t g_seed = 0x5eed;
static uint64_t g_draw = 0;

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double ORACLE_RandomValue(uint64_t seed, uint64_t draw, uint64_t col, uint64_t row)
{
    uint64_t key = mix64(seed ^ mix64((draw << 32) ^ col));
    uint64_t bits = mix64(key + row);
    return (double)(bits >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

void ORACLE_SetSeed(uint64_t seed)
{
    g_seed = seed;
    g_draw = 0;
}

/* ---- single-vector ops ------------------------------------------------ */

static void o_vec_random(void *vec, GCGE_OPS *ops)
{
    (void)ops;
    ORACLE_VEC *v = (ORACLE_VEC *)vec;
    uint64_t draw = g_draw++;
    for (int i = 0; i < v->size; ++i) v->data[i] = ORACLE_RandomValue(g_seed, draw, 0, (uint64_t)i);
}

static void csr_matvec(const ORACLE_CSR *m, const double *x, double *r)
{
    memset(r, 0, (size_t)m->nrows * sizeof(double));
    for (int i = 0; i < m->nrows; ++i)
        for (int j = m->rowptr[i]; j < m->rowptr[i + 1]; ++j) r[i] += m->vals[j] * x[m->colidx[j]];
}

static void o_mat_dot_vec(void *mat, void *x, void *r, GCGE_OPS *ops)
{
    (void)ops;
    csr_matvec((ORACLE_CSR *)mat, ((ORACLE_VEC *)x)->data, ((ORACLE_VEC *)r)->data);
}

static void o_mat_t_dot_vec(void *mat, void *x, void *r, GCGE_OPS *ops)
{
    (void)ops;
    const ORACLE_CSR *m = (ORACLE_CSR *)mat;
    double *rd = ((ORACLE_VEC *)r)->data;
    const double *xd = ((ORACLE_VEC *)x)->data;
    memset(rd, 0, (size_t)m->ncols * sizeof(double));
    for (int i = 0; i < m->nrows; ++i)
        for (int j = m->rowptr[i]; j < m->rowptr[i + 1]; ++j) rd[m->colidx[j]] += m->vals[j] * xd[i];
}

static void axpby_array(double a, const double *x, double b, double *y, int n)
{
    if (a == 0.0) {
        for (int i = 0; i < n; ++i) y[i] = b * y[i];
    } else if (b == 0.0) {
        if (a == 1.0)
            for (int i = 0; i < n; ++i) y[i] = x[i];
        else
            for (int i = 0; i < n; ++i) y[i] = a * x[i];
    } else {
        for (int i = 0; i < n; ++i) y[i] = a * x[i] + b * y[i];
    }
}

static void o_vec_axpby(GCGE_DOUBLE a, void *x, GCGE_DOUBLE b, void *y, GCGE_OPS *ops)
{
    (void)ops;
    ORACLE_VEC *xv = (ORACLE_VEC *)x, *yv = (ORACLE_VEC *)y;
    axpby_array(a, xv->data, b, yv->data, xv->size);
}

static double dot_array(const double *x, const double *y, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

static void o_vec_inner(void *x, void *y, GCGE_DOUBLE *v, GCGE_OPS *ops)
{
    (void)ops;
    *v = dot_array(((ORACLE_VEC *)x)->data, ((ORACLE_VEC *)y)->data, ((ORACLE_VEC *)x)->size);
}

static void o_vec_create_by_vec(void **d, void *s, GCGE_OPS *ops)
{
    (void)ops;
    *d = vec_new(((ORACLE_VEC *)s)->size);
}

static void o_vec_create_by_mat(void **v, void *mat, GCGE_OPS *ops)
{
    (void)ops;
    *v = vec_new(((ORACLE_CSR *)mat)->nrows);
}

static void o_vec_destroy(void **v, GCGE_OPS *ops)
{
    (void)ops;
    if (!*v) return;
    free(((ORACLE_VEC *)*v)->data);
    free(*v);
    *v = NULL;
}

/* ---- multivec: an array of vector handles, column loops ---------------- */

static void o_get(void **V, GCGE_INT j, void **x, GCGE_OPS *ops)
{
    (void)ops;
    *x = V[j];
}

static void o_restore(void **V, GCGE_INT j, void **x, GCGE_OPS *ops)
{
    (void)ops;
    V[j] = *x;
}

static void o_mv_create_by_mat(void ***mv, GCGE_INT n, void *mat, GCGE_OPS *ops)
{
    *mv = (void **)malloc((size_t)(n > 0 ? n : 1) * sizeof(void *));
    for (GCGE_INT i = 0; i < n; ++i) o_vec_create_by_mat(*mv + i, mat, ops);
}

static void o_mv_create_by_vec(void ***mv, GCGE_INT n, void *vec, GCGE_OPS *ops)
{
    *mv = (void **)malloc((size_t)(n > 0 ? n : 1) * sizeof(void *));
    for (GCGE_INT i = 0; i < n; ++i) o_vec_create_by_vec(*mv + i, vec, ops);
}

static void o_mv_create_by_mv(void ***mv, GCGE_INT n, void **init, GCGE_OPS *ops)
{
    o_mv_create_by_vec(mv, n, init[0], ops);
}

static void o_mv_destroy(void ***mv, GCGE_INT n, GCGE_OPS *ops)
{
    if (!*mv) return;
    for (GCGE_INT i = 0; i < n; ++i) o_vec_destroy(&(*mv)[i], ops);
    free(*mv);
    *mv = NULL;
}

static void o_mv_random(void **mv, GCGE_INT start, GCGE_INT n, GCGE_OPS *ops)
{
    (void)ops;
    uint64_t draw = g_draw++;
    for (GCGE_INT c = 0; c < n; ++c) {
        ORACLE_VEC *v = (ORACLE_VEC *)mv[start + c];
        for (int i = 0; i < v->size; ++i)
            v->data[i] = ORACLE_RandomValue(g_seed, draw, (uint64_t)(start + c), (uint64_t)i);
    }
}

static void o_mat_dot_mv(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                         GCGE_OPS *ops)
{
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i)
        o_mat_dot_vec(mat, x[start[0] + i], y[start[1] + i], ops);
}

static void o_mat_t_dot_mv(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                           GCGE_OPS *ops)
{
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i)
        o_mat_t_dot_vec(mat, x[start[0] + i], y[start[1] + i], ops);
}

static void o_mv_axpby(GCGE_DOUBLE a, void **x, GCGE_DOUBLE b, void **y, GCGE_INT *start,
                       GCGE_INT *end, GCGE_OPS *ops)
{
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i)
        o_vec_axpby(a, x[start[0] + i], b, y[start[1] + i], ops);
}

static void o_mv_axpby_column(GCGE_DOUBLE a, void **x, GCGE_INT cx, GCGE_DOUBLE b, void **y,
                              GCGE_INT cy, GCGE_OPS *ops)
{
    o_vec_axpby(a, x[cx], b, y[cy], ops);
}

static void o_mv_lincomb(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a,
                         GCGE_INT lda, GCGE_INT if_Vec, GCGE_DOUBLE alpha, GCGE_DOUBLE beta,
                         GCGE_OPS *ops)
{
    (void)if_Vec;
    (void)alpha;
    for (GCGE_INT iy = start[1]; iy < end[1]; ++iy) {
        const double *col = a + (size_t)(iy - start[1]) * lda;
        o_vec_axpby(col[0], x[start[0]], beta, y[iy], ops);
        for (GCGE_INT ix = start[0] + 1; ix < end[0]; ++ix)
            o_vec_axpby(col[ix - start[0]], x[ix], 1.0, y[iy], ops);
    }
}

static void o_mv_inner(void **V, void **W, GCGE_DOUBLE *a, char *is_sym, GCGE_INT *start,
                       GCGE_INT *end, GCGE_INT lda, GCGE_INT if_Vec, GCGE_OPS *ops)
{
    (void)if_Vec;
    if (strcmp(is_sym, "sym") == 0) {
        GCGE_INT len = end[1] - start[1];
        for (GCGE_INT iw = 0; iw < len; ++iw)
            for (GCGE_INT iv = 0; iv <= iw; ++iv)
                o_vec_inner(V[start[0] + iv], W[start[1] + iw], a + (size_t)iw * lda + iv, ops);
        for (GCGE_INT iw = 0; iw < len; ++iw)
            for (GCGE_INT iv = iw + 1; iv < len; ++iv)
                a[(size_t)iw * lda + iv] = a[(size_t)iv * lda + iw];
    } else {
        for (GCGE_INT iw = start[1]; iw < end[1]; ++iw)
            for (GCGE_INT iv = start[0]; iv < end[0]; ++iv)
                o_vec_inner(V[iv], W[iw], a + (size_t)(iw - start[1]) * lda + (iv - start[0]), ops);
    }
}

static void o_mv_swap(void **V1, void **V2, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *ops)
{
    (void)ops;
    for (GCGE_INT i = 0; i < end[0] - start[0]; ++i) {
        void *t = V1[start[0] + i];
        V1[start[0] + i] = V2[start[1] + i];
        V2[start[1] + i] = t;
    }
}

static void o_mv_print(void **x, GCGE_INT n, GCGE_OPS *ops)
{
    (void)ops;
    for (GCGE_INT c = 0; c < n; ++c) {
        ORACLE_VEC *v = (ORACLE_VEC *)x[c];
        for (int i = 0; i < v->size; ++i) printf("%0.4f\t", v->data[i]);
        printf("\n");
    }
}

/* Zero the rows of a Dirichlet node in the initial vectors: rows where B has
 * no entry above 1e-20 and A has exactly one (reference
 * app/csr/gcge_app_csr.c:130-183). */
static void o_dirichlet(void **vecs, GCGE_INT nev, void *A, void *B)
{
    const ORACLE_CSR *a = (ORACLE_CSR *)A, *b = (ORACLE_CSR *)B;
    if (!a || !b) return;
    const double tol = 1e-20;
    for (int i = 0; i < a->nrows; ++i) {
        int b_empty = 1;
        for (int j = b->rowptr[i]; j < b->rowptr[i + 1]; ++j)
            if (fabs(b->vals[j]) > tol) {
                b_empty = 0;
                break;
            }
        if (!b_empty) continue;
        int count = 0;
        for (int j = a->rowptr[i]; j < a->rowptr[i + 1]; ++j)
            if (fabs(a->vals[j]) > tol) count++;
        if (count == 1)
            for (GCGE_INT k = 0; k < nev; ++k) ((ORACLE_VEC *)vecs[k])->data[i] = 0.0;
    }
}

/* ---- dense helpers ------------------------------------------------------ */

/* All eigenpairs of the symmetric matrix in the upper triangle of a (lda),
 * ascending, by cyclic Jacobi rotations; vectors in evec (lde). */
void ORACLE_SymEig(int n, const double *a, int lda, double *eval, double *evec, int lde)
{
    if (n <= 0) return;
    double *A = (double *)malloc((size_t)n * n * sizeof(double));
    double *Q = (double *)calloc((size_t)n * n, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) A[i + (size_t)j * n] = A[j + (size_t)i * n] = a[i + (size_t)j * lda];
    for (int i = 0; i < n; ++i) Q[i + (size_t)i * n] = 1.0;
    double fro = 0.0;
    for (size_t i = 0; i < (size_t)n * n; ++i) fro += A[i] * A[i];
    fro = sqrt(fro);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off += A[i + (size_t)j * n] * A[i + (size_t)j * n];
        if (sqrt(2.0 * off) <= 1e-17 * fro || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p + (size_t)q * n];
                if (apq == 0.0) continue;
                double app = A[p + (size_t)p * n], aqq = A[q + (size_t)q * n];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
                    A[k + (size_t)p * n] = c * akp - s * akq;
                    A[k + (size_t)q * n] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
                    A[p + (size_t)k * n] = c * apk - s * aqk;
                    A[q + (size_t)k * n] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double qkp = Q[k + (size_t)p * n], qkq = Q[k + (size_t)q * n];
                    Q[k + (size_t)p * n] = c * qkp - s * qkq;
                    Q[k + (size_t)q * n] = s * qkp + c * qkq;
                }
            }
        }
    }
    int *order = (int *)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) order[i] = i;
    for (int i = 1; i < n; ++i) { /* insertion sort by eigenvalue */
        int k = order[i], j = i - 1;
        while (j >= 0 && A[order[j] + (size_t)order[j] * n] > A[k + (size_t)k * n]) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = k;
    }
    for (int c = 0; c < n; ++c) {
        int src = order[c];
        eval[c] = A[src + (size_t)src * n];
        for (int r = 0; r < n; ++r) evec[r + (size_t)c * lde] = Q[r + (size_t)src * n];
    }
    free(order);
    free(A);
    free(Q);
}

static void o_dense_eig(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT n, GCGE_DOUBLE *eval,
                        GCGE_DOUBLE *evec, GCGE_INT lde, GCGE_INT il, GCGE_INT iu,
                        GCGE_INT *iwork, GCGE_DOUBLE *dwork)
{
    (void)iwork;
    (void)dwork;
    if (n <= 0) return;
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * n * sizeof(double));
    ORACLE_SymEig(n, a, lda, w, z, n);
    if (il == 1 && iu == n) {
        /* dsyev path: vectors into a (lda), then n*n contiguous doubles to evec */
        for (int j = 0; j < n; ++j) {
            eval[j] = w[j];
            memcpy(a + (size_t)j * lda, z + (size_t)j * n, (size_t)n * sizeof(double));
        }
        memcpy(evec, a, (size_t)n * n * sizeof(double));
    } else {
        for (int j = il - 1; j < iu; ++j) {
            eval[j - il + 1] = w[j];
            memcpy(evec + (size_t)(j - il + 1) * lde, z + (size_t)j * n, (size_t)n * sizeof(double));
        }
    }
    free(w);
    free(z);
}

static void o_gemm(char *ta, char *tb, GCGE_INT *M, GCGE_INT *N, GCGE_INT *K, GCGE_DOUBLE *alpha,
                   GCGE_DOUBLE *a, GCGE_INT *lda, GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta,
                   GCGE_DOUBLE *c, GCGE_INT *ldc)
{
    const int at = (*ta == 'T' || *ta == 't'), bt = (*tb == 'T' || *tb == 't');
    for (int j = 0; j < *N; ++j)
        for (int i = 0; i < *M; ++i) {
            double s = 0.0;
            for (int l = 0; l < *K; ++l) {
                double av = at ? a[l + (size_t)i * *lda] : a[i + (size_t)l * *lda];
                double bv = bt ? b[j + (size_t)l * *ldb] : b[l + (size_t)j * *ldb];
                s += av * bv;
            }
            double *cij = c + i + (size_t)j * *ldc;
            *cij = (*beta == 0.0 ? 0.0 : *beta * *cij) + *alpha * s;
        }
}

static void o_symm(char *side, char *uplo, GCGE_INT *M, GCGE_INT *N, GCGE_DOUBLE *alpha,
                   GCGE_DOUBLE *a, GCGE_INT *lda, GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta,
                   GCGE_DOUBLE *c, GCGE_INT *ldc)
{
    /* side 'L', uplo 'U' is the only form the solver uses */
    (void)side;
    (void)uplo;
    for (int j = 0; j < *N; ++j)
        for (int i = 0; i < *M; ++i) {
            double s = 0.0;
            for (int k = 0; k < *M; ++k) {
                double v = (i <= k) ? a[i + (size_t)k * *lda] : a[k + (size_t)i * *lda];
                s += v * b[k + (size_t)j * *ldb];
            }
            double *cij = c + i + (size_t)j * *ldc;
            *cij = (*beta == 0.0 ? 0.0 : *beta * *cij) + *alpha * s;
        }
}

static GCGE_DOUBLE o_array_dot(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT n) { return dot_array(x, y, n); }
static GCGE_DOUBLE o_array_norm(GCGE_DOUBLE *x, GCGE_INT n) { return sqrt(dot_array(x, x, n)); }
static void o_array_axpby(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_DOUBLE b, GCGE_DOUBLE *y, GCGE_INT n)
{
    if (b != 1.0)
        for (int i = 0; i < n; ++i) y[i] *= b;
    for (int i = 0; i < n; ++i) y[i] += a * x[i];
}
static void o_array_copy(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT n) { memmove(y, x, (size_t)n * sizeof(double)); }
static void o_array_scale(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_INT n)
{
    for (int i = 0; i < n; ++i) x[i] *= a;
}

void ORACLE_SetOps(GCGE_OPS *ops)
{
    ops->VecSetRandomValue = o_vec_random;
    ops->MatDotVec = o_mat_dot_vec;
    ops->MatTransposeDotVec = o_mat_t_dot_vec;
    ops->VecAxpby = o_vec_axpby;
    ops->VecInnerProd = o_vec_inner;
    ops->VecLocalInnerProd = o_vec_inner;
    ops->VecCreateByVec = o_vec_create_by_vec;
    ops->VecCreateByMat = o_vec_create_by_mat;
    ops->VecDestroy = o_vec_destroy;
    ops->GetVecFromMultiVec = o_get;
    ops->RestoreVecForMultiVec = o_restore;
    ops->MultiVecCreateByVec = o_mv_create_by_vec;
    ops->MultiVecCreateByMat = o_mv_create_by_mat;
    ops->MultiVecCreateByMultiVec = o_mv_create_by_mv;
    ops->MultiVecDestroy = o_mv_destroy;
    ops->MultiVecSetRandomValue = o_mv_random;
    ops->MatDotMultiVec = o_mat_dot_mv;
    ops->MatTransposeDotMultiVec = o_mat_t_dot_mv;
    ops->MultiVecAxpby = o_mv_axpby;
    ops->MultiVecAxpbyColumn = o_mv_axpby_column;
    ops->MultiVecLinearComb = o_mv_lincomb;
    ops->MultiVecInnerProd = o_mv_inner;
    ops->MultiVecInnerProdLocal = o_mv_inner;
    ops->MultiVecSwap = o_mv_swap;
    ops->MultiVecPrint = o_mv_print;
    ops->SetDirichletBoundary = o_dirichlet;
    ops->DenseMatEigenSolver = o_dense_eig;
    ops->DenseMatDotDenseMat = o_gemm;
    ops->DenseSymMatDotDenseMat = o_symm;
    ops->ArrayDotArray = o_array_dot;
    ops->ArrayNorm = o_array_norm;
    ops->ArrayAXPBY = o_array_axpby;
    ops->ArrayCopy = o_array_copy;
    ops->ArrayScale = o_array_scale;
}

void ORACLE_MultiVecSetHost(void **mv, GCGE_INT start, GCGE_INT ncols, const GCGE_DOUBLE *host,
                            GCGE_INT ld)
{
    for (GCGE_INT c = 0; c < ncols; ++c) {
        ORACLE_VEC *v = (ORACLE_VEC *)mv[start + c];
        memcpy(v->data, host + (size_t)c * ld, (size_t)v->size * sizeof(double));
    }
}

void ORACLE_MultiVecGetHost(void **mv, GCGE_INT start, GCGE_INT ncols, GCGE_DOUBLE *host,
                            GCGE_INT ld)
{
    for (GCGE_INT c = 0; c < ncols; ++c) {
        ORACLE_VEC *v = (ORACLE_VEC *)mv[start + c];
        memcpy(host + (size_t)c * ld, v->data, (size_t)v->size * sizeof(double));
    }
}

/* A backend descriptor with the same layout as the driver's GCGE_BACKEND:
 * {SetOps, MultiVecSetHost, MultiVecGetHost}. */
typedef struct {
    void (*SetOps)(GCGE_OPS *);
    void (*SetHost)(void **, GCGE_INT, GCGE_INT, const GCGE_DOUBLE *, GCGE_INT);
    void (*GetHost)(void **, GCGE_INT, GCGE_INT, GCGE_DOUBLE *, GCGE_INT);
} oracle_backend;

static const oracle_backend g_backend = {ORACLE_SetOps, ORACLE_MultiVecSetHost, ORACLE_MultiVecGetHost};

const void *ORACLE_Backend(void) { return &g_backend; }

/* ---- array-level kernels (column-major, explicit leading dimensions) ---- */

void ORACLE_SpMM(const ORACLE_CSR *m, int k, const double *X, int ldx, double *Y, int ldy)
{
    for (int c = 0; c < k; ++c) csr_matvec(m, X + (size_t)c * ldx, Y + (size_t)c * ldy);
}

/* G(i, j) = V(:, i)^T W(:, j), serial sums */
void ORACLE_Gram(int n, int mv, int kw, const double *V, int ldv, const double *W, int ldw,
                 double *G, int ldg)
{
    for (int j = 0; j < kw; ++j)
        for (int i = 0; i < mv; ++i)
            G[i + (size_t)j * ldg] = dot_array(V + (size_t)i * ldv, W + (size_t)j * ldw, n);
}

/* Y(:, j) = beta Y(:, j) + sum_i V(:, i) C(i, j), accumulated column by
 * column as the default linear combination does */
void ORACLE_LinearComb(int n, int m, int k, const double *V, int ldv, const double *C, int ldc,
                       double beta, double *Y, int ldy)
{
    for (int j = 0; j < k; ++j) {
        double *y = Y + (size_t)j * ldy;
        axpby_array(C[(size_t)j * ldc], V, beta, y, n);
        for (int i = 1; i < m; ++i) axpby_array(C[i + (size_t)j * ldc], V + (size_t)i * ldv, 1.0, y, n);
    }
}

void ORACLE_RandomFill(uint64_t seed, uint64_t draw, int col0, int ncols, int n, double *X, int ldx)
{
    for (int c = 0; c < ncols; ++c)
        for (int i = 0; i < n; ++i)
            X[i + (size_t)c * ldx] = ORACLE_RandomValue(seed, draw, (uint64_t)(col0 + c), (uint64_t)i);
}

/* Block CG of the reference (GCGE_BCG, src/gcge_cg.c:302-501) on arrays:
 * k independent CG iterations on (A + shift B) x = b (B NULL: identity when
 * if_shift is 1), started from x, stopped per column when the residual norm
 * falls below rate times the initial one or below 1e-15, at most max_it - 1
 * iterations.  rhs is overwritten by the final residuals, as the reference
 * does when it runs without a separate residual block.  Bookkeeping follows
 * the reference: the scalars rho1/rho2 are stored by position in the list of
 * unconverged columns (src/gcge_cg.c:467-476), while beta reads them by
 * column index (src/gcge_cg.c:405); both agree until a column drops out. */
void ORACLE_BlockCG(const ORACLE_CSR *A, int if_shift, double shift, const ORACLE_CSR *B, int k,
                    double *rhs, int ldb, double *x, int ldx, int max_it, double rate)
{
    const int n = A->nrows;
    double *P = (double *)calloc((size_t)n * (k > 0 ? k : 1), sizeof(double));
    double *W = (double *)calloc((size_t)n * (k > 0 ? k : 1), sizeof(double));
    double *Bx = (double *)calloc((size_t)n, sizeof(double));
    double *rho1 = (double *)calloc((size_t)5 * (k > 0 ? k : 1), sizeof(double));
    double *rho2 = rho1 + k, *err = rho2 + k, *last = err + k, *ptw = last + k;
    int *unlock = (int *)calloc((size_t)(k > 0 ? k : 1), sizeof(int));
    int nu = 0;
    for (int c = 0; c < k; ++c) {
        double *xc = x + (size_t)c * ldx, *b = rhs + (size_t)c * ldb, *r = W + (size_t)c * n;
        csr_matvec(A, xc, r);
        if (if_shift == 1) {
            if (!B) {
                axpby_array(shift, xc, 1.0, r, n);
            } else {
                csr_matvec(B, xc, Bx);
                axpby_array(shift, Bx, 1.0, r, n);
            }
        }
        axpby_array(-1.0, r, 1.0, b, n);
        rho2[c] = dot_array(b, b, n);
    }
    for (int c = 0; c < k; ++c) {
        err[c] = sqrt(rho2[c]);
        last[c] = err[c];
        if (err[c] >= 1e-15) unlock[nu++] = c;
    }
    for (int it = 1; nu > 0 && it < max_it; ++it) {
        for (int id = 0; id < nu; ++id) {
            const int c = unlock[id];
            double *p = P + (size_t)c * n, *w = W + (size_t)c * n, *r = rhs + (size_t)c * ldb;
            if (it == 1)
                axpby_array(1.0, r, 0.0, p, n);
            else
                axpby_array(1.0, r, rho2[c] / rho1[c], p, n);
            csr_matvec(A, p, w);
            if (if_shift == 1) {
                if (!B) {
                    axpby_array(shift, p, 1.0, w, n);
                } else {
                    csr_matvec(B, p, Bx);
                    axpby_array(shift, Bx, 1.0, w, n);
                }
            }
            ptw[id] = dot_array(p, w, n);
        }
        for (int id = 0; id < nu; ++id) {
            const int c = unlock[id];
            const double alpha = rho2[id] / ptw[id];
            double *r = rhs + (size_t)c * ldb;
            axpby_array(alpha, P + (size_t)c * n, 1.0, x + (size_t)c * ldx, n);
            axpby_array(-alpha, W + (size_t)c * n, 1.0, r, n);
            rho1[id] = rho2[id];
            rho2[id] = dot_array(r, r, n);
        }
        for (int id = 0; id < nu; ++id) last[unlock[id]] = sqrt(fabs(rho2[id]));
        const int old = nu;
        nu = 0;
        for (int id = 0; id < old; ++id) {
            const int c = unlock[id];
            if (last[c] >= rate * err[c] && last[c] >= 1e-15) unlock[nu++] = c;
        }
    }
    free(P);
    free(W);
    free(Bx);
    free(rho1);
    free(unlock);
}
