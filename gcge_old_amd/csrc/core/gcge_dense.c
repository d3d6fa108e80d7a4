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
#define _GNU_SOURCE /* sched_getaffinity */
#include <math.h>
#include <stdint.h>
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
 * multiple eigenvalues.
 *
 * From order GCGE_DENSE_MIN_N (default 80) the work is shared by a small
 * pool of host threads (GCGE_DENSE_THREADS, default min(8, CPUs available);
 * 1 turns it off).  Every element still sees the same operations in the same
 * order, so the eigenpairs are bit-identical for every thread count:
 *   - reduction: the columns of the trailing block are split between the
 *     threads; A stays exactly symmetric (both triangles get the same
 *     rank-2 terms, x_i w_c + w_i x_c), so the matrix-vector product entry
 *     p_i is the dot product of column i with tau v, summed in the same
 *     order as the column sweep; three barriers per reflector;
 *   - QL: the calling thread runs the scalar recurrence (it needs only the
 *     tridiagonal) and publishes each sweep's rotations in a log;
 *   - meanwhile the other threads form Q (columns are independent, each gets
 *     the reflectors in descending order; dealt out cyclically), then apply
 *     the published sweeps in order to their own row ranges of it. */

#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>

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

/* one step of a spin wait; yields the CPU now and then, so that a machine
 * with fewer free cores than threads still makes progress */
static inline void spin_step(unsigned *k)
{
    if ((++*k & 1023u) == 0) {
        sched_yield();
        return;
    }
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#endif
}

/* ---- worker pool -------------------------------------------------------- */
enum { EIG_MAX_THREADS = 64 };

typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    atomic_uint gen;  /* job generation; workers run a job when it changes */
    atomic_int done;  /* workers finished with the current job */
    atomic_int bar_count, bar_sense;   /* all threads */
    atomic_int wbar_count, wbar_sense; /* the workers (threads 1 ..) */
    int nthreads;     /* including the calling thread; 0 = not started */
    unsigned gen0;    /* generation when the workers were started */
    unsigned spins;   /* spin steps of an idle worker before it sleeps */
    void (*fn)(void *, int);
    void *arg;
    /* sweep log of the QL phase, kept between calls */
    double *lcs, *lsn;
    int *llo, *lhi, *loff;
    int lcap, lscap;
} eig_pool;

static eig_pool g_pool = {.mu = PTHREAD_MUTEX_INITIALIZER, .cv = PTHREAD_COND_INITIALIZER};
/* one solve uses the pool at a time; a concurrent caller runs single-threaded */
static pthread_mutex_t g_pool_busy = PTHREAD_MUTEX_INITIALIZER;
static int g_pool_atfork;

static void *pool_worker(void *arg)
{
    const int tid = (int)(intptr_t)arg;
    unsigned seen = g_pool.gen0;
    for (;;) {
        unsigned g, spins = 0;
        while ((g = atomic_load_explicit(&g_pool.gen, memory_order_acquire)) == seen) {
            if (spins < g_pool.spins) {
                spin_step(&spins);
                continue;
            }
            pthread_mutex_lock(&g_pool.mu);
            while ((g = atomic_load_explicit(&g_pool.gen, memory_order_acquire)) == seen)
                pthread_cond_wait(&g_pool.cv, &g_pool.mu);
            pthread_mutex_unlock(&g_pool.mu);
            break;
        }
        seen = g;
        g_pool.fn(g_pool.arg, tid);
        atomic_fetch_add_explicit(&g_pool.done, 1, memory_order_release);
    }
    return NULL;
}

/* a forked child has none of the parent's threads */
static void pool_child_reset(void)
{
    pthread_mutex_init(&g_pool.mu, NULL);
    pthread_cond_init(&g_pool.cv, NULL);
    pthread_mutex_init(&g_pool_busy, NULL);
    g_pool.nthreads = 0;
}

static int env_int(const char *name, int dflt)
{
    const char *s = getenv(name);
    if (!s || !*s) return dflt;
    return atoi(s);
}

/* threads to use (>= 1); starts the workers the first time */
static int pool_threads(void)
{
    if (g_pool.nthreads > 0) return g_pool.nthreads;
    int want = env_int("GCGE_DENSE_THREADS", 0);
    if (want <= 0) {
        cpu_set_t set;
        int ncpu = 1;
        if (sched_getaffinity(0, sizeof set, &set) == 0) ncpu = CPU_COUNT(&set);
        want = ncpu < 8 ? ncpu : 8;
    }
    if (want > EIG_MAX_THREADS) want = EIG_MAX_THREADS;
    if (want < 1) want = 1;
    if (!g_pool_atfork) {
        pthread_atfork(NULL, NULL, pool_child_reset);
        g_pool_atfork = 1;
    }
    g_pool.gen0 = atomic_load(&g_pool.gen);
    g_pool.spins = (unsigned)env_int("GCGE_DENSE_SPIN", 20000);
    int started = 1;
    for (int t = 1; t < want; ++t) {
        pthread_t th;
        pthread_attr_t at;
        pthread_attr_init(&at);
        pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
        if (pthread_create(&th, &at, pool_worker, (void *)(intptr_t)t) == 0) ++started;
        pthread_attr_destroy(&at);
        if (started != t + 1) break;
    }
    g_pool.nthreads = started;
    return started;
}

static void pool_run(void (*fn)(void *, int), void *arg)
{
    g_pool.fn = fn;
    g_pool.arg = arg;
    atomic_store_explicit(&g_pool.done, 0, memory_order_relaxed);
    pthread_mutex_lock(&g_pool.mu);
    atomic_fetch_add_explicit(&g_pool.gen, 1, memory_order_release);
    pthread_cond_broadcast(&g_pool.cv);
    pthread_mutex_unlock(&g_pool.mu);
    fn(arg, 0);
    unsigned k = 0;
    while (atomic_load_explicit(&g_pool.done, memory_order_acquire) < g_pool.nthreads - 1)
        spin_step(&k);
}

/* sense-reversing barrier of T threads on (count, flag) */
static void barrier_on(atomic_int *count, atomic_int *flag, int T, int *sense)
{
    if (T <= 1) return;
    const int s = !*sense;
    *sense = s;
    if (atomic_fetch_add_explicit(count, 1, memory_order_acq_rel) == T - 1) {
        atomic_store_explicit(count, 0, memory_order_relaxed);
        atomic_store_explicit(flag, s, memory_order_release);
    } else {
        unsigned k = 0;
        while (atomic_load_explicit(flag, memory_order_acquire) != s) spin_step(&k);
    }
}

static void pool_barrier(int T, int *sense)
{
    barrier_on(&g_pool.bar_count, &g_pool.bar_sense, T, sense);
}

/* ---- dense products ------------------------------------------------------ */
/* c = alpha op(a) op(b) + beta c, all column-major (Fortran dgemm contract:
 * with beta == 0, c is not read).  Columns j0 .. j1-1 of c. */
GCGE_INLINE void gemm_cols(int ta, int tb, int M, int K, double al, const double *a, int LA,
                           const double *b, int LB, double be, double *c, int LC, int j0, int j1)
{
    for (int j = j0; j < j1; ++j) {
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

/* columns j0 .. j1-1 of the symmetric product: `full` is the expanded S x S
 * matrix, s one column of M sums */
GCGE_INLINE void symm_cols(int left, int M, int S, double al, const double *full, const double *b,
                           int LB, double be, double *c, int LC, double *s, int j0, int j1)
{
    for (int j = j0; j < j1; ++j) {
        for (int i = 0; i < M; ++i) s[i] = 0.0;
        if (left) {
            for (int k = 0; k < M; ++k) {
                const double bkj = b[k + (size_t)j * LB];
                const double *ak = full + (size_t)k * S;
                for (int i = 0; i < M; ++i) s[i] += ak[i] * bkj;
            }
        } else {
            for (int k = 0; k < S; ++k) {
                const double akj = full[k + (size_t)j * S];
                const double *bk = b + (size_t)k * LB;
                for (int i = 0; i < M; ++i) s[i] += bk[i] * akj;
            }
        }
        double *cj = c + (size_t)j * LC;
        for (int i = 0; i < M; ++i) cj[i] = (be == 0.0 ? 0.0 : be * cj[i]) + al * s[i];
    }
}

typedef struct {
    int symm, ta, tb, left, M, N, K, S, LA, LB, LC, T;
    double al, be;
    const double *a, *b;
    double *c, *s;
} prod_job;

GCGE_VECTOR_CLONES
static void prod_job_run(void *arg, int t)
{
    const prod_job *P = (const prod_job *)arg;
    const int j0 = (int)((long)P->N * t / P->T), j1 = (int)((long)P->N * (t + 1) / P->T);
    if (P->symm)
        symm_cols(P->left, P->M, P->S, P->al, P->a, P->b, P->LB, P->be, P->c, P->LC,
                  P->s + (size_t)t * P->M, j0, j1);
    else
        gemm_cols(P->ta, P->tb, P->M, P->K, P->al, P->a, P->LA, P->b, P->LB, P->be, P->c, P->LC, j0,
                  j1);
}

/* the product on the pool when it is large enough and the pool is free
 * (then the pool is held until prod_run): the threads' count, else 1.  The
 * columns of c are dealt out in contiguous ranges, each computed exactly as
 * the single-threaded loop computes it. */
static int prod_threads(long flops, int N)
{
    if (flops < env_int("GCGE_DENSE_MIN_FLOPS", 2000000) || N < 2) return 1;
    if (pthread_mutex_trylock(&g_pool_busy) != 0) return 1;
    const int T = pool_threads();
    if (T <= 1) pthread_mutex_unlock(&g_pool_busy);
    return T;
}

static void prod_run(prod_job *P)
{
    pool_run(prod_job_run, P);
    pthread_mutex_unlock(&g_pool_busy);
}

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
    if (M <= 0 || N <= 0) return;
    const int T = prod_threads(2L * M * N * (K > 0 ? K : 0), N);
    if (T > 1) {
        prod_job P = {0, ta, tb, 0, M, N, K, 0, *lda, *ldb, *ldc, T, *alpha, *beta, a, b, c, NULL};
        prod_run(&P);
        return;
    }
    gemm_cols(ta, tb, M, K, *alpha, a, *lda, b, *ldb, *beta, c, *ldc, 0, N);
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
    const int M = *nrows, N = *ncols, LA = *lda;
    const int left = lsame(side, 'L'), upper = lsame(uplo, 'U');
    if (M <= 0 || N <= 0) return;
    const int S = left ? M : N; /* order of a */
    const int T = prod_threads(2L * M * N * S, N);
    double *full = (double *)malloc(((size_t)S * S + (size_t)M * T) * sizeof(double));
    if (!full) {
        fprintf(stderr, "GCGE_Default_DenseSymMatDotDenseMat: out of memory\n");
        abort();
    }
    double *s = full + (size_t)S * S; /* one column of sums per thread */
    for (int k = 0; k < S; ++k)
        for (int i = 0; i < S; ++i)
            full[i + (size_t)k * S] = (upper ? (i <= k) : (i >= k)) ? a[i + (size_t)k * LA]
                                                                  : a[k + (size_t)i * LA];
    if (T > 1) {
        prod_job P = {1, 0, 0, left, M, N, 0, S, LA, *ldb, *ldc, T, *alpha, *beta, full, b, c, s};
        prod_run(&P);
    } else {
        symm_cols(left, M, S, *alpha, full, b, *ldb, *beta, c, *ldc, s, 0, N);
    }
    free(full);
}

/* ---- the solve ---------------------------------------------------------- */
typedef struct {
    int n, T;
    double *A, *d, *e, *tau, *p, *sn; /* caller's work array */
    double *vv, *wb;                  /* tau v; one w per thread */
    double *Z;
    int ldz;
    double *panel; /* per worker: its rows of Z, columns 64-byte aligned */
    int ldp;
    int ql_status;
    /* one cache line each: the workers poll `published` while the others
     * store their progress */
    _Alignas(64) atomic_int published;
    atomic_int finished;
    struct {
        _Alignas(64) atomic_int v;
    } consumed[EIG_MAX_THREADS];
} eig_job;

#define AC(i, j) A[(size_t)(j) * n + (i)]

/* Householder reflector of column j: d[j], e[j], tau[j]; the column below
 * the diagonal becomes v (unit leading entry stored), vv = tau v.  Returns 0
 * when the column is already reduced (tau = 0). */
GCGE_INLINE int reflector(int n, double *A, int j, double *d, double *e, double *tau, double *vv)
{
    const int m = n - j - 1;
    double *x = &AC(j + 1, j);
    const double alpha = x[0];
    const double xnorm = m > 1 ? norm2(x + 1, m - 1) : 0.0;
    d[j] = AC(j, j);
    if (xnorm == 0.0) {
        tau[j] = 0.0;
        e[j] = alpha;
        return 0;
    }
    const double beta = -copysign(hypot_safe(alpha, xnorm), alpha);
    tau[j] = (beta - alpha) / beta;
    const double scal = 1.0 / (alpha - beta);
    for (int i = 1; i < m; ++i) x[i] *= scal;
    x[0] = 1.0;
    e[j] = beta;
    for (int c = 0; c < m; ++c) vv[c] = tau[j] * x[c];
    return 1;
}

/* p_i = (A22 tau v)_i for i in [c0, c1): column i of A22 dotted with vv */
GCGE_INLINE void reduce_matvec(int n, const double *A, int j, const double *vv, double *p, int c0,
                               int c1)
{
    const int m = n - j - 1;
    for (int i = c0; i < c1; ++i) {
        const double *col = &AC(j + 1, j + 1 + i);
        double s = 0.0;
        for (int c = 0; c < m; ++c) s += col[c] * vv[c];
        p[i] = s;
    }
}

/* w = p - (tau/2)(p.v) v (the whole vector, into this thread's w), then
 * A22(:, c) -= v w_c + w v_c for c in [c0, c1) */
GCGE_INLINE void reduce_update(int n, double *A, int j, double tau_j, const double *p, double *w,
                               int c0, int c1)
{
    const int m = n - j - 1;
    const double *x = &AC(j + 1, j);
    double pv = 0.0;
    for (int i = 0; i < m; ++i) pv += p[i] * x[i];
    const double h = 0.5 * tau_j * pv;
    for (int i = 0; i < m; ++i) w[i] = p[i] - h * x[i];
    for (int c = c0; c < c1; ++c) {
        double *col = &AC(j + 1, j + 1 + c);
        const double wc = w[c], vc = x[c];
        for (int i = 0; i < m; ++i) col[i] -= x[i] * wc + w[i] * vc;
    }
}

/* Q = H_0 H_1 ... H_{n-3} applied to column c of the identity: the
 * reflectors j = min(c-1, n-3) .. 0, each touching rows j+1 .. n-1 */
GCGE_INLINE void form_q_column(int n, const double *A, const double *tau, double *Z, int ldz, int c)
{
    double *zc = Z + (size_t)c * ldz;
    for (int i = 0; i < n; ++i) zc[i] = (i == c) ? 1.0 : 0.0;
    for (int j = (c - 1 < n - 3 ? c - 1 : n - 3); j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        const int m = n - j - 1;
        const double *v = A + (size_t)j * n + j + 1;
        double *z = zc + j + 1;
        double s = z[0];
        for (int i = 1; i < m; ++i) s += v[i] * z[i];
        s *= tau[j];
        z[0] -= s;
        for (int i = 1; i < m; ++i) z[i] -= s * v[i];
    }
}
#undef AC

/* Apply the plane rotations i = hi-1 .. lo of one QL sweep (c_i = cs[i-lo],
 * s_i = sn[i-lo]) to rows 0 .. nrows-1 of columns lo .. hi of Z, as the
 * sweep would one at a time:
 *     (z_i, z_{i+1}) <- (c_i z_i - s_i z_{i+1}, s_i z_i + c_i z_{i+1}).
 * Rows are taken in chunks whose running column stays in registers, so every
 * element is loaded and stored once per sweep. */
GCGE_INLINE void apply_sweep(double *Z, int ldz, int nrows, int lo, int hi, const double *cs,
                             const double *sn)
{
    enum { CH = 16 };
    for (int k0 = 0; k0 < nrows; k0 += CH) {
        const int kn = (nrows - k0 < CH) ? nrows - k0 : CH;
        double t[CH];
        const double *zh = Z + (size_t)hi * ldz + k0;
        for (int k = 0; k < kn; ++k) t[k] = zh[k];
        for (int i = hi - 1; i >= lo; --i) {
            const double c = cs[i - lo], s = sn[i - lo];
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

static int min_consumed(eig_job *J)
{
    int m = atomic_load_explicit(&J->consumed[1].v, memory_order_acquire);
    for (int t = 2; t < J->T; ++t) {
        const int c = atomic_load_explicit(&J->consumed[t].v, memory_order_acquire);
        if (c < m) m = c;
    }
    return m;
}

/* hand sweep k (rotations lo .. hi-1) to the workers; the log is reused from
 * the start once every worker has applied everything published */
static void publish_sweep(eig_job *J, int *pos, int k, int lo, int hi, const double *cs,
                          const double *sn)
{
    const int L = hi - lo;
    if (*pos + L > g_pool.lcap || k - min_consumed(J) >= g_pool.lscap) {
        unsigned spins = 0;
        while (min_consumed(J) < k) spin_step(&spins);
        *pos = 0;
    }
    const int slot = k % g_pool.lscap;
    memcpy(g_pool.lcs + *pos, cs + lo, (size_t)L * sizeof(double));
    memcpy(g_pool.lsn + *pos, sn + lo, (size_t)L * sizeof(double));
    g_pool.llo[slot] = lo;
    g_pool.lhi[slot] = hi;
    g_pool.loff[slot] = *pos;
    *pos += L;
    atomic_store_explicit(&J->published, k + 1, memory_order_release);
}

/* Implicit-shift QL on the tridiagonal (d, e); the rotations of a sweep are
 * recorded in cs/sn (n entries each) and applied to the columns of Z
 * together, here (single thread) or by the workers.  Returns 0 on success,
 * 1 when an eigenvalue needs more than 60 sweeps. */
GCGE_INLINE int tridiag_ql(eig_job *J, int n, double *d, double *e, double *cs, double *sn)
{
    int k = 0, pos = 0, status = 0;
    for (int l = 0; l < n && !status; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1.1102230246251565e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) {
                    status = 1;
                    break;
                }
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
                if (i + 1 < m) {
                    if (J->T > 1)
                        publish_sweep(J, &pos, k++, i + 1, m, cs, sn);
                    else
                        apply_sweep(J->Z, J->ldz, n, i + 1, m, cs + i + 1, sn + i + 1);
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    if (J->T > 1) atomic_store_explicit(&J->finished, 1, memory_order_release);
    return status;
}

/* rows of Z per QL worker */
GCGE_INLINE int ql_rows(int n, int T) { return (n + T - 2) / (T - 1); }

/* worker t >= 1 of the QL phase: rows [r0, r1) of Z, copied to the worker's
 * own panel (no cache line shared with another worker), every published
 * sweep applied in order, then copied back */
GCGE_INLINE void ql_worker(eig_job *J, int t)
{
    const int n = J->n, chunk = ql_rows(n, J->T);
    const int r0 = (t - 1) * chunk < n ? (t - 1) * chunk : n;
    const int r1 = r0 + chunk < n ? r0 + chunk : n, nr = r1 - r0, ldp = J->ldp;
    double *Pn = J->panel + (size_t)(t - 1) * ldp * n;
    for (int c = 0; c < n && nr > 0; ++c)
        memcpy(Pn + (size_t)c * ldp, J->Z + (size_t)c * J->ldz + r0, (size_t)nr * sizeof(double));
    int k = 0;
    unsigned spins = 0;
    for (;;) {
        int pub = atomic_load_explicit(&J->published, memory_order_acquire);
        if (pub == k) {
            if (!atomic_load_explicit(&J->finished, memory_order_acquire)) {
                spin_step(&spins);
                continue;
            }
            pub = atomic_load_explicit(&J->published, memory_order_acquire);
            if (pub == k) break;
        }
        for (; k < pub; ++k) {
            const int slot = k % g_pool.lscap;
            const int off = g_pool.loff[slot];
            if (nr > 0)
                apply_sweep(Pn, ldp, nr, g_pool.llo[slot], g_pool.lhi[slot], g_pool.lcs + off,
                            g_pool.lsn + off);
            atomic_store_explicit(&J->consumed[t].v, k + 1, memory_order_release);
        }
    }
    for (int c = 0; c < n && nr > 0; ++c)
        memcpy(J->Z + (size_t)c * J->ldz + r0, Pn + (size_t)c * ldp, (size_t)nr * sizeof(double));
}

/* reflectors with fewer trailing rows than this are done by thread 0 alone */
enum { EIG_PAR_MIN_M = 48 };

GCGE_VECTOR_CLONES
static void eig_job_run(void *arg, int t)
{
    eig_job *J = (eig_job *)arg;
    const int n = J->n, T = J->T;
    double *A = J->A, *d = J->d, *e = J->e, *tau = J->tau;
    int sense = atomic_load_explicit(&g_pool.bar_sense, memory_order_acquire);
    int wsense = atomic_load_explicit(&g_pool.wbar_sense, memory_order_acquire);
    double *w = J->wb + (size_t)t * n;

    /* tridiagonal reduction.  Thread 0 owns the leading columns of every
     * trailing block, so after its share of update j it already holds the
     * final column j+1 and computes reflector j+1 before the barrier. */
    int j = 0;
    /* m >= T keeps column 0 of every trailing block with thread 0 */
    const int minm = EIG_PAR_MIN_M > T ? EIG_PAR_MIN_M : T;
    if (T > 1 && n - 1 >= minm) {
        if (t == 0) reflector(n, A, 0, d, e, tau, J->vv);
        pool_barrier(T, &sense);
        for (; j < n - 2; ++j) {
            const int m = n - j - 1;
            if (tau[j] != 0.0) {
                const int c0 = (int)((long)m * t / T), c1 = (int)((long)m * (t + 1) / T);
                reduce_matvec(n, A, j, J->vv, J->p, c0, c1);
                pool_barrier(T, &sense);
                reduce_update(n, A, j, tau[j], J->p, w, c0, c1);
            }
            const int next = j + 1 < n - 2 && m - 1 >= minm;
            if (t == 0 && next) reflector(n, A, j + 1, d, e, tau, J->vv);
            pool_barrier(T, &sense);
            if (!next) {
                ++j;
                break;
            }
        }
    }
    if (t == 0) {
        for (; j < n - 2; ++j)
            if (reflector(n, A, j, d, e, tau, J->vv)) {
                const int m = n - j - 1;
                reduce_matvec(n, A, j, J->vv, J->p, 0, m);
                reduce_update(n, A, j, tau[j], J->p, w, 0, m);
            }
        if (n >= 2) {
            d[n - 2] = A[(size_t)(n - 2) * n + n - 2];
            e[n - 2] = A[(size_t)(n - 2) * n + n - 1];
            tau[n - 2] = 0.0;
        }
        d[n - 1] = A[(size_t)(n - 1) * n + n - 1];
        e[n - 1] = 0.0;
    }
    pool_barrier(T, &sense);

    if (T == 1) {
        for (int c = 0; c < n; ++c) form_q_column(n, A, tau, J->Z, J->ldz, c);
        J->ql_status = tridiag_ql(J, n, d, e, J->p, J->sn);
        return;
    }
    /* thread 0 runs the QL recurrence (it needs only d and e) while the
     * workers form the orthogonal factor, columns dealt out cyclically; they
     * then apply the published sweeps to their rows */
    if (t == 0) {
        J->ql_status = tridiag_ql(J, n, d, e, J->p, J->sn);
    } else {
        for (int c = t - 1; c < n; c += T - 1) form_q_column(n, A, tau, J->Z, J->ldz, c);
        barrier_on(&g_pool.wbar_count, &g_pool.wbar_sense, T - 1, &wsense);
        ql_worker(J, t);
    }
    pool_barrier(T, &sense);
}

/* grow the pool's sweep log and per-thread buffers for order n */
static int pool_reserve(int n, int T)
{
    const int lcap = 2 * n * n + 4 * n, lscap = 4 * n + 64;
    if (lcap > g_pool.lcap) {
        free(g_pool.lcs);
        free(g_pool.lsn);
        g_pool.lcs = (double *)malloc((size_t)lcap * sizeof(double));
        g_pool.lsn = (double *)malloc((size_t)lcap * sizeof(double));
        g_pool.lcap = (g_pool.lcs && g_pool.lsn) ? lcap : 0;
    }
    if (lscap > g_pool.lscap) {
        free(g_pool.llo);
        free(g_pool.lhi);
        free(g_pool.loff);
        g_pool.llo = (int *)malloc((size_t)lscap * sizeof(int));
        g_pool.lhi = (int *)malloc((size_t)lscap * sizeof(int));
        g_pool.loff = (int *)malloc((size_t)lscap * sizeof(int));
        g_pool.lscap = (g_pool.llo && g_pool.lhi && g_pool.loff) ? lscap : 0;
    }
    (void)T;
    return g_pool.lcap >= lcap && g_pool.lscap >= lscap;
}

/* All eigenpairs of the symmetric n x n matrix whose upper triangle is in
 * a (column-major, lda).  Eigenvalues ascending in eval; eigenvector j in
 * evec[j*lde .. j*lde+n-1] (lde >= n).  work needs n*n + 5n doubles. */
void GCGE_Dense_SymEig(GCGE_DOUBLE *a, GCGE_INT lda, GCGE_INT n, GCGE_DOUBLE *eval,
                       GCGE_DOUBLE *evec, GCGE_INT lde, GCGE_DOUBLE *work)
{
    if (n <= 0) return;
    eig_job J;
    memset(&J, 0, sizeof J);
    J.n = n;
    J.A = work;
    J.d = work + (size_t)n * n;
    J.e = J.d + n;
    J.tau = J.e + n;
    J.p = J.tau + n;
    J.sn = J.p + n;
    J.Z = evec;
    J.ldz = lde;
    double *A = J.A, *d = J.d;
    /* full symmetric copy from the upper triangle */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            const double v = a[i + (size_t)j * lda];
            A[(size_t)j * n + i] = v;
            A[(size_t)i * n + j] = v;
        }
    int T = 1, locked = 0;
    if (n >= env_int("GCGE_DENSE_MIN_N", 80) && pthread_mutex_trylock(&g_pool_busy) == 0) {
        locked = 1;
        T = pool_threads();
        if (T > 1 && !pool_reserve(n, T)) T = 1;
    }
    J.T = T;
    /* tau v and the per-thread w: the caller's p is one of them when T = 1 */
    double *extra = NULL;
    if (T > 1) {
        J.ldp = (ql_rows(n, T) + 7) & ~7;
        const size_t nv = (((size_t)T + 1) * n + 7) & ~(size_t)7;
        const size_t bytes = (nv + (size_t)(T - 1) * J.ldp * n) * sizeof(double);
        extra = (double *)aligned_alloc(64, (bytes + 63) & ~(size_t)63);
        if (!extra) T = J.T = 1;
    }
    if (T > 1) {
        J.vv = extra;
        J.wb = extra + n;
        J.panel = extra + ((((size_t)T + 1) * n + 7) & ~(size_t)7);
        pool_run(eig_job_run, &J);
    } else {
        double *buf = (double *)malloc((size_t)2 * n * sizeof(double));
        if (!buf) {
            fprintf(stderr, "GCGE_Dense_SymEig: out of memory\n");
            abort();
        }
        J.vv = buf;
        J.wb = buf + n;
        eig_job_run(&J, 0);
        free(buf);
    }
    free(extra);
    if (locked) pthread_mutex_unlock(&g_pool_busy);
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
