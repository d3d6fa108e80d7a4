/*
 * ORACLE -- test infrastructure only.
 *
 * The drop-in check: the reference solver core (src/*.c, compiled unchanged
 * from its sources, see oracle/Makefile) driven through the MI355X backend
 * instead of its CSR backend.  The only differences from ref_driver.c are
 * the lines a maintainer would change in a reference program:
 *     CSR_MAT from the file            ->  GCGE_HIP_MatCreateCSR(...)
 *     GCGE_CSR_SetOps(solver->ops)     ->  GCGE_HIP_SetOps(solver->ops)
 * Same arguments and output layout as ref_driver.c.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gcge.h"

/* The backend's C ABI (include/gcge_hip.h), declared against the reference's
 * own GCGE_OPS type, whose layout the backend's table shares. */
extern int GCGE_HIP_Init(int device);
extern void GCGE_HIP_SetOps(GCGE_OPS *ops);
extern void GCGE_HIP_SetSeed(uint64_t seed);
extern void *GCGE_HIP_MatCreateCSR(int nrows, int ncols, int nnz, const int *rowptr,
                                   const int *colidx, const double *vals);
extern void GCGE_HIP_MultiVecSetHost(void **mv, int start, int ncols, const double *host, int ld);
extern void GCGE_HIP_MultiVecGetHost(void **mv, int start, int ncols, double *host, int ld);
extern void GCGE_HIP_SetLinearSolverParams(int max_it, double rate, int if_shift, double shift,
                                           void *B);
extern void GCGE_HIP_Synchronize(void);

static void *read_bin(const char *path, int *n_out)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "gcge_ref_hip: cannot open %s\n", path);
        exit(2);
    }
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, f) != 3) exit(2);
    int *rp = (int *)malloc((size_t)(hdr[0] + 1) * sizeof(int));
    int *ci = (int *)malloc((size_t)hdr[2] * sizeof(int));
    double *v = (double *)malloc((size_t)hdr[2] * sizeof(double));
    if (fread(rp, sizeof(int), (size_t)hdr[0] + 1, f) != (size_t)hdr[0] + 1) exit(2);
    if (fread(ci, sizeof(int), (size_t)hdr[2], f) != (size_t)hdr[2]) exit(2);
    if (fread(v, sizeof(double), (size_t)hdr[2], f) != (size_t)hdr[2]) exit(2);
    fclose(f);
    void *m = GCGE_HIP_MatCreateCSR(hdr[0], hdr[1], hdr[2], rp, ci, v);
    free(rp);
    free(ci);
    free(v);
    *n_out = hdr[0];
    return m;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s A.bin B.bin|- nev X0.bin|- out.bin [-gcge_* ...]\n", argv[0]);
        return 2;
    }
    if (GCGE_HIP_Init(0) != 0) {
        fprintf(stderr, "gcge_ref_hip: no GPU\n");
        return 2;
    }
    GCGE_HIP_SetSeed(1);
    int n = 0, nb = 0;
    void *A = read_bin(argv[1], &n);
    void *B = strcmp(argv[2], "-") ? read_bin(argv[2], &nb) : NULL;
    int nev = atoi(argv[3]);

    GCGE_SOLVER *s;
    GCGE_SOLVER_Create(&s);
    s->para->nev = nev;
    GCGE_PARA_SetFromCommandLine(s->para, argc - 6, argv + 6);
    nev = s->para->nev;
    GCGE_HIP_SetOps(s->ops);
    GCGE_HIP_SetLinearSolverParams(s->para->cg_max_it, s->para->cg_rate, 0, 0.0, B);
    double *eval = (double *)calloc((size_t)nev, sizeof(double));
    void **evec;
    s->ops->MultiVecCreateByMat(&evec, nev, A, s->ops);
    GCGE_SOLVER_SetMatA(s, A);
    if (B) GCGE_SOLVER_SetMatB(s, B);
    GCGE_SOLVER_SetEigenvalues(s, eval);
    GCGE_SOLVER_SetEigenvectors(s, evec);
    GCGE_SOLVER_Setup(s);
    if (strcmp(argv[4], "-")) {
        double *x0 = (double *)malloc((size_t)n * nev * sizeof(double));
        FILE *f = fopen(argv[4], "rb");
        if (!f || fread(x0, sizeof(double), (size_t)n * nev, f) != (size_t)n * nev) return 2;
        fclose(f);
        GCGE_HIP_MultiVecSetHost(evec, 0, nev, x0, n);
        free(x0);
        s->para->num_init_evec = nev;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    GCGE_SOLVER_Solve(s);
    GCGE_HIP_Synchronize();
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

    FILE *o = fopen(argv[5], "wb");
    if (!o) return 2;
    int hdr[4] = {nev, s->para->num_iter, s->para->num_conv_continuous, n};
    fwrite(hdr, sizeof(int), 4, o);
    fwrite(&secs, sizeof(double), 1, o);
    fwrite(eval, sizeof(double), (size_t)nev, o);
    fwrite(s->para->res, sizeof(double), (size_t)nev, o);
    double *X = (double *)malloc((size_t)n * nev * sizeof(double));
    GCGE_HIP_MultiVecGetHost(evec, 0, nev, X, n);
    fwrite(X, sizeof(double), (size_t)n * nev, o);
    fclose(o);
    printf("gcge_ref_hip: done in %.6f s\n", secs);
    return 0;
}
