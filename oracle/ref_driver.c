/*
 * ORACLE -- test infrastructure only.
 *
 * Command-line driver for the reference solver built from its own sources
 * (src/*.c and app/csr/gcge_app_csr.c, see oracle/Makefile).  It is this
 * project's code: it reads the inputs, hands them to the reference CSR
 * path exactly as example/csr_solver.c does (GCGE_CSR_Solver_Init,
 * GCGE_SOLVER_Solve), and writes the results in a raw binary layout.
 *
 *   gcge_ref A.bin B.bin|- nev X0.bin|- out.bin [-gcge_* flags...]
 *
 * Matrix files: int32 nrows, ncols, nnz; int32 rowptr[nrows+1];
 * int32 colidx[nnz]; float64 vals[nnz].  X0: float64 n*nev, column-major;
 * when given, the solver starts from it (num_init_evec = nev).
 * Output: int32 nev, num_iter, num_conv_continuous, n; float64 seconds;
 * float64 eval[nev]; float64 res[nev]; float64 evec[n*nev].
 * With "-gcge_print_eval 1" the reference prints every iteration's
 * eigenvalues on stdout, which the tests parse for the trajectory.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gcge.h"
#include "gcge_app_csr.h"

static CSR_MAT *read_bin(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "gcge_ref: cannot open %s\n", path);
        exit(2);
    }
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, f) != 3) exit(2);
    CSR_MAT *m = (CSR_MAT *)malloc(sizeof(CSR_MAT));
    m->N_Rows = hdr[0];
    m->N_Columns = hdr[1];
    m->N_Entries = hdr[2];
    m->RowPtr = (int *)malloc((size_t)(hdr[0] + 1) * sizeof(int));
    m->KCol = (int *)malloc((size_t)hdr[2] * sizeof(int));
    m->Entries = (double *)malloc((size_t)hdr[2] * sizeof(double));
    if (fread(m->RowPtr, sizeof(int), (size_t)hdr[0] + 1, f) != (size_t)hdr[0] + 1) exit(2);
    if (fread(m->KCol, sizeof(int), (size_t)hdr[2], f) != (size_t)hdr[2]) exit(2);
    if (fread(m->Entries, sizeof(double), (size_t)hdr[2], f) != (size_t)hdr[2]) exit(2);
    fclose(f);
    return m;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s A.bin B.bin|- nev X0.bin|- out.bin [-gcge_* ...]\n", argv[0]);
        return 2;
    }
    srand(1);
    CSR_MAT *A = read_bin(argv[1]);
    CSR_MAT *B = strcmp(argv[2], "-") ? read_bin(argv[2]) : NULL;
    int nev = atoi(argv[3]);
    int n = A->N_Rows;
    GCGE_SOLVER *s = GCGE_CSR_Solver_Init(A, B, nev, argc - 6, argv + 6);
    nev = s->para->nev;
    if (strcmp(argv[4], "-")) {
        FILE *f = fopen(argv[4], "rb");
        if (!f) return 2;
        for (int j = 0; j < nev; ++j) {
            CSR_VEC *v = (CSR_VEC *)s->evec[j];
            if (fread(v->Entries, sizeof(double), (size_t)n, f) != (size_t)n) return 2;
        }
        fclose(f);
        s->para->num_init_evec = nev;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    GCGE_SOLVER_Solve(s);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

    FILE *o = fopen(argv[5], "wb");
    if (!o) return 2;
    int hdr[4] = {nev, s->para->num_iter, s->para->num_conv_continuous, n};
    fwrite(hdr, sizeof(int), 4, o);
    fwrite(&secs, sizeof(double), 1, o);
    fwrite(s->eval, sizeof(double), (size_t)nev, o);
    fwrite(s->para->res, sizeof(double), (size_t)nev, o);
    for (int j = 0; j < nev; ++j) fwrite(((CSR_VEC *)s->evec[j])->Entries, sizeof(double), (size_t)n, o);
    fclose(o);
    printf("gcge_ref: done in %.6f s\n", secs);
    return 0;
}
