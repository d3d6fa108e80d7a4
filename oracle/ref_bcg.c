/*
 * ORACLE -- test infrastructure only.
 *
 * Runs the reference block CG (GCGE_BCG, src/gcge_cg.c:302-501) built from
 * its own sources on the reference CSR operations (app/csr), for fixtures
 * that pin the device block CG and the CPU restatement ORACLE_BlockCG.
 *
 *   gcge_ref_bcg A.bin B.bin|- k if_shift shift max_it rate rhs.bin x0.bin out.bin
 *
 * Matrix files as for gcge_ref.  rhs.bin / x0.bin: float64 n*k column-major.
 * out.bin: float64 x[n*k] (solution), then float64 r[n*k] (the right-hand
 * side block, which GCGE_BCG overwrites with the final residuals).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"
#include "gcge_app_csr.h"

static CSR_MAT *read_bin(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "gcge_ref_bcg: cannot open %s\n", path);
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

static void read_block(const char *path, void **mv, int n, int k)
{
    FILE *f = fopen(path, "rb");
    if (!f) exit(2);
    for (int j = 0; j < k; ++j)
        if (fread(((CSR_VEC *)mv[j])->Entries, sizeof(double), (size_t)n, f) != (size_t)n) exit(2);
    fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 11) {
        fprintf(stderr, "usage: %s A.bin B.bin|- k if_shift shift max_it rate rhs.bin x0.bin out.bin\n",
                argv[0]);
        return 2;
    }
    CSR_MAT *A = read_bin(argv[1]);
    CSR_MAT *B = strcmp(argv[2], "-") ? read_bin(argv[2]) : NULL;
    const int k = atoi(argv[3]);
    const int if_shift = atoi(argv[4]);
    const double shift = atof(argv[5]);
    const int max_it = atoi(argv[6]);
    const double rate = atof(argv[7]);
    const int n = A->N_Rows;

    GCGE_OPS *ops;
    GCGE_OPS_Create(&ops);
    GCGE_CSR_SetOps(ops);
    GCGE_OPS_Setup(ops);
    void **rhs, **x, **p, **w;
    ops->MultiVecCreateByMat(&rhs, k, A, ops);
    ops->MultiVecCreateByMat(&x, k, A, ops);
    ops->MultiVecCreateByMat(&p, k, A, ops);
    ops->MultiVecCreateByMat(&w, k, A, ops);
    read_block(argv[8], rhs, n, k);
    read_block(argv[9], x, n, k);
    double *dtmp = (double *)calloc((size_t)6 * k, sizeof(double));
    int *itmp = (int *)calloc((size_t)k, sizeof(int));
    int start[5] = {0, 0, 0, 0, 0};
    int end[5] = {k, k, k, k, k};
    GCGE_BCG(A, if_shift, shift, B, rhs, x, start, end, max_it, rate, ops, p, w, NULL, dtmp, itmp);

    FILE *o = fopen(argv[10], "wb");
    if (!o) return 2;
    for (int j = 0; j < k; ++j) fwrite(((CSR_VEC *)x[j])->Entries, sizeof(double), (size_t)n, o);
    for (int j = 0; j < k; ++j) fwrite(((CSR_VEC *)rhs[j])->Entries, sizeof(double), (size_t)n, o);
    fclose(o);
    ops->MultiVecDestroy(&rhs, k, ops);
    ops->MultiVecDestroy(&x, k, ops);
    ops->MultiVecDestroy(&p, k, ops);
    ops->MultiVecDestroy(&w, k, ops);
    GCGE_OPS_Free(&ops);
    free(dtmp);
    free(itmp);
    return 0;
}
