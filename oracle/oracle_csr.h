/*
 * ORACLE -- test infrastructure only (see oracle_csr.c).
 */
#ifndef ORACLE_CSR_H_
#define ORACLE_CSR_H_

#include <stdint.h>

#include "gcge_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ORACLE_CSR_ {
    int nrows, ncols, nnz;
    int *rowptr, *colidx;
    double *vals;
} ORACLE_CSR;

ORACLE_CSR *ORACLE_CSR_Create(int nrows, int ncols, int nnz, const int *rowptr, const int *colidx,
                              const double *vals);
void ORACLE_CSR_Free(ORACLE_CSR *m);
int ORACLE_CSR_Nrows(const ORACLE_CSR *m);
int ORACLE_CSR_Nnz(const ORACLE_CSR *m);
void ORACLE_CSR_Export(const ORACLE_CSR *m, int *rowptr, int *colidx, double *vals);
ORACLE_CSR *ORACLE_CSR_ReadFile(const char *path);
int ORACLE_CSR_WriteFile(const ORACLE_CSR *m, const char *path);

double ORACLE_RandomValue(uint64_t seed, uint64_t draw, uint64_t col, uint64_t row);
void ORACLE_SetSeed(uint64_t seed);
void ORACLE_SetOps(GCGE_OPS *ops);
void ORACLE_MultiVecSetHost(void **mv, GCGE_INT start, GCGE_INT ncols, const GCGE_DOUBLE *host,
                            GCGE_INT ld);
void ORACLE_MultiVecGetHost(void **mv, GCGE_INT start, GCGE_INT ncols, GCGE_DOUBLE *host,
                            GCGE_INT ld);
const void *ORACLE_Backend(void);

void ORACLE_SymEig(int n, const double *a, int lda, double *eval, double *evec, int lde);
void ORACLE_SpMM(const ORACLE_CSR *m, int k, const double *X, int ldx, double *Y, int ldy);
void ORACLE_Gram(int n, int mv, int kw, const double *V, int ldv, const double *W, int ldw,
                 double *G, int ldg);
void ORACLE_LinearComb(int n, int m, int k, const double *V, int ldv, const double *C, int ldc,
                       double beta, double *Y, int ldy);
void ORACLE_RandomFill(uint64_t seed, uint64_t draw, int col0, int ncols, int n, double *X,
                       int ldx);
void ORACLE_BlockCG(const ORACLE_CSR *A, int if_shift, double shift, const ORACLE_CSR *B, int k,
                    double *rhs, int ldb, double *x, int ldx, int max_it, double rate);

#ifdef __cplusplus
}
#endif

#endif
