/*
 * gcge_hip.h -- the MI355X (gfx950) backend of the GCG eigensolver: the
 * C ABI of libgcge_hip.so.
 *
 * The backend fills the GCGE operations table (gcge_ops.h) with device
 * implementations, the way the reference CSR backend fills it with host loops
 * (reference app/csr/gcge_app_csr.c:70-84, GCGE_CSR_SetOps).  A program
 * written for the reference switches backends by replacing
 *     GCGE_CSR_SetOps(solver->ops)          with   GCGE_HIP_SetOps(solver->ops)
 * and creating its matrices with GCGE_HIP_MatCreateCSR instead of the CSR_MAT
 * constructors (reference app/csr/gcge_app_csr.h:27-46).
 *
 * Data layout in device memory:
 *   matrices   CSR input, stored as sliced ELLPACK with slices of 64 rows
 *              (one wavefront per slice, entries of a slice column-major),
 *              int32 column indices, float64 values (and an optional float32
 *              copy of the values for the mixed-precision inner solves);
 *              matrices whose entries lie on few diagonals also get one
 *              float64 array per diagonal, which their products then use
 *              (GCGE_HIP_MatSetStorage);
 *   multivecs  one column-major block of n x ncols float64, leading dimension
 *              n rounded up to 64; column j is a vector view.
 * All device work runs on one HIP stream per process.  Operations that
 * return host values synchronise that stream; the others are asynchronous.
 * Errors: the first HIP error is recorded (GCGE_HIP_LastError) and the
 * process aborts with a message, since the operations table has no error
 * channel (every table operation returns void).
 */
#ifndef GCGE_HIP_H_
#define GCGE_HIP_H_

#include <stdint.h>

#include "gcge_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle ---------------------------------------------------------- */
/* Select the device and create the stream and buffers (idempotent).  Returns 0
 * on success, nonzero when no usable GPU is present.  No counterpart in the
 * reference (its CSR backend needs no set-up). */
int GCGE_HIP_Init(int device);
void GCGE_HIP_Finalize(void);
/* First recorded error message, or "" */
const char *GCGE_HIP_LastError(void);
/* Wait for all queued device work */
void GCGE_HIP_Synchronize(void);
/* Seed of the counter-based uniform(-1, 1) generator used by
 * VecSetRandomValue / MultiVecSetRandomValue (reference
 * app/csr/gcge_app_csr.c:284-294 draws uniform[0,1) from rand()) */
void GCGE_HIP_SetSeed(uint64_t seed);

/* ---- matrices (replace CSR_MAT, reference app/csr/gcge_app_csr.h:27-41) --- */
/* Copy a host CSR matrix (0-based int32 rowptr[nrows+1], colidx[nnz], float64
 * vals[nnz]) to the device.  Returns an opaque matrix handle for the
 * operations table, or NULL on error. */
void *GCGE_HIP_MatCreateCSR(int nrows, int ncols, int nnz, const int *rowptr, const int *colidx,
                            const double *vals);
void GCGE_HIP_MatDestroy(void *mat);
/* fp32 != 0: keep a float32 copy of the values and use it in the inner block
 * CG (mixed-precision form); fp32 == 0: float64 everywhere (default). */
void GCGE_HIP_MatSetPrecision(void *mat, int fp32);
/* Storage used by the products of this matrix (no counterpart in the
 * reference, whose CSR products are per-vector loops).  At creation a square
 * matrix whose entries lie on few diagonals (sorted offsets col - row
 * covered by at most 9 runs of three consecutive offsets, rows with
 * increasing columns, diagonals at least half full: stencil and FEM matrices
 * in their natural ordering) also gets diagonal storage, and its products
 * use it.  A matrix most of whose entries sit in sliced-ELLPACK entry rows
 * spanning 256 columns or more (scattered columns, e.g. a random symmetric
 * permutation) instead
 * runs its products of 16 columns or more with whole-row gathers from a
 * row-major copy of the column block.  storage: 0 = sliced ELLPACK with
 * element gathers (every matrix has it), 1 = diagonal (when it exists,
 * else 0), 2 = sliced ELLPACK with row gathers.  Returns the storage now
 * used (0, 1 or 2).  Results of the forms agree to rounding. */
int GCGE_HIP_MatSetStorage(void *mat, int storage);
/* number of stored diagonals and of runs (0, 0 without diagonal storage);
 * returns the storage the products use (0, 1 or 2 as above) */
int GCGE_HIP_MatStorageInfo(void *mat, int *ndiag, int *ngroups);
/* Diagonal storage walks its row blocks in one of two orders, with the same
 * result bit for bit: consecutive blocks, or, for the 27-point pattern of a
 * lexicographically numbered grid (run centres a N + b S, a, b in {-1, 0, 1},
 * S rows per plane dividing the row count), the same in-plane block of
 * consecutive planes together, so that the rows of X the plane runs share are
 * read from HBM once.  GCGE_HIP_MatSetStorage with storage 1 picks the order
 * (plane order from S = 8192 rows), 3 forces consecutive blocks, 4 plane
 * order where the pattern exists; both return 1.  Returns the plane stride
 * the products use, 0 for consecutive blocks or other storages. */
int GCGE_HIP_MatDiagonalPlane(void *mat);

/* Q1 finite-element matrices assembled on the device (no counterpart in the
 * reference, whose drivers read assembled matrices from files): the uniform
 * E^3 mesh of the unit cube, trilinear elements, Dirichlet nodes eliminated,
 * n = (E-1)^3 interior nodes numbered (ix-1) + N(iy-1) + N^2(iz-1), N = E-1.
 * which = 0: stiffness K = sum_e kappa_e h Khat (kappa: E^3 host doubles in
 * element order ex + E ey + E^2 ez, NULL for kappa = 1; entries that come out
 * exactly zero, such as every face-neighbour entry, are not stored);
 * which = 1: consistent mass M = sum_e h^3 Mhat (kappa ignored).  node_row
 * (host, n ints, optional) renumbers rows and columns: row node_row[i] holds
 * node i; it must be a permutation (else an error).  Rows list their columns
 * in increasing order.  The result equals, entry for entry and bit for bit,
 * GCGE_HIP_MatCreateCSR of the host assembly that adds each entry's element
 * terms in the order of the element offsets (x slowest) and scales once. */
void *GCGE_HIP_MatCreateQ1(int E, const double *kappa, int which, const int *node_row);
/* Copy a matrix back in CSR form (rowptr[nrows+1], colidx[nnz], vals[nnz];
 * any may be NULL), entries in the order they were given or assembled.
 * Returns nnz. */
int GCGE_HIP_MatGetCSR(void *mat, int *rowptr, int *colidx, double *vals);

/* ---- multivecs (replace the default MultiVecCreateByMat of per-vector
 *      allocations, reference src/gcge_ops.c:72-80) ----------------------- */
/* Returns a void** handle usable wherever the table expects a multivec. */
void **GCGE_HIP_MultiVecCreate(int n, int ncols);
void GCGE_HIP_MultiVecDestroy(void **mv);
/* host (column-major, leading dimension ld) <-> columns start..start+ncols-1 */
void GCGE_HIP_MultiVecSetHost(void **mv, int start, int ncols, const double *host, int ld);
void GCGE_HIP_MultiVecGetHost(void **mv, int start, int ncols, double *host, int ld);

/* ---- block kernels on multivec column ranges ----------------------------- */
/* Y(:, ys:ys+k) = A X(:, xs:xs+k)  (reference per-column loop
 * src/gcge_ops.c:115-129 over app/csr/gcge_app_csr.c:297-323) */
void GCGE_HIP_SpMM(void *mat, void **X, int xs, void **Y, int ys, int k);
/* G(i, j) = V(:, vs+i)^T W(:, ws+j), host G with leading dimension ldg
 * (reference src/gcge_ops.c:226-309, "ns" form) */
void GCGE_HIP_Gram(void **V, int vs, int m, void **W, int ws, int k, double *G, int ldg);
/* GCGE_HIP_Gram without allocating: no counterpart in the reference; for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums and the result live in a fixed 8 MiB device array that each
 * device holds for the life of the process; m * k may be at most 524288
 * (else returns 1 and sets GCGE_HIP_LastError, computing nothing).  Same
 * kernels and, while the partial sums fit, the same row ranges as
 * GCGE_HIP_Gram, so the same result bit for bit; beyond that fewer ranges
 * (the same products summed in larger groups).  G is written through a
 * plain device-to-host copy.  Every call on one device shares its array:
 * calls that overlap on one device (two host threads) race.  Returns 0. */
int GCGE_HIP_GramNoAlloc(void **V, int vs, int m, void **W, int ws, int k, double *G, int ldg);
/* Y(:, ys+j) = beta Y(:, ys+j) + sum_i V(:, vs+i) C(i, j), host C with leading
 * dimension ldc (reference src/gcge_ops.c:190-221) */
void GCGE_HIP_LinearComb(void **V, int vs, int m, const double *C, int ldc, double beta,
                         void **Y, int ys, int k);
/* Y(:, ys:ys+k) = a X(:, xs:xs+k) + b Y(:, ys:ys+k) (reference
 * src/gcge_ops.c:149-163 over app/csr/gcge_app_csr.c:330-365) */
void GCGE_HIP_Axpby(double a, void **X, int xs, double b, void **Y, int ys, int k);
/* uniform(-1, 1) in columns start..start+k-1 (one draw) */
void GCGE_HIP_RandomFill(void **X, int start, int k);

/* ---- the operations table ------------------------------------------------ */
/* Fill the table with the device operations (replaces GCGE_CSR_SetOps,
 * reference app/csr/gcge_app_csr.c:70-84).  Besides the mandatory
 * single-vector operations it provides every multivec operation as a block
 * kernel, MultiLinearSolver (block CG with the reference GCGE_BCG arithmetic,
 * src/gcge_cg.c:302-501, parameters from GCGE_HIP_SetLinearSolverParams) and
 * publishes the GCGE_OPS_EXT block extensions in ops->ops. */
void GCGE_HIP_SetOps(GCGE_OPS *ops);
/* Parameters MultiLinearSolver uses when a driver calls it without them
 * (reference defaults: cg_max_it 20, cg_rate 1e-2, no shift). */
void GCGE_HIP_SetLinearSolverParams(int max_it, double rate, int if_shift, double shift, void *B);
/* Work done since the last reset, counted as the single-vector CSR calls the
 * reference's column loops would make for it (no reference counterpart: the
 * input of a CPU cost model).  out[0..len) receives: SpMV calls, their stored
 * entries summed, their rows summed, inner products, axpbys, dense
 * eigensolves, the sum of their orders cubed, the vector length of the last
 * counted vector operation.  reset != 0 clears the counts after reading.
 * Returns the number of counts (8). */
int GCGE_HIP_WorkCounts(double *out, int len, int reset);
/* The parameters last set by GCGE_HIP_SetLinearSolverParams (outputs may be
 * NULL), so that a caller can restore them. */
void GCGE_HIP_GetLinearSolverParams(int *max_it, double *rate, int *if_shift, double *shift,
                                    void **B);
/* The table's MultiLinearSolver on columns [bs, bs+k) of b (right-hand
 * sides, overwritten by the residuals as in src/gcge_cg.c:302-501) and
 * [xs, xs+k) of x (initial guess in, solution out), with the parameters of
 * GCGE_HIP_SetLinearSolverParams.  Runs the float32 direction form when the
 * matrix was switched with GCGE_HIP_MatSetPrecision. */
void GCGE_HIP_MultiLinearSolve(void *mat, void **b, int bs, void **x, int xs, int k);
/* Descriptor {SetOps, MultiVecSetHost, MultiVecGetHost} for the one-call
 * driver GCGE_SolveWithBackend (gcge.h). */
const void *GCGE_HIP_Backend(void);

/* ---- adjoints (no reference counterpart: gradients of losses on computed
 *      eigenpairs; reference eigenpairs are B-normalised, x^T B x = 1) ---- */
/* grad[p] = sum_c w[c] (X(i, xs+c) Y(j, ys+c) + Y(i, ys+c) X(j, xs+c)) / 2 for
 * the stored entry (i, j) at position p of the CSR arrays given to
 * GCGE_HIP_MatCreateCSR (nnz host doubles).  Y NULL means Y = X.  With
 * Y = X = eigenvectors: w = dL/dlambda gives dL/dA on the pattern, w =
 * -lambda dL/dlambda gives dL/dB. */
void GCGE_HIP_PatternGradient(void *mat, void **X, int xs, void **Y, int ys, int k, const double *w,
                              double *grad);
/* Trilinear Q1 elements of the uniform E^3 mesh of the unit cube, interior
 * nodes numbered (ix-1) + N(iy-1) + N^2(iz-1), N = E-1 (Dirichlet nodes
 * eliminated), element e = ex + E ey + E^2 ez with stiffness h Khat:
 * grad[e] = sum_c w[c] x_c|e^T (h Khat) y_c|e (E^3 host doubles).  With
 * K = sum_e kappa_e h Khat and Y = X = eigenvectors, w = dL/dlambda gives
 * dL/dkappa. */
void GCGE_HIP_Q1ElementGradient(int E, void **X, int xs, void **Y, int ys, int k, const double *w,
                                double *grad);
/* The same for blocks whose rows are a permutation of the nodes (a solve of
 * P K P^T, P M P^T): node_row[i] (host, (E-1)^3 ints) is the row that holds
 * interior node i; entries out of range are rejected with an error. */
void GCGE_HIP_Q1ElementGradientRows(int E, void **X, int xs, void **Y, int ys, int k,
                                    const double *w, const int *node_row, double *grad);

/* ---- measurement --------------------------------------------------------- */
/* Per-kernel device time, measured with HIP events on the backend stream
 * while enabled (on = 1: every launch; on = 2: the SpMM launches only, whose
 * events are recorded by the dispatch itself).  name: "spmm", "gram",
 * "lincomb", "axpby", ...; returns the number of launches and the summed
 * milliseconds. */
void GCGE_HIP_ProfileEnable(int on);
int GCGE_HIP_KernelTime(const char *name, double *ms_total, int *count);
/* as GCGE_HIP_KernelTime, plus the summed algorithmic bytes of the launches
 * (SpMM: nnz*12 + 4*(n+1) + 8*k*(ncols+nrows); Gram and linear combination:
 * 8*n*(m+k), plus 8*n*k when the combination also reads Y) */
int GCGE_HIP_KernelStats(const char *name, double *ms_total, int *count, double *bytes_total);
void GCGE_HIP_ProfileReset(void);
/* create events ahead so that a profiled region creates none */
void GCGE_HIP_ProfileReserve(int events);

#ifdef __cplusplus
}
#endif

#endif /* GCGE_HIP_H_ */
