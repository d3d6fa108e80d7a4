// Gradients of eigenvalue and eigenvector losses with respect to the matrix
// entries and to per-element coefficients.
//
// For B-normalised eigenpairs (A - lambda_c B) x_c = 0 of symmetric A, B and a
// loss L(lambda), dL/dA = sum_c gbar_c x_c x_c^T and dL/dB = -sum_c gbar_c
// lambda_c x_c x_c^T; only the entries on the sparsity pattern are formed
// (pattern_gradient, weights gbar or -gbar lambda).  When A = sum_e kappa_e
// A_e (Q1 stiffness on the E^3 mesh, A_e = kappa-free element matrix),
// dL/dkappa_e = sum_c gbar_c x_c|e^T A_e x_c|e (q1_element_gradient).  Both
// take a second block Y for the bilinear form y^T A_e x of eigenvector-loss
// adjoints (Y = X for eigenvalue losses).
//
//   k_pattern_grad   one lane per sliced-ELLPACK entry position, k columns
//                    summed in registers, scattered to the CSR position;
//                    HBM-bound: nnz (4 + 4 + 8) + gathers of X and Y
//   k_q1_elem_grad   one thread per element: the 8 corner values of x and y
//                    (zero on Dirichlet nodes) and the integer element
//                    numerators 36 Khat in constant memory
#include <hip/hip_runtime.h>

#include "gcge_hip_internal.h"

namespace gcge_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

__global__ __launch_bounds__(kBlock) void k_pattern_grad(int n, int nslices,
                                                        const int *__restrict__ sptr,
                                                        const int *__restrict__ sw,
                                                        const int *__restrict__ scol,
                                                        const int *__restrict__ spos,
                                                        const double *__restrict__ X, long ldx,
                                                        const double *__restrict__ Y, long ldy,
                                                        int k, const double *__restrict__ w,
                                                        double *__restrict__ grad)
{
    const int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (slice >= nslices) return;
    const int lane = threadIdx.x % kWave;
    const int row = slice * kWave + lane;
    if (row >= n) return;
    const int off = sptr[slice], width = sw[slice];
    for (int j = 0; j < width; ++j) {
        const long pos = off + (long)j * kWave + lane;
        const int p = spos[pos];
        if (p < 0) continue;
        const int col = scol[pos];
        double s = 0.0;
        for (int c = 0; c < k; ++c) {
            const double *xc = X + (long)c * ldx, *yc = Y + (long)c * ldy;
            s += w[c] * 0.5 * (xc[row] * yc[col] + yc[row] * xc[col]);
        }
        grad[p] = s;
    }
}

// 36 * Khat of the unit cube, corners p = 4 ax + 2 ay + az
__constant__ int c_k36[64];

__global__ __launch_bounds__(kBlock) void k_q1_elem_grad(int E, double scale,
                                                        const double *__restrict__ X, long ldx,
                                                        const double *__restrict__ Y, long ldy,
                                                        int k, const double *__restrict__ w,
                                                        const int *__restrict__ node_row,
                                                        double *__restrict__ grad)
{
    const long e = (long)blockIdx.x * kBlock + threadIdx.x;
    const long ne = (long)E * E * E;
    if (e >= ne) return;
    const int ex = (int)(e % E), ey = (int)((e / E) % E), ez = (int)(e / ((long)E * E));
    const int N = E - 1;
    long node[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int ix = ex + ((p >> 2) & 1), iy = ey + ((p >> 1) & 1), iz = ez + (p & 1);
        const bool interior = ix >= 1 && ix <= N && iy >= 1 && iy <= N && iz >= 1 && iz <= N;
        node[p] = interior ? (long)(ix - 1) + (long)N * (iy - 1) + (long)N * N * (iz - 1) : -1;
        if (node_row && node[p] >= 0) node[p] = node_row[node[p]];  // row of a permuted block
    }
    double total = 0.0;
    for (int c = 0; c < k; ++c) {
        const double *xc = X + (long)c * ldx, *yc = Y + (long)c * ldy;
        double xv[8], yv[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            xv[p] = node[p] >= 0 ? xc[node[p]] : 0.0;
            yv[p] = node[p] >= 0 ? yc[node[p]] : 0.0;
        }
        double q = 0.0;
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < 8; ++r) t += (double)c_k36[p * 8 + r] * yv[r];
            q += xv[p] * t;
        }
        total += w[c] * q;
    }
    grad[e] = scale * total;
}

}  // namespace

void pattern_gradient(const Mat *A, const double *X, long ldx, const double *Y, long ldy, int k,
                      const double *d_w, double *d_grad)
{
    if (k <= 0 || A->nrows <= 0) return;
    const int nb = (A->nslices + kBlock / kWave - 1) / (kBlock / kWave);
    ProfScope prof("pattern_grad", (double)A->nnz * 16.0 + 16.0 * k * (double)A->nrows);
    hipLaunchKernelGGL(k_pattern_grad, dim3(nb), dim3(kBlock), 0, state().stream, A->nrows,
                       A->nslices, A->d_slice_ptr, A->d_slice_w, A->d_col, A->d_csrpos, X, ldx, Y,
                       ldy, k, d_w, d_grad);
    HIP_CHECK(hipGetLastError());
}

void q1_element_gradient(int E, const double *X, long ldx, const double *Y, long ldy, int k,
                         const double *d_w, double *d_grad, const int *d_node_row)
{
    if (E < 2 || k <= 0) return;
    static bool uploaded = false;
    if (!uploaded) {
        int k36[64];
        for (int p = 0; p < 8; ++p)
            for (int q = 0; q < 8; ++q) {
                const int a[3] = {(p >> 2) & 1, (p >> 1) & 1, p & 1};
                const int b[3] = {(q >> 2) & 1, (q >> 1) & 1, q & 1};
                int k1[3], m6[3];
                for (int d = 0; d < 3; ++d) {
                    k1[d] = a[d] == b[d] ? 1 : -1;
                    m6[d] = a[d] == b[d] ? 2 : 1;
                }
                k36[p * 8 + q] = k1[0] * m6[1] * m6[2] + m6[0] * k1[1] * m6[2] + m6[0] * m6[1] * k1[2];
            }
        HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_k36), k36, sizeof(k36)));
        uploaded = true;
    }
    const long ne = (long)E * E * E;
    const double h = 1.0 / E;
    ProfScope prof("q1_elem_grad", 16.0 * k * (double)(E - 1) * (E - 1) * (E - 1) + 8.0 * ne);
    hipLaunchKernelGGL(k_q1_elem_grad, dim3((unsigned)((ne + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       state().stream, E, h / 36.0, X, ldx, Y, ldy, k, d_w, d_node_row, d_grad);
    HIP_CHECK(hipGetLastError());
}

}  // namespace gcge_hip
