// Q1 finite-element matrices assembled on the device, and the sliced-ELLPACK
// helpers a device-built matrix needs.
//
// The uniform E^3 mesh of the unit cube, trilinear elements, Dirichlet nodes
// eliminated: n = N^3 interior nodes, N = E - 1, node (x, y, z) (0-based,
// x fastest).  Node i couples to the nodes at the 27 offsets d = (dx, dy, dz)
// in {-1, 0, 1}^3, slot o = 9 (dz + 1) + 3 (dy + 1) + (dx + 1).  Every element
// around i that also holds i + d contributes the same integer numerator to the
// entry: 36 Khat(d) = k_x m_y m_z + m_x k_y m_z + m_x m_y k_z and 216 Mhat(d) =
// m_x m_y m_z, with k = 1, m = 2 along an axis where d is 0 and k = -1, m = 1
// where it is not.  The stiffness entry is
//     K(i, i + d) = (sum over those elements, in the order sx, sy, sz of the
//                    element offsets, of 36Khat(d) kappa_e) * (h / 36)
// and the mass entry (count * 216Mhat(d)) * (h^3 / 216).  Face neighbours
// (one nonzero component of d) get 36Khat = 0 and are not stored in K.  The
// sums are formed without fused multiply-adds, in that order, so the values
// equal those of a host assembly that adds the element terms one by one.
//
//   k_q1_masks     one thread per node: kept slots of its row (bit o) and the
//                  count of entries above 1e-20 in magnitude
//   k_q1_fill      one thread per node: the kept entries at their positions in
//                  the row (columns increasing; rows and columns renumbered by
//                  node_row), written to sliced ELLPACK with their CSR
//                  positions, and to diagonal storage when it exists
//   k_sell_pad     the padding of every slice (zero in the row's own column)
//   k_sell_spread  per slice: entries of entry rows spanning < 256 columns
//   k_sell_to_csr  the stored entries back at their CSR positions, row counts
#include <hip/hip_runtime.h>

#include <climits>

#include "gcge_hip_internal.h"

#pragma clang fp contract(off)

namespace gcge_hip {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

__device__ __forceinline__ int slot_dx(int o) { return o % 3 - 1; }
__device__ __forceinline__ int slot_dy(int o) { return (o / 3) % 3 - 1; }
__device__ __forceinline__ int slot_dz(int o) { return o / 9 - 1; }

// numerators 36 Khat(d) and 216 Mhat(d) of slot o
__device__ __forceinline__ int q1_k36(int o)
{
    const int dx = slot_dx(o), dy = slot_dy(o), dz = slot_dz(o);
    const int kx = dx ? -1 : 1, ky = dy ? -1 : 1, kz = dz ? -1 : 1;
    const int mx = dx ? 1 : 2, my = dy ? 1 : 2, mz = dz ? 1 : 2;
    return kx * my * mz + mx * ky * mz + mx * my * kz;
}
__device__ __forceinline__ int q1_m216(int o)
{
    return (slot_dx(o) ? 1 : 2) * (slot_dy(o) ? 1 : 2) * (slot_dz(o) ? 1 : 2);
}

// The row of node i: value, kept flag and (natural) column of every slot.
struct Q1Row {
    double val[27];
    bool kept[27];
    long col[27];
};

__device__ __forceinline__ void q1_row(const Q1Args &q, long i, Q1Row &r)
{
    const int N = q.E - 1;
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((long)N * N));
    double kap[8];  // elements (x + sx, y + sy, z + sz), s = 4 sx + 2 sy + sz
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const long e = (long)(x + ((s >> 2) & 1)) + (long)q.E * (y + ((s >> 1) & 1)) +
                       (long)q.E * q.E * (z + (s & 1));
        kap[s] = q.kappa ? q.kappa[e] : 1.0;
    }
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        const int dx = slot_dx(o), dy = slot_dy(o), dz = slot_dz(o);
        const bool valid = x + dx >= 0 && x + dx < N && y + dy >= 0 && y + dy < N &&
                           z + dz >= 0 && z + dz < N;
        r.col[o] = i + dx + (long)N * dy + (long)N * N * dz;
        double v;
        if (q.which == 0) {
            const double t = (double)q1_k36(o);
            double sum = 0.0;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int sx = (s >> 2) & 1, sy = (s >> 1) & 1, sz = s & 1;
                // element s holds node i at local corner (1 - sx, 1 - sy, 1 - sz)
                // and node i + d only if every 1 - s + d is 0 or 1
                const bool in = (dx == 0 || sx == (dx > 0)) && (dy == 0 || sy == (dy > 0)) &&
                                (dz == 0 || sz == (dz > 0));
                if (in) sum = sum + t * kap[s];
            }
            v = sum * q.scale;
            r.kept[o] = valid && v != 0.0;
        } else {
            const int count = (dx ? 1 : 2) * (dy ? 1 : 2) * (dz ? 1 : 2);
            v = (double)(count * q1_m216(o)) * q.scale;
            r.kept[o] = valid;
        }
        r.val[o] = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_q1_masks(Q1Args q, int *__restrict__ mask,
                                                    int *__restrict__ big)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    const long n = (long)(q.E - 1) * (q.E - 1) * (q.E - 1);
    if (i >= n) return;
    Q1Row r;
    q1_row(q, i, r);
    int m = 0, nb = 0;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        if (r.kept[o]) {
            m |= 1 << o;
            if (fabs(r.val[o]) > 1e-20) ++nb;
        }
    }
    const long row = q.node_row ? q.node_row[i] : i;
    mask[row] = m;
    big[row] = nb;
}

__global__ __launch_bounds__(kBlock) void k_q1_fill(Q1Args q, const int *__restrict__ rowptr,
                                                   const int *__restrict__ sptr,
                                                   int *__restrict__ scol, double *__restrict__ sval,
                                                   int *__restrict__ spos, DiaOffsets dia,
                                                   double *__restrict__ d_dia, long dia_ld)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    const long n = (long)(q.E - 1) * (q.E - 1) * (q.E - 1);
    if (i >= n) return;
    Q1Row r;
    q1_row(q, i, r);
    const long row = q.node_row ? q.node_row[i] : i;
    if (q.node_row) {
#pragma unroll
        for (int o = 0; o < 27; ++o)
            if (r.kept[o]) r.col[o] = q.node_row[r.col[o]];
    }
    const long base = sptr[row / kWave] + row % kWave;
    const int p0 = rowptr[row], p1 = rowptr[row + 1];
    int before = 0;  // kept slots ahead of o (natural order: columns increase with o)
#pragma unroll
    for (int o = 0; o < 27; ++o) {
        if (!r.kept[o]) continue;
        int rank = before++;
        if (q.node_row) {
            rank = 0;
#pragma unroll
            for (int t = 0; t < 27; ++t) rank += (r.kept[t] && r.col[t] < r.col[o]) ? 1 : 0;
        }
        if (p0 + rank >= p1) continue;  // the row's length from the mask pass bounds every write
        const long pos = base + (long)rank * kWave;
        scol[pos] = (int)r.col[o];
        sval[pos] = r.val[o];
        spos[pos] = p0 + rank;
        if (d_dia) {
            const long off = r.col[o] - row;
            int s = 0;
            while (s < dia.nd - 1 && dia.off[s] != off) ++s;
            d_dia[(long)s * dia_ld + row] = r.val[o];
        }
    }
}

// one wave per slice
__global__ __launch_bounds__(kBlock) void k_sell_pad(int nrows, int ncols, int nslices,
                                                    const int *__restrict__ sptr,
                                                    const int *__restrict__ sw,
                                                    int *__restrict__ scol,
                                                    double *__restrict__ sval,
                                                    int *__restrict__ spos)
{
    const int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (slice >= nslices) return;
    const int lane = threadIdx.x % kWave;
    const int row = slice * kWave + lane;
    const int width = sw[slice];
    for (int j = 0; j < width; ++j) {
        const long pos = sptr[slice] + (long)j * kWave + lane;
        scol[pos] = row < ncols ? row : 0;
        sval[pos] = 0.0;
        spos[pos] = -1;
    }
}

__device__ __forceinline__ int wave_min(int v)
{
    for (int m = 1; m < kWave; m <<= 1) v = min(v, __shfl_xor(v, m));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
    for (int m = 1; m < kWave; m <<= 1) v = max(v, __shfl_xor(v, m));
    return v;
}

// out[2 s] = stored entries of slice s in entry rows spanning < 256 columns,
// out[2 s + 1] = stored entries of slice s
__global__ __launch_bounds__(kBlock) void k_sell_spread(int nslices, const int *__restrict__ sptr,
                                                       const int *__restrict__ sw,
                                                       const int *__restrict__ scol,
                                                       const int *__restrict__ spos,
                                                       int *__restrict__ out)
{
    const int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (slice >= nslices) return;
    const int lane = threadIdx.x % kWave;
    const int width = sw[slice];
    int tight = 0, total = 0;
    for (int j = 0; j < width; ++j) {
        const long pos = sptr[slice] + (long)j * kWave + lane;
        const bool stored = spos[pos] >= 0;
        const int c = scol[pos];
        const int lo = wave_min(stored ? c : INT_MAX), hi = wave_max(stored ? c : -1);
        const int cnt = __popcll(__ballot(stored));
        total += cnt;
        if (cnt && hi - lo < 256) tight += cnt;
    }
    if (lane == 0) {
        out[2 * slice] = tight;
        out[2 * slice + 1] = total;
    }
}

__global__ __launch_bounds__(kBlock) void k_sell_to_csr(int nrows, int nslices,
                                                       const int *__restrict__ sptr,
                                                       const int *__restrict__ sw,
                                                       const int *__restrict__ scol,
                                                       const double *__restrict__ sval,
                                                       const int *__restrict__ spos,
                                                       int *__restrict__ ocol,
                                                       double *__restrict__ oval,
                                                       int *__restrict__ rowlen)
{
    const int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (slice >= nslices) return;
    const int lane = threadIdx.x % kWave;
    const int row = slice * kWave + lane;
    if (row >= nrows) return;
    const int width = sw[slice];
    int cnt = 0;
    for (int j = 0; j < width; ++j) {
        const long pos = sptr[slice] + (long)j * kWave + lane;
        const int p = spos[pos];
        if (p < 0) continue;
        ocol[p] = scol[pos];
        oval[p] = sval[pos];
        ++cnt;
    }
    rowlen[row] = cnt;
}

inline unsigned blocks_for(long threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }
inline unsigned slice_blocks(int nslices) { return (unsigned)((nslices + 3) / 4); }

}  // namespace

void q1_row_masks(const Q1Args &q, int *d_mask, int *d_big)
{
    const long n = (long)(q.E - 1) * (q.E - 1) * (q.E - 1);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_q1_masks, dim3(blocks_for(n)), dim3(kBlock), 0, state().stream, q, d_mask,
                       d_big);
    HIP_CHECK(hipGetLastError());
}

void q1_fill(const Q1Args &q, const int *d_rowptr, const int *d_sptr, int *d_col, double *d_val,
             int *d_pos, const DiaOffsets &dia, double *d_dia, long dia_ld)
{
    const long n = (long)(q.E - 1) * (q.E - 1) * (q.E - 1);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_q1_fill, dim3(blocks_for(n)), dim3(kBlock), 0, state().stream, q,
                       d_rowptr, d_sptr, d_col, d_val, d_pos, dia, d_dia, dia_ld);
    HIP_CHECK(hipGetLastError());
}

void sell_pad(int nrows, int ncols, int nslices, const int *d_sptr, const int *d_sw, int *d_col,
              double *d_val, int *d_pos)
{
    if (nslices <= 0) return;
    hipLaunchKernelGGL(k_sell_pad, dim3(slice_blocks(nslices)), dim3(kBlock), 0, state().stream,
                       nrows, ncols, nslices, d_sptr, d_sw, d_col, d_val, d_pos);
    HIP_CHECK(hipGetLastError());
}

void sell_spread(int nslices, const int *d_sptr, const int *d_sw, const int *d_col,
                 const int *d_pos, int *d_out)
{
    if (nslices <= 0) return;
    hipLaunchKernelGGL(k_sell_spread, dim3(slice_blocks(nslices)), dim3(kBlock), 0,
                       state().stream, nslices, d_sptr, d_sw, d_col, d_pos, d_out);
    HIP_CHECK(hipGetLastError());
}

void sell_to_csr(const Mat *A, int *d_col, double *d_val, int *d_rowlen)
{
    if (A->nslices <= 0) return;
    hipLaunchKernelGGL(k_sell_to_csr, dim3(slice_blocks(A->nslices)), dim3(kBlock), 0,
                       state().stream, A->nrows, A->nslices, A->d_slice_ptr, A->d_slice_w,
                       A->d_col, A->d_val, A->d_csrpos, d_col, d_val, d_rowlen);
    HIP_CHECK(hipGetLastError());
}

}  // namespace gcge_hip
