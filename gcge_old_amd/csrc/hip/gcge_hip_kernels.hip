// Device kernels of the GCG backend for gfx950 (CDNA4, 64-wide wavefronts).
//
//   spmm        sliced-ELLPACK (64-row slices) times a column block (float64,
//               or float32 values and vectors for the inner solves); one
//               wavefront per slice, each lane one row, the columns of a
//               16-column chunk in registers, four entries' index and value
//               loads issued ahead of their gathers; entries of a slice are
//               stored column-major so
//               that every entry load is one 512-byte coalesced access, and
//               workgroups are remapped so each XCD sweeps a contiguous band
//               of rows (its L2 keeps the band of X that the stencil reuses).
//               HBM-bound: nnz*(8+4) + n*k*8*2 bytes per launch (float64),
//               nnz*(4+4) + n*k*4*2 (float32).
//   gram        V^T W for tall-skinny column-major blocks: 64x64 (or 16-wide)
//               output tiles, 32-row steps staged through LDS (transposed),
//               4x4 register blocks per thread, split over row chunks with a
//               deterministic second-pass reduction.
//   lincomb     Y = beta Y + V C: one lane per row holding 8..64 output
//               columns in registers; the coefficients of one input column
//               are a wave-uniform contiguous read, so V is read once per
//               group of up to 64 output columns.
//   axpby/swap/random/coldots/residual/bcg_*: streaming kernels.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "gcge_hip_internal.h"

namespace gcge_hip {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kXcds = 8;

// Workgroup b is dispatched to XCD b % 8.  Renumber so that the workgroups of
// one XCD cover one contiguous range of work items.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int q = nb / kXcds, r = nb % kXcds;
    const int xcd = bid % kXcds, loc = bid / kXcds;
    return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// DPP move of a double (two 32-bit halves); lanes without a source, or in
// rows outside ROW_MASK, get 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes, returned to every lane (all lanes must be active).
// Within each row of 16 lanes, row_shr:1/2/4/8 steps leave the row sum in the
// row's last lane; row_bcast:15 adds rows 0 and 2 into rows 1 and 3,
// row_bcast:31 adds lane 31 into rows 2 and 3, so lane 63 holds the total.
// Data moves in the VALU, without the LDS crossbar of lane shuffles.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0x111, 0xf>(v);  // row_shr:1
    v += dpp_move<0x112, 0xf>(v);  // row_shr:2
    v += dpp_move<0x114, 0xf>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf>(v);  // row_shr:8
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), kWave - 1);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), kWave - 1);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / kWave; ++w) s += lds[w];
    __syncthreads();
    return s;
}

// ---- SpMM ------------------------------------------------------------------------
// One chunk of CK consecutive output columns for the lane's row; the slice's
// entries are re-read per chunk (from L1/L2 after the first chunk).  VT is the
// stored value type, XT the vector and accumulation type (double, double) for
// the float64 form, (float, float) for the float32 inner solves.  Entries are
// taken four at a time: their column indices and values are loaded first, so
// the gathers of X for four entries are in flight together; each accumulator
// still adds the entries in order.
// DOT > 0 (block CG, X = P): also p.w (DOT 1) and p.p (DOT 2) of the lane's
// row, summed over the wave into dpart[c * nslices + slice] and
// dpart[(k + c) * nslices + slice].
template <int CK, int DOT = 0, typename VT, typename XT>
__device__ __forceinline__ void sell_chunk(int row, int n, int w, const int *__restrict__ cp,
                                           const VT *__restrict__ vp,
                                           const XT *__restrict__ X, long ldx,
                                           XT *__restrict__ Y, long ldy, int c0,
                                           const int *__restrict__ xmap,
                                           const int *__restrict__ ymap,
                                           double *__restrict__ dpart = nullptr, int slice = 0,
                                           int nslices = 0, int k = 0)
{
    const XT *xc[CK];
#pragma unroll
    for (int cc = 0; cc < CK; ++cc) xc[cc] = X + (long)(xmap ? xmap[c0 + cc] : c0 + cc) * ldx;
    XT acc[CK];
#pragma unroll
    for (int cc = 0; cc < CK; ++cc) acc[cc] = XT(0);
    int j = 0;
    for (; j + 4 <= w; j += 4) {
        int col[4];
        XT v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            col[u] = cp[(long)(j + u) * kWave];
            v[u] = (XT)vp[(long)(j + u) * kWave];
        }
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) {
            XT xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xv[u] = xc[cc][col[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[cc] += v[u] * xv[u];
        }
    }
    for (; j < w; ++j) {
        const int col = cp[(long)j * kWave];
        const XT v = (XT)vp[(long)j * kWave];
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) acc[cc] += v * xc[cc][col];
    }
    if (row < n) {
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) Y[(long)(ymap ? ymap[c0 + cc] : c0 + cc) * ldy + row] = acc[cc];
    }
    if constexpr (DOT > 0) {
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) {
            const double pv = row < n ? (double)xc[cc][row] : 0.0;
            const double pw = wave_sum(pv * (double)acc[cc]);
            double pp = 0.0;
            if constexpr (DOT > 1) pp = wave_sum(pv * pv);
            if (threadIdx.x % kWave == 0) {
                dpart[(long)(c0 + cc) * nslices + slice] = pw;
                if constexpr (DOT > 1) dpart[(long)(k + c0 + cc) * nslices + slice] = pp;
            }
        }
    }
}

// remainder chunk of R0 <= r < MAXCK columns, r chosen at run time
template <int R, int MAXCK, int DOT, typename VT, typename XT>
__device__ __forceinline__ void sell_remainder(int r, int row, int n, int w, const int *cp,
                                               const VT *vp, const XT *X, long ldx, XT *Y,
                                               long ldy, int c0, const int *xmap,
                                               const int *ymap, double *dpart, int slice,
                                               int nslices, int k)
{
    if constexpr (R < MAXCK) {
        if (r == R)
            sell_chunk<R, DOT>(row, n, w, cp, vp, X, ldx, Y, ldy, c0, xmap, ymap, dpart, slice,
                               nslices, k);
        else
            sell_remainder<R + 1, MAXCK, DOT>(r, row, n, w, cp, vp, X, ldx, Y, ldy, c0, xmap, ymap,
                                              dpart, slice, nslices, k);
    }
}

// SPLIT waves of a workgroup share one slice, each taking a contiguous part
// of the group's columns (chunks of MAXCK, then one remainder chunk); the
// slice's entries reach the second wave from L1.  MAXCK and SPLIT are picked
// per launch: small k does not pay the registers of wide chunks, and
// splitting the columns gives more wavefronts to hide the gather latency.
// Wide blocks are cut into groups of gcols columns, blockIdx.y the group:
// the dispatcher runs the groups as successive sweeps over the rows, so the
// rows of X that the stencil re-reads stay cached for gcols columns at a
// time instead of all k.
template <typename VT, typename XT, int MAXCK, int SPLIT, int DOT>
__global__ __launch_bounds__(kBlock) void k_sell_spmm(int n, int nslices,
                                                     const int *__restrict__ sptr,
                                                     const int *__restrict__ sw,
                                                     const int *__restrict__ scol,
                                                     const VT *__restrict__ sval,
                                                     const XT *__restrict__ X, long ldx,
                                                     XT *__restrict__ Y, long ldy, int k,
                                                     int gcols,
                                                     const int *__restrict__ xmap,
                                                     const int *__restrict__ ymap,
                                                     double *__restrict__ dpart,
                                                     const int *__restrict__ count)
{
    constexpr int kSlicesPerBlock = kBlock / kWave / SPLIT;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = threadIdx.x / kWave;
    const int slice = blk * kSlicesPerBlock + wid / SPLIT;
    if (slice >= nslices) return;
    const int part = wid % SPLIT;
    const int kc = count ? min(k, *count) : k;  // columns listed on the device
    const int gb = blockIdx.y * gcols, ge = min(kc, gb + gcols);
    if (gb >= ge) return;
    const int per = (ge - gb + SPLIT - 1) / SPLIT;
    const int cb = gb + part * per, ce = min(ge, cb + per);
    if (cb >= ce) return;
    const int lane = threadIdx.x % kWave;
    const int row = slice * kWave + lane;
    const int off = sptr[slice];
    const int w = sw[slice];
    const int *cp = scol + off + lane;
    const VT *vp = sval + off + lane;
    int c0 = cb;
    for (; c0 + MAXCK <= ce; c0 += MAXCK)
        sell_chunk<MAXCK, DOT>(row, n, w, cp, vp, X, ldx, Y, ldy, c0, xmap, ymap, dpart, slice,
                               nslices, k);
    if (c0 < ce)
        sell_remainder<1, MAXCK, DOT>(ce - c0, row, n, w, cp, vp, X, ldx, Y, ldy, c0, xmap, ymap,
                                      dpart, slice, nslices, k);
}

template <typename VT, typename XT, int MAXCK, int SPLIT, int DOT>
void launch_sell_spmm_v(const Mat *A, const VT *vals, const XT *X, long ldx, XT *Y, long ldy, int k,
                        int gcols, const int *d_xmap, const int *d_ymap, double *dpart,
                        const int *d_count, hipEvent_t ev0, hipEvent_t ev1)
{
    constexpr int spb = kBlock / kWave / SPLIT;
    const int ng = (k + gcols - 1) / gcols;
    int nb = (A->nslices + spb - 1) / spb;
    // workgroup (x, y) goes to XCD (y * nb + x) % 8: with nb a multiple of 8
    // that is x % 8 in every group, as the remap assumes (extra workgroups
    // find no slice and return)
    if (ng > 1) nb = (nb + kXcds - 1) / kXcds * kXcds;
    // hipExtLaunchKernelGGL records the (optional) events in the dispatch packet
    hipExtLaunchKernelGGL((k_sell_spmm<VT, XT, MAXCK, SPLIT, DOT>), dim3(nb, ng), dim3(kBlock), 0,
                          state().stream, ev0, ev1, 0, A->nrows, A->nslices, A->d_slice_ptr,
                          A->d_slice_w, A->d_col, vals, X, ldx, Y, ldy, k, gcols, d_xmap, d_ymap,
                          dpart, d_count);
    HIP_CHECK(hipGetLastError());
}

// columns per wave after the split: 1..4 -> MAXCK 4, 5..8 -> 8, else 16;
// GCGE_HIP_SPMM_SPLIT = 1 | 2 | 4 overrides the split (for measurement)
template <int DOT, typename VT, typename XT>
void launch_sell_spmm(const Mat *A, const VT *vals, const XT *X, long ldx, XT *Y, long ldy, int k,
                      const int *d_xmap, const int *d_ymap, double *dpart, const int *d_count,
                      hipEvent_t ev0, hipEvent_t ev1)
{
    static const int forced = [] {
        const char *e = getenv("GCGE_HIP_SPMM_SPLIT");
        return e ? atoi(e) : 0;
    }();
    static const int forced_group = [] {
        const char *e = getenv("GCGE_HIP_SPMM_GROUP");
        return e ? atoi(e) : 0;
    }();
    // measured on MI355X (Q1 128^3, k = 50..200): groups of 32 columns beat
    // one sweep over all columns from about 64 columns up (k = 100: 3.4 ms
    // against 4.3 ms on K)
    const int gcols = forced_group > 0 ? std::min(forced_group, k) : (k > 48 ? 32 : k);
    const int kg = gcols;
    int split = forced;
    // measured on MI355X (Q1 64^3, k = 1..50): one wave per slice is best up to
    // 10 columns, two up to 20, four beyond
    if (split != 1 && split != 2 && split != 4) split = kg <= 10 ? 1 : (kg <= 20 ? 2 : 4);
    const int per = (kg + split - 1) / split;
#define GCGE_SPMM_LAUNCH(MC, SP) \
    launch_sell_spmm_v<VT, XT, MC, SP, DOT>(A, vals, X, ldx, Y, ldy, k, gcols, d_xmap, d_ymap, dpart, \
                                            d_count, ev0, ev1)
    if (split == 1) {
        if (per <= 4) GCGE_SPMM_LAUNCH(4, 1); else if (per <= 8) GCGE_SPMM_LAUNCH(8, 1); else GCGE_SPMM_LAUNCH(16, 1);
    } else if (split == 2) {
        if (per <= 4) GCGE_SPMM_LAUNCH(4, 2); else if (per <= 8) GCGE_SPMM_LAUNCH(8, 2); else GCGE_SPMM_LAUNCH(16, 2);
    } else {
        if (per <= 4) GCGE_SPMM_LAUNCH(4, 4); else if (per <= 8) GCGE_SPMM_LAUNCH(8, 4); else GCGE_SPMM_LAUNCH(16, 4);
    }
#undef GCGE_SPMM_LAUNCH
}

// ---- SpMM on diagonal storage ----------------------------------------------------------
// A wavefront owns a block of 62 consecutive rows: lane L (1..62) computes
// row r0 + L - 1, lanes 0 and 63 hold the rows just outside the block.  For a
// run of offsets {c-1, c, c+1} every lane loads x[r0 + L - 1 + c] (one
// coalesced access per column); lane L takes x[row + c - 1] and
// x[row + c + 1] from lanes L - 1 and L + 1 through DPP wavefront shifts.  A
// 27-point stencil thus issues 9 loads per column and row block instead of 27
// gathers, and reads no column indices.  The row's values stay in registers
// for all columns of the wave; SPLIT waves of a workgroup share a row block
// (each takes part of the columns, the values reach the later waves from L1).
// Products are added in increasing column order, as the CSR rows list them.
// Indices outside 0..n-1 are clamped: the matrix has no entry there (zero
// value), so the clamped load only keeps the access in bounds.
constexpr int kDiaRows = kWave - 2;

__device__ __forceinline__ int dpp_prev(int src)  // lane i <- lane i - 1
{
    return __builtin_amdgcn_update_dpp(0, src, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_next(int src)  // lane i <- lane i + 1
{
    return __builtin_amdgcn_update_dpp(0, src, 0x130, 0xf, 0xf, false);
}
template <bool NEXT>
__device__ __forceinline__ double lane_shift(double src)
{
    const long long s = __double_as_longlong(src);
    const int lo = (int)(s & 0xffffffffLL), hi = (int)(s >> 32);
    const int rl = NEXT ? dpp_next(lo) : dpp_prev(lo);
    const int rh = NEXT ? dpp_next(hi) : dpp_prev(hi);
    return __longlong_as_double(((long long)rh << 32) | (unsigned int)rl);
}
template <bool NEXT>
__device__ __forceinline__ float lane_shift(float src)
{
    const int i = __float_as_int(src);
    return __int_as_float(NEXT ? dpp_next(i) : dpp_prev(i));
}

// columns [c0, c0 + cn) of the wave's row block, cn <= CK
template <int CK, int DOT, typename XT>
__device__ __forceinline__ void dia_chunk(const XT (&v)[kDiaMaxGroups][3], const DiaDesc &dd,
                                          int n, int row, bool own, int lane, int rb, int nblocks,
                                          const XT *__restrict__ X, long ldx, XT *__restrict__ Y,
                                          long ldy, int k, int c0, int cn,
                                          const int *__restrict__ xmap,
                                          const int *__restrict__ ymap, double *__restrict__ dpart)
{
    const XT *xc[CK];
#pragma unroll
    for (int cc = 0; cc < CK; ++cc) {
        const int c = c0 + min(cc, cn - 1);
        xc[cc] = X + (long)(xmap ? xmap[c] : c) * ldx;
    }
    // all loads of the chunk first (9 * CK in flight), then the products
    XT xv[kDiaMaxGroups][CK];
#pragma unroll
    for (int g = 0; g < kDiaMaxGroups; ++g) {
        const int ic = min(max(row + (g < dd.ngroups ? dd.center[g] : 0), 0), n - 1);
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) xv[g][cc] = (g < dd.ngroups && cc < cn) ? xc[cc][ic] : XT(0);
    }
    XT acc[CK];
#pragma unroll
    for (int cc = 0; cc < CK; ++cc) acc[cc] = XT(0);
#pragma unroll
    for (int g = 0; g < kDiaMaxGroups; ++g) {
        if (g < dd.ngroups) {
            const bool lo = dd.slot[g][0] >= 0, mid = dd.slot[g][1] >= 0, hi = dd.slot[g][2] >= 0;
#pragma unroll
            for (int cc = 0; cc < CK; ++cc) {
                if (lo) acc[cc] += v[g][0] * lane_shift<false>(xv[g][cc]);
                if (mid) acc[cc] += v[g][1] * xv[g][cc];
                if (hi) acc[cc] += v[g][2] * lane_shift<true>(xv[g][cc]);
            }
        }
    }
    if (own) {
#pragma unroll
        for (int cc = 0; cc < CK; ++cc)
            if (cc < cn) Y[(long)(ymap ? ymap[c0 + cc] : c0 + cc) * ldy + row] = acc[cc];
    }
    if constexpr (DOT > 0) {
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) {
            if (cc < cn) {
                const double pv = own ? (double)xc[cc][row] : 0.0;
                const double pw = wave_sum(pv * (own ? (double)acc[cc] : 0.0));
                double pp = 0.0;
                if constexpr (DOT > 1) pp = wave_sum(pv * pv);
                if (lane == 0) {
                    dpart[(long)(c0 + cc) * nblocks + rb] = pw;
                    if constexpr (DOT > 1) dpart[(long)(k + c0 + cc) * nblocks + rb] = pp;
                }
            }
        }
    }
}

// Plane order (S > 0): the rows form planes of S rows, each cut into pb row
// blocks (the last one partial); workgroup (q, g) takes in-plane block q of
// the kBlocksPerGroup consecutive planes of plane group g, q varying fastest.
// The waves of a workgroup read the same rows of X for the runs one plane
// apart close together in time, and neighbouring in-plane blocks run
// alongside as in consecutive order.
struct DiaPlanes {
    int S, pb, planes, groups;
};

template <typename VT, typename XT, int CK, int SPLIT, int DOT>
__global__ __launch_bounds__(kBlock) void k_dia_spmm(int n, int nblocks, DiaDesc dd,
                                                    const VT *__restrict__ dval, long ldd,
                                                    const XT *__restrict__ X, long ldx,
                                                    XT *__restrict__ Y, long ldy, int k, int gcols,
                                                    const int *__restrict__ xmap,
                                                    const int *__restrict__ ymap,
                                                    double *__restrict__ dpart,
                                                    const int *__restrict__ count, DiaPlanes pm)
{
    constexpr int kBlocksPerGroup = kBlock / kWave / SPLIT;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    int rb, row0, in_plane = 0;
    if (pm.S == 0) {
        rb = blk * kBlocksPerGroup + wid / SPLIT;
        if (rb >= nblocks) return;
        row0 = rb * kDiaRows;
    } else {
        const int q = blk % pm.pb, z = (blk / pm.pb) * kBlocksPerGroup + wid / SPLIT;
        if (q >= pm.pb || z >= pm.planes) return;
        rb = z * pm.pb + q;
        in_plane = q * kDiaRows;
        row0 = z * pm.S + in_plane;
    }
    const int part = wid % SPLIT;
    const int kc = count ? min(k, *count) : k;  // columns listed on the device
    const int gb = blockIdx.y * gcols, ge = min(kc, gb + gcols);
    if (gb >= ge) return;
    const int per = (ge - gb + SPLIT - 1) / SPLIT;
    const int cb = gb + part * per, ce = min(ge, cb + per);
    if (cb >= ce) return;
    const int row = row0 + lane - 1;
    const bool own = lane >= 1 && lane <= kDiaRows && row < n &&
                     (pm.S == 0 || in_plane + lane - 1 < pm.S);
    XT v[kDiaMaxGroups][3];
#pragma unroll
    for (int g = 0; g < kDiaMaxGroups; ++g)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const int s = g < dd.ngroups ? dd.slot[g][p] : -1;
            v[g][p] = (s >= 0 && own) ? (XT)dval[(long)s * ldd + row] : XT(0);
        }
    for (int c0 = cb; c0 < ce; c0 += CK)
        dia_chunk<CK, DOT>(v, dd, n, row, own, lane, rb, nblocks, X, ldx, Y, ldy, k, c0,
                           min(CK, ce - c0), xmap, ymap, dpart);
}

// Plane order (automatic choice) for layouts with a plane stride of at least
// 8192 rows; at smaller strides consecutive row blocks already share X
// through L2.
inline int plane_rows(const Mat *A)
{
    if (A->dia_plane <= 0 || A->dia_order == 1) return 0;
    return (A->dia_order == 2 || A->dia_plane >= 8192) ? A->dia_plane : 0;
}
inline int dia_plane_blocks(int S) { return (S + kDiaRows - 1) / kDiaRows; }
inline int dia_blocks(const Mat *A)
{
    const int S = plane_rows(A);
    if (S) return A->nrows / S * dia_plane_blocks(S);
    return (A->nrows + kDiaRows - 1) / kDiaRows;
}

// Column groups (blockIdx.y) of gcols columns, SPLIT waves per row block, CK
// columns per register chunk.  GCGE_HIP_DIA_GROUP / GCGE_HIP_DIA_SPLIT
// override the group width and the split (for measurement).
template <int DOT, typename VT, typename XT>
void launch_dia_spmm(const Mat *A, const VT *vals, const XT *X, long ldx, XT *Y, long ldy, int k,
                     const int *d_xmap, const int *d_ymap, double *dpart, const int *d_count,
                     hipEvent_t ev0, hipEvent_t ev1)
{
    static const int forced_group = [] {
        const char *e = getenv("GCGE_HIP_DIA_GROUP");
        return e ? atoi(e) : 0;
    }();
    static const int forced_split = [] {
        const char *e = getenv("GCGE_HIP_DIA_SPLIT");
        return e ? atoi(e) : 0;
    }();
    static const int forced_ck = [] {
        const char *e = getenv("GCGE_HIP_DIA_CK");
        return e ? atoi(e) : 0;
    }();
    const int gcols = forced_group > 0 ? std::min(forced_group, k) : k;
    // measured on MI355X (Q1 64^3 and 128^3, k = 1..100): one wave per row
    // block and chunks of 2 float64 columns (18 loads in flight at four
    // waves per SIMD) beat wider chunks, column splits and column groups
    int split = forced_split;
    if (split != 1 && split != 2 && split != 4) split = 1;
    const int per = (gcols + split - 1) / split;
    const int ng = (k + gcols - 1) / gcols;
    const int nblocks = dia_blocks(A);
    const int bpg = kBlock / kWave / split;
    DiaPlanes pm{plane_rows(A), 0, 0, 0};
    int nb = (nblocks + bpg - 1) / bpg;
    if (pm.S) {
        pm.pb = dia_plane_blocks(pm.S);
        pm.planes = A->nrows / pm.S;
        pm.groups = (pm.planes + bpg - 1) / bpg;
        nb = pm.pb * pm.groups;
    }
    if (ng > 1) nb = (nb + kXcds - 1) / kXcds * kXcds;
    const long ldd = (long)A->nslices * kWave;
#define GCGE_DIA_LAUNCH(CK, SP)                                                                      \
    hipExtLaunchKernelGGL((k_dia_spmm<VT, XT, CK, SP, DOT>), dim3(nb, ng), dim3(kBlock), 0,          \
                          state().stream, ev0, ev1, 0, A->nrows, nblocks, A->dia, vals, ldd, X, ldx,  \
                          Y, ldy, k, gcols, d_xmap, d_ymap, dpart, d_count, pm)
    const int ck_auto = sizeof(XT) == 8 ? 2 : (per <= 2 ? 2 : 4);
    const int ckw = forced_ck == 2 || forced_ck == 4 || forced_ck == 8 ? forced_ck : ck_auto;
#define GCGE_DIA_CK(SP)                                     \
    if (ckw == 2) GCGE_DIA_LAUNCH(2, SP);                   \
    else if (ckw == 4) GCGE_DIA_LAUNCH(4, SP);              \
    else GCGE_DIA_LAUNCH(8, SP);
    if (split == 1) {
        GCGE_DIA_CK(1)
    } else if (split == 2) {
        GCGE_DIA_CK(2)
    } else {
        GCGE_DIA_CK(4)
    }
#undef GCGE_DIA_CK
#undef GCGE_DIA_LAUNCH
    HIP_CHECK(hipGetLastError());
}

// ---- SpMM on diagonal storage, plane tiles (wide column blocks) ------------------------
// For a diagonal layout that is a 27-point plane stencil (offsets a N + b S + d,
// a, b, d in {-1, 0, 1}; S rows per plane, N per line).  A workgroup of 256
// threads owns in-plane rows [q0, q0 + 256) of PZ consecutive planes; thread
// t computes in-plane row q0 + t of each of them and keeps their PZ x 27
// coefficients in registers for every column.  Per chunk of CK columns the
// PZ + 2 planes the tile reads (in-plane rows q0 - N - 1 .. q0 + 256 + N) are
// staged in LDS once, the CK values of a position side by side (one vector
// LDS access per stencil read); each staged value serves the stencil
// positions of up to 27 rows of the tile.  Two workgroups per CU overlap one
// chunk's global loads with the other's products.  Positions outside the grid
// are staged as zero.  Products are added in increasing column order
// (b, then a, then d), as the CSR rows list them and as k_dia_spmm adds
// them; diagonals the matrix lacks contribute zero products (finite X), so
// both kernels give the same sums.
// Workgroup b runs on XCD b % 8; XCD x takes a contiguous range of plane
// groups, in-plane tile by in-plane tile, so that the workgroups resident on
// an XCD at one time cover neighbouring groups and tiles: the planes two
// groups share and the in-plane halos come from the XCD's L2 (at 128^3 the
// L2 misses of a 100-column float64 product fell from 1.69 to 1.17 times the
// algorithmic read bytes).
constexpr int kTileRows = 256;
struct TileDesc {
    int S, N, planes, tq, groups, nparts;  // strides, planes, in-plane tiles, plane groups,
                                           // partial sums per column (DOT)
    int slot[27];
};

template <typename XT, int CK>
struct alignas(sizeof(XT) * CK) TilePack {
    XT v[CK];
};

// Workgroup barrier for LDS hand-offs only: the wave's LDS accesses are
// complete (lgkmcnt 0) before s_barrier, and the empty asm statements keep the
// compiler from moving memory accesses across it.  Unlike __syncthreads() it
// does not wait for the wave's outstanding global loads and stores, so a
// chunk's output stores and the next chunk's loads stay in flight.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // vmcnt and expcnt unconstrained, lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// NF: the six face diagonals (offsets +-1, +-N, +-S) are absent (the
// stiffness matrix of trilinear elements: its face couplings vanish for any
// element coefficients); their coefficients take no registers and their
// products are skipped (a zero coefficient adds an exact zero to a sum that
// starts at +0, so the sums are unchanged).
__host__ __device__ constexpr bool tile_face(int j)
{
    return (j / 9 != 1) + ((j / 3) % 3 != 1) + (j % 3 != 1) == 1;
}

template <typename VT, typename XT, int PZ, int CK, int J, int DOT, bool NF>
__global__ __launch_bounds__(kTileRows, 2) void k_dia_tile(TileDesc td, const VT *__restrict__ dval,
                                                          long ldd, const XT *__restrict__ X, long ldx,
                                                          XT *__restrict__ Y, long ldy, int k,
                                                          const int *__restrict__ xmap,
                                                          const int *__restrict__ ymap,
                                                          double *__restrict__ dpart,
                                                          const int *__restrict__ count)
{
    using Pack = TilePack<XT, CK>;
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_smem[];
    Pack *lds = reinterpret_cast<Pack *>(tile_smem);
    const int xcd = blockIdx.x % kXcds, loc = blockIdx.x / kXcds;
    // XCD x: the contiguous plane groups [g0, g0 + gx), tile-major, so that
    // adjacent groups of one in-plane tile run side by side and share their
    // halo planes through the XCD's L2
    const int gq = td.groups / kXcds, gr = td.groups % kXcds;
    const int gx = gq + (xcd < gr ? 1 : 0), g0 = xcd * gq + min(xcd, gr);
    if (gx == 0) return;
    const int qt = loc / gx, zg = g0 + loc % gx;
    if (qt >= td.tq) return;
    const int W = kTileRows + 2 * td.N + 2;
    const int t = threadIdx.x;
    const int q0 = qt * kTileRows, z0 = zg * PZ;
    const int q = q0 + t;
    const long S = td.S;
    bool own[PZ];
    XT v[PZ][27];
#pragma unroll
    for (int p = 0; p < PZ; ++p) {
        own[p] = q < td.S && z0 + p < td.planes;
        const long row = (z0 + p) * S + q;
#pragma unroll
        for (int j = 0; j < 27; ++j) {
            if (NF && tile_face(j)) continue;
            const int s = td.slot[j];
            v[p][j] = (s >= 0 && own[p]) ? (XT)dval[(long)s * ldd + row] : XT(0);
        }
    }
    const int kc = count ? min(k, *count) : k;
    // staged positions: plane pl (grid plane z0 - 1 + pl), in-plane offset
    // o = t + 256 jj (in-plane row q0 - N - 1 + o)
    Pack stage[PZ + 2][J];
    auto fetch = [&](int c0) {
        const XT *col[CK];
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) {
            const int c = min(c0 + cc, kc - 1);
            col[cc] = X + (long)(xmap ? xmap[c] : c) * ldx;
        }
#pragma unroll
        for (int pl = 0; pl < PZ + 2; ++pl) {
            const int z = z0 - 1 + pl;
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const int o = t + kTileRows * jj;
                const int ip = q0 - td.N - 1 + o;
                const bool ok = o < W && z >= 0 && z < td.planes && ip >= 0 && ip < td.S;
                const long g = ok ? (long)z * S + ip : 0;
#pragma unroll
                for (int cc = 0; cc < CK; ++cc)
                    stage[pl][jj].v[cc] = (ok && c0 + cc < kc) ? col[cc][g] : XT(0);
            }
        }
    };
    for (int c0 = 0; c0 < kc; c0 += CK) {
        fetch(c0);
        lds_barrier();  // the previous chunk's products are done with the LDS
#pragma unroll
        for (int pl = 0; pl < PZ + 2; ++pl)
#pragma unroll
            for (int jj = 0; jj < J; ++jj) {
                const int o = t + kTileRows * jj;
                if (o < W) lds[pl * W + o] = stage[pl][jj];
            }
        lds_barrier();
        XT acc[PZ][CK];
#pragma unroll
        for (int p = 0; p < PZ; ++p)
#pragma unroll
            for (int cc = 0; cc < CK; ++cc) acc[p][cc] = XT(0);
        const int base = t + td.N + 1;
#pragma unroll
        for (int pl = 0; pl < PZ + 2; ++pl) {
            // the nine staged values of this plane, then their products (a
            // diagonal the matrix lacks has zero coefficients)
            Pack xv[3][3];
#pragma unroll
            for (int a = -1; a <= 1; ++a)
#pragma unroll
                for (int d = -1; d <= 1; ++d) xv[a + 1][d + 1] = lds[pl * W + base + a * td.N + d];
#pragma unroll
            for (int a = -1; a <= 1; ++a)
#pragma unroll
                for (int d = -1; d <= 1; ++d)
#pragma unroll
                    for (int b = -1; b <= 1; ++b) {
                        const int p = pl - 1 - b;
                        if (p < 0 || p >= PZ) continue;
                        const int j = 9 * (b + 1) + 3 * (a + 1) + d + 1;
                        if (NF && tile_face(j)) continue;
#pragma unroll
                        for (int cc = 0; cc < CK; ++cc) acc[p][cc] += v[p][j] * xv[a + 1][d + 1].v[cc];
                    }
            // the sums as they stand after this plane (an empty asm that reads
            // and writes them): the products are not sunk to the stores, so
            // only this plane's nine vectors are live
#pragma unroll
            for (int p = 0; p < PZ; ++p)
#pragma unroll
                for (int cc = 0; cc < CK; ++cc) asm volatile("" : "+v"(acc[p][cc]));
        }
        const int cn = min(CK, kc - c0);
#pragma unroll
        for (int cc = 0; cc < CK; ++cc) {
            if (cc >= cn) continue;
            XT *y = Y + (long)(ymap ? ymap[c0 + cc] : c0 + cc) * ldy;
#pragma unroll
            for (int p = 0; p < PZ; ++p)
                if (own[p]) y[(z0 + p) * S + q] = acc[p][cc];
        }
        if constexpr (DOT > 0) {
            const int part = (zg * td.tq + qt) * (kTileRows / kWave) + t / kWave;
            // this thread's own staged values, one 16-byte read per plane
            Pack own_x[PZ];
#pragma unroll
            for (int p = 0; p < PZ; ++p) own_x[p] = lds[(p + 1) * W + base];
#pragma unroll
            for (int cc = 0; cc < CK; ++cc) {
                if (cc >= cn) continue;
                double pw = 0.0, pp = 0.0;
#pragma unroll
                for (int p = 0; p < PZ; ++p) {
                    const double pv = own[p] ? (double)own_x[p].v[cc] : 0.0;
                    pw += pv * (own[p] ? (double)acc[p][cc] : 0.0);
                    if constexpr (DOT > 1) pp += pv * pv;
                }
                pw = wave_sum(pw);
                if constexpr (DOT > 1) pp = wave_sum(pp);
                if (t % kWave == 0) {
                    dpart[(long)(c0 + cc) * td.nparts + part] = pw;
                    if constexpr (DOT > 1) dpart[(long)(k + c0 + cc) * td.nparts + part] = pp;
                }
            }
        }
    }
}

// Plane tiles apply to matrices with a complete 27-point table from
// kTileMinCols columns (automatic order), or always (order 3).  Planes per
// tile: 4 for float32 values, 2 for float64 (register budget of the
// coefficients); GCGE_HIP_TILE_PZ and GCGE_HIP_TILE_MIN override (for
// measurement).
constexpr int kTileMinCols = 8;
inline int tile_min_cols()
{
    static const int v = [] {
        const char *e = getenv("GCGE_HIP_TILE_MIN");
        return e ? atoi(e) : kTileMinCols;
    }();
    return v;
}
template <typename XT>
inline int tile_pz()
{
    if (sizeof(XT) == 8) return 2;
    static const int v = [] {
        const char *e = getenv("GCGE_HIP_TILE_PZ");
        return (e && atoi(e) == 2) ? 2 : 4;
    }();
    return v;
}
template <typename XT>
inline bool use_tiles(const Mat *A, int k)
{
    if (!A->use_dia || !A->dia_tile_ok || A->dia_plane <= 0 || A->dia_order == 1 || A->dia_order == 2)
        return false;
    // staged planes x positions x 16 bytes within 64 KiB of LDS
    const long W = kTileRows + 2L * A->dia_N + 2;
    if (W > 3 * kTileRows || (tile_pz<XT>() + 2) * W * 16 > 65536) return false;
    return A->dia_order == 3 || k >= tile_min_cols();
}
template <typename XT>
inline TileDesc tile_desc(const Mat *A)
{
    TileDesc td{};
    td.S = A->dia_plane;
    td.N = A->dia_N;
    td.planes = A->nrows / A->dia_plane;
    td.tq = (td.S + kTileRows - 1) / kTileRows;
    td.groups = (td.planes + tile_pz<XT>() - 1) / tile_pz<XT>();
    td.nparts = td.groups * td.tq * (kTileRows / kWave);
    for (int j = 0; j < 27; ++j) td.slot[j] = A->tile27[j];
    return td;
}
template <typename XT>
inline int tile_parts(const Mat *A) { return tile_desc<XT>(A).nparts; }

template <int DOT, typename VT, typename XT>
void launch_dia_tile(const Mat *A, const VT *vals, const XT *X, long ldx, XT *Y, long ldy, int k,
                     const int *d_xmap, const int *d_ymap, double *dpart, const int *d_count,
                     hipEvent_t ev0, hipEvent_t ev1)
{
    constexpr int CK = 16 / sizeof(XT);
    const TileDesc td = tile_desc<XT>(A);
    const int W = kTileRows + 2 * td.N + 2;
    const int J = (W + kTileRows - 1) / kTileRows;
    const int pz = tile_pz<XT>();
    const int rounds = (td.groups + kXcds - 1) / kXcds;
    const dim3 grid(td.tq * rounds * kXcds);
    const size_t lds = (size_t)(pz + 2) * W * 16;
    const long ldd = (long)A->nslices * kWave;
    bool nf = true;
    for (int j = 0; j < 27; ++j)
        if (tile_face(j) && td.slot[j] >= 0) nf = false;
#define GCGE_TILE_LAUNCH(PZ, JJ)                                                                    \
    do {                                                                                            \
        if (nf)                                                                                     \
            hipExtLaunchKernelGGL((k_dia_tile<VT, XT, PZ, CK, JJ, DOT, true>), grid, dim3(kTileRows), \
                                  lds, state().stream, ev0, ev1, 0, td, vals, ldd, X, ldx, Y, ldy, k, \
                                  d_xmap, d_ymap, dpart, d_count);                                  \
        else                                                                                        \
            hipExtLaunchKernelGGL((k_dia_tile<VT, XT, PZ, CK, JJ, DOT, false>), grid,               \
                                  dim3(kTileRows), lds, state().stream, ev0, ev1, 0, td, vals, ldd, X, \
                                  ldx, Y, ldy, k, d_xmap, d_ymap, dpart, d_count);                  \
    } while (0)
    if constexpr (sizeof(XT) == 4) {
        if (pz == 4) {
            if (J <= 2) GCGE_TILE_LAUNCH(4, 2);
            else GCGE_TILE_LAUNCH(4, 3);
            HIP_CHECK(hipGetLastError());
            return;
        }
    }
    if (J <= 2) GCGE_TILE_LAUNCH(2, 2);
    else GCGE_TILE_LAUNCH(2, 3);
#undef GCGE_TILE_LAUNCH
    HIP_CHECK(hipGetLastError());
}

// ---- SpMM with whole-row gathers (matrices with scattered columns) ------------------
// The column block X (column-major, optional column map) is first copied to a
// row-major block XT (row i: the k values of row i, row stride kp), through
// 64 x 64 LDS tiles so that both sides are coalesced.  The product then runs
// with lanes along the columns: a wavefront owns 16 rows of a sliced-ELLPACK
// slice and 64 columns (grid y: column passes); for every entry of those rows
// the column index and value are wave-uniform, and the gather of the matching
// row of XT is one contiguous access of 64 values, where the column-major form
// takes a separate cache line per element.  Each lane writes its 16 outputs
// of a column as one contiguous run.  DOT: p.w (and p.p) per column, summed
// over the workgroup's rows through LDS, one partial per workgroup.
constexpr int kRowsPerWave = 16;

// (only the first min(k, *count) columns: entries of a device column list
// beyond its length are not defined)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cols_to_rows(int n, int k, const T *__restrict__ X, long ldx,
                                                        const int *__restrict__ xmap,
                                                        T *__restrict__ XT, int kp,
                                                        const int *__restrict__ count)
{
    __shared__ T tile[64][65];
    const int kc = count ? min(k, *count) : k;
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    if (c0 >= kc) return;  // the whole workgroup
    const int t = threadIdx.x % 64, g = threadIdx.x / 64;
    for (int cc = g; cc < 64; cc += kBlock / 64) {
        const int c = c0 + cc, r = r0 + t;
        tile[cc][t] = (c < kc && r < n) ? X[(long)(xmap ? xmap[c] : c) * ldx + r] : T(0);
    }
    __syncthreads();
    for (int rr = g; rr < 64; rr += kBlock / 64) {
        const int r = r0 + rr, c = c0 + t;
        if (r < n && c < kc) XT[(long)r * kp + c] = tile[t][rr];
    }
}

template <typename VT, typename XT, int DOT>
__global__ __launch_bounds__(kBlock) void k_sell_rows_spmm(int n, int nslices, int slices_per_block,
                                                          const int *__restrict__ sptr,
                                                          const int *__restrict__ sw,
                                                          const int *__restrict__ scol,
                                                          const VT *__restrict__ sval,
                                                          const XT *__restrict__ XTm, int kp,
                                                          XT *__restrict__ Y, long ldy, int k,
                                                          const int *__restrict__ ymap,
                                                          double *__restrict__ dpart, int nparts,
                                                          const int *__restrict__ count)
{
    __shared__ double red[DOT > 0 ? 2 * kBlock : 1];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int kc = count ? min(k, *count) : k;
    const int c0 = blockIdx.y * kWave;
    if (c0 >= kc) return;  // the whole workgroup: same column pass
    const int c = c0 + lane;
    const bool live = c < kc;
    const int s0 = blockIdx.x * slices_per_block, s1 = min(nslices, s0 + slices_per_block);
    double pw = 0.0, pp = 0.0;
    for (int sl = s0; sl < s1; ++sl) {
        const int w = sw[sl];
        const int base = sptr[sl] + wid * kRowsPerWave;
        XT acc[kRowsPerWave];
#pragma unroll
        for (int i = 0; i < kRowsPerWave; ++i) acc[i] = XT(0);
        for (int j = 0; j < w; ++j) {
            const int e0 = base + j * kWave;
            XT xv[kRowsPerWave];
#pragma unroll
            for (int i = 0; i < kRowsPerWave; ++i) {
                const int col = scol[e0 + i];
                xv[i] = live ? XTm[(long)col * kp + c] : XT(0);
            }
#pragma unroll
            for (int i = 0; i < kRowsPerWave; ++i) acc[i] += (XT)sval[e0 + i] * xv[i];
        }
        const int rbase = sl * kWave + wid * kRowsPerWave;
        if (live) {
            XT *y = Y + (long)(ymap ? ymap[c] : c) * ldy;
#pragma unroll
            for (int i = 0; i < kRowsPerWave; ++i)
                if (rbase + i < n) y[rbase + i] = acc[i];
        }
        if constexpr (DOT > 0) {
#pragma unroll
            for (int i = 0; i < kRowsPerWave; ++i) {
                const int r = rbase + i;
                const double pv = (live && r < n) ? (double)XTm[(long)r * kp + c] : 0.0;
                pw += pv * (double)acc[i];
                if constexpr (DOT > 1) pp += pv * pv;
            }
        }
    }
    if constexpr (DOT > 0) {
        red[threadIdx.x] = pw;
        red[kBlock + threadIdx.x] = pp;
        __syncthreads();
        if (wid == 0 && live) {
            double a = 0.0, b = 0.0;
            for (int q = 0; q < kBlock / kWave; ++q) {
                a += red[q * kWave + lane];
                b += red[kBlock + q * kWave + lane];
            }
            dpart[(long)c * nparts + blockIdx.x] = a;
            if constexpr (DOT > 1) dpart[(long)(k + c) * nparts + blockIdx.x] = b;
        }
    }
}

// The same product with the column passes of a row group in one workgroup:
// wave w takes columns 64 w .. 64 w + 63 (k <= 256) of every row group of the
// workgroup's slices, so the four 512-byte pieces of a gathered row of XT
// are read by the four waves close together in time, instead of by
// workgroups of different passes far apart.  Each row is summed as in
// k_sell_rows_spmm; the p.w partial of a column comes from the one wave that
// owns it.
template <typename VT, typename XT, int DOT>
__global__ __launch_bounds__(kBlock) void k_sell_rows_spmm_passes(
    int n, int nslices, int slices_per_block, const int *__restrict__ sptr,
    const int *__restrict__ sw, const int *__restrict__ scol, const VT *__restrict__ sval,
    const XT *__restrict__ XTm, int kp, XT *__restrict__ Y, long ldy, int k,
    const int *__restrict__ ymap, double *__restrict__ dpart, int nparts,
    const int *__restrict__ count)
{
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int kc = count ? min(k, *count) : k;
    const int c0 = wid * kWave;
    if (c0 >= kc) return;  // the whole wave: one column pass
    const int c = c0 + lane;
    const bool live = c < kc;
    XT *y = live ? Y + (long)(ymap ? ymap[c] : c) * ldy : nullptr;
    const int s0 = blockIdx.x * slices_per_block, s1 = min(nslices, s0 + slices_per_block);
    double pw = 0.0, pp = 0.0;
    for (int sl = s0; sl < s1; ++sl) {
        const int w = sw[sl];
        for (int g = 0; g < kWave / kRowsPerWave; ++g) {
            const int base = sptr[sl] + g * kRowsPerWave;
            XT acc[kRowsPerWave];
#pragma unroll
            for (int i = 0; i < kRowsPerWave; ++i) acc[i] = XT(0);
            for (int j = 0; j < w; ++j) {
                const int e0 = base + j * kWave;
                XT xv[kRowsPerWave];
#pragma unroll
                for (int i = 0; i < kRowsPerWave; ++i) {
                    const int col = scol[e0 + i];
                    xv[i] = live ? XTm[(long)col * kp + c] : XT(0);
                }
#pragma unroll
                for (int i = 0; i < kRowsPerWave; ++i) acc[i] += (XT)sval[e0 + i] * xv[i];
            }
            const int rbase = sl * kWave + g * kRowsPerWave;
            if (live) {
#pragma unroll
                for (int i = 0; i < kRowsPerWave; ++i)
                    if (rbase + i < n) y[rbase + i] = acc[i];
            }
            if constexpr (DOT > 0) {
#pragma unroll
                for (int i = 0; i < kRowsPerWave; ++i) {
                    const int r = rbase + i;
                    const double pv = (live && r < n) ? (double)XTm[(long)r * kp + c] : 0.0;
                    pw += pv * (double)acc[i];
                    if constexpr (DOT > 1) pp += pv * pv;
                }
            }
        }
    }
    if constexpr (DOT > 0) {
        if (live) {
            dpart[(long)c * nparts + blockIdx.x] = pw;
            if constexpr (DOT > 1) dpart[(long)(k + c) * nparts + blockIdx.x] = pp;
        }
    }
}

// slices per workgroup of the row-gather product: about 2048 workgroups per
// column pass
inline int rows_slices_per_block(const Mat *A) { return std::max(1, (A->nslices + 2047) / 2048); }
inline int rows_parts(const Mat *A)
{
    const int spb = rows_slices_per_block(A);
    return (A->nslices + spb - 1) / spb;
}

template <int DOT, typename VT, typename XT>
void launch_rows_spmm(const Mat *A, const VT *vals, const XT *X, long ldx, XT *Y, long ldy, int k,
                      const int *d_xmap, const int *d_ymap, double *dpart, const int *d_count,
                      hipEvent_t ev0, hipEvent_t ev1)
{
    const int n = A->nrows;
    const int kp = (k + 7) / 8 * 8;
    XT *XTm = (XT *)rowblock_buffer((size_t)A->ncols * kp * sizeof(XT));
    hipLaunchKernelGGL((k_cols_to_rows<XT>), dim3((A->ncols + 63) / 64, (k + 63) / 64), dim3(kBlock), 0,
                       state().stream, A->ncols, k, X, ldx, d_xmap, XTm, kp, d_count);
    const int spb = rows_slices_per_block(A);
    const int nparts = rows_parts(A);
    static const int passes_mode = [] {
        const char *e = getenv("GCGE_HIP_ROWS_PASSES");  // "0": one pass per workgroup (measurement)
        return e ? atoi(e) : 1;
    }();
    const int npass = (k + kWave - 1) / kWave;
    if (passes_mode != 0 && npass >= 2 && npass <= kBlock / kWave)
        hipExtLaunchKernelGGL((k_sell_rows_spmm_passes<VT, XT, DOT>), dim3(nparts), dim3(kBlock), 0,
                              state().stream, ev0, ev1, 0, n, A->nslices, spb, A->d_slice_ptr,
                              A->d_slice_w, A->d_col, vals, XTm, kp, Y, ldy, k, d_ymap, dpart,
                              nparts, d_count);
    else
        hipExtLaunchKernelGGL((k_sell_rows_spmm<VT, XT, DOT>), dim3(nparts, npass), dim3(kBlock), 0,
                              state().stream, ev0, ev1, 0, n, A->nslices, spb, A->d_slice_ptr,
                              A->d_slice_w, A->d_col, vals, XTm, kp, Y, ldy, k, d_ymap, dpart,
                              nparts, d_count);
    HIP_CHECK(hipGetLastError());
}

// the row-gather form from this many columns (fewer leave most lanes idle)
constexpr int kRowsMinCols = 16;

__global__ __launch_bounds__(kBlock) void k_to_float(size_t len, const double *__restrict__ src,
                                                    float *__restrict__ dst)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) dst[i] = (float)src[i];
}

// X(:, c) += D(:, c) for c < k (float32 correction into a float64 block)
__global__ __launch_bounds__(kBlock) void k_add_float(int n, const float *__restrict__ D, long ldd,
                                                     double *__restrict__ X, long ldx)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row < n) X[(long)c * ldx + row] += (double)D[(long)c * ldd + row];
}

// X(:, c) = 0 for c < k (float32 block)
__global__ __launch_bounds__(kBlock) void k_zero_float(int n, float *__restrict__ X, long ldx)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row < n) X[(long)blockIdx.y * ldx + row] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_to_double(size_t len, const float *__restrict__ src,
                                                     double *__restrict__ dst)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) dst[i] = (double)src[i];
}

// ---- Gram -----------------------------------------------------------------------
constexpr int kGramRows = 32;

template <int TM, int TN>
__global__ __launch_bounds__(kBlock) void k_gram(int n, int rows_per_block,
                                                const double *__restrict__ V, long ldv, int m,
                                                const double *__restrict__ W, long ldw, int k,
                                                double *__restrict__ partial)
{
    constexpr int RM = TM / 16, RN = TN / 16;
    __shared__ double Vs[kGramRows][TM + 1];
    __shared__ double Ws[kGramRows][TN + 1];
    const int bx = blockIdx.x;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.z * TN;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int r_begin = bx * rows_per_block;
    const int r_end = min(n, r_begin + rows_per_block);
    double acc[RM][RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int b = 0; b < RN; ++b) acc[a][b] = 0.0;

    for (int r0 = r_begin; r0 < r_end; r0 += kGramRows) {
        // stage V(r0:r0+32, i0:i0+TM) and W(...) transposed into LDS; lanes run
        // along rows so each column segment is one coalesced read
        for (int e = threadIdx.x; e < kGramRows * TM; e += kBlock) {
            const int r = e % kGramRows, c = e / kGramRows;
            const int gr = r0 + r, gc = i0 + c;
            Vs[r][c] = (gr < r_end && gc < m) ? V[(long)gc * ldv + gr] : 0.0;
        }
        for (int e = threadIdx.x; e < kGramRows * TN; e += kBlock) {
            const int r = e % kGramRows, c = e / kGramRows;
            const int gr = r0 + r, gc = j0 + c;
            Ws[r][c] = (gr < r_end && gc < k) ? W[(long)gc * ldw + gr] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < kGramRows; ++r) {
            double va[RM], wb[RN];
#pragma unroll
            for (int a = 0; a < RM; ++a) va[a] = Vs[r][tx + 16 * a];
#pragma unroll
            for (int b = 0; b < RN; ++b) wb[b] = Ws[r][ty + 16 * b];
#pragma unroll
            for (int a = 0; a < RM; ++a)
#pragma unroll
                for (int b = 0; b < RN; ++b) acc[a][b] += va[a] * wb[b];
        }
        __syncthreads();
    }
    double *out = partial + (long)bx * m * k;
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int b = 0; b < RN; ++b) {
            const int i = i0 + tx + 16 * a, j = j0 + ty + 16 * b;
            if (i < m && j < k) out[(long)j * m + i] = acc[a][b];
        }
}

// Small Gram (m <= MR, k <= KR, MR, KR <= 8): every thread walks rows with
// stride 256 and keeps the whole m x k block in registers; the block's sums
// go through wave shuffles and LDS to partial[bx * m * k + j * m + i].
template <int MR, int KR>
__global__ __launch_bounds__(kBlock) void k_gram_small(int n, int rows_per_block,
                                                      const double *__restrict__ V, long ldv,
                                                      int m, const double *__restrict__ W,
                                                      long ldw, int k,
                                                      double *__restrict__ partial)
{
    __shared__ double lds[kBlock / kWave][MR * KR];
    const int bx = blockIdx.x;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double acc[MR][KR];
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int b = 0; b < KR; ++b) acc[a][b] = 0.0;
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
        double v[MR], w[KR];
#pragma unroll
        for (int a = 0; a < MR; ++a) v[a] = a < m ? V[(long)a * ldv + r] : 0.0;
#pragma unroll
        for (int b = 0; b < KR; ++b) w[b] = b < k ? W[(long)b * ldw + r] : 0.0;
#pragma unroll
        for (int a = 0; a < MR; ++a)
#pragma unroll
            for (int b = 0; b < KR; ++b) acc[a][b] += v[a] * w[b];
    }
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int b = 0; b < KR; ++b) {
            const double t = wave_sum(acc[a][b]);
            if (lane == 0) lds[wid][a * KR + b] = t;
        }
    __syncthreads();
    const int e = threadIdx.x;
    if (e < MR * KR) {
        const int a = e / KR, b = e % KR;
        if (a < m && b < k) {
            double t = 0.0;
            for (int w = 0; w < kBlock / kWave; ++w) t += lds[w][e];
            partial[(long)bx * m * k + (long)b * m + a] = t;
        }
    }
}

// Gram for m up to a few hundred and k <= 16 (the W orthonormalisation and
// Rayleigh-Ritz blocks): wave w of a workgroup owns V columns
// [(blockIdx.y * 4 + w) * MR, +MR) and every W column (KR >= k); the four
// waves walk the same rows, so each W row is fetched from memory once per
// workgroup, and V is read once.  Lanes take consecutive rows.  Each wave
// reduces its MR x KR block over lanes and writes
// partial[bx * m * k + j * m + i].
template <int MR, int KR>
__global__ __launch_bounds__(kBlock) void k_gram_cols(int n, int rows_per_block,
                                                     const double *__restrict__ V, long ldv,
                                                     int m, const double *__restrict__ W,
                                                     long ldw, int k,
                                                     double *__restrict__ partial)
{
    const int bx = blockIdx.x;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int i0 = (blockIdx.y * (kBlock / kWave) + wid) * MR;
    if (i0 >= m) return;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double acc[MR][KR];
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int b = 0; b < KR; ++b) acc[a][b] = 0.0;
    const double *vcol[MR];
#pragma unroll
    for (int a = 0; a < MR; ++a) vcol[a] = V + (long)min(i0 + a, m - 1) * ldv;
    // two rows per lane per step: the loads of both are in flight together
    int r = r0 + lane;
    for (; r + kWave < r1; r += 2 * kWave) {
        double v[2][MR], w[2][KR];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int a = 0; a < MR; ++a) v[h][a] = (i0 + a < m) ? vcol[a][r + h * kWave] : 0.0;
#pragma unroll
            for (int b = 0; b < KR; ++b) w[h][b] = b < k ? W[(long)b * ldw + r + h * kWave] : 0.0;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int a = 0; a < MR; ++a)
#pragma unroll
                for (int b = 0; b < KR; ++b) acc[a][b] += v[h][a] * w[h][b];
    }
    if (r < r1) {
        double v[MR], w[KR];
#pragma unroll
        for (int a = 0; a < MR; ++a) v[a] = (i0 + a < m) ? vcol[a][r] : 0.0;
#pragma unroll
        for (int b = 0; b < KR; ++b) w[b] = b < k ? W[(long)b * ldw + r] : 0.0;
#pragma unroll
        for (int a = 0; a < MR; ++a)
#pragma unroll
            for (int b = 0; b < KR; ++b) acc[a][b] += v[a] * w[b];
    }
    double *out = partial + (long)bx * m * k;
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int b = 0; b < KR; ++b) {
            const double t = wave_sum(acc[a][b]);
            if (lane == 0 && i0 + a < m && b < k) out[(long)b * m + i0 + a] = t;
        }
}

// V^T W on the matrix cores (v_mfma_f64_16x16x4f64).  Workgroup (bx, by, bz)
// takes rows [bx * rows_per_block, ...), V columns [128 by, +128) and W
// columns [128 bz, +128).  Per stage, 16 rows of those columns go to LDS
// (row stride MP + KP + 1 doubles: V in the first MP columns, W after),
// 16 consecutive threads reading one column's 16 consecutive rows; every
// thread issues all its loads of a stage at once, and the next stage's loads
// are in flight while the matrix cores work on the current one.  The
// (16 x 16) output tiles are shared round-robin by the NW waves; each MFMA
// takes one element of V and one of W per lane from LDS (A: lane l holds
// Vt[l % 16][l / 16] of the 16 x 4 block, B: W[l / 16][l % 16]), and lane l,
// component v of the accumulator holds D[l / 16 + 4 v][l % 16] (layout checked
// against the oracle Gram on the device).  Partials as k_gram.
typedef double gcge_double4 __attribute__((ext_vector_type(4)));
constexpr int kGmRows = 16;
constexpr int kGmCols = 128;

template <int PW, int EPT, int NW>
__global__ __launch_bounds__(64 * NW) void k_gram_mfma(int n, int rows_per_block,
                                                     const double *__restrict__ V, long ldv,
                                                     int m, const double *__restrict__ W,
                                                     long ldw, int k,
                                                     double *__restrict__ partial, int bcv, int bcw)
{
    extern __shared__ double lds_gm[];
    const int i_base = blockIdx.y * bcv, j_base = blockIdx.z * bcw;
    const int mb = min(bcv, m - i_base), kb = min(bcw, k - j_base);
    const int ti = (mb + 15) / 16, tj = (kb + 15) / 16;
    const int MP = ti * 16, KP = tj * 16, npairs = ti * tj;
    const int ld = MP + KP + 1;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    const int ncols = MP + KP;
    // this thread's stage elements: row q, columns c0 + kColStep t
    constexpr int kColStep = 64 * NW / kGmRows;
    const int q = threadIdx.x % kGmRows, c0 = threadIdx.x / kGmRows;
    double buf[EPT];
    auto load_stage = [&](int rs) {
        const int r = rs + q;
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            const int c = c0 + kColStep * t;
            double v = 0.0;
            if (r < r1) {
                if (c < MP) {
                    if (c < mb) v = V[(long)(i_base + c) * ldv + r];
                } else if (c - MP < kb) {
                    v = W[(long)(j_base + c - MP) * ldw + r];
                }
            }
            buf[t] = v;
        }
    };
    gcge_double4 acc[PW];
#pragma unroll
    for (int p = 0; p < PW; ++p) acc[p] = gcge_double4{0.0, 0.0, 0.0, 0.0};
    if (r0 < r1) load_stage(r0);
    for (int rs = r0; rs < r1; rs += kGmRows) {
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            const int c = c0 + kColStep * t;
            if (c < ncols) lds_gm[q * ld + c] = buf[t];
        }
        __syncthreads();
        if (rs + kGmRows < r1) load_stage(rs + kGmRows);
#pragma unroll
        for (int kk = 0; kk < kGmRows / 4; ++kk) {
            const int qa = kk * 4 + lane / 16;
#pragma unroll
            for (int p = 0; p < PW; ++p) {
                const int pair = wid + NW * p;
                if (pair < npairs) {
                    const double a = lds_gm[qa * ld + (pair / tj) * 16 + lane % 16];
                    const double b = lds_gm[qa * ld + MP + (pair % tj) * 16 + lane % 16];
                    acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[p], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    double *out = partial + (long)blockIdx.x * m * k;
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int pair = wid + NW * p;
        if (pair < npairs) {
            const int j = j_base + (pair % tj) * 16 + lane % 16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i_base + (pair / tj) * 16 + lane / 16 + 4 * v;
                if (i < i_base + mb && j < j_base + kb) out[(long)j * m + i] = acc[p][v];
            }
        }
    }
}

template <int P, int NW>
void launch_gram_mfma_e(int ept, dim3 grid, size_t lds, int n, int rows, const double *V, long ldv,
                        int m, const double *W, long ldw, int k, double *partial, int bcv, int bcw)
{
    hipStream_t st = state().stream;
    const dim3 block(64 * NW);
#define GCGE_GM_LAUNCH(E)                                                                             \
    hipLaunchKernelGGL((k_gram_mfma<P, E, NW>), grid, block, lds, st, n, rows, V, ldv, m, W, ldw, k, \
                       partial, bcv, bcw)
    if (ept <= 2) GCGE_GM_LAUNCH(2);
    else if (ept <= 4) GCGE_GM_LAUNCH(4);
    else if (ept <= 8) GCGE_GM_LAUNCH(8);
    else GCGE_GM_LAUNCH(16);
#undef GCGE_GM_LAUNCH
}

// tile pairs per wave: 1, 2 or 8 (the 4-pair instance keeps its
// accumulators in VGPRs and drops to one or two waves per SIMD; the 8-pair
// one does not); NW = 4 or 8 waves per workgroup; output blocks of bcv x bcw
void launch_gram_mfma(int nw, int pw, int ept, dim3 grid, size_t lds, int n, int rows,
                      const double *V, long ldv, int m, const double *W, long ldw, int k,
                      double *partial, int bcv, int bcw)
{
#define GCGE_GM_ARGS ept, grid, lds, n, rows, V, ldv, m, W, ldw, k, partial, bcv, bcw
    if (nw == 8) {
        if (pw <= 1) launch_gram_mfma_e<1, 8>(GCGE_GM_ARGS);
        else if (pw <= 2) launch_gram_mfma_e<2, 8>(GCGE_GM_ARGS);
        else launch_gram_mfma_e<8, 8>(GCGE_GM_ARGS);
    } else {
        if (pw <= 1) launch_gram_mfma_e<1, 4>(GCGE_GM_ARGS);
        else if (pw <= 2) launch_gram_mfma_e<2, 4>(GCGE_GM_ARGS);
        else launch_gram_mfma_e<8, 4>(GCGE_GM_ARGS);
    }
#undef GCGE_GM_ARGS
}

// V^T W on the matrix cores for the wide projection blocks (64 <= m <= 128
// MI, 64 <= k <= 16 NJ): one workgroup of eight waves holds the whole m x k
// result of its row range, so V and W are read from HBM once.  Wave w owns
// the output row tiles w, w + 8, w + 16 (up to MI of them) against every
// column tile; per step of 4 rows it reads its column operands once and its
// row operands once from LDS for up to MI x NJ matrix-core products.  Stages
// of 16 rows go through two LDS buffers (row stride kGlLd, V columns first):
// the next stage's global loads are issued before the current stage's
// products and written to the other buffer after them, one barrier per
// stage.  MFMA operand and accumulator layout as in k_gram_mfma; partials as
// k_gram (one m x k block per row range).
constexpr int kGlStage = 16;

template <int MI, int NJ>
__global__ __launch_bounds__(512) void k_gram_lds(int n, int rows_per_block,
                                                  const double *__restrict__ V, long ldv, int m,
                                                  const double *__restrict__ W, long ldw, int k,
                                                  double *__restrict__ partial)
{
    constexpr int kLd = MI * 128 + NJ * 16 + 1;  // odd: rows start in different banks
    constexpr int kPass = 512 / kGlStage;        // columns staged per pass
    constexpr int EPT = (MI * 128 + NJ * 16 + kPass - 1) / kPass;
    __shared__ double lds[2][kGlStage][kLd];
    const int ti = (m + 15) / 16, tj = (k + 15) / 16;
    const int MP = ti * 16, ncols = MP + tj * 16;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int li = lane % 16, lq = lane / 16;
    const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    const int q = threadIdx.x % kGlStage, c0 = threadIdx.x / kGlStage;
    double buf[EPT];
    auto load = [&](int rs) {
        const int r = rs + q;
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int c = c0 + kPass * u;
            double v = 0.0;
            if (r < r1) {
                if (c < MP) {
                    if (c < m) v = V[(long)c * ldv + r];
                } else if (c - MP < k) {
                    v = W[(long)(c - MP) * ldw + r];
                }
            }
            buf[u] = v;
        }
    };
    auto stash = [&](int b) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int c = c0 + kPass * u;
            if (c < ncols) lds[b][q][c] = buf[u];
        }
    };
    gcge_double4 acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = gcge_double4{0.0, 0.0, 0.0, 0.0};
    if (r0 < r1) {
        load(r0);
        stash(0);
    }
    __syncthreads();
    int b = 0;
    for (int rs = r0; rs < r1; rs += kGlStage) {
        const bool more = rs + kGlStage < r1;
        if (more) load(rs + kGlStage);
        // NJ is the number of 16-column tiles of W (the launch picks the
        // instance), so the products of a row tile run without branches
#pragma unroll
        for (int kk = 0; kk < kGlStage / 4; ++kk) {
            const int qa = 4 * kk + lq;
            double bw[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) bw[j] = lds[b][qa][MP + 16 * j + li];
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const int it = wid + 8 * i;
                if (it < ti) {
                    const double a = lds[b][qa][16 * it + li];
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bw[j], acc[i][j], 0, 0, 0);
                }
            }
        }
        if (more) stash(b ^ 1);
        __syncthreads();
        b ^= 1;
    }
    double *out = partial + (long)blockIdx.x * m * k;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const int it = wid + 8 * i;
        if (it >= ti) continue;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int col = 16 * j + li;
            if (j >= tj || col >= k) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 16 * it + lq + 4 * v;
                if (row < m) out[(long)col * m + row] = acc[i][j][v];
            }
        }
    }
}

static int num_cus()
{
    State &s = state();
    if (!s.cus) {
        int v = 0;
        HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, s.device));
        s.cus = std::max(1, v);
    }
    return s.cus;
}

// the k_gram_lds instance for an m x k block, false when none covers it
static bool launch_gram_lds(int nbx, int n, int rows, const double *V, long ldv, int m, const double *W,
                     long ldw, int k, double *partial)
{
    const int ti = (m + 15) / 16, tj = (k + 15) / 16;
    const int mi = (ti + 7) / 8;
    hipStream_t st = state().stream;
#define GCGE_GL(MI, NJ)                                                                                  \
    hipLaunchKernelGGL((k_gram_lds<MI, NJ>), dim3(nbx), dim3(512), 0, st, n, rows, V, ldv, m, W, ldw, k, \
                       partial)
#define GCGE_GL_MI(NJ)                       \
    do {                                     \
        if (mi == 1) GCGE_GL(1, NJ);         \
        else if (mi == 2) GCGE_GL(2, NJ);    \
        else if (mi == 3) GCGE_GL(3, NJ);    \
        else return false;                   \
    } while (0)
    // one instance per W tile count: no masked products
    switch (tj) {
    case 1: GCGE_GL_MI(1); break;
    case 2: GCGE_GL_MI(2); break;
    case 3: GCGE_GL_MI(3); break;
    case 4: GCGE_GL_MI(4); break;
    case 5: GCGE_GL_MI(5); break;
    case 6: GCGE_GL_MI(6); break;
    case 7: GCGE_GL_MI(7); break;
    case 8:
        if (mi == 1) GCGE_GL(1, 8);
        else if (mi == 2) GCGE_GL(2, 8);
        else return false;
        break;
    default: return false;
    }
#undef GCGE_GL_MI
#undef GCGE_GL
    return true;
}

// out[e] = sum_b partial[b * stride + e], e < count.  A block owns OUTS
// consecutive outputs; its 256 / OUTS threads per output split the partials
// (coalesced over e), each with 4 independent running sums, and combine them
// through LDS in a fixed order, so the result does not depend on timing.
template <int OUTS>
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double *__restrict__ partial,
                                                            long stride, int nparts, long count,
                                                            double *__restrict__ out)
{
    constexpr int T = kBlock / OUTS;  // threads per output
    __shared__ double lds[kBlock];
    const int o = threadIdx.x % OUTS, t = threadIdx.x / OUTS;
    const long e = (long)blockIdx.x * OUTS + o;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (e < count) {
        int b = t;
        for (; b + 3 * T < nparts; b += 4 * T) {
            s0 += partial[(long)b * stride + e];
            s1 += partial[(long)(b + T) * stride + e];
            s2 += partial[(long)(b + 2 * T) * stride + e];
            s3 += partial[(long)(b + 3 * T) * stride + e];
        }
        for (; b < nparts; b += T) s0 += partial[(long)b * stride + e];
    }
    lds[threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int h = T / 2; h > 0; h >>= 1) {
        if (t < h) lds[threadIdx.x] += lds[threadIdx.x + h * OUTS];
        __syncthreads();
    }
    if (t == 0 && e < count) out[e] = lds[o];
}

// ---- linear combination --------------------------------------------------------------

// Ct is the coefficient matrix stored row by row (Ct[i*ldct + j] = C(i, j))
// and padded with zero columns to a multiple of KT, so the KT coefficients a
// row of V meets are one contiguous, wave-uniform read (scalar loads).  Each
// lane owns one row and KT accumulators; V is read once per group of KT
// output columns.
template <int KT>
__global__ __launch_bounds__(kBlock) void k_lincomb(int n, const double *__restrict__ V, long ldv,
                                                   int m, const double *__restrict__ Ct, int ldct,
                                                   double beta, double *__restrict__ Y, long ldy,
                                                   int k)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= n) return;
    for (int c0 = 0; c0 < k; c0 += KT) {
        double acc[KT];
#pragma unroll
        for (int cc = 0; cc < KT; ++cc) acc[cc] = 0.0;
        const double *cp = Ct + c0;
        for (int i = 0; i < m; ++i) {
            const double v = V[(long)i * ldv + row];
            const double *ci = cp + (long)i * ldct;
#pragma unroll
            for (int cc = 0; cc < KT; ++cc) acc[cc] += v * ci[cc];
        }
        const int nc = min(KT, k - c0);
#pragma unroll
        for (int cc = 0; cc < KT; ++cc) {
            if (cc < nc) {
                double *y = Y + (long)(c0 + cc) * ldy + row;
                *y = (beta == 0.0) ? acc[cc] : beta * (*y) + acc[cc];
            }
        }
    }
}

// Y = beta Y + V C on the matrix cores, computed as Y^T = C^T V^T with
// v_mfma_f64_16x16x4f64: a wavefront owns 16 RT rows and 16 TJ output columns
// (grid y: column groups); per step of 4 input columns the B operand is 16
// consecutive rows of a V column per lane group (lane l: V(r0 + l % 16,
// m0 + l / 16)) and the A operand 16 coefficients of a row of Ct; lane l,
// component v of a tile then holds Y(r0 + l % 16, j0 + l / 16 + 4 v), so 16
// lanes write 16 consecutive rows of one column.
template <int TJ, int RT>
__global__ __launch_bounds__(kBlock) void k_lincomb_mfma(int n, const double *__restrict__ V, long ldv,
                                                        int m, const double *__restrict__ Ct, int ldct,
                                                        double beta, double *__restrict__ Y, long ldy,
                                                        int k)
{
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int r0 = (blockIdx.x * (kBlock / kWave) + wid) * 16 * RT;
    if (r0 >= n) return;
    const int j0 = blockIdx.y * 16 * TJ;
    const int kk = lane / 16, li = lane % 16;
    gcge_double4 acc[TJ][RT];
#pragma unroll
    for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int q = 0; q < RT; ++q) acc[t][q] = gcge_double4{0.0, 0.0, 0.0, 0.0};
    for (int m0 = 0; m0 < m; m0 += 4) {
        const int mm = m0 + kk;
        const bool mok = mm < m;
        double a[TJ], b[RT];
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            const int j = j0 + t * 16 + li;
            a[t] = (mok && j < k) ? Ct[(long)mm * ldct + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < RT; ++q) {
            const int r = r0 + q * 16 + li;
            b[q] = (mok && r < n) ? V[(long)mm * ldv + r] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TJ; ++t)
#pragma unroll
            for (int q = 0; q < RT; ++q)
                acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[q], acc[t][q], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int q = 0; q < RT; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + t * 16 + kk + 4 * v, r = r0 + q * 16 + li;
                if (j < k && r < n) {
                    double *y = Y + (long)j * ldy + r;
                    *y = (beta == 0.0) ? acc[t][q][v] : beta * (*y) + acc[t][q][v];
                }
            }
}

// The same product with V staged through LDS: a workgroup owns 16 RT rows
// and up to 128 output columns (4 waves x TJW column tiles), so each V value
// is read from HBM once for all the columns (the wavefront kernel above
// reads V once per 64 columns, and each wave waits for its own loads).
// Stages of 16 input columns go through two LDS buffers: the next stage's
// global loads are issued before the current stage's matrix-core products
// and written to the other buffer after them, one barrier per stage.  The
// stage's coefficients C(m0 .. m0+15, jg .. jg+127) go through LDS too: one
// coalesced load per workgroup instead of one per wavefront and step.
constexpr int kLcStage = 16;
template <int TJW, int RT>
__global__ __launch_bounds__(kBlock, 2) void k_lincomb_lds(int n, const double *__restrict__ V, long ldv,
                                                          int m, const double *__restrict__ Ct,
                                                          int ldct, double beta,
                                                          double *__restrict__ Y, long ldy, int k)
{
    constexpr int R = 16 * RT;                       // rows per workgroup
    constexpr int PER = kLcStage * R / kBlock;       // staged values per thread
    constexpr int CPER = kLcStage * 128 / kBlock;    // staged coefficients per thread
    __shared__ double lds[2][kLcStage][R];
    __shared__ double ldc[2][kLcStage][128];
    const int t = threadIdx.x, lane = t % kWave, wid = t / kWave;
    const int r0 = blockIdx.x * R;
    const int jg = blockIdx.y * 128;
    const int kk = lane / 16, li = lane % 16;
    gcge_double4 acc[TJW][RT];
#pragma unroll
    for (int u = 0; u < TJW; ++u)
#pragma unroll
        for (int q = 0; q < RT; ++q) acc[u][q] = gcge_double4{0.0, 0.0, 0.0, 0.0};
    double pv[PER], pc[CPER];
    auto load = [&](int m0) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = t + kBlock * u, c = e / R, r = e % R;
            pv[u] = (m0 + c < m && r0 + r < n) ? V[(long)(m0 + c) * ldv + r0 + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < CPER; ++u) {
            const int e = t + kBlock * u, i = e / 128, j = jg + e % 128;
            pc[u] = (m0 + i < m && j < k) ? Ct[(long)(m0 + i) * ldct + j] : 0.0;
        }
    };
    auto stash = [&](int b) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = t + kBlock * u;
            lds[b][e / R][e % R] = pv[u];
        }
#pragma unroll
        for (int u = 0; u < CPER; ++u) {
            const int e = t + kBlock * u;
            ldc[b][e / 128][e % 128] = pc[u];
        }
    };
    load(0);
    stash(0);
    __syncthreads();
    int buf = 0;
    for (int m0 = 0; m0 < m; m0 += kLcStage) {
        const bool more = m0 + kLcStage < m;
        if (more) load(m0 + kLcStage);
#pragma unroll
        for (int st = 0; st < kLcStage / 4; ++st) {
            double a[TJW], b[RT];
#pragma unroll
            for (int u = 0; u < TJW; ++u) a[u] = ldc[buf][4 * st + kk][16 * (wid + 4 * u) + li];
#pragma unroll
            for (int q = 0; q < RT; ++q) b[q] = lds[buf][4 * st + kk][16 * q + li];
#pragma unroll
            for (int u = 0; u < TJW; ++u)
#pragma unroll
                for (int q = 0; q < RT; ++q)
                    acc[u][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[q], acc[u][q], 0, 0, 0);
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int u = 0; u < TJW; ++u)
#pragma unroll
        for (int q = 0; q < RT; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = jg + 16 * (wid + 4 * u) + kk + 4 * v, r = r0 + 16 * q + li;
                if (j < k && r < n) {
                    double *y = Y + (long)j * ldy + r;
                    *y = (beta == 0.0) ? acc[u][q][v] : beta * (*y) + acc[u][q][v];
                }
            }
}

// ---- streaming kernels ------------------------------------------------------------------
// y = a x + b y with the reference special cases (a == 0: x not read;
// b == 0: y not read)
__global__ __launch_bounds__(kBlock) void k_axpby(int n, double a, const double *__restrict__ X,
                                                 long ldx, double b, double *__restrict__ Y,
                                                 long ldy)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row >= n) return;
    double *y = Y + (long)c * ldy + row;
    if (a == 0.0) {
        *y = b * (*y);
    } else {
        const double x = X[(long)c * ldx + row];
        *y = (b == 0.0) ? a * x : a * x + b * (*y);
    }
}

// per-column y = a x + b y over column lists, VecAxpby rules per column
__global__ __launch_bounds__(kBlock) void k_axpby_list(int n, const double *__restrict__ a,
                                                      const double *__restrict__ X, long ldx,
                                                      const int *__restrict__ xmap,
                                                      const double *__restrict__ b,
                                                      double *__restrict__ Y, long ldy,
                                                      const int *__restrict__ ymap)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row >= n) return;
    const double ac = a[c], bc = b[c];
    double *y = Y + (long)ymap[c] * ldy + row;
    if (ac == 0.0) {
        *y = bc * (*y);
    } else {
        const double x = X[(long)xmap[c] * ldx + row];
        *y = (bc == 0.0) ? ac * x : ac * x + bc * (*y);
    }
}

__global__ __launch_bounds__(kBlock) void k_swap(int n, double *__restrict__ X, long ldx,
                                                double *__restrict__ Y, long ldy)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row >= n) return;
    double *x = X + (long)c * ldx + row;
    double *y = Y + (long)c * ldy + row;
    const double t = *x;
    *x = *y;
    *y = t;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_random(int n, double *__restrict__ X, long ldx,
                                                  int col0, uint64_t seed, uint64_t draw)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row >= n) return;
    const uint64_t key = mix64(seed ^ mix64((draw << 32) ^ (uint64_t)(col0 + c)));
    const uint64_t bits = mix64(key + (uint64_t)row);
    X[(long)c * ldx + row] = (double)(bits >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// per-column dot products, partials[c * nbx + bx], accumulated in float64
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_coldots(int n, int rows_per_block,
                                                   const TA *__restrict__ A, long lda,
                                                   const int *__restrict__ amap,
                                                   const TB *__restrict__ B, long ldb,
                                                   const int *__restrict__ bmap,
                                                   double *__restrict__ partial)
{
    __shared__ double lds[kBlock / kWave];
    const int bx = blockIdx.x, c = blockIdx.y;
    const TA *a = A + (long)(amap ? amap[c] : c) * lda;
    const TB *b = B + (long)(bmap ? bmap[c] : c) * ldb;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double s = 0.0;
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock) s += (double)a[r] * (double)b[r];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partial[(long)c * gridDim.x + bx] = s;
}

// out[c] = sum of the nbx partials of column c, one wavefront per column:
// lane l adds partials l, l + 64, ... in order, then the wavefront sum (the
// block-CG bookkeeping kernel sums r.r partials the same way)
__device__ __forceinline__ double wave_col_sum(const double *__restrict__ part, int nbx, int lane)
{
    double v = 0.0;
    for (int b = lane; b < nbx; b += kWave) v += part[b];
    return wave_sum(v);
}
__global__ __launch_bounds__(kBlock) void k_reduce_cols_wave(const double *__restrict__ partial,
                                                            int nbx, int ncols,
                                                            double *__restrict__ out)
{
    const int c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (c >= ncols) return;
    const int lane = threadIdx.x % kWave;
    const double v = wave_col_sum(partial + (long)c * nbx, nbx, lane);
    if (lane == 0) out[c] = v;
}

// out[c] = sum of the nbx partials of column c, 1024 threads per column
// (the SpMM's per-row-block partials: a few thousand per column)
constexpr int kWideBlock = 1024;
__global__ __launch_bounds__(kWideBlock) void k_reduce_rows_wide(const double *__restrict__ partial,
                                                                int nbx, double *__restrict__ out)
{
    __shared__ double lds[kWideBlock / kWave];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nbx; b += kWideBlock) s += partial[(long)c * nbx + b];
    s = wave_sum(s);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) lds[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int w = 0; w < kWideBlock / kWave; ++w) tot += lds[w];
        out[c] = tot;
    }
}

// out[c] = sum of the nbx partials of column c
__global__ __launch_bounds__(kBlock) void k_reduce_rows(const double *__restrict__ partial,
                                                       int nbx, double *__restrict__ out)
{
    __shared__ double lds[kBlock / kWave];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nbx; b += kBlock) s += partial[(long)c * nbx + b];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[c] = s;
}

// residual pieces per column: |AX - lam BX|^2, |X|^2, |BX|^2
__global__ __launch_bounds__(kBlock) void k_residual(int n, int rows_per_block,
                                                    const double *__restrict__ AX, long ldax,
                                                    const double *__restrict__ BX, long ldbx,
                                                    const double *__restrict__ X, long ldx,
                                                    const double *__restrict__ lam,
                                                    double *__restrict__ partial)
{
    __shared__ double lds[kBlock / kWave];
    const int bx = blockIdx.x, c = blockIdx.y, nbx = gridDim.x, k = gridDim.y;
    const double *ax = AX + (long)c * ldax, *bxp = BX + (long)c * ldbx, *x = X + (long)c * ldx;
    const double l = lam[c];
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double s_r = 0.0, s_x = 0.0, s_b = 0.0;
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const double bv = bxp[r];
        const double rv = ax[r] - l * bv;
        s_r += rv * rv;
        const double xv = x[r];
        s_x += xv * xv;
        s_b += bv * bv;
    }
    s_r = block_sum(s_r, lds);
    s_x = block_sum(s_x, lds);
    s_b = block_sum(s_b, lds);
    if (threadIdx.x == 0) {
        partial[(long)c * nbx + bx] = s_r;
        partial[(long)(k + c) * nbx + bx] = s_x;
        partial[(long)(2 * k + c) * nbx + bx] = s_b;
    }
}

// p = r (first) or p = r + beta p, column lists.  float64 P: unscaled.
// float32 P holds p / s with a per-column scale: p_new = (r + beta s_old p) /
// s_new, s_new chosen by the host so that |p_new| <= 1 (no overflow when
// beta is large).  R is float64, or float32 with the float32 directions.
template <typename PT, typename RT>
__global__ __launch_bounds__(kBlock) void k_bcg_p(int n, PT *__restrict__ P, long ldp,
                                                 const int *__restrict__ pmap,
                                                 const RT *__restrict__ R, long ldr,
                                                 const int *__restrict__ rmap,
                                                 const double *__restrict__ beta,
                                                 const double *__restrict__ s_old,
                                                 const double *__restrict__ s_new, int first,
                                                 const int *__restrict__ count, int rows_per_block)
{
    // a block sweeps rows_per_block rows of one column, four rows per
    // thread at a time (four loads of each operand in flight)
    const int c = blockIdx.y;
    if (count && c >= *count) return;
    PT *p = P + (long)pmap[c] * ldp;
    const RT *rr = R + (long)rmap[c] * ldr;
    const double b = s_new ? beta[c] * s_old[c] : beta[c];
    const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    for (int i0 = r0 + threadIdx.x; i0 < r1; i0 += 4 * kBlock) {
        double r[4], pv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kBlock;
            r[u] = i < r1 ? (double)rr[i] : 0.0;
            pv[u] = (!first && i < r1) ? (double)p[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * kBlock;
            if (i >= r1) continue;
            if (!s_new)
                p[i] = (PT)(first ? r[u] : r[u] + b * pv[u]);
            else
                p[i] = (PT)((first ? r[u] : r[u] + b * pv[u]) / s_new[c]);
        }
    }
}

// Four consecutive elements as one 16-byte access (float) or two (double),
// widened to double; the address is 16-byte aligned.  NT marks the float
// accesses non-temporal (streams much larger than the caches: the block-CG
// updates gain 5-10 % by it); float64 accesses stay temporal, where the
// non-temporal form was measured slower.
template <typename T>
struct Quad;
template <>
struct Quad<float> {
    typedef float f4 __attribute__((ext_vector_type(4)));
    template <bool NT = false>
    static __device__ __forceinline__ void load(const float *p, double (&v)[4])
    {
        const f4 *q = reinterpret_cast<const f4 *>(p);
        const f4 t = NT ? __builtin_nontemporal_load(q) : *q;
        v[0] = t[0];
        v[1] = t[1];
        v[2] = t[2];
        v[3] = t[3];
    }
    template <bool NT = false>
    static __device__ __forceinline__ void store(float *p, const double (&v)[4])
    {
        const f4 t = {(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
        f4 *q = reinterpret_cast<f4 *>(p);
        if (NT)
            __builtin_nontemporal_store(t, q);
        else
            *q = t;
    }
};
template <>
struct Quad<double> {
    template <bool NT = false>
    static __device__ __forceinline__ void load(const double *p, double (&v)[4])
    {
        const double2 a = *reinterpret_cast<const double2 *>(p);
        const double2 b = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = b.x;
        v[3] = b.y;
    }
    template <bool NT = false>
    static __device__ __forceinline__ void store(double *p, const double (&v)[4])
    {
        *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
    }
};

// k_bcg_p with 16-byte accesses: a thread takes four consecutive rows per
// access, two such groups at a time (rows_per_block is a multiple of 4; the
// columns are 16-byte aligned).  Element arithmetic as in k_bcg_p.
template <typename PT, typename RT, bool NT>
__global__ __launch_bounds__(kBlock) void k_bcg_p4(int n, PT *__restrict__ P, long ldp,
                                                  const int *__restrict__ pmap,
                                                  const RT *__restrict__ R, long ldr,
                                                  const int *__restrict__ rmap,
                                                  const double *__restrict__ beta,
                                                  const double *__restrict__ s_old,
                                                  const double *__restrict__ s_new, int first,
                                                  const int *__restrict__ count, int rows_per_block)
{
    const int c = blockIdx.y;
    if (count && c >= *count) return;
    PT *p = P + (long)pmap[c] * ldp;
    const RT *rr = R + (long)rmap[c] * ldr;
    const double b = s_new ? beta[c] * s_old[c] : beta[c];
    const double sn = s_new ? s_new[c] : 1.0;
    const int r0 = blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
    const int r1q = r0 + ((r1 - r0) & ~3);
    auto one = [&](double r, double pv) -> double {
        const double t = first ? r : r + b * pv;
        return s_new ? t / sn : t;
    };
    for (int i0 = r0 + 4 * threadIdx.x; i0 < r1q; i0 += 8 * kBlock) {
        const int i1 = i0 + 4 * kBlock;
        const bool two = i1 < r1q;
        double ra[4], pa[4] = {0.0, 0.0, 0.0, 0.0}, rb[4], pb[4] = {0.0, 0.0, 0.0, 0.0};
        Quad<RT>::template load<NT>(rr + i0, ra);
        if (!first) Quad<PT>::template load<NT>(p + i0, pa);
        if (two) {
            Quad<RT>::template load<NT>(rr + i1, rb);
            if (!first) Quad<PT>::template load<NT>(p + i1, pb);
        }
        double oa[4], ob[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            oa[u] = one(ra[u], pa[u]);
            ob[u] = two ? one(rb[u], pb[u]) : 0.0;
        }
        Quad<PT>::template store<NT>(p + i0, oa);
        if (two) Quad<PT>::template store<NT>(p + i1, ob);
    }
    for (int i = r1q + threadIdx.x; i < r1; i += kBlock)
        p[i] = (PT)one((double)rr[i], first ? 0.0 : (double)p[i]);
}

// p.w and p.p per column (the float32 directions of the block CG),
// partials[c * nbx + bx] and partials[(k + c) * nbx + bx]
template <typename PT>
__global__ __launch_bounds__(kBlock) void k_coldots_pw(int n, int rows_per_block,
                                                      const PT *__restrict__ P, long ldp,
                                                      const int *__restrict__ pmap,
                                                      const PT *__restrict__ W, long ldw,
                                                      const int *__restrict__ wmap,
                                                      double *__restrict__ partial)
{
    __shared__ double lds[kBlock / kWave];
    const int bx = blockIdx.x, c = blockIdx.y, nbx = gridDim.x, k = gridDim.y;
    const PT *p = P + (long)pmap[c] * ldp;
    const PT *w = W + (long)wmap[c] * ldw;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double s_pw = 0.0, s_pp = 0.0;
    for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const double pv = (double)p[r];
        s_pw += pv * (double)w[r];
        s_pp += pv * pv;
    }
    s_pw = block_sum(s_pw, lds);
    s_pp = block_sum(s_pp, lds);
    if (threadIdx.x == 0) {
        partial[(long)c * nbx + bx] = s_pw;
        partial[(long)(k + c) * nbx + bx] = s_pp;
    }
}

// x += alpha p; r -= alpha w; partial r.r, with alpha = rho2[c] / ptw[c]
// (float64 P), or alpha = rho2[c] / (scale[c] ptw[c]) for the scaled float32
// directions (p = scale p_stored, ptw the stored p . A p_stored); x float64,
// r float64 or float32 (with float32 directions); arithmetic and r.r in float64
template <typename PT, typename RT, typename XT>
__global__ __launch_bounds__(kBlock) void k_bcg_xr(int n, int rows_per_block,
                                                  XT *__restrict__ X, long ldx,
                                                  const int *__restrict__ xmap,
                                                  RT *__restrict__ R, long ldr,
                                                  const int *__restrict__ rmap,
                                                  const PT *__restrict__ P, long ldp,
                                                  const int *__restrict__ pmap,
                                                  const PT *__restrict__ W, long ldw,
                                                  const int *__restrict__ wmap,
                                                  const double *__restrict__ rho2,
                                                  const double *__restrict__ ptw,
                                                  const double *__restrict__ scale,
                                                  double *__restrict__ partial,
                                                  const int *__restrict__ count)
{
    __shared__ double lds[kBlock / kWave];
    const int bx = blockIdx.x, c = blockIdx.y;
    if (count && c >= *count) return;
    const double alpha = scale ? rho2[c] / (scale[c] * ptw[c]) : rho2[c] / ptw[c];
    XT *x = X + (long)xmap[c] * ldx;
    RT *r = R + (long)rmap[c] * ldr;
    const PT *p = P + (long)pmap[c] * ldp;
    const PT *w = W + (long)wmap[c] * ldw;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    double s = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += kBlock) {
        x[i] = (XT)(alpha * (double)p[i] + (double)x[i]);
        const RT rn = (RT)(-alpha * (double)w[i] + (double)r[i]);
        r[i] = rn;
        s += (double)rn * (double)rn;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partial[(long)c * gridDim.x + bx] = s;
}

// k_bcg_xr with 16-byte accesses (four consecutive rows per access;
// rows_per_block a multiple of 4, columns 16-byte aligned).  Element
// arithmetic as in k_bcg_xr; the r.r partial of a block sums the same rows in
// another order.
template <typename PT, typename RT, typename XT, bool NT>
__global__ __launch_bounds__(kBlock) void k_bcg_xr4(int n, int rows_per_block,
                                                   XT *__restrict__ X, long ldx,
                                                   const int *__restrict__ xmap,
                                                   RT *__restrict__ R, long ldr,
                                                   const int *__restrict__ rmap,
                                                   const PT *__restrict__ P, long ldp,
                                                   const int *__restrict__ pmap,
                                                   const PT *__restrict__ W, long ldw,
                                                   const int *__restrict__ wmap,
                                                   const double *__restrict__ rho2,
                                                   const double *__restrict__ ptw,
                                                   const double *__restrict__ scale,
                                                   double *__restrict__ partial,
                                                   const int *__restrict__ count)
{
    __shared__ double lds[kBlock / kWave];
    const int bx = blockIdx.x, c = blockIdx.y;
    if (count && c >= *count) return;
    const double alpha = scale ? rho2[c] / (scale[c] * ptw[c]) : rho2[c] / ptw[c];
    XT *x = X + (long)xmap[c] * ldx;
    RT *r = R + (long)rmap[c] * ldr;
    const PT *p = P + (long)pmap[c] * ldp;
    const PT *w = W + (long)wmap[c] * ldw;
    const int r0 = bx * rows_per_block, r1 = min(n, r0 + rows_per_block);
    const int r1q = r0 + ((r1 - r0) & ~3);
    double s = 0.0;
    for (int i0 = r0 + 4 * threadIdx.x; i0 < r1q; i0 += 4 * kBlock) {
        double pa[4], xa[4], wa[4], ra[4], xo[4], ro[4];
        Quad<PT>::template load<NT>(p + i0, pa);
        Quad<XT>::template load<NT>(x + i0, xa);
        Quad<PT>::template load<NT>(w + i0, wa);
        Quad<RT>::template load<NT>(r + i0, ra);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            xo[u] = (double)(XT)(alpha * pa[u] + xa[u]);
            ro[u] = (double)(RT)(-alpha * wa[u] + ra[u]);
            s += ro[u] * ro[u];
        }
        Quad<XT>::template store<NT>(x + i0, xo);
        Quad<RT>::template store<NT>(r + i0, ro);
    }
    for (int i = r1q + threadIdx.x; i < r1; i += kBlock) {
        x[i] = (XT)(alpha * (double)p[i] + (double)x[i]);
        const RT rn = (RT)(-alpha * (double)w[i] + (double)r[i]);
        r[i] = rn;
        s += (double)rn * (double)rn;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partial[(long)c * gridDim.x + bx] = s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_cols(int n, const double *__restrict__ alpha,
                                                     const T *__restrict__ X, long ldx,
                                                     const int *__restrict__ xmap,
                                                     T *__restrict__ Y, long ldy,
                                                     const int *__restrict__ ymap)
{
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (row >= n) return;
    Y[(long)(ymap ? ymap[c] : c) * ldy + row] += (T)alpha[c] * X[(long)(xmap ? xmap[c] : c) * ldx + row];
}

// Block-CG bookkeeping (reference src/gcge_cg.c:395-490 order of
// operations), one workgroup; the state lives in LDS during the launch.
__global__ __launch_bounds__(kBlock) void k_bcg_control(BcgControl d, int init, int finished_step,
                                                       int prepare_step)
{
    // the host form of this arithmetic has no fused multiply-adds: the same
    // roundings here keep both forms bit-identical
#pragma clang fp contract(off)
    __shared__ double s_rho1[kBcgControlMax], s_rho2[kBcgControlMax], s_err[kBcgControlMax],
        s_lerr[kBcgControlMax], s_scol[kBcgControlMax], s_pn[kBcgControlMax];
    __shared__ int s_unlock[kBcgControlMax];
    __shared__ double s_rr[kBcgControlMax];
    __shared__ int s_count;
    const int t = threadIdx.x, len = d.len;
    for (int i = t; i < len; i += kBlock) {
        s_rho1[i] = d.rho1[i];
        if (init) {
            const double r = d.rr[i];
            s_rho2[i] = r;
            s_err[i] = sqrt(r);
            s_lerr[i] = s_err[i];
            s_scol[i] = 1.0;
            s_pn[i] = 0.0;
            s_unlock[i] = i;
        } else {
            s_rho2[i] = d.rho2[i];
            s_err[i] = d.err[i];
            s_lerr[i] = d.last_err[i];
            s_scol[i] = d.s_col[i];
            s_pn[i] = d.pn_col[i];
            s_unlock[i] = d.unlock[i];
        }
    }
    if (t == 0) s_count = init ? len : *d.count;
    if (finished_step > 0 && d.rr_part) {
        // r.r of the step's columns from the partial sums, one wavefront per
        // column, issued with the state loads above (all len columns: the
        // partial buffer holds len of them; only the first count are used)
        const int lane = t % kWave;
        for (int id = t / kWave; id < len; id += kBlock / kWave) {
            const double v = wave_col_sum(d.rr_part + (long)id * d.rr_nbx, d.rr_nbx, lane);
            if (lane == 0) s_rr[id] = v;
        }
    }
    __syncthreads();
    const int nu = s_count;
    if (finished_step > 0) {
        // rho1 <- rho2, rho2 <- r.r, |p| for the scaled directions, then
        // last_err from the new rho2 (each position touches its own entries)
        for (int id = t; id < nu; id += kBlock) {
            const int idx = s_unlock[id];
            const int q = d.by_column ? idx : id;
            s_rho1[q] = s_rho2[q];
            s_rho2[q] = d.rr_part ? s_rr[id] : d.rr[id];
            if (d.scaled) s_pn[idx] = sqrt(d.ptw[len + id]);
            s_lerr[idx] = sqrt(fabs(s_rho2[q]));
        }
        if (t == 0) d.hist[finished_step - 1] = nu;
        __syncthreads();
    }
    if (init || finished_step > 0) {
        // keep the columns that have not reached the rate (wave 0, in order)
        if (t < kWave) {
            int out = 0;
            const double rate = d.rate;
            for (int base = 0; base < nu; base += kWave) {
                const int id = base + t;
                int idx = 0;
                bool keep = false;
                if (id < nu) {
                    idx = s_unlock[id];
                    keep = init ? s_err[idx] >= 1e-15
                                : (s_lerr[idx] >= rate * s_err[idx] && s_lerr[idx] >= 1e-15);
                }
                const unsigned long long m = __ballot(keep);
                const int pos = out + __popcll(m & ((1ull << t) - 1ull));
                if (keep) s_unlock[pos] = idx;
                out += __popcll(m);
            }
            if (t == 0) s_count = out;
        }
        __syncthreads();
    }
    const int nn = s_count;
    if (prepare_step > 0) {
        for (int id = t; id < nn; id += kBlock) {
            const int idx = s_unlock[id];
            d.maps[id] = d.p0 + idx;
            d.maps[len + id] = d.r0 + idx;
            d.maps[2 * len + id] = d.x0 + idx;
            d.maps[3 * len + id] = d.w0 + idx;
            const double beta = prepare_step == 1 ? 0.0 : s_rho2[idx] / s_rho1[idx];  // by column
            d.beta[id] = beta;
            d.alpha_num[id] = s_rho2[d.by_column ? idx : id];  // reference: by position
            if (d.scaled) {
                double snew = prepare_step == 1 ? s_lerr[idx]
                                                : s_lerr[idx] + fabs(beta) * s_scol[idx] * s_pn[idx];
                if (!(snew > 0.0) && !isnan(snew)) snew = 1.0;
                d.s_old[id] = s_scol[idx];
                d.s_new[id] = snew;
            }
        }
        __syncthreads();
        if (d.scaled)
            for (int id = t; id < nn; id += kBlock) s_scol[s_unlock[id]] = d.s_new[id];
        __syncthreads();
    }
    for (int i = t; i < len; i += kBlock) {
        d.rho1[i] = s_rho1[i];
        d.rho2[i] = s_rho2[i];
        d.err[i] = s_err[i];
        d.last_err[i] = s_lerr[i];
        d.s_col[i] = s_scol[i];
        d.pn_col[i] = s_pn[i];
        d.unlock[i] = s_unlock[i];
    }
    if (t == 0) *d.count = nn;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// out[e] = sum over nparts partial rows (row stride `stride`) of entry e
void reduce_partials(const double *partial, long stride, int nparts, long count, double *out)
{
    hipStream_t st = state().stream;
    if (count <= 1)
        hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(kBlock), 0, st, partial, stride, nparts, count, out);
    else if (count <= 4)
        hipLaunchKernelGGL(k_reduce_partials<4>, dim3(1), dim3(kBlock), 0, st, partial, stride, nparts, count, out);
    else if (count <= 4096)
        // 16 threads per output: at least count / 16 workgroups share the
        // partials (64 outputs per workgroup left a 10 x 70 block of 200
        // partials to 11 workgroups, 15 us of mostly latency)
        hipLaunchKernelGGL(k_reduce_partials<16>, dim3(ceil_div(count, 16)), dim3(kBlock), 0, st, partial, stride,
                           nparts, count, out);
    else
        hipLaunchKernelGGL(k_reduce_partials<64>, dim3(ceil_div(count, 64)), dim3(kBlock), 0, st, partial, stride,
                           nparts, count, out);
    HIP_CHECK(hipGetLastError());
}

// rows per block for column reductions: enough blocks to fill the chip
// Block-CG streaming updates (direction, x/r): about GCGE_HIP_STREAM_BLOCKS
// blocks over all columns (default 32768; more, shorter blocks stream faster
// than the 2048 of the reductions), at least 4 x 256 rows each and a multiple
// of that; float32 accesses non-temporal unless GCGE_HIP_STREAM_NT is 0.
struct StreamCfg {
    long blocks = 32768;
    bool nt = true;
};

inline const StreamCfg &stream_cfg()
{
    static const StreamCfg cfg = [] {
        StreamCfg c;
        if (const char *e = getenv("GCGE_HIP_STREAM_BLOCKS")) c.blocks = std::max(1L, atol(e));
        if (const char *e = getenv("GCGE_HIP_STREAM_NT")) c.nt = atoi(e) != 0;
        return c;
    }();
    return cfg;
}

inline int stream_rows(int n, int k)
{
    const long per_col = std::max<long>(1, stream_cfg().blocks / std::max(1, k));
    const long rows = std::max<long>(4L * kBlock, (n + per_col - 1) / per_col);
    return (int)((rows + 4 * kBlock - 1) / (4 * kBlock) * (4 * kBlock));
}

inline int reduction_rows(int n, int k)
{
    const long target_blocks = 2048;
    long per_col = std::max<long>(1, target_blocks / std::max(1, k));
    long rows = std::max<long>(kBlock * 4, (n + per_col - 1) / per_col);
    return (int)((rows + kBlock - 1) / kBlock * kBlock);
}

}  // namespace

int dia_plane_rows(const Mat *A) { return plane_rows(A); }

// ---- launchers ---------------------------------------------------------------------------

void spmm(const Mat *A, const double *X, long ldx, double *Y, long ldy, int k, const int *d_xmap,
          const int *d_ymap)
{
    if (k <= 0 || A->nrows <= 0) return;
    // algorithmic bytes: values + column indices + row pointers once, X read
    // once, Y written once
    const double bytes = (double)A->nnz * (8 + 4) + 4.0 * (A->nrows + 1) +
                         8.0 * (double)k * (A->ncols + A->nrows);
    ProfScope prof("spmm", bytes, true);
    if (A->use_dia && use_tiles<double>(A, k))
        launch_dia_tile<0>(A, A->d_dia, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                           prof.start_event(), prof.stop_event());
    else if (A->use_dia)
        launch_dia_spmm<0>(A, A->d_dia, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                           prof.start_event(), prof.stop_event());
    else if (A->use_rows && k >= kRowsMinCols)
        launch_rows_spmm<0>(A, A->d_val, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                            prof.start_event(), prof.stop_event());
    else
        launch_sell_spmm<0>(A, A->d_val, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                            prof.start_event(), prof.stop_event());
}

void spmm(const Mat *A, const float *X, long ldx, float *Y, long ldy, int k, const int *d_xmap,
          const int *d_ymap)
{
    if (k <= 0 || A->nrows <= 0) return;
    if (!A->d_val32) fail_msg("float32 SpMM on a matrix without float32 values");
    const double bytes = (double)A->nnz * (4 + 4) + 4.0 * (A->nrows + 1) +
                         4.0 * (double)k * (A->ncols + A->nrows);
    ProfScope prof("spmm32", bytes, true);
    if (A->use_dia && A->d_dia32 && use_tiles<float>(A, k))
        launch_dia_tile<0>(A, A->d_dia32, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                           prof.start_event(), prof.stop_event());
    else if (A->use_dia && A->d_dia32)
        launch_dia_spmm<0>(A, A->d_dia32, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                           prof.start_event(), prof.stop_event());
    else if (A->use_rows && k >= kRowsMinCols)
        launch_rows_spmm<0>(A, A->d_val32, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                            prof.start_event(), prof.stop_event());
    else
        launch_sell_spmm<0>(A, A->d_val32, X, ldx, Y, ldy, k, d_xmap, d_ymap, nullptr, nullptr,
                            prof.start_event(), prof.stop_event());
}

void spmm_dot(const Mat *A, const double *P, long ldp, double *W, long ldw, int k,
              const int *d_pmap, const int *d_wmap, double *d_out, const int *d_count,
              int *prof_slot)
{
    if (k <= 0 || A->nrows <= 0) return;
    // P read and W written (the p.w epilogue reuses the P values the product reads)
    const double bytes = (double)A->nnz * (8 + 4) + 4.0 * (A->nrows + 1) +
                         8.0 * (double)k * ((double)A->ncols + A->nrows);
    const bool rows = !A->use_dia && A->use_rows && k >= kRowsMinCols;
    const bool tiles = A->use_dia && use_tiles<double>(A, k);
    const int nparts = tiles ? tile_parts<double>(A)
                             : (A->use_dia ? dia_blocks(A) : (rows ? rows_parts(A) : A->nslices));
    double *dpart = partial_buffer((size_t)nparts * k);
    {
        ProfScope prof("spmm", bytes, true);
        if (prof_slot) *prof_slot = prof.slot;
        if (tiles)
            launch_dia_tile<1>(A, A->d_dia, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                               prof.start_event(), prof.stop_event());
        else if (A->use_dia)
            launch_dia_spmm<1>(A, A->d_dia, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                               prof.start_event(), prof.stop_event());
        else if (rows)
            launch_rows_spmm<1>(A, A->d_val, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                                prof.start_event(), prof.stop_event());
        else
            launch_sell_spmm<1>(A, A->d_val, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                                prof.start_event(), prof.stop_event());
    }
    hipLaunchKernelGGL(k_reduce_rows_wide, dim3(k), dim3(kWideBlock), 0, state().stream, dpart,
                       nparts, d_out);
    HIP_CHECK(hipGetLastError());
}

void spmm_dot(const Mat *A, const float *P, long ldp, float *W, long ldw, int k, const int *d_pmap,
              const int *d_wmap, double *d_out, const int *d_count, int *prof_slot)
{
    if (k <= 0 || A->nrows <= 0) return;
    if (!A->d_val32) fail_msg("float32 SpMM on a matrix without float32 values");
    const double bytes = (double)A->nnz * (4 + 4) + 4.0 * (A->nrows + 1) +
                         4.0 * (double)k * ((double)A->ncols + A->nrows);
    const bool dia = A->use_dia && A->d_dia32;
    const bool rows = !dia && A->use_rows && k >= kRowsMinCols;
    const bool tiles = dia && use_tiles<float>(A, k);
    const int nparts = tiles ? tile_parts<float>(A)
                             : (dia ? dia_blocks(A) : (rows ? rows_parts(A) : A->nslices));
    double *dpart = partial_buffer((size_t)nparts * 2 * k);
    {
        ProfScope prof("spmm32", bytes, true);
        if (prof_slot) *prof_slot = prof.slot;
        if (tiles)
            launch_dia_tile<2>(A, A->d_dia32, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                               prof.start_event(), prof.stop_event());
        else if (dia)
            launch_dia_spmm<2>(A, A->d_dia32, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                               prof.start_event(), prof.stop_event());
        else if (rows)
            launch_rows_spmm<2>(A, A->d_val32, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                                prof.start_event(), prof.stop_event());
        else
            launch_sell_spmm<2>(A, A->d_val32, P, ldp, W, ldw, k, d_pmap, d_wmap, dpart, d_count,
                                prof.start_event(), prof.stop_event());
    }
    hipLaunchKernelGGL(k_reduce_rows_wide, dim3(2 * k), dim3(kWideBlock), 0, state().stream, dpart,
                       nparts, d_out);
    HIP_CHECK(hipGetLastError());
}

void bcg_control(const BcgControl &c, int init, int finished_step, int prepare_step)
{
    if (c.len > kBcgControlMax) fail_msg("block CG bookkeeping: too many columns for one workgroup");
    hipLaunchKernelGGL(k_bcg_control, dim3(1), dim3(kBlock), 0, state().stream, c, init,
                       finished_step, prepare_step);
    HIP_CHECK(hipGetLastError());
}

void to_float(const double *src, float *dst, size_t len)
{
    if (!len) return;
    hipLaunchKernelGGL(k_to_float, dim3(ceil_div((long)len, kBlock)), dim3(kBlock), 0,
                       state().stream, len, src, dst);
    HIP_CHECK(hipGetLastError());
}

void add_float(int n, const float *D, long ldd, double *X, long ldx, int k)
{
    if (k <= 0 || n <= 0) return;
    ProfScope prof("convert");
    hipLaunchKernelGGL(k_add_float, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream,
                       n, D, ldd, X, ldx);
    HIP_CHECK(hipGetLastError());
}

void zero_float(int n, float *X, long ldx, int k)
{
    if (k <= 0 || n <= 0) return;
    hipLaunchKernelGGL(k_zero_float, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream,
                       n, X, ldx);
    HIP_CHECK(hipGetLastError());
}

void to_double(const float *src, double *dst, size_t len)
{
    if (!len) return;
    hipLaunchKernelGGL(k_to_double, dim3(ceil_div((long)len, kBlock)), dim3(kBlock), 0,
                       state().stream, len, src, dst);
    HIP_CHECK(hipGetLastError());
}

// rocBLAS dgemm for the large Gram blocks and linear combinations instead of
// the kernels here (GCGE_HIP_BLAS=1: a comparison for measurement; the
// kernels here are the default)
static bool use_blas()
{
    static const bool v = [] {
        const char *e = getenv("GCGE_HIP_BLAS");
        return e && atoi(e) == 1;
    }();
    return v;
}
static rocblas_handle blas_handle()
{
    static rocblas_handle h = nullptr;
    if (!h) {
        if (rocblas_create_handle(&h) != rocblas_status_success) fail_msg("rocblas_create_handle failed");
        if (rocblas_set_stream(h, state().stream) != rocblas_status_success) fail_msg("rocblas_set_stream failed");
    }
    return h;
}

// The row ranges of a Gram launch: as many as `want`, or fewer when the
// caller's partial buffer (cap doubles, 0: the backend's grown buffer)
// cannot hold that many m x k blocks; rows per range rounded up to `gran`.
static int gram_ranges(int n, int want, long mk, size_t cap, int gran, int *rows)
{
    int nbx = std::max(1, want);
    if (cap && (size_t)nbx * mk > cap) nbx = (int)std::max<size_t>(1, cap / mk);
    int r = ceil_div(n, nbx);
    r = (r + gran - 1) / gran * gran;
    *rows = r;
    return std::max(1, ceil_div(n, r));
}

int gram_impl(int n, const double *V, long ldv, int m, const double *W, long ldw, int k,
              double *d_out, double *partial, size_t partial_cap, bool profile)
{
    State &s = state();
    const long mk = (long)m * k;
    if (partial && (size_t)mk > partial_cap) return -1;
    if (!partial && use_blas() && m > 16 && k > 16) {
        ProfScope prof(profile ? "gram" : nullptr, 8.0 * (double)n * (m + k));
        const double one = 1.0, zero = 0.0;
        if (rocblas_dgemm(blas_handle(), rocblas_operation_transpose, rocblas_operation_none, m, k, n, &one,
                          V, ldv, W, ldw, &zero, d_out, m) != rocblas_status_success)
            fail_msg("rocblas_dgemm failed");
        return 1;
    }
    const char *pname = profile ? "gram" : nullptr;
    const char *rname = profile ? "gram_reduce" : nullptr;
    const double gbytes = 8.0 * (double)n * (m + k);
    int rows = 0, nbx = 0;
    if (m <= 8 && k <= 8) {
        // register-blocked path for the small blocks of the orthonormalisation
        nbx = gram_ranges(n, std::min(ceil_div(n, 4 * kBlock), 512), mk, partial_cap, 1, &rows);
        if (!partial) partial = partial_buffer((size_t)nbx * mk);
        {
            ProfScope prof(pname, gbytes);
            const int MR = m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : 8;
            const int KR = k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8;
#define GCGE_GRAM_SMALL(A, B)                                                                        \
    if (MR == A && KR == B)                                                                          \
        hipLaunchKernelGGL((k_gram_small<A, B>), dim3(nbx), dim3(kBlock), 0, s.stream, n, rows, V, ldv, \
                           m, W, ldw, k, partial);
            GCGE_GRAM_SMALL(1, 1) GCGE_GRAM_SMALL(1, 2) GCGE_GRAM_SMALL(1, 4) GCGE_GRAM_SMALL(1, 8)
            GCGE_GRAM_SMALL(2, 1) GCGE_GRAM_SMALL(2, 2) GCGE_GRAM_SMALL(2, 4) GCGE_GRAM_SMALL(2, 8)
            GCGE_GRAM_SMALL(4, 1) GCGE_GRAM_SMALL(4, 2) GCGE_GRAM_SMALL(4, 4) GCGE_GRAM_SMALL(4, 8)
            GCGE_GRAM_SMALL(8, 1) GCGE_GRAM_SMALL(8, 2) GCGE_GRAM_SMALL(8, 4) GCGE_GRAM_SMALL(8, 8)
#undef GCGE_GRAM_SMALL
            HIP_CHECK(hipGetLastError());
        }
        ProfScope prof(rname);
        reduce_partials(partial, mk, nbx, mk, d_out);
        return nbx;
    }
    static const int gram_mode = [] {
        const char *e = getenv("GCGE_HIP_GRAM");  // "fma": the vector-FMA kernels (measurement)
        return (e && strcmp(e, "fma") == 0) ? 1 : 0;
    }();
    // smallest m for the whole-block kernel, any k (GCGE_HIP_GRAM_LDS; 0: the
    // k_gram_mfma blocks only, for measurement).  Measured at n = 2 048 383
    // (wall time of the call): 300 x 48 1.81 against 2.81 ms, 300 x 16 1.89
    // against 2.03, 200 x 8 0.87 against 1.39, 100 x 24 0.65 against 0.88;
    // at n = 250 047, 60 x 10 took 68 against 57 us, so small m stay on the
    // other kernels
    static const int gram_lds_min = [] {
        const char *e = getenv("GCGE_HIP_GRAM_LDS");
        return e ? atoi(e) : 64;
    }();
    const int gti = ceil_div(m, 16), gtj = ceil_div(k, 16);
    if (gram_mode == 0 && gram_lds_min > 0 && m >= gram_lds_min &&
        ((gtj <= 7 && gti <= 24) || (gtj == 8 && gti <= 16))) {
        // one workgroup per CU (eight waves, up to 127 KiB of LDS), the whole
        // m x k block each, about 500 stages of 16 rows per workgroup at 128^3
        nbx = gram_ranges(n, std::min(ceil_div(n, 4 * kGlStage), num_cus()), mk, partial_cap, kGlStage,
                          &rows);
        if (!partial) partial = partial_buffer((size_t)nbx * mk);
        {
            ProfScope prof(pname, gbytes);
            if (!launch_gram_lds(nbx, n, rows, V, ldv, m, W, ldw, k, partial)) fail_msg("k_gram_lds shape");
            HIP_CHECK(hipGetLastError());
        }
        ProfScope prof(rname);
        reduce_partials(partial, mk, nbx, mk, d_out);
        return nbx;
    }
    if (gram_mode == 0) {
        // matrix cores: 128 x 128 output blocks (grid y, z), tile pairs shared
        // by four or eight waves (one block of up to 304 x 128 outputs, W read
        // once, measured slower at 300 x 100: 5.85 against 4.77 ms)
        const int bcv = kGmCols, bcw = kGmCols;
        const int gy = ceil_div(m, bcv), gz = ceil_div(k, bcw);
        const int ti = ceil_div(std::min(m, bcv), 16), tj = ceil_div(std::min(k, bcw), 16);
        // eight waves per workgroup beyond 32 output tiles (four would hold 9..16
        // accumulator tiles each and run at one wave per SIMD)
        const int nw = ti * tj > 32 ? 8 : 4;
        const int pw = ceil_div(ti * tj, nw);
        const int ncols = 16 * (ti + tj);
        const int ept = ceil_div(kGmRows * ncols, 64 * nw);
        const size_t lds = (size_t)kGmRows * (ncols + 1) * sizeof(double);
        // about four workgroups per CU in all, at least two stages each, and
        // partials of at most a tenth of the input bytes (the reduction pass
        // reads them all)
        const long cap_parts = std::max(1L, (long)n * (m + k) / (10L * mk));
        int want = std::max(1, std::min(ceil_div(n, 2 * kGmRows), std::max(1, 1024 / (gy * gz))));
        want = (int)std::min<long>(want, cap_parts);
        nbx = gram_ranges(n, want, mk, partial_cap, kGmRows, &rows);
        if (!partial) partial = partial_buffer((size_t)nbx * mk);
        {
            ProfScope prof(pname, gbytes);
            dim3 grid(nbx, gy, gz);
            launch_gram_mfma(nw, pw, ept, grid, lds, n, rows, V, ldv, m, W, ldw, k, partial, bcv, bcw);
            HIP_CHECK(hipGetLastError());
        }
        ProfScope prof(rname);
        reduce_partials(partial, mk, nbx, mk, d_out);
        return nbx;
    }
    if (k <= 16) {
        // column-block register path: 8 x 8 or 4 x 16 accumulators per lane
        const int MR = k <= 8 ? 8 : 4;
        const int colblocks = ceil_div(m, MR);
        const int gy = ceil_div(colblocks, kBlock / kWave);
        nbx = gram_ranges(n, std::min(ceil_div(n, 4 * kWave), std::max(1, 1024 / gy)), mk, partial_cap, 1,
                          &rows);
        if (!partial) partial = partial_buffer((size_t)nbx * mk);
        {
            ProfScope prof(pname, gbytes);
            if (k <= 8)
                hipLaunchKernelGGL((k_gram_cols<8, 8>), dim3(nbx, gy), dim3(kBlock), 0, s.stream, n, rows, V,
                                   ldv, m, W, ldw, k, partial);
            else
                hipLaunchKernelGGL((k_gram_cols<4, 16>), dim3(nbx, gy), dim3(kBlock), 0, s.stream, n, rows, V,
                                   ldv, m, W, ldw, k, partial);
            HIP_CHECK(hipGetLastError());
        }
        ProfScope prof(rname);
        reduce_partials(partial, mk, nbx, mk, d_out);
        return nbx;
    }
    const bool wide_m = m > 16, wide_k = k > 16;
    const int TM = wide_m ? 64 : 16, TN = wide_k ? 64 : 16;
    const int tm = ceil_div(m, TM), tn = ceil_div(k, TN);
    const int tiles = tm * tn;
    // about two workgroups per CU over all tiles, at least 256 rows each
    nbx = gram_ranges(n, std::min(ceil_div(n, kBlock), std::max(1, 512 / tiles)), mk, partial_cap,
                      kGramRows, &rows);
    if (!partial) partial = partial_buffer((size_t)nbx * mk);
    {
        ProfScope prof(pname, gbytes);
        dim3 grid(nbx, tm, tn);
        if (wide_m && wide_k)
            hipLaunchKernelGGL((k_gram<64, 64>), grid, dim3(kBlock), 0, s.stream, n, rows, V, ldv, m, W, ldw, k, partial);
        else if (wide_m)
            hipLaunchKernelGGL((k_gram<64, 16>), grid, dim3(kBlock), 0, s.stream, n, rows, V, ldv, m, W, ldw, k, partial);
        else if (wide_k)
            hipLaunchKernelGGL((k_gram<16, 64>), grid, dim3(kBlock), 0, s.stream, n, rows, V, ldv, m, W, ldw, k, partial);
        else
            hipLaunchKernelGGL((k_gram<16, 16>), grid, dim3(kBlock), 0, s.stream, n, rows, V, ldv, m, W, ldw, k, partial);
        HIP_CHECK(hipGetLastError());
    }
    ProfScope prof(rname);
    reduce_partials(partial, mk, nbx, mk, d_out);
    return nbx;
}

void gram(int n, const double *V, long ldv, int m, const double *W, long ldw, int k, double *G,
          int ldg, bool sync_result)
{
    if (m <= 0 || k <= 0) return;
    if (n <= 0) {
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < m; ++i) G[(long)j * ldg + i] = 0.0;
        return;
    }
    State &s = state();
    const long mk = (long)m * k;
    double *d_out = small_buffer2((size_t)mk);
    gram_impl(n, V, ldv, m, W, ldw, k, d_out, nullptr, 0, true);
    double *h = result_buffer((size_t)mk);
    HIP_CHECK(hipMemcpyAsync(h, d_out, mk * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    if (!sync_result) return;
    HIP_CHECK(hipStreamSynchronize(s.stream));
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i) G[(long)j * ldg + i] = h[(long)j * m + i];
}

int lincomb_group(int k) { return k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64; }

// the kernel switches of lincomb_dev (measurement): GCGE_HIP_LINCOMB=fma
// for the vector kernel, GCGE_HIP_LINCOMB_LDS=0 for the wavefront kernels
static int lincomb_mode()
{
    static const int mode = [] {
        const char *e = getenv("GCGE_HIP_LINCOMB");
        return (e && strcmp(e, "fma") == 0) ? 1 : 0;
    }();
    return mode;
}
static int lincomb_lds_mode()
{
    static const int v = [] {
        const char *e = getenv("GCGE_HIP_LINCOMB_LDS");
        return e ? atoi(e) : 1;
    }();
    return v;
}

bool lincomb_in_place_ok(int m, int k)
{
    return k > 32 && k <= 128 && m >= 32 && !use_blas() && lincomb_mode() == 0 &&
           lincomb_lds_mode() != 0;
}

// Y = V C where the columns of Y may overlap those of V (one block of
// columns, same rows and leading dimension).  The LDS-staged kernel loads
// all m input columns of a workgroup's 128 rows in its stage loop, which
// ends in a workgroup barrier, and writes those rows only after it; with
// k <= 128 a single workgroup owns each row block, so the output may
// overwrite the input (the same kernel, the same arithmetic as the
// out-of-place call).
void lincomb_in_place(int n, const double *V, double *Y, long ld, int m, const double *d_Ct, int ldct,
                      int k)
{
    if (n <= 0 || k <= 0) return;
    if (!lincomb_in_place_ok(m, k)) fail_msg("lincomb_in_place: shape not supported");
    ProfScope prof("lincomb", 8.0 * (double)n * (m + k));
    const dim3 g(ceil_div(n, 128), 1), block(kBlock);
    hipStream_t st = state().stream;
    if (k > 64)
        hipLaunchKernelGGL((k_lincomb_lds<2, 8>), g, block, 0, st, n, V, ld, m, d_Ct, ldct, 0.0, Y, ld, k);
    else
        hipLaunchKernelGGL((k_lincomb_lds<1, 8>), g, block, 0, st, n, V, ld, m, d_Ct, ldct, 0.0, Y, ld, k);
    HIP_CHECK(hipGetLastError());
}

void lincomb_dev(int n, const double *V, long ldv, int m, const double *d_Ct, int ldct,
                 double beta, double *Y, long ldy, int k)
{
    if (k <= 0 || n <= 0) return;
    if (m <= 0) {  // empty combination: y = beta y
        axpby(n, 0.0, Y, ldy, beta, Y, ldy, k);
        return;
    }
    ProfScope prof("lincomb", 8.0 * (double)n * (m + k) + (beta == 0.0 ? 0.0 : 8.0 * (double)n * k));
    if (use_blas() && m > 16 && k > 16) {
        // Y = beta Y + V C, C^T stored column-major in d_Ct (leading dimension ldct)
        const double one = 1.0;
        if (rocblas_dgemm(blas_handle(), rocblas_operation_none, rocblas_operation_transpose, n, k, m, &one, V,
                          ldv, d_Ct, ldct, &beta, Y, ldy) != rocblas_status_success)
            fail_msg("rocblas_dgemm failed");
        return;
    }
    const dim3 grid(ceil_div(n, kBlock)), block(kBlock);
    hipStream_t st = state().stream;
    const int mode = lincomb_mode();
    // matrix cores for 9..16 and more than 32 columns: 64 rows x 16 columns
    // or 32 rows x 64 columns per wavefront; measured on MI355X (n = 250047):
    // m = 70, k = 10: 27.7 us against 34.1 (vector kernel); m = 90, k = 50:
    // 113.7 against 136.3; m = 80, k = 30: 57.5 against 47.9, so the vector
    // kernel keeps 17..32
    const int lds_mode = lincomb_lds_mode();
    if (mode == 0 && lds_mode && k > 32 && m >= 32) {
        // V staged through LDS: 128 rows x up to 128 columns per workgroup
        const int tjw = std::min(k, 128) > 64 ? 2 : 1;
        const dim3 g(ceil_div(n, 128), ceil_div(k, 128));
        if (tjw == 2)
            hipLaunchKernelGGL((k_lincomb_lds<2, 8>), g, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y,
                               ldy, k);
        else
            hipLaunchKernelGGL((k_lincomb_lds<1, 8>), g, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y,
                               ldy, k);
        HIP_CHECK(hipGetLastError());
        return;
    }
    if (mode == 0 && k > 8 && (k <= 16 || k > 32)) {
        const int wpb = kBlock / kWave;
        if (k <= 16) {
            hipLaunchKernelGGL((k_lincomb_mfma<1, 4>), dim3(ceil_div(n, 64 * wpb)), block, 0, st, n, V, ldv, m,
                               d_Ct, ldct, beta, Y, ldy, k);
        } else {
            hipLaunchKernelGGL((k_lincomb_mfma<4, 2>), dim3(ceil_div(n, 32 * wpb), ceil_div(k, 64)), block, 0, st,
                               n, V, ldv, m, d_Ct, ldct, beta, Y, ldy, k);
        }
        HIP_CHECK(hipGetLastError());
        return;
    }
    switch (lincomb_group(k)) {
    case 8: hipLaunchKernelGGL(k_lincomb<8>, grid, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y, ldy, k); break;
    case 16: hipLaunchKernelGGL(k_lincomb<16>, grid, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y, ldy, k); break;
    case 32: hipLaunchKernelGGL(k_lincomb<32>, grid, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y, ldy, k); break;
    default: hipLaunchKernelGGL(k_lincomb<64>, grid, block, 0, st, n, V, ldv, m, d_Ct, ldct, beta, Y, ldy, k); break;
    }
    HIP_CHECK(hipGetLastError());
}

void axpby(int n, double a, const double *X, long ldx, double b, double *Y, long ldy, int k)
{
    if (k <= 0 || n <= 0) return;
    ProfScope prof("axpby");
    hipLaunchKernelGGL(k_axpby, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream, n,
                       a, X, ldx, b, Y, ldy);
    HIP_CHECK(hipGetLastError());
}

void axpby_list(int n, const double *d_a, const double *X, long ldx, const int *d_xmap,
                const double *d_b, double *Y, long ldy, const int *d_ymap, int k)
{
    if (k <= 0 || n <= 0) return;
    ProfScope prof("axpby");
    hipLaunchKernelGGL(k_axpby_list, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream,
                       n, d_a, X, ldx, d_xmap, d_b, Y, ldy, d_ymap);
    HIP_CHECK(hipGetLastError());
}

void swap_cols(int n, double *X, long ldx, double *Y, long ldy, int k)
{
    if (k <= 0 || n <= 0) return;
    ProfScope prof("swap");
    hipLaunchKernelGGL(k_swap, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream, n, X,
                       ldx, Y, ldy);
    HIP_CHECK(hipGetLastError());
}

void random_fill(int n, double *X, long ldx, int col0, int k, uint64_t seed, uint64_t draw)
{
    if (k <= 0 || n <= 0) return;
    ProfScope prof("random");
    hipLaunchKernelGGL(k_random, dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0, state().stream, n,
                       X, ldx, col0, seed, draw);
    HIP_CHECK(hipGetLastError());
}

template <typename TA, typename TB>
static void coldots_t(int n, const TA *A, long lda, const int *d_amap, const TB *B, long ldb,
                      const int *d_bmap, int k, double *d_out)
{
    if (k <= 0) return;
    const int rows = reduction_rows(n, k);
    const int nbx = std::max(1, ceil_div(n, rows));
    double *partial = partial_buffer((size_t)nbx * k);
    ProfScope prof("coldots");
    hipLaunchKernelGGL((k_coldots<TA, TB>), dim3(nbx, k), dim3(kBlock), 0, state().stream, n, rows,
                       A, lda, d_amap, B, ldb, d_bmap, partial);
    hipLaunchKernelGGL(k_reduce_rows, dim3(k), dim3(kBlock), 0, state().stream, partial, nbx, d_out);
    HIP_CHECK(hipGetLastError());
}

void coldots(int n, const double *A, long lda, const int *d_amap, const double *B, long ldb,
             const int *d_bmap, int k, double *d_out)
{
    coldots_t(n, A, lda, d_amap, B, ldb, d_bmap, k, d_out);
}

void residual_sq(int n, const double *AX, long ldax, const double *BX, long ldbx, const double *X,
                 long ldx, const double *d_lam, int k, double *d_out)
{
    if (k <= 0) return;
    const int rows = reduction_rows(n, k);
    const int nbx = std::max(1, ceil_div(n, rows));
    double *partial = partial_buffer((size_t)nbx * 3 * k);
    ProfScope prof("residual");
    hipLaunchKernelGGL(k_residual, dim3(nbx, k), dim3(kBlock), 0, state().stream, n, rows, AX, ldax,
                       BX, ldbx, X, ldx, d_lam, partial);
    hipLaunchKernelGGL(k_reduce_rows, dim3(3 * k), dim3(kBlock), 0, state().stream, partial, nbx,
                       d_out);
    HIP_CHECK(hipGetLastError());
}

// every column of a block with leading dimension ld starts 16-byte aligned
template <typename T>
static bool quad_ok(const T *base, long ld)
{
    return ((uintptr_t)base % 16) == 0 && ((size_t)ld * sizeof(T)) % 16 == 0;
}

template <typename PT, typename RT>
static void bcg_p_t(int n, PT *P, long ldp, const int *d_pmap, const RT *R, long ldr,
                    const int *d_rmap, const double *d_beta, const double *d_sold,
                    const double *d_snew, int first, int k, const int *d_count)
{
    if (k <= 0) return;
    ProfScope prof("bcg_p");
    const int rows = stream_rows(n, k);
    constexpr bool f32 = std::is_same<PT, float>::value;
    if (quad_ok(P, ldp) && quad_ok(R, ldr)) {
        if (f32 && stream_cfg().nt)
            hipLaunchKernelGGL((k_bcg_p4<PT, RT, f32>), dim3(ceil_div(n, rows), k), dim3(kBlock), 0,
                               state().stream, n, P, ldp, d_pmap, R, ldr, d_rmap, d_beta, d_sold,
                               d_snew, first, d_count, rows);
        else
            hipLaunchKernelGGL((k_bcg_p4<PT, RT, false>), dim3(ceil_div(n, rows), k), dim3(kBlock), 0,
                               state().stream, n, P, ldp, d_pmap, R, ldr, d_rmap, d_beta, d_sold,
                               d_snew, first, d_count, rows);
    } else
        hipLaunchKernelGGL((k_bcg_p<PT, RT>), dim3(ceil_div(n, rows), k), dim3(kBlock), 0,
                           state().stream, n, P, ldp, d_pmap, R, ldr, d_rmap, d_beta, d_sold, d_snew,
                           first, d_count, rows);
    HIP_CHECK(hipGetLastError());
}

void bcg_p_update(int n, double *P, long ldp, const int *d_pmap, const double *R, long ldr,
                  const int *d_rmap, const double *d_beta, int first, int k, const int *d_count)
{
    bcg_p_t(n, P, ldp, d_pmap, R, ldr, d_rmap, d_beta, (const double *)nullptr,
            (const double *)nullptr, first, k, d_count);
}

void bcg_p_update(int n, float *P, long ldp, const int *d_pmap, const float *R, long ldr,
                  const int *d_rmap, const double *d_beta, const double *d_sold,
                  const double *d_snew, int first, int k, const int *d_count)
{
    bcg_p_t(n, P, ldp, d_pmap, R, ldr, d_rmap, d_beta, d_sold, d_snew, first, k, d_count);
}

void coldots_pw(int n, const float *P, long ldp, const int *d_pmap, const float *W, long ldw,
                const int *d_wmap, int k, double *d_out)
{
    if (k <= 0) return;
    const int rows = reduction_rows(n, k);
    const int nbx = std::max(1, ceil_div(n, rows));
    double *partial = partial_buffer((size_t)nbx * 2 * k);
    ProfScope prof("coldots");
    hipLaunchKernelGGL((k_coldots_pw<float>), dim3(nbx, k), dim3(kBlock), 0, state().stream, n,
                       rows, P, ldp, d_pmap, W, ldw, d_wmap, partial);
    hipLaunchKernelGGL(k_reduce_rows, dim3(2 * k), dim3(kBlock), 0, state().stream, partial, nbx,
                       d_out);
    HIP_CHECK(hipGetLastError());
}

template <typename PT, typename RT, typename XT>
static void bcg_xr_t(int n, XT *X, long ldx, const int *d_xmap, RT *R, long ldr,
                     const int *d_rmap, const PT *P, long ldp, const int *d_pmap, const PT *W,
                     long ldw, const int *d_wmap, const double *d_rho2, const double *d_ptw,
                     const double *d_scale, int k, double *d_rr_out, const int *d_count,
                     const double **rr_part, int *rr_nbx)
{
    if (k <= 0) return;
    const int rows = stream_rows(n, k);
    const int nbx = std::max(1, ceil_div(n, rows));
    double *partial = partial_buffer((size_t)nbx * k);
    ProfScope prof("bcg_xr");
    // rows is a multiple of 4
    constexpr bool f32 = std::is_same<PT, float>::value;
    if (quad_ok(X, ldx) && quad_ok(R, ldr) && quad_ok(P, ldp) && quad_ok(W, ldw)) {
        if (f32 && stream_cfg().nt)
            hipLaunchKernelGGL((k_bcg_xr4<PT, RT, XT, f32>), dim3(nbx, k), dim3(kBlock), 0,
                               state().stream, n, rows, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap,
                               W, ldw, d_wmap, d_rho2, d_ptw, d_scale, partial, d_count);
        else
            hipLaunchKernelGGL((k_bcg_xr4<PT, RT, XT, false>), dim3(nbx, k), dim3(kBlock), 0,
                               state().stream, n, rows, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap,
                               W, ldw, d_wmap, d_rho2, d_ptw, d_scale, partial, d_count);
    } else
        hipLaunchKernelGGL((k_bcg_xr<PT, RT, XT>), dim3(nbx, k), dim3(kBlock), 0, state().stream, n,
                           rows, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap, W, ldw, d_wmap, d_rho2,
                           d_ptw, d_scale, partial, d_count);
    if (rr_part) {
        *rr_part = partial;
        *rr_nbx = nbx;
    } else {
        hipLaunchKernelGGL(k_reduce_cols_wave, dim3(ceil_div(k, kBlock / kWave)), dim3(kBlock), 0,
                           state().stream, partial, nbx, k, d_rr_out);
    }
    HIP_CHECK(hipGetLastError());
}

void bcg_xr_update(int n, double *X, long ldx, const int *d_xmap, double *R, long ldr,
                   const int *d_rmap, const double *P, long ldp, const int *d_pmap,
                   const double *W, long ldw, const int *d_wmap, const double *d_rho2,
                   const double *d_ptw, int k, double *d_rr_out, const int *d_count,
                   const double **rr_part, int *rr_nbx)
{
    bcg_xr_t(n, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap, W, ldw, d_wmap, d_rho2, d_ptw,
             (const double *)nullptr, k, d_rr_out, d_count, rr_part, rr_nbx);
}

void bcg_xr_update(int n, double *X, long ldx, const int *d_xmap, float *R, long ldr,
                   const int *d_rmap, const float *P, long ldp, const int *d_pmap, const float *W,
                   long ldw, const int *d_wmap, const double *d_rho2, const double *d_ptw,
                   const double *d_scale, int k, double *d_rr_out, const int *d_count,
                   const double **rr_part, int *rr_nbx)
{
    bcg_xr_t(n, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap, W, ldw, d_wmap, d_rho2, d_ptw,
             d_scale, k, d_rr_out, d_count, rr_part, rr_nbx);
}

void bcg_xr_update(int n, float *X, long ldx, const int *d_xmap, float *R, long ldr,
                   const int *d_rmap, const float *P, long ldp, const int *d_pmap, const float *W,
                   long ldw, const int *d_wmap, const double *d_rho2, const double *d_ptw,
                   const double *d_scale, int k, double *d_rr_out, const int *d_count,
                   const double **rr_part, int *rr_nbx)
{
    bcg_xr_t(n, X, ldx, d_xmap, R, ldr, d_rmap, P, ldp, d_pmap, W, ldw, d_wmap, d_rho2, d_ptw,
             d_scale, k, d_rr_out, d_count, rr_part, rr_nbx);
}

template <typename T>
static void axpy_cols_t(int n, const double *d_alpha, const T *X, long ldx, const int *d_xmap,
                        T *Y, long ldy, const int *d_ymap, int k)
{
    if (k <= 0) return;
    ProfScope prof("axpy_cols");
    hipLaunchKernelGGL((k_axpy_cols<T>), dim3(ceil_div(n, kBlock), k), dim3(kBlock), 0,
                       state().stream, n, d_alpha, X, ldx, d_xmap, Y, ldy, d_ymap);
    HIP_CHECK(hipGetLastError());
}

void axpy_cols(int n, const double *d_alpha, const double *X, long ldx, const int *d_xmap,
               double *Y, long ldy, const int *d_ymap, int k)
{
    axpy_cols_t(n, d_alpha, X, ldx, d_xmap, Y, ldy, d_ymap, k);
}

void axpy_cols(int n, const double *d_alpha, const float *X, long ldx, const int *d_xmap,
               float *Y, long ldy, const int *d_ymap, int k)
{
    axpy_cols_t(n, d_alpha, X, ldx, d_xmap, Y, ldy, d_ymap, k);
}

}  // namespace gcge_hip
