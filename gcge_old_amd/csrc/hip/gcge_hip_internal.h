// Internal declarations shared by the device kernels and the backend glue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gcge_hip {

// ---- error handling --------------------------------------------------------
void fail(const char *what, hipError_t err, const char *file, int line);
#define HIP_CHECK(expr)                                                       \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) ::gcge_hip::fail(#expr, e_, __FILE__, __LINE__); \
    } while (0)
void fail_msg(const char *msg);

// ---- process-wide state ------------------------------------------------------
struct State {
    int device = -1;
    hipStream_t stream = nullptr;
    // device scratch, grown on demand
    double *d_partial = nullptr;  // reduction partials
    size_t partial_cap = 0;       // doubles
    double *d_small = nullptr;    // small results / coefficient matrices
    size_t small_cap = 0;         // doubles
    double *d_small2 = nullptr;
    size_t small2_cap = 0;
    int *d_index = nullptr;       // column index lists
    size_t index_cap = 0;         // ints
    void *d_rowblk = nullptr;     // row-major copy of a column block (irregular products)
    size_t rowblk_cap = 0;        // bytes
    double *d_bcg = nullptr;      // block-CG bookkeeping state
    size_t bcg_cap = 0;           // doubles
    int *d_bcgi = nullptr;
    size_t bcgi_cap = 0;          // ints
    // pinned host staging
    double *h_result = nullptr;
    size_t result_cap = 0;
    static const int kRing = 8;
    double *h_ring[kRing] = {};
    size_t ring_cap[kRing] = {};
    hipEvent_t ring_event[kRing] = {};
    int ring_next = 0;
    uint64_t seed = 0x5eed;
    uint64_t draw = 0;
    int cus = 0;  // compute units of the device (0: not queried yet)
    bool profiling = false;
    int profile_mode = 0;  // 1: every launch, 2: SpMM launches only
};
State &state();
int ensure_init();

double *partial_buffer(size_t doubles);
double *small_buffer(size_t doubles);
double *small_buffer2(size_t doubles);
int *index_buffer(size_t ints);
void *rowblock_buffer(size_t bytes);
double *result_buffer(size_t doubles);
// copy host data into a device buffer through the pinned ring (async, safe
// against reuse of the staging slot)
void upload_async(void *dst, const void *src, size_t bytes);

// ---- profiling -----------------------------------------------------------------
// Brackets one launch with HIP events on the backend stream while profiling
// is enabled; `bytes` is the launch's algorithmic HBM traffic.  With
// packet = true the caller passes start_event()/stop_event() to
// hipExtLaunchKernelGGL, which records them in the dispatch packet itself
// (no separate event packets around the kernel, so the interval is the
// kernel's own execution time).
// set the algorithmic bytes of a recorded launch (slot from ProfScope)
void prof_set_bytes(int slot, double bytes);
struct ProfScope {
    int slot;
    bool packet;
    explicit ProfScope(const char *name, double bytes = 0.0, bool packet = false);
    hipEvent_t start_event() const;
    hipEvent_t stop_event() const;
    ~ProfScope();
};

// ---- matrices --------------------------------------------------------------------
// Diagonal storage: offsets col - row of the entries, sorted, grouped in runs
// {c - 1, c, c + 1} around a centre c; slot[g][p] is the diagonal holding
// offset c - 1 + p, or -1 when the matrix has no such offset.
constexpr int kDiaMaxGroups = 9;
struct DiaDesc {
    int ngroups;  // 0: the matrix has no diagonal storage
    int center[kDiaMaxGroups];
    int slot[kDiaMaxGroups][3];
};

struct Mat {
    uint32_t magic;
    int nrows, ncols, nnz;
    int nslices;
    int *d_slice_ptr;   // nslices + 1
    int *d_slice_w;     // nslices
    int *d_col;         // sliced ELLPACK column indices
    double *d_val;      // float64 values
    float *d_val32;     // float32 copy (mixed precision), or null
    size_t sell_len;    // padded entries
    int fp32;           // use d_val32 in the inner solves
    int *d_csrpos;      // CSR position of each sliced-ELLPACK entry, -1 for padding
    // diagonal storage (square matrices whose offsets fit kDiaMaxGroups runs
    // of three): d_dia[s * nslices * 64 + row] = A(row, row + offset s), zero
    // where the row has no such entry; used by the SpMM when use_dia is set
    DiaDesc dia;
    int dia_plane;  // rows per plane of a 27-point pattern (centers a N + b S,
                    // a, b in {-1, 0, 1}, S = dia_plane dividing nrows), else 0
    int dia_order;  // row blocks of the diagonal products: 0 automatic (plane
                    // tiles from kTileMinCols columns, else plane order from a
                    // plane stride of 8192 rows), 1 consecutive, 2 plane order
                    // whenever dia_plane > 0, 3 plane tiles whenever possible
    // the same layout as a 27-point plane stencil (dia_plane > 0): line
    // stride N and, for offset a N + b S + d (a, b, d in {-1, 0, 1}), the
    // diagonal holding it in tile27[9 (b + 1) + 3 (a + 1) + d + 1] (-1: none);
    // dia_tile_ok when every stored diagonal has its place there
    int dia_N;
    int dia_tile_ok;
    int tile27[27];
    int ndiag;
    double *d_dia;
    float *d_dia32;
    int use_dia;
    // scattered columns (fewer than half of the stored entries sit in
    // sliced-ELLPACK entry rows spanning under 256 columns): wide products
    // gather whole rows of a row-major copy of X instead of single elements
    // (use_rows)
    int irregular;
    int use_rows;
};
constexpr uint32_t kMatMagic = 0x4d415431u;  // "MAT1"

// ---- vectors and multivecs ---------------------------------------------------------
struct MultiVec;
struct Vec {
    double *data;
    int n;
    int owner;          // 1: data allocated by this vector
    MultiVec *parent;   // for column views
    int col;
};
struct MultiVec {
    uint32_t magic;
    int n, ncols;
    long ld;
    double *data;
    Vec *views;
};
constexpr uint32_t kMultiVecMagic = 0x4d564531u;  // "MVE1"
constexpr int kAlign = 64;

inline long padded_ld(int n) { return ((long)n + kAlign - 1) / kAlign * kAlign; }
MultiVec *as_mv(void **handle);
Vec *as_vec(void *handle);

// rows per plane the diagonal products use (0: consecutive row blocks)
int dia_plane_rows(const Mat *A);

// ---- kernel launchers (stream-ordered, asynchronous) ---------------------------------
// Y(:, ymap[c]) = A X(:, xmap[c]) for c < k; maps may be null (identity).
void spmm(const Mat *A, const double *X, long ldx, double *Y, long ldy, int k,
          const int *d_xmap = nullptr, const int *d_ymap = nullptr);
// float32 form on the float32 copy of the values (Mat::d_val32)
void spmm(const Mat *A, const float *X, long ldx, float *Y, long ldy, int k,
          const int *d_xmap = nullptr, const int *d_ymap = nullptr);
void to_float(const double *src, float *dst, size_t len);
void to_double(const float *src, double *dst, size_t len);
// X(:, c) += D(:, c), c < k (float32 into float64); D = 0 (float32 block)
void add_float(int n, const float *D, long ldd, double *X, long ldx, int k);
void zero_float(int n, float *X, long ldx, int k);
// block-CG product: W(:, wmap[c]) = A P(:, pmap[c]) and, from the same pass,
// d_out[c] = P(:, pmap[c]) . W(:, wmap[c]); the float32 form also gives
// d_out[k + c] = P(:, pmap[c]) . P(:, pmap[c]) (sums in float64)
// d_count (optional): the launch covers columns c < min(k, *d_count), the
// list length kept on the device (block CG); prof_slot (optional) receives
// the profile record of the product for a later byte correction
void spmm_dot(const Mat *A, const double *P, long ldp, double *W, long ldw, int k,
              const int *d_pmap, const int *d_wmap, double *d_out,
              const int *d_count = nullptr, int *prof_slot = nullptr);
void spmm_dot(const Mat *A, const float *P, long ldp, float *W, long ldw, int k, const int *d_pmap,
              const int *d_wmap, double *d_out, const int *d_count = nullptr,
              int *prof_slot = nullptr);
// device d_out (m x k, leading dimension m) = V^T W; the partial sums of the
// row ranges go to `partial` (caller memory of partial_cap doubles: fewer
// ranges when needed) or, when partial is null, to the backend's grown
// buffer.  profile = false records no profiling events.  Returns the number
// of row ranges, or -1 when partial_cap cannot hold one m x k block.
int gram_impl(int n, const double *V, long ldv, int m, const double *W, long ldw, int k,
              double *d_out, double *partial, size_t partial_cap, bool profile);
// host G(m x k, ldg) = V^T W
void gram(int n, const double *V, long ldv, int m, const double *W, long ldw, int k, double *G,
          int ldg, bool sync_result = true);
// Y = beta Y + V C with Ct = C stored row by row on the device (ldct >= k,
// a multiple of lincomb_group(k), padding columns zero)
int lincomb_group(int k);
void lincomb_dev(int n, const double *V, long ldv, int m, const double *d_Ct, int ldct,
                 double beta, double *Y, long ldy, int k);
// Y = V C with Y's columns allowed to overlap V's (same block of columns),
// for the shapes lincomb_in_place_ok accepts
bool lincomb_in_place_ok(int m, int k);
void lincomb_in_place(int n, const double *V, double *Y, long ld, int m, const double *d_Ct, int ldct,
                      int k);
void axpby(int n, double a, const double *X, long ldx, double b, double *Y, long ldy, int k);
void swap_cols(int n, double *X, long ldx, double *Y, long ldy, int k);
// Y(:, ymap[c]) = a[c] X(:, xmap[c]) + b[c] Y(:, ymap[c]), scalars and maps on the device
void axpby_list(int n, const double *d_a, const double *X, long ldx, const int *d_xmap,
                const double *d_b, double *Y, long ldy, const int *d_ymap, int k);
void random_fill(int n, double *X, long ldx, int col0, int k, uint64_t seed, uint64_t draw);
// per-column dots out[c] = A(:, amap[c]) . B(:, bmap[c]) into device memory d_out
void coldots(int n, const double *A, long lda, const int *d_amap, const double *B, long ldb,
             const int *d_bmap, int k, double *d_out);
// residual norms of columns c < k: |AX - lam BX|^2, |X|^2, |BX|^2 -> d_out[3k]
void residual_sq(int n, const double *AX, long ldax, const double *BX, long ldbx, const double *X,
                 long ldx, const double *d_lam, int k, double *d_out);
// block CG pieces (column lists on the device); the direction block P, its
// product W and the residual block R are float64 or float32; x is float64,
// or, in the device-driven float32 form, a float32 correction that is added
// to the float64 initial guess once at the end.  The float32
// directions are stored scaled, p = s p_stored with per-column s (d_sold,
// d_snew, d_scale: one entry per listed column).
void bcg_p_update(int n, double *P, long ldp, const int *d_pmap, const double *R, long ldr,
                  const int *d_rmap, const double *d_beta, int first, int k,
                  const int *d_count = nullptr);
void bcg_p_update(int n, float *P, long ldp, const int *d_pmap, const float *R, long ldr,
                  const int *d_rmap, const double *d_beta, const double *d_sold,
                  const double *d_snew, int first, int k, const int *d_count = nullptr);
// d_out[c] = p_c . w_c, d_out[k + c] = p_c . p_c (float32 blocks, float64 sums)
void coldots_pw(int n, const float *P, long ldp, const int *d_pmap, const float *W, long ldw,
                const int *d_wmap, int k, double *d_out);
void bcg_xr_update(int n, double *X, long ldx, const int *d_xmap, double *R, long ldr,
                   const int *d_rmap, const double *P, long ldp, const int *d_pmap,
                   const double *W, long ldw, const int *d_wmap, const double *d_rho2,
                   const double *d_ptw, int k, double *d_rr_out, const int *d_count = nullptr,
                   const double **rr_part = nullptr, int *rr_nbx = nullptr);
void bcg_xr_update(int n, double *X, long ldx, const int *d_xmap, float *R, long ldr,
                   const int *d_rmap, const float *P, long ldp, const int *d_pmap, const float *W,
                   long ldw, const int *d_wmap, const double *d_rho2, const double *d_ptw,
                   const double *d_scale, int k, double *d_rr_out, const int *d_count = nullptr,
                   const double **rr_part = nullptr, int *rr_nbx = nullptr);
// the same with a float32 x (the correction of the initial guess)
void bcg_xr_update(int n, float *X, long ldx, const int *d_xmap, float *R, long ldr,
                   const int *d_rmap, const float *P, long ldp, const int *d_pmap, const float *W,
                   long ldw, const int *d_wmap, const double *d_rho2, const double *d_ptw,
                   const double *d_scale, int k, double *d_rr_out, const int *d_count = nullptr,
                   const double **rr_part = nullptr, int *rr_nbx = nullptr);
// (r.r per column: rr_part given, the partial sums rr_part[c * rr_nbx + b]
// are left for the bookkeeping kernel instead of being summed into d_rr_out;
// both sum them the same way, lane-strided over one wavefront per column)
// Block-CG bookkeeping on the device (one workgroup): the per-column scalars
// and the list of unconverged columns of the reference loop
// (src/gcge_cg.c:395-490), kept in device memory between steps so that the
// steps need no host round trip.  rho1/rho2 are indexed as the reference
// indexes them (by position in the list, or by column with by_column).
struct BcgControl {
    double *rho1, *rho2, *err, *last_err, *s_col, *pn_col;  // len each
    double *beta, *alpha_num, *s_old, *s_new;               // by position, read by the steps
    const double *ptw;   // p.w of the step [len], p.p [len + id] (scaled directions)
    const double *rr;    // r.r of the step, or the initial r.r (init)
    const double *rr_part;  // or, for a finished step, its partial sums
    int rr_nbx;             // (rr_nbx per column, from bcg_xr_update)
    int *unlock;         // unconverged columns [len]
    int *count;          // number of entries of unlock
    int *maps;           // p, r, x, w column lists [4 len]
    int *hist;           // columns of each step (for the profile byte counts)
    int len, p0, r0, x0, w0, by_column, scaled;
    double rate;
};
constexpr int kBcgControlMax = 1024;  // columns handled by the device bookkeeping
// init: rho2 = rr, err = last_err = sqrt(rr), unlock = columns with err >= 1e-15;
// finished_step > 0: take the r.r of that step and drop converged columns;
// prepare_step > 0: lists and scalars of that step
void bcg_control(const BcgControl &c, int init, int finished_step, int prepare_step);

void axpy_cols(int n, const double *d_alpha, const double *X, long ldx, const int *d_xmap,
               double *Y, long ldy, const int *d_ymap, int k);
void axpy_cols(int n, const double *d_alpha, const float *X, long ldx, const int *d_xmap,
               float *Y, long ldy, const int *d_ymap, int k);

// ---- adjoint (gcge_hip_adjoint.hip) ----------------------------------------------------
// d_grad[p] = sum_c w[c] (X(r, c) Y(q, c) + Y(r, c) X(q, c)) / 2 for the stored
// entry (r, q) at CSR position p of A (d_w: k weights on the device)
void pattern_gradient(const Mat *A, const double *X, long ldx, const double *Y, long ldy, int k,
                      const double *d_w, double *d_grad);
// Q1 elements of the E^3 mesh: d_grad[e] = (h / 36) sum_c w[c] x_c|e^T K36 y_c|e;
// d_node_row (optional): the row of X and Y that holds each interior node
void q1_element_gradient(int E, const double *X, long ldx, const double *Y, long ldy, int k,
                         const double *d_w, double *d_grad, const int *d_node_row = nullptr);

// ---- Q1 assembly and sliced-ELLPACK helpers (gcge_hip_assembly.hip) -----------------------
// which = 0: stiffness, entries (sum of 36 Khat kappa_e) * scale, exact zeros
// dropped; which = 1: mass, (count * 216 Mhat) * scale.  kappa: E^3 device
// doubles (element ex + E ey + E^2 ez) or null (kappa = 1); node_row: device
// map node -> row (null: natural order).
struct Q1Args {
    int E, which;
    double scale;
    const double *kappa;
    const int *node_row;
};
// offsets col - row of diagonal storage, sorted (nd <= 27)
struct DiaOffsets {
    int nd;
    int off[27];
};
// per row: bit o of d_mask = slot o kept, d_big = kept entries with |a| > 1e-20
void q1_row_masks(const Q1Args &q, int *d_mask, int *d_big);
// kept entries of each row at their sliced-ELLPACK positions (columns
// increasing, d_pos = CSR position), and in diagonal storage when d_dia
void q1_fill(const Q1Args &q, const int *d_rowptr, const int *d_sptr, int *d_col, double *d_val,
             int *d_pos, const DiaOffsets &dia, double *d_dia, long dia_ld);
// padding of every slice: zero in the row's own column, CSR position -1
void sell_pad(int nrows, int ncols, int nslices, const int *d_sptr, const int *d_sw, int *d_col,
              double *d_val, int *d_pos);
// d_out[2 s], d_out[2 s + 1]: entries of slice s in entry rows spanning
// < 256 columns, and all its stored entries
void sell_spread(int nslices, const int *d_sptr, const int *d_sw, const int *d_col,
                 const int *d_pos, int *d_out);
// stored entries back at their CSR positions, entries per row
void sell_to_csr(const Mat *A, int *d_col, double *d_val, int *d_rowlen);

}  // namespace gcge_hip
