// Internal declarations shared by the translation units of
// libmldainv_hip.so.  Not part of the public ABI (include/mld_ainv_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mld_ainv_hip.h"

namespace ainv {

// Error carrying an AINV_ERR_* code; converted to an info code at the
// C-ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define AINV_HIP_CHECK(expr)                                                \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::ainv::Error(AINV_ERR_HIP, std::string(#expr) + ": " +         \
                                            hipGetErrorString(e_));         \
  } while (0)

// Host CSR, 0-based, rows sorted by column.
struct HostCsr {
  int32_t n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  std::vector<double> val;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// Owned device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b);
  void release();
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Device CSR, FP64 values (the build output kept for export / re-mold).
struct DevCsr {
  int32_t n = 0;
  int64_t nnz = 0;
  DevBuf ptr, idx, val;
};

// Sliced ELLPACK ("HLL"): slices of 64 consecutive rows, one per
// wavefront; inside a slice entries are stored column-major
// (entry j of row 64*s + l at sptr[s] + 64*j + l) so that the 64 lanes of a
// wave read consecutive words.  Padding entries carry value 0 and the row's
// own index as column.  rlen holds each stored row's true length.
// Optionally the rows are sorted by length within windows of `sigma` rows
// (a multiple of 64) before they are cut into slices, so that a slice holds
// rows of like length and pads little (SELL-C-sigma, Kreutzer et al., SIAM
// J. Sci. Comput. 36(5), 2014): perm then maps stored row i (slice i / 64,
// lane i % 64) to the matrix row it holds, n for the rows past n of the last
// slice.  Entries keep their order within a row, so every row sum is formed
// as in the unsorted layout.
constexpr int kSlice = 64;
// slices per workgroup of the apply kernels (one wavefront each), and the
// XCDs a launch's workgroups are dealt to round-robin
constexpr int kSliceGroup = 4;
constexpr int kXcdCount = 8;
struct Hll {
  int32_t n = 0;
  int32_t ncols = 0;  // columns (rows of the x it multiplies)
  int32_t nslices = 0;
  int64_t nnz = 0;     // true entries
  int64_t padded = 0;  // stored entries including padding
  int32_t max_width = 0;
  int prec = AINV_PREC_F64;
  DevBuf sptr;  // int64 [nslices + 1]
  DevBuf rlen;  // int32 [n]
  DevBuf col;   // int32 [padded]
  DevBuf val;   // VT    [padded]
  // optional CSR copy (row pointers, columns, VT values) for the
  // one-wavefront-per-row mapping AINV_MAP_WIDE (long rows on few rows)
  DevBuf cptr, cidx, cval;
  // optional order in which the run mapping's workgroups take their groups
  // of 4 slices (int32 [ceil(nslices / 4)], a permutation within each XCD's
  // range; empty: linear), set by hll_set_order: strips of `band` groups
  // taken across `period`-group planes
  DevBuf gorder;
  int32_t order_period = 0, order_band = 0;
  // row order of sorted slices (int32 [nslices * 64]; empty: rows in order)
  // and its window in rows (0: unsorted)
  DevBuf perm;
  int32_t sigma = 0;
  int64_t unsorted_padded = 0;  // stored entries the unsorted layout needs
  // slice widths (entries per row slot): mean and percentiles, the load
  // balance of the slices' waves
  double width_mean = 0;
  int32_t width_p50 = 0, width_p99 = 0;
  int64_t bytes() const {
    return (int64_t)(sptr.bytes + rlen.bytes + col.bytes + val.bytes +
                     cptr.bytes + cidx.bytes + cval.bytes + gorder.bytes +
                     perm.bytes);
  }
};

int value_bytes(int prec);

// ---- device-side helpers (hll.hip) --------------------------------------
void csr_upload(const HostCsr &h, DevCsr &d, hipStream_t s);
void csr_download(const DevCsr &d, HostCsr &h, hipStream_t s);
// d_i = 1 / (1/d_i + sum_k W(i,k) x_k Z(k,i)) with zt = Z^T (hll.hip)
void update_pivots(const DevCsr &w, const DevCsr &zt, const double *x,
                   double *d, hipStream_t s);

// MatrixMarket text I/O (mmio.cpp)
void mm_write_csr(const std::string &path, const HostCsr &a, const char *head);
void mm_write_vector(const std::string &path, const std::vector<double> &v,
                     const char *head);
HostCsr mm_read_csr(const std::string &path);
std::vector<double> mm_read_vector(const std::string &path);
// Transpose of a device CSR (rows sorted by column in the result).
void csr_transpose(const DevCsr &a, DevCsr &t, hipStream_t s);
// HLL from device CSR, values converted to `prec`.  ncols: column count of
// a rectangular matrix (padding entries point at a column below it); < 0
// for a square one.
// Median over slices of a middle row's reach (the larger of its distance
// to its first and to its last column), in groups of 4 slices (hll.hip).
int32_t hll_reach_groups(const Hll &h, hipStream_t s);
// Slice-group order of the run mapping: the groups of each of the 8 XCD
// ranges taken in strips -- `band` consecutive groups of every
// `period`-group plane of the range, then the next band -- so that the x
// rows a band's planes share (the factor's clusters of diagonals one and
// two planes apart) are gathered while still in the XCD's L2.  period or
// band <= 0: linear order (hll.hip).
void hll_set_order(Hll &h, int32_t period, int32_t band, hipStream_t s);
// sigma: window of the row sort in rows (a multiple of 64; 0: rows in
// order; kSortAuto: the window hll_pick_sigma picks from the row lengths).
constexpr int32_t kSortAuto = -1;
void hll_from_csr(const DevCsr &a, int prec, Hll &h, hipStream_t s,
                  int32_t ncols = -1, int32_t sigma = 0);
// Window of the row sort for rows of the given lengths: 0 when sorting
// saves less than a fifth of the unsorted layout's stored entries or the
// layout pads less than a quarter; else the smallest of 256, 1024 and 4096
// rows whose stored entries are within 5 % of the best of the three.
int32_t hll_pick_sigma(const std::vector<int32_t> &len);
// Stored entries of slices of 64 rows of these lengths, the rows sorted by
// length (longest first, ties in row order) within windows of sigma rows
// (0: in order); perm (optional) receives the stored order, n past the
// last row.
int64_t hll_sorted_padding(const std::vector<int32_t> &len, int32_t sigma,
                           std::vector<int32_t> *perm = nullptr);
// {"nnz", "padded", "unsorted_padded", "sort_window", "slices",
//  "max_width", "width_mean", "width_p50", "width_p99"} of a layout.
std::string hll_layout_json(const Hll &h);
// Attaches the CSR copy of `a` (host, 0-based, rows sorted) to h, values in
// h's precision, for AINV_MAP_WIDE.
void hll_attach_csr(const HostCsr &a, Hll &h, hipStream_t s);
// Plain 16-byte-vector copy kernel (bandwidth reference).
void stream_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
// Convert n doubles to the storage type of prec (d vector).
void convert_vector(const double *src, int64_t n, int prec, DevBuf &dst,
                    hipStream_t s);

// ---- kernels launched by the apply / Krylov code (apply.hip) -----------
// out = d o (A x)          (EPI scale, out type by prec: F64 double, else
// float) or
// y   = alpha * A x + beta*y (EPI axpby)  for k interleaved columns.
// map: AINV_MAP_SPLIT (lanes split rows x columns, any ld),
// AINV_MAP_ROW (one row per lane, k <= 32) or AINV_MAP_COOP /
// AINV_MAP_COOP128 (the lanes of a row group load each x row together;
// blocks whose x rows exceed 256 / 128 bytes run as column chunks, k <= 64);
// all but split need ldx = ldo = K = pow2 >= k and 16-byte aligned vectors,
// otherwise the split mapping runs.
constexpr int AINV_MAP_SPLIT = 0;
constexpr int AINV_MAP_ROW = 1;
constexpr int AINV_MAP_COOP = 2;
constexpr int AINV_MAP_COOP128 = 3;
// the cooperative mappings with four waves sharing each slice (more waves
// for matrices of few slices; the fused system-matrix product runs them as
// COOP / COOP128)
constexpr int AINV_MAP_COOP_Q = 4;
constexpr int AINV_MAP_COOP128_Q = 5;
// row groups of consecutive rows that hand gathered x rows on along runs of
// consecutive diagonals, gathering only what they do not hold
constexpr int AINV_MAP_RUN = 6;
// one wavefront per row over the matrix's CSR copy (Hll::cptr), for the
// multilevel products whose rows are long and few
constexpr int AINV_MAP_WIDE = 7;
constexpr int kNumMaps = 8;
// mappings of the factor applies (no CSR copy)
constexpr int kNumApplyMaps = 7;
// mappings the fused system-matrix product distinguishes (timed by the
// autotuner in this order)
constexpr int kNumSpmvMaps = 5;
constexpr int kSpmvMaps[kNumSpmvMaps] = {AINV_MAP_SPLIT, AINV_MAP_ROW,
                                         AINV_MAP_COOP, AINV_MAP_COOP128,
                                         AINV_MAP_RUN};
// The two apply passes may use different mappings: a packed pair.
inline int map_pair(int w, int z) { return w | (z << 4); }
inline int map_w(int pair) { return pair & 15; }
inline int map_z(int pair) { return (pair >> 4) & 15; }
// kw (device pointer, may be NULL): active width of the block; the
// cooperative mappings neither read nor write columns at or beyond *kw.
void hll_spmm_scale(const Hll &a, const void *d, const void *x, bool x_is_f32,
                    int64_t ldx, void *out, int64_t ldo, int k, hipStream_t s,
                    int map = AINV_MAP_SPLIT, const int32_t *kw = nullptr);
void hll_spmm_axpby(const Hll &a, double alpha, const void *x, bool x_is_f32,
                    int64_t ldx, double beta, double *y, int64_t ldy, int k,
                    hipStream_t s, int map = AINV_MAP_SPLIT,
                    const int32_t *kw = nullptr);
// FP32 y from FP32 x (FP32 / BF16 factors only)
void hll_spmm_axpby(const Hll &a, double alpha, const void *x, bool x_is_f32,
                    int64_t ldx, double beta, float *y, int64_t ldy, int k,
                    hipStream_t s, int map = AINV_MAP_SPLIT,
                    const int32_t *kw = nullptr);
// y = alpha A x + beta z (x, z, y distinct with leading dimension ld): the
// residual b - A u of the multilevel cycle in one pass.  FP64 vectors with
// an FP64 or FP32 A, or FP32 vectors with an FP32 A.
void hll_spmm_resid(const Hll &a, double alpha, const double *x, double beta,
                    const double *z, double *y, int64_t ld, int k,
                    hipStream_t s, int map = AINV_MAP_SPLIT,
                    const int32_t *kw = nullptr);
void hll_spmm_resid(const Hll &a, double alpha, const float *x, double beta,
                    const float *z, float *y, int64_t ld, int k,
                    hipStream_t s, int map = AINV_MAP_SPLIT,
                    const int32_t *kw = nullptr);
// Block partial sums of the Krylov reductions: part[(q*kRedBlocks + b) *
// kMaxCols + c] for pair q, block b, column c.
constexpr int kRedBlocks = 2048;
// grid of the persistent (grid-stride) system-matrix kernels
constexpr int kPersistBlocks = 1024;
constexpr int kMaxCols = 64;
// y = A x (FP64 HLL, k interleaved columns, ld = k) fused with the dot
// products of its result: np = 1 gives (w, y), np = 2 gives (y, y) and
// (y, w), per column, as block partials.  Returns the number of blocks.
int hll_spmv_dot(const Hll &a, const double *x, double *y, int k, int64_t ld,
                 const double *w, int np, double *part, int map,
                 hipStream_t s, const int32_t *kw = nullptr);
// the same product of an FP32 x (the Krylov engine's FP32 preconditioned
// vectors); y, w and the sums stay FP64
int hll_spmv_dot(const Hll &a, const float *x, double *y, int k, int64_t ld,
                 const double *w, int np, double *part, int map,
                 hipStream_t s, const int32_t *kw = nullptr);

// ---- GPU row inverses (invkt.hip) ----------------------------------------
// Row-wise approximate inverse of a unit upper triangular matrix given by
// its strictly upper part `u` (device CSR).  kind AINV_KIND_INVK uses the
// level-of-fill rule with fill; AINV_KIND_INVT the threshold + cap rule.
void gpu_sparse_inverse(int kind, const DevCsr &u, int32_t fill, double thresh,
                        DevCsr &z, hipStream_t s, std::string *report);

// ---- host builds (build_host.cpp) ----------------------------------------
void host_iluk(const HostCsr &a, int32_t fill, HostCsr &l, HostCsr &u,
               std::vector<double> &d);
void host_ilut(const HostCsr &a, int32_t fill, double thresh, HostCsr &l,
               HostCsr &u, std::vector<double> &d);
// AINV: returns stored W (= W^T, CSR), Z (CSR) and d.
// Host threads of the build phase: AINV_BUILD_THREADS, else the CPUs this
// process may run on (at most 16).
int build_threads();
// How a biconjugation ran: the threads it used and whether its Z and W
// passes ran pipelined over several threads (else the column loop).
struct AinvBuildInfo {
  int threads = 1;
  bool pipelined = false;
};
void host_ainv(const HostCsr &a, int alg, int32_t fill_in, double thresh,
               HostCsr &w, HostCsr &z, std::vector<double> &d,
               int threads = 0, AinvBuildInfo *info = nullptr);

double now_seconds();

// ---- generators (gen.cpp) ---------------------------------------------------
void gen_pde3d(int32_t N, int hconv, HostCsr &a);
void gen_aniso2d(int32_t N, double eps, HostCsr &a);
void gen_random(int32_t n, uint64_t seed, double median, double sigma,
                int32_t cap, int32_t band, int32_t nlong, double dom,
                HostCsr &a);
void gen_normal(uint64_t seed, int64_t count, double *out);

// ---- error plumbing at the C ABI ------------------------------------------
// Message of the last failure on this thread (capi.cpp).
std::string &last_error();

// Runs f, mapping exceptions to info codes and recording the message.
template <class F> int guard(F &&f) {
  try {
    f();
    last_error().clear();
    return AINV_SUCCESS;
  } catch (const Error &e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    last_error() = "host allocation failed";
    return AINV_ERR_ALLOC;
  } catch (const std::exception &e) {
    last_error() = e.what();
    return AINV_ERR_INTERNAL;
  } catch (...) {
    last_error() = "unknown failure";
    return AINV_ERR_INTERNAL;
  }
}

}  // namespace ainv
