// Sparse-times-block kernels on the HLL layout: the two SpMVs of the
// preconditioner apply (mld_d_base_ainv_solver_apply_vect.f90:112-119) and
// the A*x of the Krylov loop.
//
// Mapping: one wavefront per 64-row slice.  The 64 lanes are split into
// 64/K row groups of K lanes; lane (g, c) accumulates row g of the current
// pass for right-hand side c.  K = 1 is the classic one-row-per-lane HLL
// SpMV (factor words coalesced across the wave); for K >= 8 the K lanes of
// a row read x[col*ldx + 0..K) as one contiguous segment and the factor
// word is a broadcast.  Epilogues:
//   SCALE: out[r, c] = d[r] * acc      (first SpMV fused with d o .)
//   AXPBY: y[r, c] = alpha * acc + beta * y[r, c], y not read if beta == 0
// A second mapping ("row"): one row per lane for any K, each lane loading
// the K values x[col*K .. col*K+K) of an entry as 16-byte vectors (ld = K,
// K a power of two).  It issues about K/2 times fewer load instructions per
// entry than the split mapping at the cost of K accumulators per lane; the
// autotuner picks the faster of the two per factor and K.
// The algorithmic traffic per launch is (nnz)(value + 4) bytes of factor,
// 4 bytes of row length per row, the gathered x and the output; no kernel
// here re-reads a factor word.
#include "ainv_internal.h"

namespace ainv {
namespace {

constexpr int kWavesPerBlock = kSliceGroup;
constexpr int kBlock = kWavesPerBlock * kSlice;

// Workgroups are dispatched round-robin over the XCDs (workgroup i runs on
// XCD i mod 8 on MI355X), each XCD with its own L2.  Neighbouring slices
// gather mostly the same x rows (the factor patterns are banded), so the
// kernels renumber workgroups to give each XCD a contiguous range of
// slices: an x row is then fetched into one L2 rather than into all eight.
#ifndef AINV_XCD_REMAP
#define AINV_XCD_REMAP 1
#endif
constexpr int kXcds = kXcdCount;

__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t g) {
#if AINV_XCD_REMAP
  const int32_t x = b % kXcds, l = b / kXcds;
  const int32_t per = g / kXcds, rem = g % kXcds;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
#else
  (void)g;
  return b;
#endif
}

// Persistent kernels (fixed grid, a multiple of kXcds blocks): the slice
// groups of kWavesPerBlock slices are split into kXcds contiguous ranges;
// the blocks of XCD x stride through range x together.
struct GroupRange {
  int32_t first, end, stride;
};
__device__ __forceinline__ GroupRange xcd_groups(int32_t nslices) {
  const int32_t ng = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
#if AINV_XCD_REMAP
  const int32_t x = blockIdx.x % kXcds, nb = gridDim.x / kXcds;
  return GroupRange{(int32_t)((int64_t)ng * x / kXcds) +
                        (int32_t)(blockIdx.x / kXcds),
                    (int32_t)((int64_t)ng * (x + 1) / kXcds), nb};
#else
  return GroupRange{(int32_t)blockIdx.x, ng, (int32_t)gridDim.x};
#endif
}

int persistent_grid(int32_t nslices) {
  int g = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
  if (g > kPersistBlocks) g = kPersistBlocks;
  if (g < 1) g = 1;
#if AINV_XCD_REMAP
  g = (g + kXcds - 1) / kXcds * kXcds;  // kPersistBlocks is a multiple of 8
#endif
  return g;
}

template <typename VT> struct Val;
template <> struct Val<double> {
  using acc = double;
  __device__ static double get(double v) { return v; }
};
// FP32 factors accumulate in FP32 (the factors carry FP32 rounding
// already; the Krylov recurrences around the apply stay FP64).  FP32
// accumulators take half the registers of FP64 ones, which lets the
// 32-column gather kernels run 3 waves per SIMD.
template <> struct Val<float> {
  using acc = float;
  __device__ static float get(float v) { return v; }
};
template <> struct Val<uint16_t> {  // BF16 value, FP32 accumulation
  using acc = float;
  __device__ static float get(uint16_t v) {
    return __uint_as_float((uint32_t)v << 16);
  }
};

template <typename A, typename XT> __device__ inline A cvt(XT x) {
  return (A)x;
}

// Dot product of one stored row with x (column c of k interleaved):
// entries at p0, p0 + 64, ... (HLL column-major slice).  Groups of four
// entries are software-pipelined: the column and value words of group g+1
// are requested before the gathers of group g are consumed, so each wave
// keeps about twelve loads in flight instead of waiting twice per entry.
#ifndef AINV_ROW_GROUP
#define AINV_ROW_GROUP 4
#endif
constexpr int kGroup = AINV_ROW_GROUP;

template <typename VT, typename XT>
__device__ __forceinline__ typename Val<VT>::acc row_dot(
    int64_t p0, int32_t len, const int32_t *__restrict__ col,
    const VT *__restrict__ val, const XT *__restrict__ x, int64_t ldx, int c,
    bool col_ok) {
  using A = typename Val<VT>::acc;
  // no branch inside the loop body (a branch makes the compiler drain every
  // outstanding load at the block boundary): lanes of an idle column read
  // column 0 and discard the sum, and the last group re-reads its own words
  const int cc = col_ok ? c : 0;
  A acc0 = 0;
  int32_t j = 0;
  if (len >= kGroup) {
    int64_t q = p0;
    int32_t ci[kGroup];
    A vi[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      ci[u] = col[q + u * kSlice];
      vi[u] = Val<VT>::get(val[q + u * kSlice]);
    }
    for (;;) {
      A xi[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u)
        xi[u] = cvt<A>(x[(int64_t)ci[u] * ldx + cc]);
      const int32_t jn = j + kGroup;
      const bool more = jn + kGroup - 1 < len;
      q = more ? p0 + (int64_t)jn * kSlice : q;
      int32_t cn[kGroup];
      A vn[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        cn[u] = col[q + u * kSlice];
        vn[u] = Val<VT>::get(val[q + u * kSlice]);
      }
      // one sum in entry order, as every other mapping forms it: a
      // column's result does not depend on the mapping or block width
#pragma unroll
      for (int u = 0; u < kGroup; ++u) acc0 += vi[u] * xi[u];
      j = jn;
      if (!more) break;
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        ci[u] = cn[u];
        vi[u] = vn[u];
      }
    }
  }
  for (; j < len; ++j) {
    const int64_t q = p0 + (int64_t)j * kSlice;
    acc0 += Val<VT>::get(val[q]) * cvt<A>(x[(int64_t)col[q] * ldx + cc]);
  }
  return acc0;
}

// RESID: out[r, c] = alpha * acc + beta * z[r, c], z passed as the d
// argument with the leading dimension of out
enum { EPI_SCALE = 0, EPI_AXPBY = 1, EPI_RESID = 2 };

// Matrix row held by stored row i of a layout whose rows are sorted by
// length within windows (perm: stored row -> row, n past the last row;
// null: rows in order).  Kernels walk stored rows and write matrix rows.
__device__ __forceinline__ int32_t hll_row(const int32_t *__restrict__ perm,
                                           int32_t i) {
  return perm ? perm[i] : i;
}

// The same for the kernels whose lanes store rows of their slice other than
// their own: lane l loads the row of slice row l once per slice (pv, one
// coalesced word, requested before the slice's entries) and a store takes
// slice row j's by a shuffle; rows in order take no load or shuffle.  Call
// with the whole wave active.
__device__ __forceinline__ int32_t slice_rows(const int32_t *__restrict__ perm,
                                              int32_t s, int lane) {
  return perm ? perm[(int64_t)s * kSlice + lane] : 0;
}
__device__ __forceinline__ int32_t slice_row(const int32_t *__restrict__ perm,
                                             int32_t pv, int32_t s, int j) {
  return perm ? __shfl(pv, j) : s * kSlice + j;
}

template <typename VT, typename XT, typename DT, typename OT, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_hll_spmm(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ rlen, const int32_t *__restrict__ col,
    const VT *__restrict__ val, const XT *__restrict__ x, int64_t ldx, int k,
    const DT *__restrict__ dvec, OT *__restrict__ out, int64_t ldo,
    double alpha, double beta, const int32_t *__restrict__ perm) {
  using A = typename Val<VT>::acc;
  constexpr int G = kSlice / K;  // rows per pass
  const int lane = threadIdx.x & (kSlice - 1);
  const int32_t s = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock +
                    (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int64_t base = sptr[s];
  const int c = (K == 1) ? 0 : (lane % K);
  const int g = (K == 1) ? lane : (lane / K);
  const bool col_ok = c < k;
  for (int pass = 0; pass < K; ++pass) {
    const int ri = pass * G + g;
    const int32_t r = s * kSlice + ri;
    if (r >= n) break;  // rows beyond n only occur in the last passes
    const int32_t len = rlen[r];
    const A acc = row_dot<VT, XT>(base + ri, len, col, val, x, ldx, c, col_ok);
    if (!col_ok) continue;
    const int32_t o = hll_row(perm, r);
    if constexpr (EPI == EPI_SCALE) {
      out[(int64_t)o * ldo + c] = (OT)((double)dvec[o] * (double)acc);
    } else if constexpr (EPI == EPI_RESID) {
      out[(int64_t)o * ldo + c] =
          (OT)(alpha * (double)acc + beta * (double)dvec[(int64_t)o * ldo + c]);
    } else {
      double v = alpha * (double)acc;
      OT *yp = out + (int64_t)o * ldo + c;
      if (beta != 0.0) v += beta * (double)(*yp);
      *yp = (OT)v;
    }
  }
}

// K values of one x row as 16-byte vectors when the row is 16-byte aligned
template <typename XT, int K, typename A>
__device__ __forceinline__ void load_xrow(const XT *__restrict__ p, A *out) {
  constexpr int B = K * (int)sizeof(XT);
  if constexpr (B % 16 == 0) {
    constexpr int per = 16 / (int)sizeof(XT);
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int t = 0; t < B / 16; ++t) {
      const uint4 w = q[t];
      XT e[per];
      __builtin_memcpy(e, &w, 16);
#pragma unroll
      for (int u = 0; u < per; ++u) out[t * per + u] = (A)e[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < K; ++u) out[u] = (A)p[u];
  }
}

// Entries of one lane's row, walked over the slice's padded width (a
// wave-uniform trip count: padding entries are (own row, 0) and add
// nothing), in groups of P entries: the x rows of group g are requested
// before the column / value words of group g+1, so each wave keeps the
// gathers of one group and the index loads of the next in flight together
// instead of waiting for two round trips per group.  P keeps the gathered
// values of a group within about 32 registers per lane.
// Waves per SIMD the row kernels are compiled for: at most 4 lets a lane
// hold a whole group of gathered rows (up to 128 registers).
#ifndef AINV_ROW_WAVES
#define AINV_ROW_WAVES 4
#endif
#if AINV_ROW_WAVES > 0
#define AINV_ROW_ATTR \
  __attribute__((amdgpu_waves_per_eu(RowWaves<K>::v, RowWaves<K>::v)))
#else
#define AINV_ROW_ATTR
#endif
#ifndef AINV_ROW_PIPE
#define AINV_ROW_PIPE 1
#endif

#ifndef AINV_ROW_GROUP_BYTES
#define AINV_ROW_GROUP_BYTES 128
#endif
template <typename A, int K> struct RowGroup {
  static constexpr int raw = AINV_ROW_GROUP_BYTES / (K * (int)sizeof(A));
  static constexpr int v = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
};
// K >= 32 needs 2K accumulator registers per lane: compile for 2 waves per
// SIMD (256 registers) instead of spilling.
template <int K> struct RowWaves {
  static constexpr int v = K >= 32 ? 2 : AINV_ROW_WAVES;
};

template <typename VT, typename XT, int K>
__device__ __forceinline__ void row_acc(int64_t p0, int32_t width,
                                        const int32_t *__restrict__ col,
                                        const VT *__restrict__ val,
                                        const XT *__restrict__ x,
                                        typename Val<VT>::acc (&acc)[K]) {
  using A = typename Val<VT>::acc;
  constexpr int P = RowGroup<A, K>::v;
  const int32_t nfull = width / P * P;
  if (nfull > 0) {
    int32_t ci[P];
    A vi[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      ci[u] = col[p0 + (int64_t)u * kSlice];
      vi[u] = Val<VT>::get(val[p0 + (int64_t)u * kSlice]);
#if AINV_ROW_PIPE
      // opaque copies: otherwise the compiler merges these loads with the
      // prefetches of the loop into one load at the top of each iteration,
      // which undoes the pipelining
      asm volatile("" : "+v"(ci[u]));
      asm volatile("" : "+v"(vi[u]));
#endif
    }
    for (int32_t j = 0;;) {
      A xv[P][K];
#pragma unroll
      for (int u = 0; u < P; ++u)
        load_xrow<XT, K>(x + (int64_t)ci[u] * K, xv[u]);
      const int32_t jn = j + P;
      const bool more = jn < nfull;
      // last group: re-read its own words rather than branch around the load
      const int64_t q = p0 + (int64_t)(more ? jn : j) * kSlice;
      int32_t cn[P];
      VT vn[P];  // converted after the sums: no wait on them before
#pragma unroll
      for (int u = 0; u < P; ++u) {
        cn[u] = col[q + (int64_t)u * kSlice];
        vn[u] = val[q + (int64_t)u * kSlice];
      }
#pragma unroll
      for (int u = 0; u < P; ++u)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] += vi[u] * xv[u][c];
      if (!more) break;
      j = jn;
#pragma unroll
      for (int u = 0; u < P; ++u) {
        ci[u] = cn[u];
        vi[u] = Val<VT>::get(vn[u]);
      }
    }
  }
  for (int32_t j = nfull; j < width; ++j) {
    const int64_t q = p0 + (int64_t)j * kSlice;
    const A v = Val<VT>::get(val[q]);
    A xv[K];
    load_xrow<XT, K>(x + (int64_t)col[q] * K, xv);
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] += v * xv[c];
  }
}

// Row mapping: lane l of the wave owns row 64 s + l and all K columns.
template <typename VT, typename XT, typename DT, typename OT, int K, int EPI>
__global__ __launch_bounds__(kBlock) AINV_ROW_ATTR void k_hll_spmm_row(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ perm) {
  using A = typename Val<VT>::acc;
  const int lane = threadIdx.x & (kSlice - 1);
  // wave-uniform: scalar loads of the slice bounds, scalar loop control
  const int32_t s = __builtin_amdgcn_readfirstlane(
      xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6));
  if (s >= nslices) return;
  const int64_t base = sptr[s];
  const int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
  A acc[K];
#pragma unroll
  for (int u = 0; u < K; ++u) acc[u] = 0;
  row_acc<VT, XT, K>(base + lane, width, col, val, x, acc);
  if (s * kSlice + lane >= n) return;
  const int32_t r = hll_row(perm, s * kSlice + lane);
  OT *o = out + (int64_t)r * K;
  if constexpr (EPI == EPI_SCALE) {
    const double d = (double)dvec[r];
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (u < k) o[u] = (OT)(d * (double)acc[u]);
  } else if constexpr (EPI == EPI_RESID) {
    const DT *zr = dvec + (int64_t)r * K;
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (u < k) o[u] = (OT)(alpha * (double)acc[u] + beta * (double)zr[u]);
  } else {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      if (u >= k) continue;
      double v = alpha * (double)acc[u];
      if (beta != 0.0) v += beta * (double)o[u];
      o[u] = (OT)v;
    }
  }
}

template <typename VT, typename XT, typename DT, typename OT, int EPI>
void launch_row(const Hll &a, const XT *x, int k, const DT *d, OT *out,
                double alpha, double beta, hipStream_t s) {
  if (a.nslices == 0) return;
  dim3 grid((a.nslices + kWavesPerBlock - 1) / kWavesPerBlock), block(kBlock);
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const VT *vl = a.val.as<VT>();
#define AINV_LAUNCH(KK)                                                      \
  hipLaunchKernelGGL((k_hll_spmm_row<VT, XT, DT, OT, KK, EPI>), grid, block, \
                     0, s, a.n, a.nslices, sp, cl, vl, x, k, d, out, alpha,  \
                     beta, a.perm.as<int32_t>())
  if (k <= 2)
    AINV_LAUNCH(2);
  else if (k <= 4)
    AINV_LAUNCH(4);
  else if (k <= 8)
    AINV_LAUNCH(8);
  else if (k <= 16)
    AINV_LAUNCH(16);
  else
    AINV_LAUNCH(32);
#undef AINV_LAUNCH
  AINV_HIP_CHECK(hipGetLastError());
}

template <typename VT, typename XT, typename DT, typename OT, int EPI>
void launch_k(const Hll &a, const XT *x, int64_t ldx, int k, const DT *d,
              OT *out, int64_t ldo, double alpha, double beta, hipStream_t s) {
  if (a.nslices == 0) return;
  dim3 grid((a.nslices + kWavesPerBlock - 1) / kWavesPerBlock), block(kBlock);
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *rl = a.rlen.as<int32_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const VT *vl = a.val.as<VT>();
#define AINV_LAUNCH(KK)                                                    \
  hipLaunchKernelGGL((k_hll_spmm<VT, XT, DT, OT, KK, EPI>), grid, block, 0, \
                     s, a.n, a.nslices, sp, rl, cl, vl, x, ldx, k, d, out,  \
                     ldo, alpha, beta, a.perm.as<int32_t>())
  if (k <= 1)
    AINV_LAUNCH(1);
  else if (k <= 2)
    AINV_LAUNCH(2);
  else if (k <= 4)
    AINV_LAUNCH(4);
  else if (k <= 8)
    AINV_LAUNCH(8);
  else if (k <= 16)
    AINV_LAUNCH(16);
  else if (k <= 32)
    AINV_LAUNCH(32);
  else
    AINV_LAUNCH(64);
#undef AINV_LAUNCH
  AINV_HIP_CHECK(hipGetLastError());
}

// System-matrix product fused with the dot products of the Krylov step
// that follows it, so that the product is not read back: grid-stride over
// slices with a fixed grid (at most kPersistBlocks blocks), per-column partial
// sums reduced across the row groups of a wave by shuffles and across the
// waves of a block in LDS, in a fixed order.
template <typename XT, int K, int NP>
__global__ __launch_bounds__(kBlock) void k_hll_spmv_dot(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ rlen, const int32_t *__restrict__ col,
    const double *__restrict__ val, const XT *__restrict__ x, int k,
    int64_t ld, double *__restrict__ y, const double *__restrict__ w,
    double *__restrict__ part, const int32_t *__restrict__ perm) {
  constexpr int G = kSlice / K;
  const int lane = threadIdx.x & (kSlice - 1);
  const int wave = threadIdx.x >> 6;
  const int c = (K == 1) ? 0 : (lane % K);
  const int g = (K == 1) ? lane : (lane / K);
  const bool col_ok = c < k;
  double d0 = 0.0, d1 = 0.0;
  const GroupRange gr = xcd_groups(nslices);
  for (int32_t grp = gr.first; grp < gr.end; grp += gr.stride) {
    const int32_t s = grp * kWavesPerBlock + wave;
    if (s >= nslices) break;
    const int64_t base = sptr[s];
    for (int pass = 0; pass < K; ++pass) {
      const int ri = pass * G + g;
      const int32_t r = s * kSlice + ri;
      if (r >= n) break;
      const int32_t len = rlen[r];
      const double yv =
          row_dot<double, XT>(base + ri, len, col, val, x, ld, c, col_ok);
      if (!col_ok) continue;
      const int64_t e = (int64_t)hll_row(perm, r) * ld + c;
      y[e] = yv;
      if constexpr (NP == 1) {
        d0 += w[e] * yv;
      } else {
        d0 += yv * yv;
        d1 += yv * w[e];
      }
    }
  }
  // lanes holding the same column: c, c + K, c + 2K, ...
  for (int off = K; off < kSlice; off <<= 1) {
    d0 += __shfl_xor(d0, off);
    if constexpr (NP == 2) d1 += __shfl_xor(d1, off);
  }
  __shared__ double red[2][kWavesPerBlock][kSlice];
  if (g == 0) {
    red[0][wave][c] = d0;
    red[1][wave][c] = d1;
  }
  __syncthreads();
  if (threadIdx.x < K && threadIdx.x < k) {
    double t0 = 0.0, t1 = 0.0;
    for (int wv = 0; wv < kWavesPerBlock; ++wv) {
      t0 += red[0][wv][threadIdx.x];
      t1 += red[1][wv][threadIdx.x];
    }
    part[((int64_t)0 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] = t0;
    if constexpr (NP == 2)
      part[((int64_t)1 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] =
          t1;
  }
}

// Row mapping of the fused SpMV + dots (ld = K): each lane keeps K
// partial sums for its rows and the wave reduces them once at the end.
template <typename XT, int K, int NP>
__global__ __launch_bounds__(kBlock) AINV_ROW_ATTR void k_hll_spmv_dot_row(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    const XT *__restrict__ x, int k, double *__restrict__ y,
    const double *__restrict__ w, double *__restrict__ part,
    const int32_t *__restrict__ perm) {
  const int lane = threadIdx.x & (kSlice - 1);
  const int wave = threadIdx.x >> 6;
  double d0[K], d1[K];
#pragma unroll
  for (int u = 0; u < K; ++u) d0[u] = d1[u] = 0.0;
  const GroupRange gr = xcd_groups(nslices);
  for (int32_t grp = gr.first; grp < gr.end; grp += gr.stride) {
    const int32_t s =
        __builtin_amdgcn_readfirstlane(grp * kWavesPerBlock + wave);
    if (s >= nslices) break;
    const int64_t base = sptr[s];
    const int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
    double acc[K];
#pragma unroll
    for (int u = 0; u < K; ++u) acc[u] = 0.0;
    row_acc<double, XT, K>(base + lane, width, col, val, x, acc);
    if (s * kSlice + lane >= n) continue;
    const int32_t r = hll_row(perm, s * kSlice + lane);
    double wr[K];
    load_xrow<double, K>(w + (int64_t)r * K, wr);
    double *yo = y + (int64_t)r * K;
#pragma unroll
    for (int u = 0; u < K; ++u) {
      if (u >= k) continue;
      yo[u] = acc[u];
      if constexpr (NP == 1) {
        d0[u] += wr[u] * acc[u];
      } else {
        d0[u] += acc[u] * acc[u];
        d1[u] += acc[u] * wr[u];
      }
    }
  }
  __shared__ double red[2][kWavesPerBlock][K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    double a0 = d0[u], a1 = d1[u];
    for (int off = 1; off < kSlice; off <<= 1) {
      a0 += __shfl_xor(a0, off);
      if constexpr (NP == 2) a1 += __shfl_xor(a1, off);
    }
    if (lane == 0) {
      red[0][wave][u] = a0;
      red[1][wave][u] = a1;
    }
  }
  __syncthreads();
  if (threadIdx.x < K && (int)threadIdx.x < k) {
    double t0 = 0.0, t1 = 0.0;
    for (int wv = 0; wv < kWavesPerBlock; ++wv) {
      t0 += red[0][wv][threadIdx.x];
      t1 += red[1][wv][threadIdx.x];
    }
    part[((int64_t)0 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] = t0;
    if constexpr (NP == 2)
      part[((int64_t)1 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] =
          t1;
  }
}

// Cooperative row mapping: the L lanes of a row group load one x row
// together, lane (row g, chunk c) the VB bytes at offset c * VB, so the four
// lanes of a quad read one contiguous 64-byte piece (one cache line) instead
// of four pieces of four different rows: the L1 looks up about a quarter of
// the lines the row mapping does per gathered byte.  The column and value
// words of a slice position are loaded once per wave (coalesced, lane l =
// row l) and handed to the lanes of each row group by shuffles, which run on
// the LDS pipe, not the L1.  Lane (g, c) accumulates columns
// c*CPL .. c*CPL+CPL of rows g, g + R, ..., g + (L-1) R of the slice.
template <typename XT, int K> struct Coop {
  static constexpr int RB = K * (int)sizeof(XT);    // bytes of one x row
  static constexpr int VB = RB < 16 ? RB : 16;      // bytes per lane load
  static constexpr int L = RB / VB;                 // lanes per row
  static constexpr int CPL = VB / (int)sizeof(XT);  // columns per lane
  static constexpr int R = kSlice / L;              // rows per pass
};

// Column chunks.  A block of K columns whose x rows are wider than
// CB bytes (AINV_MAP_COOP: 256, AINV_MAP_COOP128: 128) is processed as NCH chunks of KC columns by separate
// workgroups, each XCD taking one chunk for a contiguous range of slices: the
// x rows an XCD gathers between the first and the last use of a row (the
// matrix bandwidth times the row width) then fit in its L2, at the price of
// reading the factor words once per chunk.
template <typename XT, int K, int CB> struct CoopChunk {
  static constexpr int cap = CB / (int)sizeof(XT);
  static constexpr int KC = K < cap ? K : cap;
  static constexpr int NCH = K / KC;
};
// Waves per SIMD of the gather kernels: chunks of 32 columns or more hold
// 32+ accumulators per lane, so fewer waves instead of spilling -- 3 with
// FP32 sums (FP32 / BF16 factors), 2 with FP64 sums.
template <typename VT, typename XT, int K, int CB> struct CoopWaves {
  static constexpr int v =
      CoopChunk<XT, K, CB>::KC >= 32
          ? (sizeof(typename Val<VT>::acc) == 4 ? 3 : 2)
          : AINV_ROW_WAVES;
};
// The fused system-matrix product keeps FP64 sums of its FP64 matrix: at
// 32-column chunks it needs the registers of 2 waves per SIMD (measured at
// 3: 1.04 against 0.80 ms per launch at 160^3); its budget is DotWaves.
#if AINV_ROW_WAVES > 0
#define AINV_COOP_ATTR(XT)                                                  \
  __attribute__((amdgpu_waves_per_eu(CoopWaves<VT, XT, K, CB>::v,           \
                                     CoopWaves<VT, XT, K, CB>::v)))
#else
#define AINV_COOP_ATTR(XT)
#endif

// Slice positions per trip of the cooperative loop.  For chunks of 32
// columns one position puts 8 KB (FP32 x) of loads per wave in flight; two
// positions per trip double that at 2 waves per SIMD, which measured 3-4 %
// faster than one position at 2 or 4 waves (4 waves per SIMD spill, and so
// do two positions of 64-column chunks; 16-column chunks ran 5-33 % slower
// with two).  The fused system-matrix product + dots keeps one position per
// trip: two measured 2 % slower there.
template <typename A, int KC> struct CoopGroup {
  static constexpr int v = KC == 32 ? 2 : RowGroup<A, KC>::v;
};

template <typename T> __device__ __forceinline__ T shfl_raw(T v, int src) {
  if constexpr (sizeof(T) == 2) {
    return (T)__shfl((int)v, src);
  } else {
    return __shfl(v, src);
  }
}

// x rows are K apart; xs points at this lane's first column of the chunk.
// A lane whose columns are all at or beyond the active width (live false)
// takes part in the shuffles and loads row 0 of its columns instead of the
// gathered rows (a select, not a branch, so the loads keep their schedule;
// all such lanes hit the same lines), and its sums are not stored.
// NQ of the L row passes, from pass q0 on (a wave may share its slice with
// other waves, each taking some of the passes).
template <typename VT, typename XT, int K, int KC, int NQ = Coop<XT, KC>::L,
          int P = CoopGroup<typename Val<VT>::acc, KC>::v>
__device__ __forceinline__ void coop_acc(
    int64_t p0, int32_t width, int lane, int rg, int q0,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ xs, bool live,
    typename Val<VT>::acc (&acc)[NQ][Coop<XT, KC>::CPL]) {
  using A = typename Val<VT>::acc;
  using CP = Coop<XT, KC>;
  constexpr int L = NQ, CPL = CP::CPL, R = CP::R;
  const int32_t nfull = width / P * P;
  const int64_t pl = p0 + lane;
  if (nfull > 0) {
    int32_t ci[P];
    VT vi[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      ci[u] = col[pl + (int64_t)u * kSlice];
      vi[u] = val[pl + (int64_t)u * kSlice];
#if AINV_ROW_PIPE
      asm volatile("" : "+v"(ci[u]));
#endif
    }
    for (int32_t j = 0;;) {
      A xv[P][L][CPL];
#pragma unroll
      for (int u = 0; u < P; ++u)
#pragma unroll
        for (int q = 0; q < L; ++q) {
          const int32_t c = __shfl(ci[u], (q0 + q) * R + rg);
          load_xrow<XT, CPL>(live ? xs + (int64_t)c * K : xs, xv[u][q]);
        }
      const int32_t jn = j + P;
      const bool more = jn < nfull;
      const int64_t qn = pl + (int64_t)(more ? jn : j) * kSlice;
      int32_t cn[P];
      VT vn[P];
#pragma unroll
      for (int u = 0; u < P; ++u) {
        cn[u] = col[qn + (int64_t)u * kSlice];
        vn[u] = val[qn + (int64_t)u * kSlice];
      }
#pragma unroll
      for (int u = 0; u < P; ++u)
#pragma unroll
        for (int q = 0; q < L; ++q) {
          const A v = Val<VT>::get(shfl_raw(vi[u], (q0 + q) * R + rg));
#pragma unroll
          for (int c = 0; c < CPL; ++c) acc[q][c] += v * xv[u][q][c];
        }
      if (!more) break;
      j = jn;
#pragma unroll
      for (int u = 0; u < P; ++u) {
        ci[u] = cn[u];
        vi[u] = vn[u];
      }
    }
  }
  for (int32_t j = nfull; j < width; ++j) {
    const int32_t c0 = col[pl + (int64_t)j * kSlice];
    const VT v0 = val[pl + (int64_t)j * kSlice];
#pragma unroll
    for (int q = 0; q < L; ++q) {
      const int32_t c = __shfl(c0, (q0 + q) * R + rg);
      const A v = Val<VT>::get(shfl_raw(v0, (q0 + q) * R + rg));
      A xv[CPL];
      load_xrow<XT, CPL>(live ? xs + (int64_t)c * K : xs, xv);
#pragma unroll
      for (int u = 0; u < CPL; ++u) acc[q][u] += v * xv[u];
    }
  }
}

// Waves of a block that share one slice in the quarter-slice variant
// (each takes L / QW of the slice's row passes): more waves for matrices of
// few slices, whose waves are otherwise bound by the latency of walking
// long rows.  QW = 1: each wave has a slice of its own.
template <typename XT, int K, int CB, int QS> struct CoopSplit {
  static constexpr int L = Coop<XT, CoopChunk<XT, K, CB>::KC>::L;
  static constexpr int QW = QS < L ? QS : L;
  static constexpr int NQ = L / QW;
  static constexpr int SPB = kWavesPerBlock / QW;  // slices per block
};

// Epilogue of one lane's CPL columns c0.. of row r (columns at or beyond
// the active width kw are not written).
template <typename DT, typename OT, int K, int CPL, int EPI, typename A>
__device__ __forceinline__ void lane_store(int32_t r, int c0, int kw,
                                           const A (&acc)[CPL],
                                           const DT *__restrict__ dvec,
                                           OT *__restrict__ out, double alpha,
                                           double beta) {
  OT *o = out + (int64_t)r * K + c0;
  if constexpr (EPI == EPI_SCALE) {
    const double d = (double)dvec[r];
#pragma unroll
    for (int u = 0; u < CPL; ++u)
      if (c0 + u < kw) o[u] = (OT)(d * (double)acc[u]);
  } else if constexpr (EPI == EPI_RESID) {
    const DT *zr = dvec + (int64_t)r * K + c0;
#pragma unroll
    for (int u = 0; u < CPL; ++u)
      if (c0 + u < kw)
        o[u] = (OT)(alpha * (double)acc[u] + beta * (double)zr[u]);
  } else {
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
      if (c0 + u >= kw) continue;
      double v = alpha * (double)acc[u];
      if (beta != 0.0) v += beta * (double)o[u];
      o[u] = (OT)v;
    }
  }
}

template <typename VT, typename XT, typename DT, typename OT, int K, int CB,
          int EPI, int QS>
__device__ __forceinline__ void coop_spmm_body(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ kwp, const int32_t *__restrict__ perm) {
  using A = typename Val<VT>::acc;
  using CH = CoopChunk<XT, K, CB>;
  constexpr int KC = CH::KC;
  using CP = Coop<XT, KC>;
  using SP = CoopSplit<XT, K, CB, QS>;
  constexpr int NQ = SP::NQ;
  const int lane = threadIdx.x & (kSlice - 1);
  const int sub = lane % CP::L, rg = lane / CP::L;
  // grid = NCH x slice groups; each XCD runs one chunk of a slice range
  const int32_t ngrp = (nslices + SP::SPB - 1) / SP::SPB;
  const int32_t lb = __builtin_amdgcn_readfirstlane(
      xcd_block(blockIdx.x, gridDim.x));
  const int chunk = lb / ngrp;
  // active width: columns at or beyond it are neither read nor written
  const int kw = kwp ? min(k, *kwp) : k;
  if (chunk * KC >= kw) return;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int32_t s = lb % ngrp * SP::SPB + wv / SP::QW;
  if (s >= nslices) return;
  const int q0 = (wv % SP::QW) * NQ;  // first row pass of this wave
  const int64_t base = sptr[s];
  const int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
  const int32_t pv = slice_rows(perm, s, lane);
  A acc[NQ][CP::CPL];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int u = 0; u < CP::CPL; ++u) acc[q][u] = 0;
  const int c0 = chunk * KC + sub * CP::CPL;  // first column of this lane
  coop_acc<VT, XT, K, KC, NQ>(base, width, lane, rg, q0, col, val, x + c0,
                              c0 < kw, acc);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int j = (q0 + q) * CP::R + rg;  // slice row
    const int32_t r = slice_row(perm, pv, s, j);
    if (s * kSlice + j < n)
      lane_store<DT, OT, K, CP::CPL, EPI>(r, c0, kw, acc[q], dvec, out, alpha,
                                          beta);
  }
}

template <typename VT, typename XT, typename DT, typename OT, int K, int CB,
          int EPI>
__global__ __launch_bounds__(kBlock) AINV_COOP_ATTR(XT) void k_hll_spmm_coop(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ kwp, const int32_t *__restrict__ perm) {
  coop_spmm_body<VT, XT, DT, OT, K, CB, EPI, 1>(n, nslices, sptr, col, val, x,
                                                k, dvec, out, alpha, beta,
                                                kwp, perm);
}

// quarter-slice variant (AINV_MAP_COOP_Q / AINV_MAP_COOP128_Q)
template <typename VT, typename XT, typename DT, typename OT, int K, int CB,
          int EPI>
__global__ __launch_bounds__(kBlock) AINV_COOP_ATTR(XT) void k_hll_spmm_coopq(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ kwp, const int32_t *__restrict__ perm) {
  coop_spmm_body<VT, XT, DT, OT, K, CB, EPI, 4>(n, nslices, sptr, col, val, x,
                                                k, dvec, out, alpha, beta,
                                                kwp, perm);
}

// ---------------------------------------------------------------- runs
// Run-reuse mapping (AINV_MAP_RUN).  Factors of stencil-like matrices hold
// runs of consecutive diagonals: entry j + 1 of row r reads the x row that
// entry j of row r + 1 read (col[r, j+1] = col[r+1, j]).  Here the L lanes
// of a row group own L consecutive rows of the slice (row group g: rows
// gL .. gL + L - 1, one pass each), so the x row a pass needs at entry
// j + 1 is the one its next pass (or, for the group's last pass, the next
// group's first pass: one shuffle) held at entry j.  Each pass compares the
// column it needs with the column held there and takes the held values
// when they match; only the others are gathered.  The gathers go through a
// raw buffer resource over x: a lane that takes held values gets an
// offset beyond the buffer, reads nothing and makes no memory request, so
// inside a run only the slice's last row gathers.  Sums are formed in the
// same entry order from the same values as in the cooperative mapping.
constexpr uint32_t kNoLoad = 0xfffffff0u;  // out-of-range buffer offset
// factor words fetched RC entries ahead
constexpr int kRunChunk = 8;

template <int VB>
__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                         uint32_t (&w)[VB / 4]) {
  if constexpr (VB == 16) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    w[0] = v[0];
    w[1] = v[1];
    w[2] = v[2];
    w[3] = v[3];
  } else if constexpr (VB == 8) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    w[0] = v[0];
    w[1] = v[1];
  } else {
    w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
  }
}

// Largest byte size of an x block the run mapping addresses (32-bit buffer
// offsets, kNoLoad beyond it).
constexpr int64_t kRunMaxBytes = (int64_t)kNoLoad;

// Column chunks of the run mapping: at most 128 bytes of an x row per
// workgroup (8 lanes per row), so a lane holds the values, gathers and sums
// of 8 rows x 16 bytes; wider rows (64 FP32 / 32 FP64 columns) run as 2-4
// chunks of the same kernel instead of spilling 16-lane row groups.
constexpr int kRunChunkBytes = 128;
// Waves per SIMD: rows of 128 bytes hold 8 x 16-byte values, gathers and
// sums per lane -- 3 waves with FP32 sums, 2 with FP64 sums; narrower rows
// fit the row kernels' budget.
template <typename VT, typename XT, int K> struct RunWaves {
  static constexpr int rb =
      CoopChunk<XT, K, kRunChunkBytes>::KC * (int)sizeof(XT);
  static constexpr int v =
      rb >= 128 ? (sizeof(typename Val<VT>::acc) == 4 ? 3 : 2)
                : AINV_ROW_WAVES;
};
#if AINV_ROW_WAVES > 0
#define AINV_RUN_ATTR                                                       \
  __attribute__((amdgpu_waves_per_eu(RunWaves<VT, XT, K>::v,                \
                                     RunWaves<VT, XT, K>::v)))
#else
#define AINV_RUN_ATTR
#endif

// Sums of rows rb .. rb + L - 1 of one slice (entries of lane l at
// base + 64 j + l, rb = L * (lane / L)) for this lane's CPL columns of a
// chunk, x rows handed on along runs of consecutive diagonals.  x is read
// through the buffer resource xr (lane_off: the lane's byte offset in a
// row); live false: the lane's columns are beyond the active width and it
// only takes part in the shuffles.
template <typename VT, typename XT, int K, int KC>
__device__ __forceinline__ void run_acc(
    int64_t base, int32_t width, int lane, bool live,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    __amdgpu_buffer_rsrc_t xr, uint32_t lane_off,
    typename Val<VT>::acc (&acc)[Coop<XT, KC>::L][Coop<XT, KC>::CPL]) {
  using A = typename Val<VT>::acc;
  using CP = Coop<XT, KC>;
  constexpr int L = CP::L, CPL = CP::CPL, VB = CP::VB;
  constexpr int NW = VB / 4;  // dwords per lane load
  static_assert(CPL * (int)sizeof(XT) == VB, "lane load holds CPL values");
  const int rb = lane / L * L;                  // first slice row of the group
  const bool last_group = lane / L == CP::R - 1;  // rows up to 63: no next
  uint32_t xv[L][NW];  // x values of each pass at the current entry
  int32_t held[L];     // their columns (-1: nothing held)
#pragma unroll
  for (int q = 0; q < L; ++q) {
    held[q] = -1;
#pragma unroll
    for (int u = 0; u < CPL; ++u) acc[q][u] = 0;
#pragma unroll
    for (int u = 0; u < NW; ++u) xv[q][u] = 0;
  }
  const int64_t pl = base + lane;
  // factor words of RC entries at a time, the next RC requested while the
  // current ones are used (entries past the width re-read the last one)
  constexpr int RC = kRunChunk;
  const int32_t last = width > 0 ? width - 1 : 0;
  int32_t cwc[RC];
  VT vwc[RC];
#pragma unroll
  for (int t = 0; t < RC; ++t) {
    const int64_t q = pl + (int64_t)min(t, last) * kSlice;
    cwc[t] = width > 0 ? col[q] : 0;
    vwc[t] = width > 0 ? val[q] : VT(0);
  }
  for (int32_t j0 = 0; j0 < width; j0 += RC) {
    int32_t cwn[RC];
    VT vwn[RC];
#pragma unroll
    for (int t = 0; t < RC; ++t) {
      const int64_t q = pl + (int64_t)min(j0 + RC + t, last) * kSlice;
      cwn[t] = col[q];
      vwn[t] = val[q];
    }
#pragma unroll
    for (int t = 0; t < RC; ++t) {
    if (j0 + t >= width) break;
    const int32_t cw = cwc[t];
    const VT vw = vwc[t];
    // candidates: pass q + 1's values (the next group's pass 0 for the
    // last pass)
    const int32_t hn = __shfl(held[0], lane + L);
    uint32_t xn[NW];
#pragma unroll
    for (int u = 0; u < NW; ++u) xn[u] = __shfl(xv[0][u], lane + L);
    int32_t need[L];
    bool take[L];
    uint32_t ld[L][NW];
#pragma unroll
    for (int q = 0; q < L; ++q) {
      need[q] = __shfl(cw, rb + q);
      const int32_t cand = q + 1 < L ? held[q + 1] : (last_group ? -1 : hn);
      take[q] = need[q] == cand;
      const bool load = !take[q] && live;
      const uint32_t off =
          load ? (uint32_t)need[q] * (uint32_t)(K * (int)sizeof(XT)) + lane_off
               : kNoLoad;
      buf_load<VB>(xr, off, ld[q]);
    }
#pragma unroll
    for (int q = 0; q < L; ++q) {
#pragma unroll
      for (int u = 0; u < NW; ++u) {
        const uint32_t c = q + 1 < L ? xv[q + 1][u] : xn[u];
        xv[q][u] = take[q] ? c : ld[q][u];
      }
      held[q] = need[q];
      const A v = Val<VT>::get(shfl_raw(vw, rb + q));
      XT e[CPL];
      __builtin_memcpy(e, xv[q], VB);
#pragma unroll
      for (int u = 0; u < CPL; ++u) acc[q][u] += v * cvt<A>(e[u]);
    }
    }
#pragma unroll
    for (int t = 0; t < RC; ++t) {
      cwc[t] = cwn[t];
      vwc[t] = vwn[t];
    }
  }
}

// Buffer resource over an x block of ncols rows of K values (callers keep
// its byte size below kRunMaxBytes).
template <typename XT, int K>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t run_rsrc(const XT *x,
                                                           int32_t ncols) {
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<XT *>(x), 0, (int)((int64_t)ncols * K * (int64_t)sizeof(XT)),
      0x00020000);
}

template <typename VT, typename XT, typename DT, typename OT, int K, int CB,
          int EPI>
__global__ __launch_bounds__(kBlock) AINV_RUN_ATTR void k_hll_spmm_run(
    int32_t n, int32_t ncols, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ kwp, const int32_t *__restrict__ gorder,
    const int32_t *__restrict__ perm) {
  using A = typename Val<VT>::acc;
  using CH = CoopChunk<XT, K, CB>;
  constexpr int KC = CH::KC;
  using CP = Coop<XT, KC>;
  constexpr int L = CP::L, CPL = CP::CPL;
  const int lane = threadIdx.x & (kSlice - 1);
  const int sub = lane % L;
  const int rb = lane / L * L;  // first slice row of the row group
  const int kw = kwp ? min(k, *kwp) : k;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int32_t ngrp = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
  const int32_t lb = __builtin_amdgcn_readfirstlane(
      xcd_block(blockIdx.x, gridDim.x));
  const int chunk = lb / ngrp;
  if (chunk * KC >= kw) return;
  // slice group of this workgroup: linear, or the factor's strip order
  const int32_t grp = gorder ? gorder[lb % ngrp] : lb % ngrp;
  const int32_t s = grp * kWavesPerBlock + wv;
  if (s >= nslices) return;
  const int64_t base = sptr[s];
  const int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
  const int c0 = chunk * KC + sub * CPL;  // first column of this lane
  const int32_t pv = slice_rows(perm, s, lane);
  A acc[L][CPL];
  run_acc<VT, XT, K, KC>(base, width, lane, c0 < kw, col, val,
                         run_rsrc<XT, K>(x, ncols),
                         (uint32_t)(c0 * (int)sizeof(XT)), acc);
#pragma unroll
  for (int q = 0; q < L; ++q) {
    const int32_t r = slice_row(perm, pv, s, rb + q);
    if (s * kSlice + rb + q < n)
      lane_store<DT, OT, K, CPL, EPI>(r, c0, kw, acc[q], dvec, out, alpha,
                                      beta);
  }
}

template <typename VT, typename XT, typename DT, typename OT, int EPI>
void launch_run(const Hll &a, const XT *x, int k, const DT *d, OT *out,
                double alpha, double beta, hipStream_t s, const int32_t *kw) {
  if (a.nslices == 0) return;
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const VT *vl = a.val.as<VT>();
  const int ngrp = (a.nslices + kWavesPerBlock - 1) / kWavesPerBlock;
#define AINV_LAUNCH(KK)                                                     \
  hipLaunchKernelGGL((k_hll_spmm_run<VT, XT, DT, OT, KK, kRunChunkBytes,   \
                                     EPI>),                                 \
                     dim3(ngrp * CoopChunk<XT, KK, kRunChunkBytes>::NCH),   \
                     dim3(kBlock),                                          \
                     0, s, a.n, a.ncols, a.nslices, sp, cl, vl, x, k, d, out, \
                     alpha, beta, kw, a.gorder.as<int32_t>(),               \
                     a.perm.as<int32_t>())
  if (k <= 2)
    AINV_LAUNCH(2);
  else if (k <= 4)
    AINV_LAUNCH(4);
  else if (k <= 8)
    AINV_LAUNCH(8);
  else if (k <= 16)
    AINV_LAUNCH(16);
  else if (k <= 32)
    AINV_LAUNCH(32);
  else
    AINV_LAUNCH(64);
#undef AINV_LAUNCH
  AINV_HIP_CHECK(hipGetLastError());
}

// Fused system-matrix product + dots in the cooperative mapping (ld = K).
// With NCH chunks, the blocks of XCD x work on chunk x / (8 / NCH) and
// stride through a contiguous part of the slice groups.  Lanes with the same
// lane chunk hold partial sums of the same columns; they are reduced by
// shuffles across the row groups, then across the waves in LDS; a block
// writes zero partial sums for the columns of the other chunks.
// RUN: the run mapping's lane layout and hand-on of x rows along runs of
// consecutive diagonals (run_acc) instead of the cooperative one; the same
// row sums, the dots summed over a lane's rows in another order.
// (run: FP64 sums of 8 rows per lane at 128-byte rows, of 4 rows at 64)
template <typename XT, int K, int CB, int RUN> struct DotWaves {
  static constexpr int rb = RunWaves<double, XT, K>::rb;
  static constexpr int v =
      RUN ? (rb >= 128 ? 2 : rb == 64 ? 3 : AINV_ROW_WAVES)
          : CoopWaves<double, XT, K, CB>::v;
};
// SORT: a row-sorted layout (perm set); rows in order compile without the
// row map, which costs the short rows of a stencil matrix ~4 % here.
template <typename XT, int K, int CB, int NP, int RUN = 0, int SORT = 0>
__global__ __launch_bounds__(kBlock)
#if AINV_ROW_WAVES > 0
__attribute__((amdgpu_waves_per_eu(DotWaves<XT, K, CB, RUN>::v,
                                   DotWaves<XT, K, CB, RUN>::v)))
#endif
void k_hll_spmv_dot_coop(int32_t n, int32_t ncols, int32_t nslices,
                    const int64_t *__restrict__ sptr,
                    const int32_t *__restrict__ col,
                    const double *__restrict__ val,
                    const XT *__restrict__ x, int k,
                    double *__restrict__ y, const double *__restrict__ w,
                    double *__restrict__ part,
                    const int32_t *__restrict__ kwp,
                    const int32_t *__restrict__ perm) {
  using CH = CoopChunk<XT, K, CB>;
  constexpr int KC = CH::KC, NCH = CH::NCH;
  using CP = Coop<XT, KC>;
  constexpr int L = CP::L, CPL = CP::CPL, R = CP::R;
  const int lane = threadIdx.x & (kSlice - 1);
  const int wave = threadIdx.x >> 6;
  const int sub = lane % L, rg = lane / L;
  // chunk and slice-group range of this block
  const int32_t ng = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
#if AINV_XCD_REMAP
  constexpr int per = kXcds / NCH;  // XCDs per chunk (grid: multiple of 8)
  const int xcd = blockIdx.x % kXcds;
  const int chunk = xcd / per, sx = xcd % per;
  const int32_t g_first = (int32_t)((int64_t)ng * sx / per) +
                          (int32_t)(blockIdx.x / kXcds);
  const int32_t g_end = (int32_t)((int64_t)ng * (sx + 1) / per);
  const int32_t g_stride = gridDim.x / kXcds;
#else
  const int chunk = blockIdx.x % NCH;
  const int32_t g_first = blockIdx.x / NCH, g_end = ng;
  const int32_t g_stride = gridDim.x / NCH;
#endif
  const int c0 = chunk * KC + sub * CPL;
  // active width: a chunk wholly beyond it only writes its zero partials
  const int kw = kwp ? min(k, *kwp) : k;
  const int32_t g_stop = chunk * KC >= kw ? g_first : g_end;
  double d0[CPL], d1[CPL];
#pragma unroll
  for (int u = 0; u < CPL; ++u) d0[u] = d1[u] = 0.0;
  for (int32_t grp = g_first; grp < g_stop; grp += g_stride) {
    const int32_t s =
        __builtin_amdgcn_readfirstlane(grp * kWavesPerBlock + wave);
    if (s >= nslices) break;
    const int64_t base = sptr[s];
    const int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
    const int32_t pv = SORT ? perm[(int64_t)s * kSlice + lane] : 0;
    double acc[L][CPL];
#pragma unroll
    for (int q = 0; q < L; ++q)
#pragma unroll
      for (int u = 0; u < CPL; ++u) acc[q][u] = 0.0;
    if constexpr (RUN)
      run_acc<double, XT, K, KC>(base, width, lane, c0 < kw, col, val,
                                 run_rsrc<XT, K>(x, ncols),
                                 (uint32_t)(c0 * (int)sizeof(XT)), acc);
    else
      coop_acc<double, XT, K, KC, Coop<XT, KC>::L, RowGroup<double, KC>::v>(
          base, width, lane, rg, 0, col, val, x + c0, c0 < kw, acc);
#pragma unroll
    for (int q = 0; q < L; ++q) {
      const int j = RUN ? rg * L + q : q * R + rg;  // slice row
      const int32_t r = SORT ? __shfl(pv, j) : s * kSlice + j;
      if (s * kSlice + j >= n) continue;
      double wr[CPL];
      load_xrow<double, CPL>(c0 < kw ? w + (int64_t)r * K + c0 : w + c0, wr);
      double *yo = y + (int64_t)r * K + c0;
#pragma unroll
      for (int u = 0; u < CPL; ++u) {
        if (c0 + u >= kw) continue;
        yo[u] = acc[q][u];
        if constexpr (NP == 1) {
          d0[u] += wr[u] * acc[q][u];
        } else {
          d0[u] += acc[q][u] * acc[q][u];
          d1[u] += acc[q][u] * wr[u];
        }
      }
    }
  }
  __shared__ double red[2][kWavesPerBlock][KC];
#pragma unroll
  for (int u = 0; u < CPL; ++u) {
    double a0 = d0[u], a1 = d1[u];
    for (int off = L; off < kSlice; off <<= 1) {
      a0 += __shfl_xor(a0, off);
      if constexpr (NP == 2) a1 += __shfl_xor(a1, off);
    }
    if (rg == 0) {
      red[0][wave][sub * CPL + u] = a0;
      red[1][wave][sub * CPL + u] = a1;
    }
  }
  __syncthreads();
  if (threadIdx.x < K && (int)threadIdx.x < k) {
    const int cl = (int)threadIdx.x - chunk * KC;  // column within the chunk
    double t0 = 0.0, t1 = 0.0;
    if (cl >= 0 && cl < KC)
      for (int wv = 0; wv < kWavesPerBlock; ++wv) {
        t0 += red[0][wv][cl];
        t1 += red[1][wv][cl];
      }
    part[((int64_t)0 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] = t0;
    if constexpr (NP == 2)
      part[((int64_t)1 * kRedBlocks + blockIdx.x) * kMaxCols + threadIdx.x] =
          t1;
  }
}

template <typename VT, typename XT, typename DT, typename OT, int CB, int EPI,
          int QS = 1>
void launch_coop(const Hll &a, const XT *x, int k, const DT *d, OT *out,
                 double alpha, double beta, hipStream_t s,
                 const int32_t *kw) {
  if (a.nslices == 0) return;
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const VT *vl = a.val.as<VT>();
#define AINV_LAUNCH(KK)                                                     \
  do {                                                                      \
    constexpr int spb = CoopSplit<XT, KK, CB, QS>::SPB;                     \
    const int ngrp = (a.nslices + spb - 1) / spb;                           \
    const dim3 grid(ngrp * CoopChunk<XT, KK, CB>::NCH);                     \
    if (QS == 1)                                                            \
      hipLaunchKernelGGL((k_hll_spmm_coop<VT, XT, DT, OT, KK, CB, EPI>),    \
                         grid, dim3(kBlock), 0, s, a.n, a.nslices, sp, cl,  \
                         vl, x, k, d, out, alpha, beta, kw,                 \
                         a.perm.as<int32_t>());                             \
    else                                                                    \
      hipLaunchKernelGGL((k_hll_spmm_coopq<VT, XT, DT, OT, KK, CB, EPI>),   \
                         grid, dim3(kBlock), 0, s, a.n, a.nslices, sp, cl,  \
                         vl, x, k, d, out, alpha, beta, kw,                 \
                         a.perm.as<int32_t>());                             \
  } while (0)
  if (k <= 2)
    AINV_LAUNCH(2);
  else if (k <= 4)
    AINV_LAUNCH(4);
  else if (k <= 8)
    AINV_LAUNCH(8);
  else if (k <= 16)
    AINV_LAUNCH(16);
  else if (k <= 32)
    AINV_LAUNCH(32);
  else
    AINV_LAUNCH(64);
#undef AINV_LAUNCH
  AINV_HIP_CHECK(hipGetLastError());
}

// One wavefront per row over the CSR copy (AINV_MAP_WIDE), for matrices of
// few, long rows (the small levels of a multilevel hierarchy): lane (g, c)
// of the G = 64 / K groups of K lanes takes column c of entries g, g + G,
// ... of the row; the wave loads 64 (column, value) words at a time,
// coalesced, and hands them out by shuffles (uniform trip counts: the row
// is the wave's); each group's K lanes gather one contiguous x row.  The
// groups' sums are combined by xor shuffles in a fixed order.  Rows of
// ld = K, x and out with ld = K.
template <typename VT, typename XT, typename DT, typename OT, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_csr_wide(
    int32_t n, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ idx, const VT *__restrict__ val,
    const XT *__restrict__ x, int k, const DT *__restrict__ dvec,
    OT *__restrict__ out, double alpha, double beta,
    const int32_t *__restrict__ kwp) {
  using A = typename Val<VT>::acc;
  constexpr int G = kSlice / K;
  const int lane = threadIdx.x & (kSlice - 1);
  const int c = lane % K, g = lane / K;
  const int32_t r = __builtin_amdgcn_readfirstlane(
      (int32_t)blockIdx.x * kWavesPerBlock + (int32_t)(threadIdx.x >> 6));
  if (r >= n) return;
  const int kw = kwp ? min(k, *kwp) : k;
  const bool live = c < kw;
  const int64_t beg = ptr[r], end = ptr[r + 1];
  A acc = 0;
  for (int64_t j0 = beg; j0 < end; j0 += kSlice) {
    const int64_t jl = j0 + lane;
    const int32_t cl = jl < end ? idx[jl] : 0;
    const VT vl = jl < end ? val[jl] : VT(0);
    const int cnt = (int)min((int64_t)kSlice, end - j0);
    for (int t = 0; t < cnt; t += G) {
      const int te = t + g;  // < 64: t < 64 and g < G
      const int32_t cc = __shfl(cl, te);
      const A v = Val<VT>::get(shfl_raw(vl, te));
      const A xv = cvt<A>(x[(int64_t)cc * K + (live ? c : 0)]);
      if (te < cnt) acc += v * xv;
    }
  }
#pragma unroll
  for (int o = K; o < kSlice; o <<= 1) acc += __shfl_xor(acc, o);
  if (g != 0 || !live) return;
  OT *o = out + (int64_t)r * K + c;
  if constexpr (EPI == EPI_SCALE) {
    *o = (OT)((double)dvec[r] * (double)acc);
  } else if constexpr (EPI == EPI_RESID) {
    *o = (OT)(alpha * (double)acc + beta * (double)dvec[(int64_t)r * K + c]);
  } else {
    double v = alpha * (double)acc;
    if (beta != 0.0) v += beta * (double)(*o);
    *o = (OT)v;
  }
}

template <typename VT, typename XT, typename DT, typename OT, int EPI>
void launch_wide(const Hll &a, const XT *x, int k, const DT *d, OT *out,
                 double alpha, double beta, hipStream_t s,
                 const int32_t *kw) {
  if (!a.cptr.p)
    throw Error(AINV_ERR_INTERNAL, "the wide mapping needs the CSR copy");
  if (a.n == 0) return;
  const dim3 grid((unsigned)((a.n + kWavesPerBlock - 1) / kWavesPerBlock));
  const int64_t *cp = a.cptr.as<int64_t>();
  const int32_t *ci = a.cidx.as<int32_t>();
  const VT *cv = a.cval.as<VT>();
#define AINV_LAUNCH(KK)                                                     \
  hipLaunchKernelGGL((k_csr_wide<VT, XT, DT, OT, KK, EPI>), grid,           \
                     dim3(kBlock), 0, s, a.n, cp, ci, cv, x, k, d, out,     \
                     alpha, beta, kw)
  if (k <= 1)
    AINV_LAUNCH(1);
  else if (k <= 2)
    AINV_LAUNCH(2);
  else if (k <= 4)
    AINV_LAUNCH(4);
  else if (k <= 8)
    AINV_LAUNCH(8);
  else if (k <= 16)
    AINV_LAUNCH(16);
  else if (k <= 32)
    AINV_LAUNCH(32);
  else
    AINV_LAUNCH(64);
#undef AINV_LAUNCH
  AINV_HIP_CHECK(hipGetLastError());
}

// kw (device, may be NULL): active width of the block, honoured by the
// cooperative and wide mappings (the others compute every column).
template <typename VT, typename XT, typename DT, typename OT, int EPI>
void launch_lane(int map, const Hll &a, const XT *x, int k, const DT *d,
                 OT *out, double alpha, double beta, hipStream_t s,
                 const int32_t *kw) {
  if (map == AINV_MAP_COOP)
    launch_coop<VT, XT, DT, OT, 256, EPI>(a, x, k, d, out, alpha, beta, s, kw);
  else if (map == AINV_MAP_COOP128)
    launch_coop<VT, XT, DT, OT, 128, EPI>(a, x, k, d, out, alpha, beta, s, kw);
  else if (map == AINV_MAP_COOP_Q)
    launch_coop<VT, XT, DT, OT, 256, EPI, 4>(a, x, k, d, out, alpha, beta, s,
                                             kw);
  else if (map == AINV_MAP_COOP128_Q)
    launch_coop<VT, XT, DT, OT, 128, EPI, 4>(a, x, k, d, out, alpha, beta, s,
                                             kw);
  else if (map == AINV_MAP_RUN && a.ncols > 0 &&
           (int64_t)a.ncols * 64 * (int64_t)sizeof(XT) < kRunMaxBytes)
    launch_run<VT, XT, DT, OT, EPI>(a, x, k, d, out, alpha, beta, s, kw);
  else if (map == AINV_MAP_RUN)  // x too large for 32-bit buffer offsets
    launch_coop<VT, XT, DT, OT, 256, EPI>(a, x, k, d, out, alpha, beta, s, kw);
  else if (map == AINV_MAP_WIDE)
    launch_wide<VT, XT, DT, OT, EPI>(a, x, k, d, out, alpha, beta, s, kw);
  else
    launch_row<VT, XT, DT, OT, EPI>(a, x, k, d, out, alpha, beta, s);
}

// The row kernels read whole x rows as 16-byte vectors.
bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

}  // namespace

template <typename XT, int K, int CB, int NP, int RUN = 0>
void launch_dot_coop(int g, hipStream_t s, const Hll &a, const XT *x, int k,
                     double *y, const double *w, double *part,
                     const int32_t *kw) {
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const double *vl = a.val.as<double>();
  if (a.perm.p)
    hipLaunchKernelGGL((k_hll_spmv_dot_coop<XT, K, CB, NP, RUN, 1>), dim3(g),
                       dim3(kBlock), 0, s, a.n, a.ncols, a.nslices, sp, cl,
                       vl, x, k, y, w, part, kw, a.perm.as<int32_t>());
  else
    hipLaunchKernelGGL((k_hll_spmv_dot_coop<XT, K, CB, NP, RUN, 0>), dim3(g),
                       dim3(kBlock), 0, s, a.n, a.ncols, a.nslices, sp, cl,
                       vl, x, k, y, w, part, kw, nullptr);
}

template <typename XT>
int spmv_dot_impl(const Hll &a, const XT *x, double *y, int k, int64_t ld,
                  const double *w, int np, double *part, int map,
                  hipStream_t s, const int32_t *kw) {
  if (a.prec != AINV_PREC_F64)
    throw Error(AINV_ERR_INTERNAL, "system matrix must be FP64");
  // the fused product + dots has no quarter-slice or wide variant; the run
  // mapping addresses x through 32-bit buffer offsets
  if (map == AINV_MAP_COOP_Q ||
      (map == AINV_MAP_RUN &&
       (int64_t)a.ncols * 64 * (int64_t)sizeof(XT) >= kRunMaxBytes))
    map = AINV_MAP_COOP;
  if (map == AINV_MAP_COOP128_Q || map == AINV_MAP_WIDE)
    map = AINV_MAP_COOP128;
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  const int g = persistent_grid(a.nslices);
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *rl = a.rlen.as<int32_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const double *vl = a.val.as<double>();
  const int32_t *pm = a.perm.as<int32_t>();
  const bool coop = map == AINV_MAP_COOP || map == AINV_MAP_COOP128 ||
                    map == AINV_MAP_RUN;
  if ((map == AINV_MAP_ROW && k > 1 && k <= 32) ||
      (coop && k > 1 && k <= 64)) {
    int K = 2;
    while (K < k) K <<= 1;
    if (ld != K || !aligned16(x) || !aligned16(y) || !aligned16(w))
      throw Error(AINV_ERR_INTERNAL,
                  "row mapping needs ld = K and 16-byte aligned vectors");
#define AINV_LR(KK, NPP)                                                     \
  do {                                                                       \
    if (map == AINV_MAP_RUN)                                                 \
      launch_dot_coop<XT, KK, kRunChunkBytes, NPP, 1>(g, s, a, x, k, y, w,   \
                                                      part, kw);             \
    else if (map == AINV_MAP_COOP)                                           \
      launch_dot_coop<XT, KK, 256, NPP>(g, s, a, x, k, y, w, part, kw);      \
    else if (map == AINV_MAP_COOP128)                                        \
      launch_dot_coop<XT, KK, 128, NPP>(g, s, a, x, k, y, w, part, kw);      \
    else if constexpr (KK >= 16 && NPP == 2)                                 \
      /* 3K FP64 sums per lane do not fit: the cooperative kernel instead */ \
      launch_dot_coop<XT, KK, 128, NPP>(g, s, a, x, k, y, w, part, kw);      \
    else                                                                     \
      hipLaunchKernelGGL((k_hll_spmv_dot_row<XT, KK, NPP>), dim3(g),         \
                         dim3(kBlock), 0, s, a.n, a.nslices, sp, cl, vl, x,  \
                         k, y, w, part, pm);                                 \
  } while (0)
// K = 64 exists only in the cooperative mapping; FP32 rows of 64 columns
// (256 bytes) run as two 128-byte chunks in both variants (16 lanes per row
// would spill the FP64 sums)
#define AINV_LC(KK, NPP)                                                     \
  do {                                                                       \
    if (map == AINV_MAP_RUN)                                                 \
      launch_dot_coop<XT, KK, kRunChunkBytes, NPP, 1>(g, s, a, x, k, y, w,   \
                                                      part, kw);             \
    else if constexpr (sizeof(XT) == 4)                                      \
      launch_dot_coop<XT, KK, 128, NPP>(g, s, a, x, k, y, w, part, kw);      \
    else if (map == AINV_MAP_COOP128)                                        \
      launch_dot_coop<XT, KK, 128, NPP>(g, s, a, x, k, y, w, part, kw);      \
    else                                                                     \
      launch_dot_coop<XT, KK, 256, NPP>(g, s, a, x, k, y, w, part, kw);      \
  } while (0)
#define AINV_LRK(NPP)   \
  if (K == 2)           \
    AINV_LR(2, NPP);    \
  else if (K == 4)      \
    AINV_LR(4, NPP);    \
  else if (K == 8)      \
    AINV_LR(8, NPP);    \
  else if (K == 16)     \
    AINV_LR(16, NPP);   \
  else if (K == 32)     \
    AINV_LR(32, NPP);   \
  else                  \
    AINV_LC(64, NPP);
    if (np == 1) {
      AINV_LRK(1)
    } else {
      AINV_LRK(2)
    }
#undef AINV_LRK
#undef AINV_LC
#undef AINV_LR
    AINV_HIP_CHECK(hipGetLastError());
    return g;
  }
#define AINV_LD(KK, NPP)                                                     \
  hipLaunchKernelGGL((k_hll_spmv_dot<XT, KK, NPP>), dim3(g), dim3(kBlock), 0, s, \
                     a.n, a.nslices, sp, rl, cl, vl, x, k, ld, y, w, part, \
                     a.perm.as<int32_t>())
#define AINV_LDK(NPP)        \
  if (k <= 1)                \
    AINV_LD(1, NPP);         \
  else if (k <= 2)           \
    AINV_LD(2, NPP);         \
  else if (k <= 4)           \
    AINV_LD(4, NPP);         \
  else if (k <= 8)           \
    AINV_LD(8, NPP);         \
  else if (k <= 16)          \
    AINV_LD(16, NPP);        \
  else if (k <= 32)          \
    AINV_LD(32, NPP);        \
  else                       \
    AINV_LD(64, NPP);
  if (np == 1) {
    AINV_LDK(1)
  } else {
    AINV_LDK(2)
  }
#undef AINV_LDK
#undef AINV_LD
  AINV_HIP_CHECK(hipGetLastError());
  return g;
}

int hll_spmv_dot(const Hll &a, const double *x, double *y, int k, int64_t ld,
                 const double *w, int np, double *part, int map,
                 hipStream_t s, const int32_t *kw) {
  return spmv_dot_impl<double>(a, x, y, k, ld, w, np, part, map, s, kw);
}

int hll_spmv_dot(const Hll &a, const float *x, double *y, int k, int64_t ld,
                 const double *w, int np, double *part, int map,
                 hipStream_t s, const int32_t *kw) {
  return spmv_dot_impl<float>(a, x, y, k, ld, w, np, part, map, s, kw);
}

// Row mapping needs ld = K (a power of two) and 16-byte aligned x and
// output.
static bool is_coop(int map) {
  return map == AINV_MAP_COOP || map == AINV_MAP_COOP128 ||
         map == AINV_MAP_COOP_Q || map == AINV_MAP_COOP128_Q ||
         map == AINV_MAP_RUN;
}

static bool use_row(int map, int k, int64_t ldx, int64_t ldo, const void *x,
                    const void *o) {
  // row: k <= 32; coop and wide: k <= 64 (column chunks); wide also k = 1
  const bool wide = map == AINV_MAP_WIDE;
  if ((map != AINV_MAP_ROW && !is_coop(map) && !wide) || k < (wide ? 1 : 2) ||
      k > (is_coop(map) || wide ? 64 : 32))
    return false;
  int K = wide && k == 1 ? 1 : 2;
  while (K < k) K <<= 1;
  return ldx == K && ldo == K && aligned16(x) && aligned16(o);
}

void hll_spmm_scale(const Hll &a, const void *d, const void *x, bool x_is_f32,
                    int64_t ldx, void *out, int64_t ldo, int k, hipStream_t s,
                    int map, const int32_t *kw) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (x_is_f32) {
    // FP32 x (FP32 / BF16 factors only): FP32 t
    if (a.prec == AINV_PREC_F64)
      throw Error(AINV_ERR_INTERNAL, "FP64 factor with FP32 x");
    const float *xf = static_cast<const float *>(x);
    const float *df = static_cast<const float *>(d);
    float *of = static_cast<float *>(out);
    if (use_row(map, k, ldx, ldo, x, out)) {
      if (a.prec == AINV_PREC_F32)
        launch_lane<float, float, float, float, EPI_SCALE>(map, a, xf, k, df,
                                                           of, 1.0, 0.0, s, kw);
      else
        launch_lane<uint16_t, float, float, float, EPI_SCALE>(
            map, a, xf, k, df, of, 1.0, 0.0, s, kw);
    } else if (a.prec == AINV_PREC_F32) {
      launch_k<float, float, float, float, EPI_SCALE>(a, xf, ldx, k, df, of,
                                                      ldo, 1.0, 0.0, s);
    } else {
      launch_k<uint16_t, float, float, float, EPI_SCALE>(a, xf, ldx, k, df, of,
                                                         ldo, 1.0, 0.0, s);
    }
    return;
  }
  const double *xd = static_cast<const double *>(x);
  if (use_row(map, k, ldx, ldo, x, out)) {
    switch (a.prec) {
      case AINV_PREC_F64:
        launch_lane<double, double, double, double, EPI_SCALE>(
            map, a, xd, k, static_cast<const double *>(d), static_cast<double *>(out),
            1.0, 0.0, s, kw);
        break;
      case AINV_PREC_F32:
        launch_lane<float, double, float, float, EPI_SCALE>(
            map, a, xd, k, static_cast<const float *>(d), static_cast<float *>(out),
            1.0, 0.0, s, kw);
        break;
      default:
        launch_lane<uint16_t, double, float, float, EPI_SCALE>(
            map, a, xd, k, static_cast<const float *>(d), static_cast<float *>(out),
            1.0, 0.0, s, kw);
    }
    return;
  }
  switch (a.prec) {
    case AINV_PREC_F64:
      launch_k<double, double, double, double, EPI_SCALE>(
          a, xd, ldx, k, static_cast<const double *>(d),
          static_cast<double *>(out), ldo, 1.0, 0.0, s);
      break;
    case AINV_PREC_F32:
      launch_k<float, double, float, float, EPI_SCALE>(
          a, xd, ldx, k, static_cast<const float *>(d),
          static_cast<float *>(out), ldo, 1.0, 0.0, s);
      break;
    default:
      launch_k<uint16_t, double, float, float, EPI_SCALE>(
          a, xd, ldx, k, static_cast<const float *>(d),
          static_cast<float *>(out), ldo, 1.0, 0.0, s);
  }
}

void hll_spmm_axpby(const Hll &a, double alpha, const void *x, bool x_is_f32,
                    int64_t ldx, double beta, double *y, int64_t ldy, int k,
                    hipStream_t s, int map, const int32_t *kw) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (use_row(map, k, ldx, ldy, x, y)) {
    switch (a.prec) {
      case AINV_PREC_F64:
        if (x_is_f32) throw Error(AINV_ERR_INTERNAL, "F64 factor with FP32 t");
        launch_lane<double, double, double, double, EPI_AXPBY>(
            map, a, static_cast<const double *>(x), k, (const double *)nullptr, y,
            alpha, beta, s, kw);
        break;
      case AINV_PREC_F32:
        if (x_is_f32)
          launch_lane<float, float, float, double, EPI_AXPBY>(
              map, a, static_cast<const float *>(x), k, (const float *)nullptr, y,
              alpha, beta, s, kw);
        else
          launch_lane<float, double, float, double, EPI_AXPBY>(
              map, a, static_cast<const double *>(x), k, (const float *)nullptr, y,
              alpha, beta, s, kw);
        break;
      default:
        if (x_is_f32)
          launch_lane<uint16_t, float, float, double, EPI_AXPBY>(
              map, a, static_cast<const float *>(x), k, (const float *)nullptr, y,
              alpha, beta, s, kw);
        else
          launch_lane<uint16_t, double, float, double, EPI_AXPBY>(
              map, a, static_cast<const double *>(x), k, (const float *)nullptr, y,
              alpha, beta, s, kw);
    }
    return;
  }
  switch (a.prec) {
    case AINV_PREC_F64:
      if (x_is_f32) throw Error(AINV_ERR_INTERNAL, "F64 factor with FP32 t");
      launch_k<double, double, double, double, EPI_AXPBY>(
          a, static_cast<const double *>(x), ldx, k, (const double *)nullptr,
          y, ldy, alpha, beta, s);
      break;
    case AINV_PREC_F32:
      if (x_is_f32)
        launch_k<float, float, float, double, EPI_AXPBY>(
            a, static_cast<const float *>(x), ldx, k, (const float *)nullptr, y,
            ldy, alpha, beta, s);
      else
        launch_k<float, double, float, double, EPI_AXPBY>(
            a, static_cast<const double *>(x), ldx, k, (const float *)nullptr,
            y, ldy, alpha, beta, s);
      break;
    default:
      if (x_is_f32)
        launch_k<uint16_t, float, float, double, EPI_AXPBY>(
            a, static_cast<const float *>(x), ldx, k, (const float *)nullptr, y,
            ldy, alpha, beta, s);
      else
        launch_k<uint16_t, double, float, double, EPI_AXPBY>(
            a, static_cast<const double *>(x), ldx, k, (const float *)nullptr,
            y, ldy, alpha, beta, s);
  }
}

// y = alpha A x + beta z with FP64 vectors (A FP64 or FP32; z and y
// distinct, all three with the leading dimension ld)
void hll_spmm_resid(const Hll &a, double alpha, const double *x, double beta,
                    const double *z, double *y, int64_t ld, int k,
                    hipStream_t s, int map, const int32_t *kw) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  const bool row = use_row(map, k, ld, ld, x, y) && aligned16(z);
  if (a.prec == AINV_PREC_F64) {
    if (row)
      launch_lane<double, double, double, double, EPI_RESID>(
          map, a, x, k, z, y, alpha, beta, s, kw);
    else
      launch_k<double, double, double, double, EPI_RESID>(a, x, ld, k, z, y,
                                                          ld, alpha, beta, s);
  } else if (a.prec == AINV_PREC_F32) {
    if (row)
      launch_lane<float, double, double, double, EPI_RESID>(
          map, a, x, k, z, y, alpha, beta, s, kw);
    else
      launch_k<float, double, double, double, EPI_RESID>(a, x, ld, k, z, y,
                                                         ld, alpha, beta, s);
  } else {
    throw Error(AINV_ERR_INTERNAL, "residual product of a BF16 matrix");
  }
}

// the same with FP32 vectors (FP32 matrix, FP64 accumulation)
void hll_spmm_resid(const Hll &a, double alpha, const float *x, double beta,
                    const float *z, float *y, int64_t ld, int k,
                    hipStream_t s, int map, const int32_t *kw) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (a.prec != AINV_PREC_F32)
    throw Error(AINV_ERR_INTERNAL, "FP32 residual product needs an FP32 matrix");
  if (use_row(map, k, ld, ld, x, y) && aligned16(z))
    launch_lane<float, float, float, float, EPI_RESID>(map, a, x, k, z, y,
                                                       alpha, beta, s, kw);
  else
    launch_k<float, float, float, float, EPI_RESID>(a, x, ld, k, z, y, ld,
                                                    alpha, beta, s);
}

// FP32 output y = alpha Z t + beta y from an FP32 t (FP32 / BF16 factors):
// the preconditioner output the Krylov engine keeps in FP32.
void hll_spmm_axpby(const Hll &a, double alpha, const void *x, bool x_is_f32,
                    int64_t ldx, double beta, float *y, int64_t ldy, int k,
                    hipStream_t s, int map, const int32_t *kw) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (a.prec == AINV_PREC_F64 || !x_is_f32)
    throw Error(AINV_ERR_INTERNAL, "FP32 output needs FP32 t and factors");
  const float *xf = static_cast<const float *>(x);
  if (use_row(map, k, ldx, ldy, x, y)) {
    if (a.prec == AINV_PREC_F32)
      launch_lane<float, float, float, float, EPI_AXPBY>(
          map, a, xf, k, (const float *)nullptr, y, alpha, beta, s, kw);
    else
      launch_lane<uint16_t, float, float, float, EPI_AXPBY>(
          map, a, xf, k, (const float *)nullptr, y, alpha, beta, s, kw);
  } else if (a.prec == AINV_PREC_F32) {
    launch_k<float, float, float, float, EPI_AXPBY>(
        a, xf, ldx, k, (const float *)nullptr, y, ldy, alpha, beta, s);
  } else {
    launch_k<uint16_t, float, float, float, EPI_AXPBY>(
        a, xf, ldx, k, (const float *)nullptr, y, ldy, alpha, beta, s);
  }
}

}  // namespace ainv
