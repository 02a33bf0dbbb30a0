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
// The algorithmic traffic per launch is (nnz)(value + 4) bytes of factor,
// 4 bytes of row length per row, the gathered x and the output; no kernel
// here re-reads a factor word.
#include "ainv_internal.h"

namespace ainv {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWavesPerBlock * kSlice;

template <typename VT> struct Val;
template <> struct Val<double> {
  using acc = double;
  __device__ static double get(double v) { return v; }
};
template <> struct Val<float> {
  using acc = double;
  __device__ static double get(float v) { return (double)v; }
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
  A acc0 = 0, acc1 = 0;
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
#pragma unroll
      for (int u = 0; u < kGroup; u += 2) {
        acc0 += vi[u] * xi[u];
        if (u + 1 < kGroup) acc1 += vi[u + 1] * xi[u + 1];
      }
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
  return acc0 + acc1;
}

enum { EPI_SCALE = 0, EPI_AXPBY = 1 };

template <typename VT, typename XT, typename DT, typename OT, int K, int EPI>
__global__ __launch_bounds__(kBlock) void k_hll_spmm(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ rlen, const int32_t *__restrict__ col,
    const VT *__restrict__ val, const XT *__restrict__ x, int64_t ldx, int k,
    const DT *__restrict__ dvec, OT *__restrict__ out, int64_t ldo,
    double alpha, double beta) {
  using A = typename Val<VT>::acc;
  constexpr int G = kSlice / K;  // rows per pass
  const int lane = threadIdx.x & (kSlice - 1);
  const int32_t s = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
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
    if constexpr (EPI == EPI_SCALE) {
      out[(int64_t)r * ldo + c] = (OT)((double)dvec[r] * (double)acc);
    } else {
      double v = alpha * (double)acc;
      OT *yp = out + (int64_t)r * ldo + c;
      if (beta != 0.0) v += beta * (double)(*yp);
      *yp = (OT)v;
    }
  }
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
                     ldo, alpha, beta)
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
// slices with a fixed grid (at most kRedBlocks blocks), per-column partial
// sums reduced across the row groups of a wave by shuffles and across the
// waves of a block in LDS, in a fixed order.
template <int K, int NP>
__global__ __launch_bounds__(kBlock) void k_hll_spmv_dot(
    int32_t n, int32_t nslices, const int64_t *__restrict__ sptr,
    const int32_t *__restrict__ rlen, const int32_t *__restrict__ col,
    const double *__restrict__ val, const double *__restrict__ x, int k,
    double *__restrict__ y, const double *__restrict__ w,
    double *__restrict__ part) {
  constexpr int G = kSlice / K;
  const int lane = threadIdx.x & (kSlice - 1);
  const int wave = threadIdx.x >> 6;
  const int c = (K == 1) ? 0 : (lane % K);
  const int g = (K == 1) ? lane : (lane / K);
  const bool col_ok = c < k;
  double d0 = 0.0, d1 = 0.0;
  for (int32_t s = blockIdx.x * kWavesPerBlock + wave; s < nslices;
       s += gridDim.x * kWavesPerBlock) {
    const int64_t base = sptr[s];
    for (int pass = 0; pass < K; ++pass) {
      const int ri = pass * G + g;
      const int32_t r = s * kSlice + ri;
      if (r >= n) break;
      const int32_t len = rlen[r];
      const double yv =
          row_dot<double, double>(base + ri, len, col, val, x, k, c, col_ok);
      if (!col_ok) continue;
      const int64_t e = (int64_t)r * k + c;
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

}  // namespace

int hll_spmv_dot(const Hll &a, const double *x, double *y, int k,
                 const double *w, int np, double *part, hipStream_t s) {
  if (a.prec != AINV_PREC_F64)
    throw Error(AINV_ERR_INTERNAL, "system matrix must be FP64");
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  int g = (a.nslices + kWavesPerBlock - 1) / kWavesPerBlock;
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  const int64_t *sp = a.sptr.as<int64_t>();
  const int32_t *rl = a.rlen.as<int32_t>();
  const int32_t *cl = a.col.as<int32_t>();
  const double *vl = a.val.as<double>();
#define AINV_LD(KK, NPP)                                                     \
  hipLaunchKernelGGL((k_hll_spmv_dot<KK, NPP>), dim3(g), dim3(kBlock), 0, s, \
                     a.n, a.nslices, sp, rl, cl, vl, x, k, y, w, part)
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

void hll_spmm_scale(const Hll &a, const void *d, const void *x, bool x_is_f32,
                    int64_t ldx, void *out, int64_t ldo, int k, hipStream_t s) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (x_is_f32) throw Error(AINV_ERR_INTERNAL, "scale pass reads FP64 x");
  const double *xd = static_cast<const double *>(x);
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
                    hipStream_t s) {
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
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

}  // namespace ainv
