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
#pragma unroll 1
  for (int pass = 0; pass < K; ++pass) {
    const int ri = pass * G + g;
    const int32_t r = s * kSlice + ri;
    if (r >= n) break;  // rows beyond n only occur in the last passes
    const int32_t len = rlen[r];
    const int64_t p0 = base + ri;
    A acc0 = 0, acc1 = 0;
    int32_t j = 0;
    // two independent chains keep more gathers in flight
    for (; j + 1 < len; j += 2) {
      const int64_t pa = p0 + (int64_t)j * kSlice;
      const int64_t pb = pa + kSlice;
      const int32_t ca = col[pa], cb = col[pb];
      const A va = Val<VT>::get(val[pa]), vb = Val<VT>::get(val[pb]);
      if (col_ok) {
        acc0 += va * cvt<A>(x[(int64_t)ca * ldx + c]);
        acc1 += vb * cvt<A>(x[(int64_t)cb * ldx + c]);
      }
    }
    if (j < len) {
      const int64_t pa = p0 + (int64_t)j * kSlice;
      const int32_t ca = col[pa];
      const A va = Val<VT>::get(val[pa]);
      if (col_ok) acc0 += va * cvt<A>(x[(int64_t)ca * ldx + c]);
    }
    const A acc = acc0 + acc1;
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

}  // namespace

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
