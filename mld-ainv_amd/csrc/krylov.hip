// Preconditioned BiCGSTAB and CG on the GPU for k interleaved right-hand
// sides, each column with its own recurrence and stopping test (istop = 2,
// ||r||_2 <= eps ||b||_2), restating PSBLAS psb_dbicgstab / psb_dcg as the
// reference drivers call them (tests/pdegen/ppde3d.f90:280,
// tests/fileread/df_sample.f90:368-369).
//
// Everything stays FP64 and on the device: vector updates are fused with
// the dot products that follow them (one pass over the vectors per fused
// step), per-column scalars (rho, alpha, omega, ...) are computed by a
// one-block kernel from deterministic block partial sums and read back by
// the next kernel from device memory.  The host only reads the number of
// still-active columns once per iteration to decide whether to continue.
#include <cmath>

#include "objects.h"

namespace ainv {
namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 1024;
constexpr int kMaxPairs = 3;
constexpr int kMaxCols = 64;

// per-column scalars of the recurrences
struct Col {
  double rho, rho_old, rho_next, alpha, omega, beta, bn, err;
  int32_t active, iter;
};

struct Vecs {
  double *x, *b, *r, *q, *p, *v, *f, *s, *z, *t;
};

enum Op {
  // BiCGSTAB
  OP_INIT_BICG,  // dots (b,b) (r,r) (q,r)
  OP_P_BICG,     // p = r (it 1) or r + beta (p - omega v)
  OP_DOT_QV,     // dot (q,v)
  OP_S_BICG,     // s = r - alpha v
  OP_DOT_TT_TS,  // dots (t,t) (t,s)
  OP_XR_BICG,    // x += alpha f + omega z ; r = s - omega t ; (r,r) (q,r)
  // CG
  OP_INIT_CG,  // dots (b,b) (r,r)
  OP_DOT_RZ,   // dot (r,z)
  OP_P_CG,     // p = z (it 1) or z + beta p
  OP_DOT_PQ,   // dot (p,q)   (q holds A p)
  OP_XR_CG,    // x += alpha p ; r -= alpha q ; (r,r)
};

template <int OP> struct NPairs { static constexpr int v = 0; };
template <> struct NPairs<OP_INIT_BICG> { static constexpr int v = 3; };
template <> struct NPairs<OP_DOT_QV> { static constexpr int v = 1; };
template <> struct NPairs<OP_DOT_TT_TS> { static constexpr int v = 2; };
template <> struct NPairs<OP_XR_BICG> { static constexpr int v = 2; };
template <> struct NPairs<OP_INIT_CG> { static constexpr int v = 2; };
template <> struct NPairs<OP_DOT_RZ> { static constexpr int v = 1; };
template <> struct NPairs<OP_DOT_PQ> { static constexpr int v = 1; };
template <> struct NPairs<OP_XR_CG> { static constexpr int v = 1; };

// One fused pass over the n x k vectors.  Thread (rr, c) of a block works on
// column c of rows rr, rr + R, ...; partial sums per column are reduced in
// LDS in a fixed tree order and written per block.
template <int OP, int K>
__global__ __launch_bounds__(kThreads) void k_vec(int32_t n, int k, Vecs v,
                                                  const Col *__restrict__ col,
                                                  int it,
                                                  double *__restrict__ part) {
  constexpr int NP = NPairs<OP>::v;
  constexpr int R = kThreads / K;
  const int c = threadIdx.x % K;
  const int rr = threadIdx.x / K;
  const bool cok = c < k && col[c < k ? c : 0].active;
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int q = 0; q < (NP > 0 ? NP : 1); ++q) acc[q] = 0.0;
  if (cok) {
    const Col cs = col[c];
    for (int64_t r = (int64_t)blockIdx.x * R + rr; r < n;
         r += (int64_t)gridDim.x * R) {
      const int64_t e = r * k + c;
      if constexpr (OP == OP_INIT_BICG) {
        acc[0] += v.b[e] * v.b[e];
        acc[1] += v.r[e] * v.r[e];
        acc[2] += v.q[e] * v.r[e];
      } else if constexpr (OP == OP_P_BICG) {
        if (it == 1)
          v.p[e] = v.r[e];
        else
          v.p[e] = v.r[e] + cs.beta * (v.p[e] + (-cs.omega) * v.v[e]);
      } else if constexpr (OP == OP_DOT_QV) {
        acc[0] += v.q[e] * v.v[e];
      } else if constexpr (OP == OP_S_BICG) {
        v.s[e] = v.r[e] + (-cs.alpha) * v.v[e];
      } else if constexpr (OP == OP_DOT_TT_TS) {
        acc[0] += v.t[e] * v.t[e];
        acc[1] += v.t[e] * v.s[e];
      } else if constexpr (OP == OP_XR_BICG) {
        v.x[e] = v.x[e] + cs.alpha * v.f[e] + cs.omega * v.z[e];
        const double rn = v.s[e] + (-cs.omega) * v.t[e];
        v.r[e] = rn;
        acc[0] += rn * rn;
        acc[1] += v.q[e] * rn;
      } else if constexpr (OP == OP_INIT_CG) {
        acc[0] += v.b[e] * v.b[e];
        acc[1] += v.r[e] * v.r[e];
      } else if constexpr (OP == OP_DOT_RZ) {
        acc[0] += v.r[e] * v.z[e];
      } else if constexpr (OP == OP_P_CG) {
        if (it == 1)
          v.p[e] = v.z[e];
        else
          v.p[e] = v.z[e] + cs.beta * v.p[e];
      } else if constexpr (OP == OP_DOT_PQ) {
        acc[0] += v.p[e] * v.q[e];
      } else if constexpr (OP == OP_XR_CG) {
        v.x[e] = v.x[e] + cs.alpha * v.p[e];
        const double rn = v.r[e] + (-cs.alpha) * v.q[e];
        v.r[e] = rn;
        acc[0] += rn * rn;
      }
    }
  }
  if constexpr (NP > 0) {
    __shared__ double red[NP][kThreads];
#pragma unroll
    for (int q = 0; q < NP; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = R / 2; st >= 1; st >>= 1) {
      if (rr < st) {
#pragma unroll
        for (int q = 0; q < NP; ++q)
          red[q][threadIdx.x] += red[q][threadIdx.x + st * K];
      }
      __syncthreads();
    }
    if (rr == 0 && c < k) {
#pragma unroll
      for (int q = 0; q < NP; ++q)
        part[((int64_t)q * kRedBlocks + blockIdx.x) * kMaxCols + c] =
            red[q][threadIdx.x];
    }
  }
}

// Scalar step of the recurrences: thread c sums its column's partials over
// the blocks in order and updates the column state.
enum Fin {
  FIN_INIT,        // bn, err, active, rho_next (BiCGSTAB) from init dots
  FIN_RHO_BICG,    // rho = rho_next, beta, breakdown on rho == 0
  FIN_ALPHA_BICG,  // alpha = rho / (q,v)
  FIN_OMEGA,       // omega = (t,s)/(t,t)
  FIN_CONV_BICG,   // err, convergence, rho_next = (q,r)
  FIN_INIT_CG,
  FIN_RHO_CG,    // rho = (r,z), beta
  FIN_ALPHA_CG,  // alpha = rho / (p,q)
  FIN_CONV_CG,
};

__device__ double sum_part(const double *part, int q, int c, int nblk) {
  double s = 0.0;
  for (int b = 0; b < nblk; ++b)
    s += part[((int64_t)q * kRedBlocks + b) * kMaxCols + c];
  return s;
}

__global__ void k_fin(int fin, int k, int nblk, int it, double eps,
                      const double *__restrict__ part, Col *__restrict__ col,
                      int32_t *__restrict__ nactive) {
  __shared__ int32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int c = threadIdx.x;
  if (c < k) {
    Col cs = col[c];
    if (fin == FIN_INIT || fin == FIN_INIT_CG) {
      const double bb = sum_part(part, 0, c, nblk);
      const double rr = sum_part(part, 1, c, nblk);
      cs.bn = sqrt(bb);
      const double rn = sqrt(rr);
      cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
      cs.active = !(cs.err <= eps);
      cs.iter = 0;
      cs.rho = 0.0;
      cs.rho_old = 0.0;
      cs.alpha = cs.omega = cs.beta = 0.0;
      if (fin == FIN_INIT) cs.rho_next = sum_part(part, 2, c, nblk);
    } else if (cs.active) {
      switch (fin) {
        case FIN_RHO_BICG:
          cs.iter = it;
          cs.rho_old = cs.rho;
          cs.rho = cs.rho_next;
          if (cs.rho == 0.0) {
            cs.active = 0;
          } else if (it > 1) {
            cs.beta = (cs.rho / cs.rho_old) * (cs.alpha / cs.omega);
          }
          break;
        case FIN_ALPHA_BICG: {
          const double sigma = sum_part(part, 0, c, nblk);
          if (sigma == 0.0)
            cs.active = 0;
          else
            cs.alpha = cs.rho / sigma;
          break;
        }
        case FIN_OMEGA: {
          const double tt = sum_part(part, 0, c, nblk);
          const double ts = sum_part(part, 1, c, nblk);
          if (tt == 0.0) {
            cs.active = 0;
          } else {
            cs.omega = ts / tt;
            if (cs.omega == 0.0) cs.active = 0;
          }
          break;
        }
        case FIN_CONV_BICG: {
          const double rn = sqrt(sum_part(part, 0, c, nblk));
          cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
          cs.rho_next = sum_part(part, 1, c, nblk);
          if (cs.err <= eps) cs.active = 0;
          break;
        }
        case FIN_RHO_CG: {
          cs.iter = it;
          cs.rho_old = cs.rho;
          cs.rho = sum_part(part, 0, c, nblk);
          if (it > 1) cs.beta = cs.rho / cs.rho_old;
          break;
        }
        case FIN_ALPHA_CG: {
          const double sigma = sum_part(part, 0, c, nblk);
          if (sigma == 0.0)
            cs.active = 0;
          else
            cs.alpha = cs.rho / sigma;
          break;
        }
        case FIN_CONV_CG: {
          const double rn = sqrt(sum_part(part, 0, c, nblk));
          cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
          if (cs.err <= eps) cs.active = 0;
          break;
        }
      }
    }
    col[c] = cs;
    if (cs.active) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *nactive = cnt;
}

int pow2_at_least(int k) {
  int K = 1;
  while (K < k) K <<= 1;
  return K;
}

int blocks_for(int32_t n, int K) {
  int64_t R = kThreads / K;
  int64_t g = (n + R - 1) / R;
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return (int)g;
}

template <int OP>
void vec_op(int32_t n, int k, const Vecs &v, const Col *col, int it,
            double *part, hipStream_t s) {
  const int K = pow2_at_least(k);
  const int g = blocks_for(n, K);
#define AINV_VL(KK)                                                        \
  hipLaunchKernelGGL((k_vec<OP, KK>), dim3(g), dim3(kThreads), 0, s, n, k, \
                     v, col, it, part)
  switch (K) {
    case 1: AINV_VL(1); break;
    case 2: AINV_VL(2); break;
    case 4: AINV_VL(4); break;
    case 8: AINV_VL(8); break;
    case 16: AINV_VL(16); break;
    case 32: AINV_VL(32); break;
    default: AINV_VL(64); break;
  }
#undef AINV_VL
  AINV_HIP_CHECK(hipGetLastError());
}

void fin(int which, int32_t n, int k, int it, double eps, const double *part,
         Col *col, int32_t *nact, hipStream_t s) {
  const int g = blocks_for(n, pow2_at_least(k));
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kMaxCols), 0, s, which, k, g, it,
                     eps, part, col, nact);
  AINV_HIP_CHECK(hipGetLastError());
}

void precond(ainv_solver *m, const double *x, double *y, int32_t n, int k,
             hipStream_t s) {
  if (m)
    solver_apply(m, 1.0, x, 0.0, y, 'N', k, k, k, s);
  else
    AINV_HIP_CHECK(hipMemcpyAsync(y, x, (size_t)n * k * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
}

}  // namespace

void krylov_solve(const std::string &method, ainv_matrix *a, ainv_solver *m,
                  const double *b, double *x, int k, double eps, int itmax,
                  int32_t *iter, double *err, hipStream_t s) {
  const int32_t n = a->n;
  if (k < 1 || k > kMaxCols) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (m && m->n != n)
    throw Error(AINV_ERR_ARG, "preconditioner and matrix sizes differ");
  const bool cg = (method == "CG");
  if (!cg && method != "BICGSTAB")
    throw Error(AINV_ERR_ARG, "unknown Krylov method " + method +
                                  " (BICGSTAB, CG available)");
  const size_t vb = (size_t)n * k * sizeof(double);
  DevBuf work, colbuf, part, nact;
  const int nvec = cg ? 4 : 8;
  work.alloc(vb * nvec);
  colbuf.alloc(sizeof(Col) * kMaxCols);
  part.alloc(sizeof(double) * kMaxPairs * kRedBlocks * kMaxCols);
  nact.alloc(sizeof(int32_t));
  AINV_HIP_CHECK(hipMemsetAsync(colbuf.p, 0, colbuf.bytes, s));
  double *w0 = work.as<double>();
  const int64_t nk = (int64_t)n * k;
  Vecs v{};
  v.x = x;
  v.b = const_cast<double *>(b);
  v.r = w0;
  v.p = w0 + nk;
  v.z = w0 + 2 * nk;
  v.q = w0 + 3 * nk;  // CG: q = A p ; BiCGSTAB: rtilde
  if (!cg) {
    v.v = w0 + 4 * nk;
    v.f = w0 + 5 * nk;
    v.s = w0 + 6 * nk;
    v.t = w0 + 7 * nk;
  }
  Col *col = colbuf.as<Col>();
  double *pt = part.as<double>();
  int32_t *na = nact.as<int32_t>();
  // r = b - A x
  AINV_HIP_CHECK(hipMemcpyAsync(v.r, b, vb, hipMemcpyDeviceToDevice, s));
  hll_spmm_axpby(a->hll, -1.0, x, false, k, 1.0, v.r, k, k, s);
  // column state starts all-active so the init pass runs every column
  std::vector<Col> init(kMaxCols);
  for (auto &c0 : init) c0 = Col{0, 0, 0, 0, 0, 0, 0, 0, 1, 0};
  AINV_HIP_CHECK(hipMemcpyAsync(col, init.data(), sizeof(Col) * kMaxCols,
                                hipMemcpyHostToDevice, s));
  if (!cg) {
    AINV_HIP_CHECK(hipMemcpyAsync(v.q, v.r, vb, hipMemcpyDeviceToDevice, s));
    vec_op<OP_INIT_BICG>(n, k, v, col, 0, pt, s);
    fin(FIN_INIT, n, k, 0, eps, pt, col, na, s);
  } else {
    vec_op<OP_INIT_CG>(n, k, v, col, 0, pt, s);
    fin(FIN_INIT_CG, n, k, 0, eps, pt, col, na, s);
  }
  int32_t active = 0;
  AINV_HIP_CHECK(hipMemcpyAsync(&active, na, sizeof(int32_t),
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  int it = 0;
  while (active > 0 && it < itmax) {
    ++it;
    if (!cg) {
      fin(FIN_RHO_BICG, n, k, it, eps, pt, col, na, s);
      vec_op<OP_P_BICG>(n, k, v, col, it, pt, s);
      precond(m, v.p, v.f, n, k, s);
      hll_spmm_axpby(a->hll, 1.0, v.f, false, k, 0.0, v.v, k, k, s);
      vec_op<OP_DOT_QV>(n, k, v, col, it, pt, s);
      fin(FIN_ALPHA_BICG, n, k, it, eps, pt, col, na, s);
      vec_op<OP_S_BICG>(n, k, v, col, it, pt, s);
      precond(m, v.s, v.z, n, k, s);
      hll_spmm_axpby(a->hll, 1.0, v.z, false, k, 0.0, v.t, k, k, s);
      vec_op<OP_DOT_TT_TS>(n, k, v, col, it, pt, s);
      fin(FIN_OMEGA, n, k, it, eps, pt, col, na, s);
      vec_op<OP_XR_BICG>(n, k, v, col, it, pt, s);
      fin(FIN_CONV_BICG, n, k, it, eps, pt, col, na, s);
    } else {
      precond(m, v.r, v.z, n, k, s);
      vec_op<OP_DOT_RZ>(n, k, v, col, it, pt, s);
      fin(FIN_RHO_CG, n, k, it, eps, pt, col, na, s);
      vec_op<OP_P_CG>(n, k, v, col, it, pt, s);
      hll_spmm_axpby(a->hll, 1.0, v.p, false, k, 0.0, v.q, k, k, s);
      vec_op<OP_DOT_PQ>(n, k, v, col, it, pt, s);
      fin(FIN_ALPHA_CG, n, k, it, eps, pt, col, na, s);
      vec_op<OP_XR_CG>(n, k, v, col, it, pt, s);
      fin(FIN_CONV_CG, n, k, it, eps, pt, col, na, s);
    }
    AINV_HIP_CHECK(hipMemcpyAsync(&active, na, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
  }
  std::vector<Col> hc(k);
  AINV_HIP_CHECK(hipMemcpyAsync(hc.data(), col, sizeof(Col) * k,
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    // a column still active at itmax ran every iteration
    iter[c] = hc[c].active ? it : hc[c].iter;
    err[c] = hc[c].err;
  }
}

}  // namespace ainv
