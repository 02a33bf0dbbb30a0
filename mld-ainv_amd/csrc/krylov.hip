// Preconditioned BiCGSTAB and CG on the GPU for k interleaved right-hand
// sides, each column with its own recurrence and stopping test (istop = 2,
// ||r||_2 <= eps ||b||_2), restating PSBLAS psb_dbicgstab / psb_dcg as the
// reference drivers call them (tests/pdegen/ppde3d.f90:280,
// tests/fileread/df_sample.f90:368-369).
//
// Everything stays FP64 and on the device.  Vector updates are fused with
// the dot products that follow them, and the dot products of A*v with the
// system-matrix product itself (one pass each).  Per-column scalars (rho,
// alpha, omega, the iteration count, eps) live in device memory and are
// updated by one block per column from deterministic block partial sums,
// so no kernel argument changes from one iteration to the next: one
// iteration is captured once as a hipGraph and replayed.  The host reads
// the number of still-active columns through pinned memory one iteration
// behind the device, so the GPU never waits for it; the iteration launched
// after the last column converged changes nothing (every kernel skips
// inactive columns and their counts stop).
//
// Workspaces and captured graphs are cached per (matrix, method,
// preconditioner, k) and rebuilt when the preconditioner's factors change.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>

#include "objects.h"

namespace ainv {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPairs = 3;
constexpr int kPlansPerMatrix = 12;

// per-column state of the recurrences
struct Col {
  double rho, rho_old, rho_next, alpha, omega, beta, bn, err, eps;
  int32_t active, iter;
  // A BiCGSTAB column admitted into the served form (x0 = 0) starts with
  // only r~ written into the workspace: until its first x / r update
  // (lazy = 1) the passes read r = b from its right-hand side (src, row
  // stride sstride), take x = 0, and an extraction gives zeros, so that an
  // admission writes one interleaved column instead of four (scattered
  // 8-byte writes into n rows, each a partial-line update in HBM).
  int32_t lazy, pad_;
  const double *src;
  int64_t sstride;
};

struct Vecs {
  double *x, *b, *r, *q, *p, *v, *f, *s, *z, *t;
  // FP32 / BF16 factors: FP32 copies of the preconditioner's inputs,
  // written by the passes that form them (BiCGSTAB: p and s; CG: r), and
  // the preconditioned vectors kept in FP32 (BiCGSTAB: f = M p, z = M s,
  // whose x update and products A f, A z read the same rounded values; CG:
  // z = M r, read by (r, z) and p = z + beta p)
  float *p32, *s32, *r32, *f32, *z32;
};

enum Op {
  OP_INIT_BICG,  // dots (b,b) (r,r) (q,r)
  OP_P_BICG,     // p = r (first step) or r + beta (p - omega v)
  OP_S_BICG,     // s = r - alpha v
  OP_XR_BICG,    // x += alpha f + omega z ; r = s - omega t ; (r,r) (q,r)
  OP_INIT_CG,    // dots (b,b) (r,r)
  OP_DOT_RZ,     // dot (r,z)
  OP_P_CG,       // p = z (first step) or z + beta p
  OP_XR_CG,      // x += alpha p ; r -= alpha q ; (r,r)
  OP_U_CGS,      // u = r + beta q ; p = u + beta (q + beta p) (first: r)
  OP_Q_CGS,      // q = u - alpha vhat ; z = u + q
  OP_XR_CGS,     // x += alpha uhat ; r -= alpha qhat ; (r,r) (rtilde,r)
  OP_DOT_ZRT,    // BiCG: (z, rtilde)
  OP_P_BICGM,    // BiCG: p = z + beta p ; ptilde = ztilde + beta ptilde
  OP_XR_BICGM,   // BiCG: x += alpha p ; r -= alpha q ; rtilde -= alpha qt
};

template <int OP> struct NPairs { static constexpr int v = 0; };
template <> struct NPairs<OP_INIT_BICG> { static constexpr int v = 3; };
template <> struct NPairs<OP_XR_BICG> { static constexpr int v = 2; };
template <> struct NPairs<OP_INIT_CG> { static constexpr int v = 2; };
template <> struct NPairs<OP_DOT_RZ> { static constexpr int v = 1; };
template <> struct NPairs<OP_XR_CG> { static constexpr int v = 1; };
template <> struct NPairs<OP_XR_CGS> { static constexpr int v = 2; };
template <> struct NPairs<OP_DOT_ZRT> { static constexpr int v = 1; };
template <> struct NPairs<OP_XR_BICGM> { static constexpr int v = 1; };

// One fused pass over the n x k vectors.  Thread (rr, c) of a block works on
// column c of rows rr, rr + R, ...; partial sums per column are reduced in
// LDS in a fixed tree order and written per block.
template <int OP, int K>
__global__ __launch_bounds__(kThreads) void k_vec(int32_t n, int k, Vecs v,
                                                  const Col *__restrict__ col,
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
      const int64_t e = r * K + c;
      if constexpr (OP == OP_INIT_BICG) {
        acc[0] += v.b[e] * v.b[e];
        acc[1] += v.r[e] * v.r[e];
        acc[2] += v.q[e] * v.r[e];
      } else if constexpr (OP == OP_P_BICG) {
        const double re = cs.lazy ? cs.src[r * cs.sstride] : v.r[e];
        const double pn =
            cs.iter == 1 ? re
                         : re + cs.beta * (v.p[e] + (-cs.omega) * v.v[e]);
        v.p[e] = pn;
        if (v.p32) v.p32[e] = (float)pn;
      } else if constexpr (OP == OP_S_BICG) {
        const double re = cs.lazy ? cs.src[r * cs.sstride] : v.r[e];
        const double sn = re + (-cs.alpha) * v.v[e];
        v.s[e] = sn;
        if (v.s32) v.s32[e] = (float)sn;
      } else if constexpr (OP == OP_XR_BICG) {
        const double fe = v.f32 ? (double)v.f32[e] : v.f[e];
        const double ze = v.z32 ? (double)v.z32[e] : v.z[e];
        const double xe = cs.lazy ? 0.0 : v.x[e];
        v.x[e] = xe + cs.alpha * fe + cs.omega * ze;
        const double rn = v.s[e] + (-cs.omega) * v.t[e];
        v.r[e] = rn;
        acc[0] += rn * rn;
        acc[1] += v.q[e] * rn;
      } else if constexpr (OP == OP_INIT_CG) {
        acc[0] += v.b[e] * v.b[e];
        acc[1] += v.r[e] * v.r[e];
        if (v.r32) v.r32[e] = (float)v.r[e];
      } else if constexpr (OP == OP_DOT_RZ) {
        acc[0] += v.r[e] * (v.z32 ? (double)v.z32[e] : v.z[e]);
      } else if constexpr (OP == OP_P_CG) {
        const double ze = v.z32 ? (double)v.z32[e] : v.z[e];
        if (cs.iter == 1)
          v.p[e] = ze;
        else
          v.p[e] = ze + cs.beta * v.p[e];
      } else if constexpr (OP == OP_XR_CG) {
        v.x[e] = v.x[e] + cs.alpha * v.p[e];
        const double rn = v.r[e] + (-cs.alpha) * v.q[e];
        v.r[e] = rn;
        if (v.r32) v.r32[e] = (float)rn;
        acc[0] += rn * rn;
      } else if constexpr (OP == OP_U_CGS) {
        // u in s, the q of CGS in t
        if (cs.iter == 1) {
          v.s[e] = v.r[e];
          v.p[e] = v.r[e];
        } else {
          const double u = v.r[e] + cs.beta * v.t[e];
          v.s[e] = u;
          v.p[e] = u + cs.beta * (v.t[e] + cs.beta * v.p[e]);
        }
      } else if constexpr (OP == OP_Q_CGS) {
        const double qn = v.s[e] + (-cs.alpha) * v.v[e];
        v.t[e] = qn;
        v.z[e] = v.s[e] + qn;
      } else if constexpr (OP == OP_DOT_ZRT) {
        acc[0] += v.z[e] * v.q[e];
      } else if constexpr (OP == OP_P_BICGM) {
        // ztilde in f, ptilde in s
        if (cs.iter == 1) {
          v.p[e] = v.z[e];
          v.s[e] = v.f[e];
        } else {
          v.p[e] = v.z[e] + cs.beta * v.p[e];
          v.s[e] = v.f[e] + cs.beta * v.s[e];
        }
      } else if constexpr (OP == OP_XR_BICGM) {
        // q of BiCG in v, qtilde in t, rtilde in q
        v.x[e] = v.x[e] + cs.alpha * v.p[e];
        const double rn = v.r[e] + (-cs.alpha) * v.v[e];
        v.r[e] = rn;
        v.q[e] = v.q[e] + (-cs.alpha) * v.t[e];
        acc[0] += rn * rn;
      } else if constexpr (OP == OP_XR_CGS) {
        v.x[e] = v.x[e] + cs.alpha * v.f[e];
        const double rn = v.r[e] + (-cs.alpha) * v.v[e];
        v.r[e] = rn;
        acc[0] += rn * rn;
        acc[1] += v.q[e] * rn;
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

// Scalar step of the recurrences, one block per column: the block sums its
// column's partials (strided, then a fixed tree) and thread 0 updates the
// column state.  nactive: zeroed at the start of an iteration (RHO steps),
// incremented by every column still active at its end (INIT / CONV).
enum Fin {
  FIN_INIT,        // bn, err, active, rho_next (BiCGSTAB)
  FIN_RHO_BICG,    // iter++, rho = rho_next, beta; breakdown on rho == 0
  FIN_ALPHA_BICG,  // alpha = rho / (q,v)
  FIN_OMEGA,       // omega = (t,s)/(t,t)
  FIN_CONV_BICG,   // err, convergence, rho_next = (q,r)
  FIN_INIT_CG,
  FIN_RHO_CG,    // iter++, rho = (r,z), beta
  FIN_ALPHA_CG,  // alpha = rho / (p,q)
  FIN_CONV_CG,
  FIN_RHO_CGS,   // iter++, rho = rho_next, beta = rho / rho_old
  FIN_RHO_BICGM, // BiCG: iter++, rho = (z, rtilde), beta = rho / rho_old
};

__host__ __device__ inline int fin_pairs(int fin) {
  switch (fin) {
    case FIN_INIT: return 3;
    case FIN_INIT_CG:
    case FIN_OMEGA:
    case FIN_CONV_BICG: return 2;
    case FIN_RHO_BICG:
    case FIN_RHO_CGS: return 0;
    case FIN_RHO_BICGM: return 1;
    default: return 1;
  }
}

__global__ __launch_bounds__(kThreads) void k_fin(
    int fin, int nblk, const double *__restrict__ part, Col *__restrict__ col,
    int32_t *__restrict__ nactive, const int32_t *__restrict__ mask) {
  const int c = blockIdx.x;
  // an admission initializes only the columns it fills
  if (mask && !mask[c]) return;
  const int np = fin_pairs(fin);
  __shared__ double red[kMaxPairs][kThreads];
  for (int q = 0; q < np; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads)
      s += part[((int64_t)q * kRedBlocks + b) * kMaxCols + c];
    red[q][threadIdx.x] = s;
  }
  __syncthreads();
  for (int st = kThreads / 2; st >= 1; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < np; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if ((fin == FIN_RHO_BICG || fin == FIN_RHO_CG || fin == FIN_RHO_CGS ||
       fin == FIN_RHO_BICGM) &&
      c == 0)
    *nactive = 0;
  Col cs = col[c];
  const double s0 = red[0][0], s1 = red[1][0], s2 = red[2][0];
  if (fin == FIN_INIT || fin == FIN_INIT_CG) {
    cs.bn = sqrt(s0);
    const double rn = sqrt(s1);
    cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
    cs.active = !(cs.err <= cs.eps);
    cs.iter = 0;
    cs.rho = cs.rho_old = 0.0;
    cs.alpha = cs.omega = cs.beta = 0.0;
    cs.rho_next = (fin == FIN_INIT) ? s2 : 0.0;
  } else if (cs.active) {
    switch (fin) {
      case FIN_RHO_BICG:
        cs.iter += 1;
        cs.rho_old = cs.rho;
        cs.rho = cs.rho_next;
        if (cs.rho == 0.0)
          cs.active = 0;
        else if (cs.iter > 1)
          cs.beta = (cs.rho / cs.rho_old) * (cs.alpha / cs.omega);
        break;
      case FIN_ALPHA_BICG:
      case FIN_ALPHA_CG:
        if (s0 == 0.0)
          cs.active = 0;
        else
          cs.alpha = cs.rho / s0;
        break;
      case FIN_OMEGA:
        if (s0 == 0.0) {
          cs.active = 0;
        } else {
          cs.omega = s1 / s0;
          if (cs.omega == 0.0) cs.active = 0;
        }
        break;
      case FIN_CONV_BICG: {
        cs.lazy = 0;  // the x / r update of this iteration wrote x and r
        const double rn = sqrt(s0);
        cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
        cs.rho_next = s1;
        if (cs.err <= cs.eps) cs.active = 0;
        break;
      }
      case FIN_RHO_CGS:
        cs.iter += 1;
        cs.rho_old = cs.rho;
        cs.rho = cs.rho_next;
        if (cs.rho == 0.0)
          cs.active = 0;
        else if (cs.iter > 1)
          cs.beta = cs.rho / cs.rho_old;
        break;
      case FIN_RHO_BICGM:
        cs.iter += 1;
        cs.rho_old = cs.rho;
        cs.rho = s0;
        if (cs.rho == 0.0)
          cs.active = 0;
        else if (cs.iter > 1)
          cs.beta = cs.rho / cs.rho_old;
        break;
      case FIN_RHO_CG:
        cs.iter += 1;
        cs.rho_old = cs.rho;
        cs.rho = s0;
        if (cs.iter > 1) cs.beta = cs.rho / cs.rho_old;
        break;
      case FIN_CONV_CG: {
        const double rn = sqrt(s0);
        cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
        if (cs.err <= cs.eps) cs.active = 0;
        break;
      }
    }
  }
  col[c] = cs;
  if ((fin == FIN_INIT || fin == FIN_INIT_CG || fin == FIN_CONV_BICG ||
       fin == FIN_CONV_CG) &&
      cs.active) {
    atomicAdd(nactive, 1);
    atomicMax(nactive + 1, c + 1);  // width for the next iteration
  }
}

// First node of an iteration: the width summed since the last roll (the
// columns active after the previous iteration and those admitted since)
// becomes this iteration's active width.
__global__ void k_kw_roll(int32_t *__restrict__ na) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    na[2] = na[1];
    na[1] = 0;
  }
}

int pow2_at_least(int k) {
  int K = 1;
  while (K < k) K <<= 1;
  return K;
}

int vec_blocks(int32_t n, int k) {
  int64_t R = kThreads / pow2_at_least(k);
  int64_t g = (n + R - 1) / R;
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return (int)g;
}

template <int OP>
int vec_op(int32_t n, int k, const Vecs &v, const Col *col, double *part,
           hipStream_t s) {
  const int K = pow2_at_least(k);
  // passes with dots run on the full grid, the plain updates on half of it
  // (measured: more waves help the former and slow the latter)
  const int g = NPairs<OP>::v > 0 ? vec_blocks(n, k)
                                  : std::max(1, std::min(vec_blocks(n, k),
                                                         kRedBlocks / 2));
#define AINV_VL(KK)                                                        \
  hipLaunchKernelGGL((k_vec<OP, KK>), dim3(g), dim3(kThreads), 0, s, n, k, \
                     v, col, part)
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
  return g;
}

void fin(int which, int k, int nblk, const double *part, Col *col,
         int32_t *nact, hipStream_t s, const int32_t *mask = nullptr) {
  hipLaunchKernelGGL(k_fin, dim3(k), dim3(kThreads), 0, s, which, nblk, part,
                     col, nact, mask);
  AINV_HIP_CHECK(hipGetLastError());
}

// Slot admission.  Entry j of the list fills slot slot[j] with the
// right-hand side src[j][row * stride[j]], x0 = 0, so r0 = b and
// (BiCGSTAB) rtilde = r0: one thread per row writes every admitted slot of
// its row, so each row's cache lines are written once per admission round.
struct AdmitList {
  int cnt;
  int slot[kMaxCols];
  int stride[kMaxCols];
  const double *src[kMaxCols];
};

// One thread per (row, admitted column): the lanes of a wave write the
// entries of a few rows, each row's admitted columns by neighbouring lanes.
// lazy (BiCGSTAB): r~ = b only (x, r and b are taken from the right-hand
// side until the first update).
__global__ void k_admit(int32_t n, int K, AdmitList L, int lazy,
                        double *__restrict__ b, double *__restrict__ x,
                        double *__restrict__ r, double *__restrict__ q,
                        float *__restrict__ r32) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * L.cnt) return;
  const int64_t row = t / L.cnt;
  const int j = (int)(t - row * L.cnt);
  const double v = L.src[j][row * L.stride[j]];
  const int64_t e = row * K + L.slot[j];
  if (lazy) {
    q[e] = v;
    return;
  }
  b[e] = v;
  x[e] = 0.0;
  r[e] = v;
  if (q) q[e] = v;
  if (r32) r32[e] = (float)v;
}

// Column states and admission mask of an admission, written on the device
// from the list passed by value (kMaxCols threads), so that admitting needs
// no host staging and no wait for the step in flight.
__global__ void k_admit_state(AdmitList L, double eps, int lazy,
                              Col *__restrict__ col,
                              int32_t *__restrict__ mask) {
  const int t = threadIdx.x;
  int adm = -1;
  for (int j = 0; j < L.cnt; ++j)
    if (L.slot[j] == t) adm = j;
  mask[t] = adm >= 0;
  if (adm >= 0) {
    Col c{};
    c.active = 1;
    c.eps = eps;
    c.lazy = lazy;
    c.src = L.src[adm];
    c.sstride = L.stride[adm];
    col[t] = c;
  }
}

// With x0 = 0 every dot of the init step of an admitted column is ||b||^2:
// block partial sums of it for entry blockIdx.y, written as the np pairs
// the init step reads (fixed-order tree, as the other reductions).
__global__ __launch_bounds__(kThreads) void k_admit_norms(
    int32_t n, AdmitList L, int np, double *__restrict__ part) {
  const int j = blockIdx.y;
  const double *src = L.src[j];
  const int64_t st = L.stride[j];
  double acc = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x; row < n;
       row += (int64_t)gridDim.x * kThreads) {
    const double v = src[row * st];
    acc += v * v;
  }
  __shared__ double red[kThreads];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < np; ++q)
      part[((int64_t)q * kRedBlocks + blockIdx.x) * kMaxCols + L.slot[j]] =
          red[0];
}

// Columns slot[j] of a K-column vector to the contiguous n-vectors dst[j],
// one thread per (row, column) so each wave reads a few rows once.
struct ExtractList {
  int cnt;
  int slot[kMaxCols];
  double *dst[kMaxCols];
};
// (a column that never reached its first update holds x0 = 0)
__global__ void k_extract_cols(int32_t n, int K, ExtractList L,
                               const double *__restrict__ x,
                               const Col *__restrict__ col) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * L.cnt) return;
  const int64_t row = t / L.cnt;
  const int j = (int)(t - row * L.cnt);
  L.dst[j][row] = col[L.slot[j]].lazy ? 0.0 : x[row * K + L.slot[j]];
}

// Copy columns [c0, c0+kc) of a K-column vector to dst (n x kc).
__global__ void k_extract(int32_t n, int K, int c0, int kc,
                          const double *__restrict__ x,
                          double *__restrict__ dst,
                          const Col *__restrict__ col) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * kc) return;
  const int64_t row = t / kc;
  const int j = (int)(t % kc);
  dst[t] = col[c0 + j].lazy ? 0.0 : x[row * K + c0 + j];
}

}  // namespace

struct KrylovPlan {
  int method = KM_BICGSTAB;
  bool cg = false;  // CG's vector set and init; otherwise BiCGSTAB's
  const ainv_solver *m = nullptr;
  uint64_t gen = 0;
  int k = 0;       // columns in use
  int K = 1;       // leading dimension of the workspace, pow2 >= k
  int apply_map = 0;  // packed pair of the two passes' mappings
  int spmv_map = AINV_MAP_SPLIT;
  int32_t n = 0;
  int holders = 0;  // open slot batches using this plan (never evicted)
  DevBuf work, colbuf, part, nact, tbuf;
  Vecs v{};
  Col *col = nullptr;
  double *pt = nullptr;
  int32_t *na = nullptr;  // [0] active count, [1] width being summed,
                          // [2] active width of this iteration
  const int32_t *kwu = nullptr;  // na + 2 for the one-graph methods
  int32_t *pinned = nullptr;  // host copies of nactive, two slots
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // slot batches: per-column states copied back each step, admission mask
  Col *pinned_cols = nullptr;  // two slots of kMaxCols states
  DevBuf mask;
  // RGMRES: restart length, Arnoldi basis (irst + 1 vectors of n x K),
  // per-column states and small dense arrays, one graph per Arnoldi step
  // of a cycle plus one for the cycle's end
  int irst = 0;
  DevBuf basis, gmcol, gmarr;
  DevBuf work32;  // BiCGSTAB: FP32 p and s for FP32 / BF16 factors
  std::vector<hipGraphExec_t> gm_exec;

  void release_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (auto &e : gm_exec)
      if (e) (void)hipGraphExecDestroy(e);
    gm_exec.clear();
  }
  ~KrylovPlan() {
    release_graph();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (pinned) (void)hipHostFree(pinned);
    if (pinned_cols) (void)hipHostFree(pinned_cols);
  }
};

void destroy_plan(KrylovPlan *p) { delete p; }

namespace {

// FP32 vectors of a plan start 256 bytes apart at least (aligned loads)
int64_t x32_stride(const KrylovPlan *p) {
  return ((int64_t)p->n * p->K + 63) / 64 * 64;
}

void alloc_x32(KrylovPlan *p, const ainv_solver *m, hipStream_t s) {
  if ((p->method != KM_BICGSTAB && p->method != KM_CG) || !m ||
      m->prec == AINV_PREC_F64 || p->work32.p)
    return;
  const int nv = p->method == KM_BICGSTAB ? 4 : 2;
  p->work32.alloc((size_t)nv * x32_stride(p) * sizeof(float));
  // padding columns are read by the row-mapped kernels: keep them 0
  AINV_HIP_CHECK(hipMemsetAsync(p->work32.p, 0, p->work32.bytes, s));
}

// Preconditioner workspace of a plan; the multilevel cycle's is cleared
// (its vectors' padding columns are read by the row-mapped kernels).
void alloc_tbuf(KrylovPlan *p, const ainv_solver *m, hipStream_t s) {
  const size_t need = solver_t_bytes(m, p->K);
  if (p->tbuf.bytes >= need) return;
  p->tbuf.alloc(need);
  if (m->kind == AINV_KIND_ML)
    AINV_HIP_CHECK(hipMemsetAsync(p->tbuf.p, 0, need, s));
}

// The plan's vectors with the FP32 copies the method writes when its
// preconditioner has FP32 / BF16 factors (none otherwise).
Vecs x32_vecs(const KrylovPlan *p) {
  Vecs w = p->v;
  if (p->m && p->m->prec != AINV_PREC_F64 && p->work32.p) {
    float *f = p->work32.as<float>();
    if (p->method == KM_BICGSTAB) {
      const int64_t nk = x32_stride(p);
      w.p32 = f;
      w.s32 = f + nk;
      w.f32 = f + 2 * nk;
      w.z32 = f + 3 * nk;
    } else if (p->method == KM_CG) {
      w.r32 = f;
      w.z32 = f + x32_stride(p);
    }
  }
  return w;
}

// The buffers the host reads after an event (pinned) must be host
// allocations and the ones kernels write (col, na) device allocations;
// checked once per plan, when they are made.
void check_buffer(const void *ptr, bool host, const char *what) {
  if (!ptr) return;
  hipPointerAttribute_t at{};
  AINV_HIP_CHECK(hipPointerGetAttributes(&at, ptr));
  const bool is_host = at.type == hipMemoryTypeHost;
  if (is_host != host)
    throw Error(AINV_ERR_INTERNAL, std::string("Krylov plan: ") + what +
                                       (host ? " is not pinned host memory"
                                             : " is not device memory"));
}

void check_host_device(const KrylovPlan *p) {
  check_buffer(p->pinned, true, "pinned count");
  check_buffer(p->pinned_cols, true, "pinned column states");
  check_buffer(p->col, false, "column states");
  check_buffer(p->na, false, "active counts");
}

KrylovPlan *get_plan(ainv_matrix *a, int method, const ainv_solver *m, int k,
                     hipStream_t s, int irst = 0) {
  for (size_t i = 0; i < a->plans.size(); ++i) {
    KrylovPlan *p = a->plans[i].get();
    if (p->method == method && p->m == m && p->k == k && p->irst == irst) {
      if (m && p->gen != m->generation) {
        p->release_graph();  // factors changed: recapture
        alloc_tbuf(p, m, s);
        alloc_x32(p, m, s);
        p->gen = m->generation;
      }
      if (i != 0) std::rotate(a->plans.begin(), a->plans.begin() + i,
                              a->plans.begin() + i + 1);
      return a->plans.front().get();
    }
  }
  std::unique_ptr<KrylovPlan, PlanDeleter> up(new KrylovPlan());
  KrylovPlan *p = up.get();
  p->method = method;
  p->irst = irst;
  p->cg = method == KM_CG;
  const bool cg = p->cg;
  p->m = m;
  p->gen = m ? m->generation : 0;
  p->k = k;
  p->K = 1;
  while (p->K < k) p->K <<= 1;
  p->n = a->n;
  const int64_t nk = (int64_t)a->n * p->K;
  // x, b + recurrence vectors (RGMRES: x, b, w, z and the basis apart)
  const int nvec = (method == KM_GMRES || method == KM_BICGSTABL || cg) ? 6 : 10;
  p->work.alloc((size_t)nk * nvec * sizeof(double));
  // padding columns (k..K) are read by the row-mapped kernels: keep them 0
  AINV_HIP_CHECK(hipMemsetAsync(p->work.p, 0, p->work.bytes, s));
  double *w0 = p->work.as<double>();
  p->v.x = w0;
  p->v.b = w0 + nk;
  p->v.r = w0 + 2 * nk;
  p->v.p = w0 + 3 * nk;
  p->v.z = w0 + 4 * nk;
  p->v.q = w0 + 5 * nk;  // CG: A p ; BiCGSTAB: rtilde
  if (!cg && method != KM_GMRES && method != KM_BICGSTABL) {
    p->v.v = w0 + 6 * nk;
    p->v.f = w0 + 7 * nk;
    p->v.s = w0 + 8 * nk;
    p->v.t = w0 + 9 * nk;
  }
  p->colbuf.alloc(sizeof(Col) * kMaxCols);
  p->part.alloc(sizeof(double) * kMaxPairs * kRedBlocks * kMaxCols);
  p->nact.alloc(3 * sizeof(int32_t));
  AINV_HIP_CHECK(hipMemsetAsync(p->nact.p, 0, 3 * sizeof(int32_t), s));
  if (m) alloc_tbuf(p, m, s);
  alloc_x32(p, m, s);
  p->col = p->colbuf.as<Col>();
  p->pt = p->part.as<double>();
  p->na = p->nact.as<int32_t>();
  // the one-graph methods skip the columns beyond the active width
  if (method <= KM_BICG) p->kwu = p->na + 2;
  AINV_HIP_CHECK(hipHostMalloc((void **)&p->pinned, 2 * sizeof(int32_t),
                               hipHostMallocDefault));
  for (auto &e : p->ev)
    AINV_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  check_host_device(p);
  a->plans.insert(a->plans.begin(), std::move(up));
  // evict the least recently used plan that no open slot batch holds (a
  // batch keeps launching its plan's graph and copying its column states
  // until it is closed)
  if ((int)a->plans.size() > kPlansPerMatrix) {
    for (size_t i = a->plans.size(); i-- > 1;)
      if (a->plans[i]->holders == 0) {
        a->plans.erase(a->plans.begin() + i);
        break;
      }
  }
  return p;
}

void precond(const KrylovPlan *p, const double *x, double *y, hipStream_t s) {
  if (p->m && p->m->kind == AINV_KIND_ML)
    ml_apply_into(p->m, x, y, p->k, p->K, p->tbuf.p, s, p->kwu);
  else if (p->m)
    solver_apply_into(p->m, x, y, p->k, p->K, p->tbuf.p, s, p->apply_map,
                      p->kwu);
  else
    AINV_HIP_CHECK(hipMemcpyAsync(y, x, (size_t)p->n * p->K * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
}

// y = M x, FP32 x and FP32 y (FP32 / BF16 factors)
void precond32(const KrylovPlan *p, const float *x, float *y, hipStream_t s) {
  const ainv_solver *m = p->m;
  if (m->kind == AINV_KIND_ML) {
    ml_apply_into32(m, x, y, p->k, p->K, p->tbuf.p, s, p->kwu);
    return;
  }
  hll_spmm_scale(m->hw, m->dv.p, x, true, p->K, p->tbuf.p, p->K, p->k, s,
                 map_w(p->apply_map), p->kwu);
  hll_spmm_axpby(m->hz, 1.0, p->tbuf.p, true, p->K, 0.0, y, p->K, p->k, s,
                 map_z(p->apply_map), p->kwu);
}

void precond_t(const KrylovPlan *p, const double *x, double *y,
               hipStream_t s) {
  if (p->m)
    solver_apply_into_t(p->m, x, y, p->k, p->K, p->tbuf.p, s, p->apply_map,
                        p->kwu);
  else
    AINV_HIP_CHECK(hipMemcpyAsync(y, x, (size_t)p->n * p->K * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
}

}  // namespace

void matrix_drop_plans(ainv_matrix *a) {
  for (const auto &p : a->plans)
    if (p->holders > 0)
      throw Error(AINV_ERR_STATE, "matrix has Krylov work in flight");
  a->plans.clear();
  std::fill(std::begin(a->spmv_map), std::end(a->spmv_map), (int8_t)-1);
  std::memset(a->spmv_ms, 0, sizeof a->spmv_ms);
}

namespace {

// Layouts a method needs besides the plan's own (built outside any capture):
// BiCG applies M^T and A^T.
void prepare_method(KrylovPlan *p, ainv_matrix *a, ainv_solver *m) {
  if (p->method != KM_BICG) return;
  if (m) solver_ensure_transposed(m);
  if (!a->have_t) {
    DevCsr at;
    csr_transpose(a->csr, at, a->stream);
    hll_from_csr(at, AINV_PREC_F64, a->hll_t, a->stream, -1, a->slice_sort);
    AINV_HIP_CHECK(hipStreamSynchronize(a->stream));
    a->have_t = true;
  }
}

// Autotuner: time the kernel mappings (split, row, coop, coop128, their
// quarter-slice variants and run) of each apply pass and of the fused
// system-matrix product (split, row, coop, coop128, run) on this plan's buffers
// and keep the fastest of each; results are cached on the solver / matrix
// per K.
double time_ms(hipStream_t s, const std::function<void()> &f, int reps) {
  hipEvent_t e0, e1;
  AINV_HIP_CHECK(hipEventCreate(&e0));
  AINV_HIP_CHECK(hipEventCreate(&e1));
  f();  // warm-up
  AINV_HIP_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) f();
  AINV_HIP_CHECK(hipEventRecord(e1, s));
  AINV_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  AINV_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

// Timed trial of the run mapping's slice-group order on one factor layout
// (f runs its pass): linear order against strips of P / 20, P / 10 and
// P / 5 groups over planes of P = R / 2 and R groups, R the factor's
// median reach (rows from a row to its farthest column, in groups of 4
// slices: the stencil factors' far cluster of diagonals lies two planes
// back).  A strip order is kept when it beats linear order by 2 % (the
// trial's own spread); results are the same bits in every order.
// Returns the trial as JSON ("name": {...}).
std::string tune_order(Hll &h, const std::function<void()> &f,
                       const char *name, hipStream_t s) {
  const int32_t R = hll_reach_groups(h, s);
  std::vector<std::pair<int32_t, int32_t>> cand{{0, 0}};
  for (const int32_t P : {R / 2, R}) {
    if (P < 4) continue;
    for (const int32_t div : {20, 10, 5}) {
      const std::pair<int32_t, int32_t> c{P, std::max<int32_t>(1, P / div)};
      if (std::find(cand.begin(), cand.end(), c) == cand.end())
        cand.push_back(c);
    }
  }
  std::string js = std::string("\"") + name + "\": {\"reach_groups\": " +
                   std::to_string(R) + ", \"ms\": [";
  double lin = 0.0, best = 0.0;
  std::pair<int32_t, int32_t> pick{0, 0};
  for (size_t i = 0; i < cand.size(); ++i) {
    hll_set_order(h, cand[i].first, cand[i].second, s);
    const double t = time_ms(s, f, 3);
    char b[96];
    std::snprintf(b, sizeof b, "%s[%d, %d, %.6f]", i ? ", " : "",
                  cand[i].first, cand[i].second, t);
    js += b;
    if (i == 0) {
      lin = best = t;
    } else if (t < best && t < lin * 0.98) {
      best = t;
      pick = cand[i];
    }
  }
  hll_set_order(h, pick.first, pick.second, s);
  return js + "]}";
}

void tune_plan(KrylovPlan *p, ainv_matrix *a, ainv_solver *m, hipStream_t s) {
  const int K = p->K;
  const int li = log2_pow2(K);
  // a mapping forced on the preconditioner (APPLY_MAP) also holds for the
  // system-matrix product of the solves it is used in
  const int forced = m ? m->map_mode : -1;
  if (K == 1 || p->k < 2) {
    p->apply_map = map_pair(AINV_MAP_SPLIT, AINV_MAP_SPLIT);
    p->spmv_map = AINV_MAP_SPLIT;
    // one column: the multilevel products may still run one wavefront per
    // long row
    if (K == 1 && m && m->kind == AINV_KIND_ML) {
      const Vecs w = x32_vecs(p);
      if (w.p32)
        ml_tune(m, 1, p->k, w.p32, w.f32, true, p->tbuf.p, s);
      else if (w.r32)
        ml_tune(m, 1, p->k, w.r32, w.z32, true, p->tbuf.p, s);
      else
        ml_tune(m, 1, p->k, p->v.b, p->v.z, false, p->tbuf.p, s);
    }
    return;
  }
  if (forced >= 0) {
    p->apply_map = map_pair(forced, forced);
    p->spmv_map = forced;
    // a forced run mapping still has its slice-group order tried
    if (forced != AINV_MAP_RUN || !m || m->kind == AINV_KIND_ML) return;
  }
  const Vecs &v = p->v;
  if (m && m->kind == AINV_KIND_ML) {
    // tuned on the vectors the cycle will run on
    const Vecs w = x32_vecs(p);
    if (w.p32)
      ml_tune(m, K, p->k, w.p32, w.f32, true, p->tbuf.p, s);
    else if (w.r32)
      ml_tune(m, K, p->k, w.r32, w.z32, true, p->tbuf.p, s);
    else
      ml_tune(m, K, p->k, v.b, v.z, false, p->tbuf.p, s);
  } else if (m) {
    // the two passes are timed separately, on the vectors the iteration
    // hands the preconditioner: with FP32 / BF16 factors BiCGSTAB and CG
    // apply it to FP32 copies (p / s, r) into FP32 outputs (f / z); t keeps
    // leading dimension K
    const bool f32t = m->prec != AINV_PREC_F64;
    const Vecs w32 = x32_vecs(p);
    const float *x32 = w32.p32 ? w32.p32 : w32.r32;
    float *y32 = w32.p32 ? w32.f32 : w32.z32;
    auto pass_w = [&](int map) {
      if (x32)
        hll_spmm_scale(m->hw, m->dv.p, x32, true, K, p->tbuf.p, K, p->k, s,
                       map);
      else
        hll_spmm_scale(m->hw, m->dv.p, v.b, false, K, p->tbuf.p, K, p->k, s,
                       map);
    };
    auto pass_z = [&](int map) {
      if (x32)
        hll_spmm_axpby(m->hz, 1.0, p->tbuf.p, true, K, 0.0, y32, K, p->k, s,
                       map);
      else
        hll_spmm_axpby(m->hz, 1.0, p->tbuf.p, f32t, K, 0.0, v.z, K, p->k, s,
                       map);
    };
    if (forced < 0 && m->apply_map[li] < 0) {
      int bw = 0, bz = 0;
      for (int map = 0; map < kNumApplyMaps; ++map) {
        m->apply_ms[li][0][map] = time_ms(s, [&] { pass_w(map); }, 3);
        m->apply_ms[li][1][map] = time_ms(s, [&] { pass_z(map); }, 3);
        if (m->apply_ms[li][0][map] < m->apply_ms[li][0][bw]) bw = map;
        if (m->apply_ms[li][1][map] < m->apply_ms[li][1][bz]) bz = map;
      }
      m->apply_map[li] = (int16_t)map_pair(bw, bz);
    }
    if (forced < 0) p->apply_map = m->apply_map[li];
    // the run mapping's slice-group order, once per layout, on the first
    // plan of 32 or more columns that runs it
    const bool run_w = map_w(p->apply_map) == AINV_MAP_RUN;
    const bool run_z = map_z(p->apply_map) == AINV_MAP_RUN;
    if (K >= 32 && m->order_period == 0 && !m->order_tuned &&
        (run_w || run_z)) {
      std::string js;
      if (run_w)
        js += tune_order(m->hw, [&] { pass_w(AINV_MAP_RUN); }, "w", s);
      if (run_z)
        js += std::string(js.empty() ? "" : ", ") +
              tune_order(m->hz, [&] { pass_z(AINV_MAP_RUN); }, "z", s);
      m->order_report = "{\"K\": " + std::to_string(K) + ", " + js + "}";
      m->order_tuned = true;
    }
  }
  if (forced >= 0) return;
  if (a->spmv_map[li] < 0) {
    // BiCGSTAB with FP32 / BF16 factors multiplies the FP32 preconditioned
    // vectors (f, z): time that form
    const float *f32 = p->method == KM_BICGSTAB ? x32_vecs(p).f32 : nullptr;
    int best = kSpmvMaps[0];
    for (const int map : kSpmvMaps) {
      a->spmv_ms[li][map] = time_ms(s, [&] {
        if (f32)
          hll_spmv_dot(a->hll, f32, v.z, p->k, K, v.r, 1, p->pt, map, s);
        else
          hll_spmv_dot(a->hll, v.b, v.z, p->k, K, v.r, 1, p->pt, map, s);
      }, 3);
      if (a->spmv_ms[li][map] < a->spmv_ms[li][best]) best = map;
    }
    a->spmv_map[li] = (int8_t)best;
  }
  p->spmv_map = a->spmv_map[li];
}

void enqueue_iteration(const KrylovPlan *p, ainv_matrix *a, hipStream_t s) {
  const int32_t n = p->n;
  const int k = p->k, K = p->K, sm = p->spmv_map;
  const Vecs &v = p->v;
  const int32_t *kw = p->kwu;
  hipLaunchKernelGGL(k_kw_roll, dim3(1), dim3(64), 0, s, p->na);
  AINV_HIP_CHECK(hipGetLastError());
  if (p->method == KM_BICG) {
    // vectors: q = rtilde, f = ztilde, s = ptilde, v = A p, t = A^T ptilde
    precond(p, v.r, v.z, s);
    precond_t(p, v.q, v.f, s);
    int g = vec_op<OP_DOT_ZRT>(n, k, v, p->col, p->pt, s);
    fin(FIN_RHO_BICGM, k, g, p->pt, p->col, p->na, s);
    vec_op<OP_P_BICGM>(n, k, v, p->col, p->pt, s);
    g = hll_spmv_dot(a->hll, v.p, v.v, k, K, v.s, 1, p->pt, sm, s, kw);
    hll_spmm_axpby(a->hll_t, 1.0, v.s, false, K, 0.0, v.t, K, k, s, sm, kw);
    fin(FIN_ALPHA_BICG, k, g, p->pt, p->col, p->na, s);
    g = vec_op<OP_XR_BICGM>(n, k, v, p->col, p->pt, s);
    fin(FIN_CONV_CG, k, g, p->pt, p->col, p->na, s);
  } else if (p->method == KM_CGS) {
    // vectors: q = rtilde, s = u, t = q of CGS, f = phat then uhat,
    // v = vhat then qhat, z = u + q
    fin(FIN_RHO_CGS, k, 0, p->pt, p->col, p->na, s);
    vec_op<OP_U_CGS>(n, k, v, p->col, p->pt, s);
    precond(p, v.p, v.f, s);
    int g = hll_spmv_dot(a->hll, v.f, v.v, k, K, v.q, 1, p->pt, sm, s, kw);
    fin(FIN_ALPHA_BICG, k, g, p->pt, p->col, p->na, s);
    vec_op<OP_Q_CGS>(n, k, v, p->col, p->pt, s);
    precond(p, v.z, v.f, s);
    hll_spmm_axpby(a->hll, 1.0, v.f, false, K, 0.0, v.v, K, k, s, sm, kw);
    g = vec_op<OP_XR_CGS>(n, k, v, p->col, p->pt, s);
    fin(FIN_CONV_BICG, k, g, p->pt, p->col, p->na, s);
  } else if (!p->cg) {
    // with FP32 / BF16 factors the preconditioner reads FP32 copies of p
    // and s (half the gathered bytes of its first pass)
    const Vecs w = x32_vecs(p);
    const bool x32 = w.p32 != nullptr;
    fin(FIN_RHO_BICG, k, 0, p->pt, p->col, p->na, s);
    vec_op<OP_P_BICG>(n, k, w, p->col, p->pt, s);
    int g;
    if (x32) {
      precond32(p, w.p32, w.f32, s);
      g = hll_spmv_dot(a->hll, w.f32, v.v, k, K, v.q, 1, p->pt, sm, s, kw);
    } else {
      precond(p, v.p, v.f, s);
      g = hll_spmv_dot(a->hll, v.f, v.v, k, K, v.q, 1, p->pt, sm, s, kw);
    }
    fin(FIN_ALPHA_BICG, k, g, p->pt, p->col, p->na, s);
    vec_op<OP_S_BICG>(n, k, w, p->col, p->pt, s);
    if (x32) {
      precond32(p, w.s32, w.z32, s);
      g = hll_spmv_dot(a->hll, w.z32, v.t, k, K, v.s, 2, p->pt, sm, s, kw);
    } else {
      precond(p, v.s, v.z, s);
      g = hll_spmv_dot(a->hll, v.z, v.t, k, K, v.s, 2, p->pt, sm, s, kw);
    }
    fin(FIN_OMEGA, k, g, p->pt, p->col, p->na, s);
    g = vec_op<OP_XR_BICG>(n, k, w, p->col, p->pt, s);
    fin(FIN_CONV_BICG, k, g, p->pt, p->col, p->na, s);
  } else {
    const Vecs w = x32_vecs(p);
    if (w.r32)
      precond32(p, w.r32, w.z32, s);
    else
      precond(p, v.r, v.z, s);
    int g = vec_op<OP_DOT_RZ>(n, k, w, p->col, p->pt, s);
    fin(FIN_RHO_CG, k, g, p->pt, p->col, p->na, s);
    vec_op<OP_P_CG>(n, k, w, p->col, p->pt, s);
    g = hll_spmv_dot(a->hll, v.p, v.q, k, K, v.p, 1, p->pt, sm, s, kw);
    fin(FIN_ALPHA_CG, k, g, p->pt, p->col, p->na, s);
    g = vec_op<OP_XR_CG>(n, k, w, p->col, p->pt, s);
    fin(FIN_CONV_CG, k, g, p->pt, p->col, p->na, s);
  }
}

// user vectors (ld k) <-> workspace columns (ld K); padding columns zero
__global__ void k_pack(int32_t n, int k, int K, const double *__restrict__ src,
                       double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * K) return;
  const int64_t row = t / K;
  const int c = (int)(t % K);
  dst[t] = c < k ? src[row * k + c] : 0.0;
}

__global__ void k_unpack(int32_t n, int k, int K, const double *__restrict__ src,
                         double *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t row = t / k;
  const int c = (int)(t % k);
  dst[t] = src[row * K + c];
}

void capture(KrylovPlan *p, ainv_matrix *a, hipStream_t s) {
  if (p->exec) return;
  tune_plan(p, a, const_cast<ainv_solver *>(p->m), s);
  AINV_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_iteration(p, a, s);
  } catch (...) {
    hipGraph_t g0 = nullptr;
    (void)hipStreamEndCapture(s, &g0);
    if (g0) (void)hipGraphDestroy(g0);
    throw;
  }
  AINV_HIP_CHECK(hipStreamEndCapture(s, &p->graph));
  AINV_HIP_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
}

}  // namespace

int krylov_method_id(const std::string &name) {
  if (name == "BICGSTAB") return KM_BICGSTAB;
  if (name == "CG") return KM_CG;
  if (name == "CGS") return KM_CGS;
  if (name == "BICG") return KM_BICG;
  if (name == "RGMRES" || name == "GMRES") return KM_GMRES;
  if (name == "BICGSTABL") return KM_BICGSTABL;
  throw Error(AINV_ERR_ARG, "unknown Krylov method " + name +
                                " (BICGSTAB, CG, CGS, BICG, RGMRES, BICGSTABL "
                                "available)");
}

// ---------------------------------------------------------- slot batches
struct SlotBatch {
  ainv_matrix *a = nullptr;
  KrylovPlan *p = nullptr;
  hipStream_t s = nullptr;
  int K = 0;
  int64_t steps = 0;
};

SlotBatch *batch_open(ainv_matrix *a, ainv_solver *m, int method, int K,
                      hipStream_t s) {
  if (K < 1 || K > kMaxCols) throw Error(AINV_ERR_ARG, "slots must be 1..64");
  if (method == KM_GMRES || method == KM_BICGSTABL)
    throw Error(AINV_ERR_ARG,
                "the served form runs BICGSTAB, CG, CGS and BICG (RGMRES and "
                "BICGSTABL run in cycles of several steps)");
  if (m && m->n != a->n)
    throw Error(AINV_ERR_ARG, "preconditioner and matrix sizes differ");
  KrylovPlan *p = get_plan(a, method, m, K, s);
  prepare_method(p, a, m);
  if (!p->pinned_cols) {
    AINV_HIP_CHECK(hipHostMalloc((void **)&p->pinned_cols,
                                 2 * kMaxCols * sizeof(Col),
                                 hipHostMallocDefault));
    check_buffer(p->pinned_cols, true, "pinned column states");
  }
  if (!p->mask.p) p->mask.alloc(kMaxCols * sizeof(int32_t));
  // every slot empty: inactive columns are skipped by every kernel
  std::vector<Col> init(kMaxCols);
  for (auto &c0 : init) c0 = Col{};
  AINV_HIP_CHECK(hipMemcpyAsync(p->col, init.data(), sizeof(Col) * kMaxCols,
                                hipMemcpyHostToDevice, s));
  AINV_HIP_CHECK(hipMemsetAsync(p->na, 0, 3 * sizeof(int32_t), s));
  AINV_HIP_CHECK(hipMemsetAsync(p->work.p, 0, p->work.bytes, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  capture(p, a, s);
  auto *b = new SlotBatch();
  b->a = a;
  b->p = p;
  ++p->holders;
  b->s = s;
  b->K = p->K;  // slot count and leading dimension (pow2 >= K)
  return b;
}

void batch_close(SlotBatch *b) {
  if (b && b->s) (void)hipStreamSynchronize(b->s);
  if (b && b->p) --b->p->holders;
  delete b;
}

namespace {
// Writes the admitted columns, their column states and admission mask, and
// runs the init step (bn, err, active; BiCGSTAB rho) for them only.
void admit(SlotBatch *b, const AdmitList &L, double eps) {
  KrylovPlan *p = b->p;
  const int32_t n = p->n;
  if (L.cnt == 0) return;
  // BiCGSTAB reads r = b and x = 0 from the right-hand side until its
  // first update; the other methods get their columns written here
  const int lazy = p->method == KM_BICGSTAB && !p->cg;
  hipLaunchKernelGGL(k_admit,
                     dim3((unsigned)(((int64_t)n * L.cnt + 255) / 256)),
                     dim3(256), 0,
                     b->s, n, b->K, L, lazy, p->v.b, p->v.x, p->v.r,
                     p->cg ? nullptr : p->v.q, x32_vecs(p).r32);
  AINV_HIP_CHECK(hipGetLastError());
  const int g = vec_blocks(n, 1);
  const int fin_id = p->cg ? FIN_INIT_CG : FIN_INIT;
  hipLaunchKernelGGL(k_admit_norms, dim3(g, L.cnt), dim3(kThreads), 0, b->s,
                     n, L, fin_pairs(fin_id), p->pt);
  AINV_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_admit_state, dim3(1), dim3(kMaxCols), 0, b->s, L, eps,
                     lazy, p->col, p->mask.as<int32_t>());
  AINV_HIP_CHECK(hipGetLastError());
  fin(fin_id, b->K, g, p->pt, p->col, p->na, b->s, p->mask.as<int32_t>());
}
}  // namespace

void batch_admit(SlotBatch *b, int c0, int kc, const double *src, double eps) {
  if (c0 < 0 || kc < 1 || c0 + kc > b->K)
    throw Error(AINV_ERR_ARG, "admission outside the slot range");
  AdmitList L{};
  L.cnt = kc;
  for (int j = 0; j < kc; ++j) {
    L.slot[j] = c0 + j;
    L.stride[j] = kc;
    L.src[j] = src + j;
  }
  admit(b, L, eps);
}

void batch_admit_cols(SlotBatch *b, const std::vector<int> &slots,
                      const std::vector<const double *> &srcs, double eps) {
  if (slots.size() != srcs.size())
    throw Error(AINV_ERR_INTERNAL, "admission lists differ in length");
  if ((int)slots.size() > b->K)
    throw Error(AINV_ERR_ARG, "more admissions than slots");
  AdmitList L{};
  L.cnt = (int)slots.size();
  for (int j = 0; j < L.cnt; ++j) {
    if (slots[j] < 0 || slots[j] >= b->K)
      throw Error(AINV_ERR_ARG, "admission outside the slot range");
    L.slot[j] = slots[j];
    L.stride[j] = 1;
    L.src[j] = srcs[j];
  }
  admit(b, L, eps);
}

void batch_step(SlotBatch *b) {
  KrylovPlan *p = b->p;
  AINV_HIP_CHECK(hipGraphLaunch(p->exec, b->s));
  const int slot = (int)(b->steps & 1);
  AINV_HIP_CHECK(hipMemcpyAsync(p->pinned_cols + slot * kMaxCols, p->col,
                                sizeof(Col) * b->K, hipMemcpyDeviceToHost,
                                b->s));
  AINV_HIP_CHECK(hipEventRecord(p->ev[slot], b->s));
  ++b->steps;
}

// States after the step before the most recent one (no wait on the most
// recent, which may still run); with wait_last, after the most recent.
int64_t batch_next_step(const SlotBatch *b) { return b->steps; }

int64_t batch_poll(SlotBatch *b, bool wait_last, std::vector<SlotState> &out) {
  KrylovPlan *p = b->p;
  out.assign(b->K, SlotState{});
  int64_t which = wait_last ? b->steps - 1 : b->steps - 2;
  if (which < 0) return -1;
  const int slot = (int)(which & 1);
  AINV_HIP_CHECK(hipEventSynchronize(p->ev[slot]));
  const Col *c = p->pinned_cols + slot * kMaxCols;
  for (int j = 0; j < b->K; ++j) {
    out[j].active = c[j].active;
    out[j].iter = c[j].iter;
    out[j].err = c[j].err;
    out[j].valid = true;
  }
  return which;
}

void batch_extract_cols(SlotBatch *b, const std::vector<int> &slots,
                        const std::vector<double *> &dsts) {
  if (slots.size() != dsts.size() || slots.size() > (size_t)kMaxCols)
    throw Error(AINV_ERR_INTERNAL, "bad extraction list");
  if (slots.empty()) return;
  ExtractList L{};
  L.cnt = (int)slots.size();
  for (int j = 0; j < L.cnt; ++j) {
    if (slots[j] < 0 || slots[j] >= b->K)
      throw Error(AINV_ERR_INTERNAL, "extraction outside the slot range");
    L.slot[j] = slots[j];
    L.dst[j] = dsts[j];
  }
  const int64_t tot = (int64_t)b->p->n * L.cnt;
  hipLaunchKernelGGL(k_extract_cols, dim3((unsigned)((tot + 255) / 256)),
                     dim3(256), 0, b->s, b->p->n, b->K, L, b->p->v.x,
                     b->p->col);
  AINV_HIP_CHECK(hipGetLastError());
}

void batch_extract(SlotBatch *b, int c0, int kc, double *dst) {
  const int64_t tot = (int64_t)b->p->n * kc;
  hipLaunchKernelGGL(k_extract, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                     0, b->s, b->p->n, b->K, c0, kc, b->p->v.x, dst,
                     b->p->col);
  AINV_HIP_CHECK(hipGetLastError());
}

void gmres_solve(ainv_matrix *a, ainv_solver *m, const double *b, double *x,
                 int k, double eps, int itmax, int irst, int32_t *iter,
                 double *err, hipStream_t s, int flags);
void bicgstabl_solve(ainv_matrix *a, ainv_solver *m, const double *b,
                     double *x, int k, double eps, int itmax, int l,
                     int32_t *iter, double *err, hipStream_t s, int flags);

void krylov_solve(const std::string &method, ainv_matrix *a, ainv_solver *m,
                  const double *b, double *x, int k, double eps, int itmax,
                  int32_t *iter, double *err, hipStream_t s, int flags,
                  int irst) {
  const int32_t n = a->n;
  if (k < 1 || k > kMaxCols) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (m && m->n != n)
    throw Error(AINV_ERR_ARG, "preconditioner and matrix sizes differ");
  const int mid = krylov_method_id(method);
  if (mid == KM_GMRES) {
    gmres_solve(a, m, b, x, k, eps, itmax, irst, iter, err, s, flags);
    return;
  }
  if (mid == KM_BICGSTABL) {
    bicgstabl_solve(a, m, b, x, k, eps, itmax, irst, iter, err, s, flags);
    return;
  }
  const bool cg = mid == KM_CG;
  KrylovPlan *p = get_plan(a, mid, m, k, s);
  prepare_method(p, a, m);
  const int K = p->K;
  const size_t vb = (size_t)n * K * sizeof(double);
  const Vecs &v = p->v;
  const unsigned gp = (unsigned)(((int64_t)n * K + 255) / 256);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, b, v.b);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, x, v.x);
  AINV_HIP_CHECK(hipGetLastError());
  // column state: all active for the init pass, eps per column
  std::vector<Col> init(kMaxCols);
  for (auto &c0 : init) {
    c0 = Col{};
    c0.active = 1;
    c0.eps = eps;
  }
  AINV_HIP_CHECK(hipMemcpyAsync(p->col, init.data(), sizeof(Col) * kMaxCols,
                                hipMemcpyHostToDevice, s));
  AINV_HIP_CHECK(hipMemsetAsync(p->na, 0, 3 * sizeof(int32_t), s));
  // r = b - A x
  AINV_HIP_CHECK(hipMemcpyAsync(v.r, v.b, vb, hipMemcpyDeviceToDevice, s));
  hll_spmm_axpby(a->hll, -1.0, v.x, false, K, 1.0, v.r, K, k, s);
  int g;
  if (!cg) {
    AINV_HIP_CHECK(hipMemcpyAsync(v.q, v.r, vb, hipMemcpyDeviceToDevice, s));
    g = vec_op<OP_INIT_BICG>(n, k, v, p->col, p->pt, s);
    fin(FIN_INIT, k, g, p->pt, p->col, p->na, s);
  } else {
    g = vec_op<OP_INIT_CG>(n, k, x32_vecs(p), p->col, p->pt, s);
    fin(FIN_INIT_CG, k, g, p->pt, p->col, p->na, s);
  }
  AINV_HIP_CHECK(hipMemcpyAsync(p->pinned, p->na, sizeof(int32_t),
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  const bool eager = (flags & AINV_KRYLOV_EAGER) != 0;
  // eager and captured runs use the same kernel mappings, so that they
  // compute the same sums in the same order
  if (p->pinned[0] > 0 && itmax > 0) tune_plan(p, a, m, s);
  if (p->pinned[0] > 0 && itmax > 0 && !eager) capture(p, a, s);
  if (p->pinned[0] > 0) {
    // launch iteration it, then look at the count of iteration it - 1
    for (int it = 1; it <= itmax; ++it) {
      if (eager)
        enqueue_iteration(p, a, s);
      else
        AINV_HIP_CHECK(hipGraphLaunch(p->exec, s));
      const int slot = it & 1;
      AINV_HIP_CHECK(hipMemcpyAsync(p->pinned + slot, p->na, sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
      AINV_HIP_CHECK(hipEventRecord(p->ev[slot], s));
      if (it >= 2) {
        AINV_HIP_CHECK(hipEventSynchronize(p->ev[slot ^ 1]));
        if (p->pinned[slot ^ 1] == 0) break;
      }
    }
  }
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)(((int64_t)n * k + 255) / 256)),
                     dim3(256), 0, s, n, k, K, v.x, x);
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<Col> hc(k);
  AINV_HIP_CHECK(hipMemcpyAsync(hc.data(), p->col, sizeof(Col) * k,
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    iter[c] = hc[c].iter;
    err[c] = hc[c].err;
  }
}

// ------------------------------------------------------------------ RGMRES
// Restarted GMRES with right preconditioning (Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., alg. 9.5), modified Gram-Schmidt Arnoldi
// and Givens rotations, psb_krylov's 'RGMRES' with restart irst
// (tests/fileread/runs/dfs.inp:10), for k interleaved columns.  The basis
// v_0..v_m lives in HBM as m+1 blocks of n x K; the Hessenberg columns,
// rotations, rotated right-hand side and solution coefficients of each
// column are small per-column arrays updated by one block per column.
// Step i of a cycle is one captured graph: z = M v_i, w = A z with
// (w, v_0) fused, i+1 passes w -= h_ji v_j each fused with the next dot
// ((v_{j+1}, w) or (w, w)), the rotations, v_{i+1} = w / h_{i+1,i}.  The
// end graph solves H y = rs per column, x += M (V y) and forms the next
// cycle's true residual.
namespace {

struct GmCol {
  double bn, beta, err, eps, hn;
  int32_t active, done, iter, len;
};

enum GmOp {
  GM_INIT,      // dots (b,b) (v0,v0)
  GM_SCALE0,    // v0 *= 1 / beta (active)
  GM_AXPY_DOT,  // w -= h_ji v_j ; (v_{j+1}, w) or, at j == i, (w, w)
  GM_SCALE,     // v_{i+1} = (1 / h_{i+1,i}) w (active)
  GM_COMBINE,   // w = sum_{j < len} y_j v_j (len > 0)
  GM_XUPD,      // x += z (len > 0)
  GM_RCOPY,     // v0 = b
  GM_RNORM,     // (v0, v0) for the columns that restart
};

struct GmArgs {
  double *x, *b, *w, *z, *v;  // v: basis vector 0
  int64_t vs;                 // elements from one basis vector to the next
  const GmCol *col;
  const double *arr;          // per-column arrays, stride astr
  int astr, m, j, i;
};

// per-column array offsets: H column-major (m+1) x m, cs, sn, rs, y
__host__ __device__ inline int gm_h(int m, int r, int c) { return c * (m + 1) + r; }
__host__ __device__ inline int gm_cs(int m) { return (m + 1) * m; }
__host__ __device__ inline int gm_sn(int m) { return (m + 1) * m + m; }
__host__ __device__ inline int gm_rs(int m) { return (m + 1) * m + 2 * m; }
__host__ __device__ inline int gm_y(int m) { return (m + 1) * m + 3 * m + 1; }
inline int gm_stride(int m) { return ((m + 1) * m + 4 * m + 1 + 7) & ~7; }

template <int OP> struct GmPairs { static constexpr int v = 0; };
template <> struct GmPairs<GM_INIT> { static constexpr int v = 2; };
template <> struct GmPairs<GM_AXPY_DOT> { static constexpr int v = 1; };
template <> struct GmPairs<GM_RNORM> { static constexpr int v = 1; };

template <int OP, int K>
__global__ __launch_bounds__(kThreads) void k_gm_vec(int32_t n, int k, GmArgs g,
                                                     double *__restrict__ part) {
  constexpr int NP = GmPairs<OP>::v;
  constexpr int R = kThreads / K;
  const int c = threadIdx.x % K;
  const int rr = threadIdx.x / K;
  bool cok = c < k;
  GmCol cs{};
  if (cok) {
    cs = g.col[c];
    if constexpr (OP == GM_SCALE0 || OP == GM_AXPY_DOT || OP == GM_SCALE)
      cok = cs.active != 0;
    else if constexpr (OP == GM_COMBINE || OP == GM_XUPD)
      cok = cs.len > 0;
    else if constexpr (OP == GM_RNORM)
      cok = cs.len > 0 && !cs.done;
  }
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int q = 0; q < (NP > 0 ? NP : 1); ++q) acc[q] = 0.0;
  if (cok) {
    const double *ar = g.arr + (int64_t)c * g.astr;
    double sc = 0.0;
    if constexpr (OP == GM_SCALE0) sc = 1.0 / cs.beta;
    if constexpr (OP == GM_SCALE) sc = 1.0 / cs.hn;
    if constexpr (OP == GM_AXPY_DOT) sc = ar[gm_h(g.m, g.j, g.i)];
    const double *vj = g.v + (int64_t)g.j * g.vs;
    for (int64_t r = (int64_t)blockIdx.x * R + rr; r < n;
         r += (int64_t)gridDim.x * R) {
      const int64_t e = r * K + c;
      if constexpr (OP == GM_INIT) {
        acc[0] += g.b[e] * g.b[e];
        acc[1] += g.v[e] * g.v[e];
      } else if constexpr (OP == GM_SCALE0) {
        g.v[e] = sc * g.v[e];
      } else if constexpr (OP == GM_AXPY_DOT) {
        const double wn = g.w[e] + (-sc) * vj[e];
        g.w[e] = wn;
        acc[0] += (g.j < g.i ? vj[g.vs + e] : wn) * wn;
      } else if constexpr (OP == GM_SCALE) {
        if (cs.hn != 0.0) g.v[(int64_t)(g.i + 1) * g.vs + e] = sc * g.w[e];
      } else if constexpr (OP == GM_COMBINE) {
        const double *y = ar + gm_y(g.m);
        double u = 0.0;
        for (int j = 0; j < cs.len; ++j) u = u + y[j] * g.v[(int64_t)j * g.vs + e];
        g.w[e] = u;
      } else if constexpr (OP == GM_XUPD) {
        g.x[e] = g.x[e] + g.z[e];
      } else if constexpr (OP == GM_RCOPY) {
        g.v[e] = g.b[e];
      } else if constexpr (OP == GM_RNORM) {
        acc[0] += g.v[e] * g.v[e];
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

template <int OP>
int gm_vec(int32_t n, int k, int K, const GmArgs &g, double *part,
           hipStream_t s) {
  const int nb = vec_blocks(n, k);
#define AINV_GL(KK)                                                         \
  hipLaunchKernelGGL((k_gm_vec<OP, KK>), dim3(nb), dim3(kThreads), 0, s, n, \
                     k, g, part)
  switch (K) {
    case 1: AINV_GL(1); break;
    case 2: AINV_GL(2); break;
    case 4: AINV_GL(4); break;
    case 8: AINV_GL(8); break;
    case 16: AINV_GL(16); break;
    case 32: AINV_GL(32); break;
    default: AINV_GL(64); break;
  }
#undef AINV_GL
  AINV_HIP_CHECK(hipGetLastError());
  return nb;
}

enum GmFin {
  FG_INIT,     // bn, beta, err, active / done, rs = beta e_1
  FG_H,        // h_ji = sum (j == 0 resets the active count)
  FG_NORM,     // h_{i+1,i}, rotations, rs, err, convergence
  FG_SOLVE,    // y = H^-1 rs (len > 0); resets the active count
  FG_RESTART,  // beta, err of the true residual; rs = beta e_1; len = 0
};

__global__ __launch_bounds__(kThreads) void k_gm_fin(
    int fin, int j, int i, int nblk, const double *__restrict__ part,
    GmCol *__restrict__ col, double *__restrict__ arr, int astr, int m,
    int32_t *__restrict__ nactive) {
  const int c = blockIdx.x;
  const int np = fin == FG_INIT ? 2 : (fin == FG_SOLVE ? 0 : 1);
  __shared__ double red[2][kThreads];
  for (int q = 0; q < np; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads)
      s += part[((int64_t)q * kRedBlocks + b) * kMaxCols + c];
    red[q][threadIdx.x] = s;
  }
  __syncthreads();
  for (int st = kThreads / 2; st >= 1; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < np; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (((fin == FG_H && j == 0) || fin == FG_SOLVE) && c == 0) *nactive = 0;
  GmCol cs = col[c];
  double *ar = arr + (int64_t)c * astr;
  double *rs = ar + gm_rs(m);
  const double s0 = red[0][0], s1 = red[1][0];
  switch (fin) {
    case FG_INIT:
      cs.bn = sqrt(s0);
      cs.beta = sqrt(s1);
      cs.err = (cs.bn > 0.0) ? cs.beta / cs.bn : cs.beta;
      cs.iter = 0;
      cs.len = 0;
      cs.hn = 0.0;
      cs.done = cs.err <= cs.eps;
      cs.active = !cs.done;
      for (int q = 0; q <= m; ++q) rs[q] = 0.0;
      rs[0] = cs.beta;
      break;
    case FG_H:
      if (cs.active) ar[gm_h(m, j, i)] = s0;
      break;
    case FG_NORM: {
      if (!cs.active) break;
      cs.iter += 1;
      double *cv = ar + gm_cs(m), *sv = ar + gm_sn(m);
      const double hn = sqrt(s0);
      cs.hn = hn;
      ar[gm_h(m, i + 1, i)] = hn;
      for (int q = 0; q < i; ++q) {
        const double t0 = ar[gm_h(m, q, i)], t1 = ar[gm_h(m, q + 1, i)];
        ar[gm_h(m, q, i)] = cv[q] * t0 + sv[q] * t1;
        ar[gm_h(m, q + 1, i)] = -sv[q] * t0 + cv[q] * t1;
      }
      const double hii = ar[gm_h(m, i, i)];
      const double gg = sqrt(hii * hii + hn * hn);
      if (gg == 0.0) {  // breakdown: stop with the steps so far
        cs.len = i;
        cs.active = 0;
        cs.done = 1;
        break;
      }
      cv[i] = hii / gg;
      sv[i] = hn / gg;
      ar[gm_h(m, i, i)] = gg;
      ar[gm_h(m, i + 1, i)] = 0.0;
      rs[i + 1] = -sv[i] * rs[i];
      rs[i] = cv[i] * rs[i];
      cs.len = i + 1;
      cs.err = (cs.bn > 0.0) ? fabs(rs[i + 1]) / cs.bn : fabs(rs[i + 1]);
      if (cs.err <= cs.eps) {
        cs.active = 0;
        cs.done = 1;
      }
      break;
    }
    case FG_SOLVE: {
      double *y = ar + gm_y(m);
      for (int q = cs.len - 1; q >= 0; --q) {
        double t = rs[q];
        for (int l = q + 1; l < cs.len; ++l) t = t - ar[gm_h(m, q, l)] * y[l];
        y[q] = t / ar[gm_h(m, q, q)];
      }
      break;
    }
    case FG_RESTART:
      if (cs.len > 0 && !cs.done) {
        cs.beta = sqrt(s0);
        cs.err = (cs.bn > 0.0) ? cs.beta / cs.bn : cs.beta;
        cs.done = cs.err <= cs.eps;
        cs.active = !cs.done;
        for (int q = 0; q <= m; ++q) rs[q] = 0.0;
        rs[0] = cs.beta;
      }
      cs.len = 0;
      break;
  }
  col[c] = cs;
  if ((fin == FG_INIT || fin == FG_NORM || fin == FG_RESTART) && cs.active)
    atomicAdd(nactive, 1);
}

void gm_fin(const KrylovPlan *p, int which, int j, int i, int nblk,
            hipStream_t s) {
  hipLaunchKernelGGL(k_gm_fin, dim3(p->k), dim3(kThreads), 0, s, which, j, i,
                     nblk, p->pt, p->gmcol.as<GmCol>(), p->gmarr.as<double>(),
                     gm_stride(p->irst), p->irst, p->na);
  AINV_HIP_CHECK(hipGetLastError());
}

GmArgs gm_args(const KrylovPlan *p, int j, int i) {
  GmArgs g{};
  g.x = p->v.x;
  g.b = p->v.b;
  g.w = p->v.r;
  g.z = p->v.z;
  g.v = p->basis.as<double>();
  g.vs = (int64_t)p->n * p->K;
  g.col = p->gmcol.as<GmCol>();
  g.arr = p->gmarr.as<double>();
  g.astr = gm_stride(p->irst);
  g.m = p->irst;
  g.j = j;
  g.i = i;
  return g;
}

void gm_step(const KrylovPlan *p, ainv_matrix *a, int i, hipStream_t s) {
  const int32_t n = p->n;
  const int k = p->k, K = p->K;
  double *v0 = p->basis.as<double>();
  const int64_t vs = (int64_t)n * K;
  precond(p, v0 + i * vs, p->v.z, s);
  int g = hll_spmv_dot(a->hll, p->v.z, p->v.r, k, K, v0, 1, p->pt,
                       p->spmv_map, s);
  gm_fin(p, FG_H, 0, i, g, s);
  for (int j = 0; j <= i; ++j) {
    g = gm_vec<GM_AXPY_DOT>(n, k, K, gm_args(p, j, i), p->pt, s);
    gm_fin(p, j < i ? FG_H : FG_NORM, j + 1, i, g, s);
  }
  gm_vec<GM_SCALE>(n, k, K, gm_args(p, 0, i), p->pt, s);
}

// the true residual of x into v_0 and its norm (columns that restart)
void gm_residual(const KrylovPlan *p, ainv_matrix *a, hipStream_t s,
                 bool init) {
  const int32_t n = p->n;
  const int k = p->k, K = p->K;
  const GmArgs g0 = gm_args(p, 0, 0);
  gm_vec<GM_RCOPY>(n, k, K, g0, p->pt, s);
  hll_spmm_axpby(a->hll, -1.0, p->v.x, false, K, 1.0, g0.v, K, k, s,
                 p->spmv_map);
  const int g = init ? gm_vec<GM_INIT>(n, k, K, g0, p->pt, s)
                     : gm_vec<GM_RNORM>(n, k, K, g0, p->pt, s);
  gm_fin(p, init ? FG_INIT : FG_RESTART, 0, 0, g, s);
  gm_vec<GM_SCALE0>(n, k, K, g0, p->pt, s);
}

void gm_end(const KrylovPlan *p, ainv_matrix *a, hipStream_t s) {
  const int32_t n = p->n;
  const int k = p->k, K = p->K;
  const GmArgs g0 = gm_args(p, 0, 0);
  gm_fin(p, FG_SOLVE, 0, 0, 0, s);
  gm_vec<GM_COMBINE>(n, k, K, g0, p->pt, s);
  precond(p, p->v.r, p->v.z, s);
  gm_vec<GM_XUPD>(n, k, K, g0, p->pt, s);
  gm_residual(p, a, s, false);
}

void gm_capture(KrylovPlan *p, ainv_matrix *a, hipStream_t s) {
  if (!p->gm_exec.empty()) return;
  for (int i = 0; i <= p->irst; ++i) {
    hipGraph_t gr = nullptr;
    AINV_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      if (i < p->irst)
        gm_step(p, a, i, s);
      else
        gm_end(p, a, s);
    } catch (...) {
      hipGraph_t g0 = nullptr;
      (void)hipStreamEndCapture(s, &g0);
      if (g0) (void)hipGraphDestroy(g0);
      throw;
    }
    AINV_HIP_CHECK(hipStreamEndCapture(s, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    AINV_HIP_CHECK(e);
    p->gm_exec.push_back(ex);
  }
}

}  // namespace

void gmres_solve(ainv_matrix *a, ainv_solver *m, const double *b, double *x,
                 int k, double eps, int itmax, int irst, int32_t *iter,
                 double *err, hipStream_t s, int flags) {
  const int32_t n = a->n;
  const int mr = irst <= 0 ? 10 : irst;
  if (mr > 500) throw Error(AINV_ERR_ARG, "irst must be <= 500");
  KrylovPlan *p = get_plan(a, KM_GMRES, m, k, s, mr);
  const int K = p->K;
  const int64_t vs = (int64_t)n * K;
  if (!p->basis.p) {
    p->basis.alloc((size_t)(mr + 1) * vs * sizeof(double));
    // padding columns are read by the row-mapped kernels: keep them 0
    AINV_HIP_CHECK(hipMemsetAsync(p->basis.p, 0, p->basis.bytes, s));
    p->gmcol.alloc(sizeof(GmCol) * kMaxCols);
    p->gmarr.alloc(sizeof(double) * (size_t)gm_stride(mr) * kMaxCols);
  }
  const unsigned gp = (unsigned)((vs + 255) / 256);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, b, p->v.b);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, x, p->v.x);
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<GmCol> init(kMaxCols);
  for (auto &c0 : init) {
    c0 = GmCol{};
    c0.eps = eps;
  }
  AINV_HIP_CHECK(hipMemcpyAsync(p->gmcol.p, init.data(),
                                sizeof(GmCol) * kMaxCols,
                                hipMemcpyHostToDevice, s));
  AINV_HIP_CHECK(hipMemsetAsync(p->na, 0, 3 * sizeof(int32_t), s));
  const bool eager = (flags & AINV_KRYLOV_EAGER) != 0;
  // the init step uses the tuned system-matrix mapping, as the graphs do
  tune_plan(p, a, m, s);
  gm_residual(p, a, s, true);
  AINV_HIP_CHECK(hipMemcpyAsync(p->pinned, p->na, sizeof(int32_t),
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  int active = p->pinned[0];
  if (active > 0 && itmax > 0 && !eager) gm_capture(p, a, s);
  int it = 0;
  while (active > 0 && it < itmax) {
    // one cycle: launch step i, then look at the count after step i - 1
    for (int i = 0; i < mr && it < itmax; ++i, ++it) {
      if (eager)
        gm_step(p, a, i, s);
      else
        AINV_HIP_CHECK(hipGraphLaunch(p->gm_exec[i], s));
      const int slot = i & 1;
      AINV_HIP_CHECK(hipMemcpyAsync(p->pinned + slot, p->na, sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
      AINV_HIP_CHECK(hipEventRecord(p->ev[slot], s));
      if (i >= 1) {
        AINV_HIP_CHECK(hipEventSynchronize(p->ev[slot ^ 1]));
        if (p->pinned[slot ^ 1] == 0) {
          ++it;
          break;
        }
      }
    }
    if (eager)
      gm_end(p, a, s);
    else
      AINV_HIP_CHECK(hipGraphLaunch(p->gm_exec[mr], s));
    AINV_HIP_CHECK(hipMemcpyAsync(p->pinned, p->na, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
    active = p->pinned[0];
  }
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)(((int64_t)n * k + 255) / 256)),
                     dim3(256), 0, s, n, k, K, p->v.x, x);
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<GmCol> hc(k);
  AINV_HIP_CHECK(hipMemcpyAsync(hc.data(), p->gmcol.p, sizeof(GmCol) * k,
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    iter[c] = hc[c].iter;
    err[c] = hc[c].err;
  }
}

// --------------------------------------------------------------- BiCGStab(l)
// Sleijpen and Fokkema, ETNA 1 (1993) 11-32, alg. 3.1, on the right-
// preconditioned operator A M (x = x0 + M yhat at exit), psb_krylov's
// 'BICGSTABL' with l = irst (tests/pdegen/runs/mld3d.inp:1, :10), for k
// interleaved columns.  rhat_0..l and uhat_0..l are 2 (l + 1) blocks of
// n x K; a whole cycle (l BiCG steps, the Gram-Schmidt minimal-residual
// step, the update and the istop = 2 test) has a fixed structure and is one
// captured graph.  The dots of each A M product ride on it
// (hll_spmv_dot), every Gram-Schmidt update carries the next dot.
namespace {

struct BlCol {
  double bn, err, eps, rho0, rho_next, alpha, omega, beta;
  int32_t active, iter, brk, pad;
};

enum BlOp {
  BL_INIT,   // rt = r0 ; yhat = 0 ; u0 = 0 ; dots (b,b) (r0,r0) (r0,rt)
  BL_U,      // u_i = r_i - beta u_i, i <= j
  BL_R,      // r_i -= alpha u_{i+1}, i <= j ; yhat += alpha u_0
  BL_DOT,    // (r_j, r_i)
  BL_MGS,    // r_j -= tau_ij r_i ; (r_j, r_{i+1}) or, at i + 1 == j,
             // (r_j, r_j) (r_0, r_j)
  BL_SIG1,   // (r_1, r_1) (r_0, r_1)
  BL_FINAL,  // yhat, r_0, u_0 updates ; (r_0, r_0) (r_0, rt)
  BL_XUPD,   // x += z
};

struct BlArgs {
  double *x, *b, *rt, *y, *z, *r, *u;  // r, u: blocks 0 of rhat, uhat
  int64_t vs;
  const BlCol *col;
  const double *arr;
  int astr, l, i, j;
};

// per-column arrays: tau (l+1) x (l+1) row-major, sig, gp, g, gpp (l+1 each)
__host__ __device__ inline int bl_tau(int l, int i, int j) { return i * (l + 1) + j; }
__host__ __device__ inline int bl_sig(int l) { return (l + 1) * (l + 1); }
__host__ __device__ inline int bl_gp(int l) { return (l + 1) * (l + 2); }
__host__ __device__ inline int bl_g(int l) { return (l + 1) * (l + 3); }
__host__ __device__ inline int bl_gpp(int l) { return (l + 1) * (l + 4); }
inline int bl_stride(int l) { return ((l + 1) * (l + 5) + 7) & ~7; }

template <int OP> struct BlPairs { static constexpr int v = 0; };
template <> struct BlPairs<BL_INIT> { static constexpr int v = 3; };
template <> struct BlPairs<BL_DOT> { static constexpr int v = 1; };
template <> struct BlPairs<BL_MGS> { static constexpr int v = 2; };
template <> struct BlPairs<BL_SIG1> { static constexpr int v = 2; };
template <> struct BlPairs<BL_FINAL> { static constexpr int v = 2; };

template <int OP, int K>
__global__ __launch_bounds__(kThreads) void k_bl_vec(int32_t n, int k, BlArgs g,
                                                     double *__restrict__ part) {
  constexpr int NP = BlPairs<OP>::v;
  constexpr int R = kThreads / K;
  const int c = threadIdx.x % K;
  const int rr = threadIdx.x / K;
  bool cok = c < k;
  BlCol cs{};
  if (cok) {
    cs = g.col[c];
    if constexpr (OP != BL_INIT && OP != BL_XUPD) cok = cs.active != 0;
  }
  double acc[NP > 0 ? NP : 1];
#pragma unroll
  for (int q = 0; q < (NP > 0 ? NP : 1); ++q) acc[q] = 0.0;
  if (cok) {
    const double *ar = g.arr + (int64_t)c * g.astr;
    const int l = g.l, i0 = g.i, j0 = g.j;
    double tau = 0.0;
    if constexpr (OP == BL_MGS) tau = ar[bl_tau(l, i0, j0)];
    for (int64_t r = (int64_t)blockIdx.x * R + rr; r < n;
         r += (int64_t)gridDim.x * R) {
      const int64_t e = r * K + c;
      if constexpr (OP == BL_INIT) {
        const double r0 = g.r[e];
        g.rt[e] = r0;
        g.y[e] = 0.0;
        g.u[e] = 0.0;
        acc[0] += g.b[e] * g.b[e];
        acc[1] += r0 * r0;
        acc[2] += r0 * r0;
      } else if constexpr (OP == BL_U) {
        for (int i = 0; i <= j0; ++i) {
          const int64_t o = (int64_t)i * g.vs + e;
          g.u[o] = g.r[o] - cs.beta * g.u[o];
        }
      } else if constexpr (OP == BL_R) {
        for (int i = 0; i <= j0; ++i) {
          const int64_t o = (int64_t)i * g.vs + e;
          g.r[o] = g.r[o] - cs.alpha * g.u[o + g.vs];
        }
        g.y[e] = g.y[e] + cs.alpha * g.u[e];
      } else if constexpr (OP == BL_DOT) {
        acc[0] += g.r[(int64_t)j0 * g.vs + e] * g.r[(int64_t)i0 * g.vs + e];
      } else if constexpr (OP == BL_MGS) {
        const int64_t oj = (int64_t)j0 * g.vs + e;
        const double rj = g.r[oj] - tau * g.r[(int64_t)i0 * g.vs + e];
        g.r[oj] = rj;
        if (i0 + 1 < j0) {
          acc[0] += rj * g.r[(int64_t)(i0 + 1) * g.vs + e];
        } else {
          acc[0] += rj * rj;
          acc[1] += g.r[e] * rj;
        }
      } else if constexpr (OP == BL_SIG1) {
        const double r1 = g.r[g.vs + e];
        acc[0] += r1 * r1;
        acc[1] += g.r[e] * r1;
      } else if constexpr (OP == BL_FINAL) {
        const double *gp = ar + bl_gp(l), *gg = ar + bl_g(l),
                     *gpp = ar + bl_gpp(l);
        double yv = g.y[e] + gg[1] * g.r[e];
        double rv = g.r[e] - gp[l] * g.r[(int64_t)l * g.vs + e];
        double uv = g.u[e] - gg[l] * g.u[(int64_t)l * g.vs + e];
        for (int jj = 1; jj <= l - 1; ++jj) {
          const int64_t o = (int64_t)jj * g.vs + e;
          uv = uv - gg[jj] * g.u[o];
          yv = yv + gpp[jj] * g.r[o];
          rv = rv - gp[jj] * g.r[o];
        }
        g.y[e] = yv;
        g.r[e] = rv;
        g.u[e] = uv;
        acc[0] += rv * rv;
        acc[1] += rv * g.rt[e];
      } else if constexpr (OP == BL_XUPD) {
        g.x[e] = g.x[e] + g.z[e];
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

template <int OP>
int bl_vec(int32_t n, int k, int K, const BlArgs &g, double *part,
           hipStream_t s) {
  const int nb = vec_blocks(n, k);
#define AINV_BL(KK)                                                         \
  hipLaunchKernelGGL((k_bl_vec<OP, KK>), dim3(nb), dim3(kThreads), 0, s, n, \
                     k, g, part)
  switch (K) {
    case 1: AINV_BL(1); break;
    case 2: AINV_BL(2); break;
    case 4: AINV_BL(4); break;
    case 8: AINV_BL(8); break;
    case 16: AINV_BL(16); break;
    case 32: AINV_BL(32); break;
    default: AINV_BL(64); break;
  }
#undef AINV_BL
  AINV_HIP_CHECK(hipGetLastError());
  return nb;
}

enum BlFin {
  FB_INIT,   // bn, err, active; rho0 = 1, alpha = 0, omega = 1
  FB_BETA,   // (j == 0: rho0 = -omega rho0) rho1, beta, rho0 = rho1
  FB_ALPHA,  // alpha = rho0 / (u_{j+1}, rt)
  FB_TAU,    // tau_ij = (r_j, r_i) / sig_i
  FB_SIG,    // sig_j, gamma'_j
  FB_GAMMA,  // gamma, gamma'', omega
  FB_CONV,   // iter += l, err, convergence, rho for the next cycle
};

__global__ __launch_bounds__(kThreads) void k_bl_fin(
    int fin, int i, int j, int nblk, const double *__restrict__ part,
    BlCol *__restrict__ col, double *__restrict__ arr, int astr, int l,
    int32_t *__restrict__ nactive) {
  const int c = blockIdx.x;
  int np = 1;
  if (fin == FB_INIT) np = 3;
  else if (fin == FB_SIG || fin == FB_CONV) np = 2;
  else if (fin == FB_GAMMA || (fin == FB_BETA && j == 0)) np = 0;
  __shared__ double red[3][kThreads];
  for (int q = 0; q < np; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads)
      s += part[((int64_t)q * kRedBlocks + b) * kMaxCols + c];
    red[q][threadIdx.x] = s;
  }
  __syncthreads();
  for (int st = kThreads / 2; st >= 1; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < np; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (fin == FB_BETA && j == 0 && c == 0) *nactive = 0;
  BlCol cs = col[c];
  double *ar = arr + (int64_t)c * astr;
  const double s0 = np > 0 ? red[0][0] : 0.0, s1 = np > 1 ? red[1][0] : 0.0,
               s2 = np > 2 ? red[2][0] : 0.0;
  if (fin == FB_INIT) {
    cs.bn = sqrt(s0);
    const double rn = sqrt(s1);
    cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
    cs.active = !(cs.err <= cs.eps);
    cs.iter = 0;
    cs.brk = 0;
    cs.rho0 = 1.0;
    cs.alpha = 0.0;
    cs.omega = 1.0;
    cs.beta = 0.0;
    cs.rho_next = s2;
  } else if (cs.active) {
    double *sig = ar + bl_sig(l), *gp = ar + bl_gp(l), *gg = ar + bl_g(l),
           *gpp = ar + bl_gpp(l);
    switch (fin) {
      case FB_BETA: {
        if (j == 0) cs.rho0 = -cs.omega * cs.rho0;
        const double rho1 = j == 0 ? cs.rho_next : s0;
        if (cs.rho0 == 0.0) {
          cs.active = 0;
          cs.brk = 1;
          break;
        }
        cs.beta = cs.alpha * rho1 / cs.rho0;
        cs.rho0 = rho1;
        break;
      }
      case FB_ALPHA:
        if (s0 == 0.0) {
          cs.active = 0;
          cs.brk = 1;
        } else {
          cs.alpha = cs.rho0 / s0;
        }
        break;
      case FB_TAU:
        ar[bl_tau(l, i, j)] = s0 / sig[i];
        break;
      case FB_SIG:
        sig[j] = s0;
        if (s0 == 0.0) {
          cs.active = 0;
          cs.brk = 1;
        } else {
          gp[j] = s1 / s0;
        }
        break;
      case FB_GAMMA: {
        gg[l] = gp[l];
        cs.omega = gg[l];
        for (int jj = l - 1; jj >= 1; --jj) {
          double t = gp[jj];
          for (int ii = jj + 1; ii <= l; ++ii) t = t - ar[bl_tau(l, jj, ii)] * gg[ii];
          gg[jj] = t;
        }
        for (int jj = 1; jj <= l - 1; ++jj) {
          double t = gg[jj + 1];
          for (int ii = jj + 1; ii <= l - 1; ++ii)
            t = t + ar[bl_tau(l, jj, ii)] * gg[ii + 1];
          gpp[jj] = t;
        }
        break;
      }
      case FB_CONV: {
        cs.iter += l;
        const double rn = sqrt(s0);
        cs.err = (cs.bn > 0.0) ? rn / cs.bn : rn;
        cs.rho_next = s1;
        if (cs.err <= cs.eps) cs.active = 0;
        break;
      }
    }
  }
  col[c] = cs;
  if ((fin == FB_INIT || fin == FB_CONV) && cs.active) atomicAdd(nactive, 1);
}

void bl_fin(const KrylovPlan *p, int which, int i, int j, int nblk,
            hipStream_t s) {
  hipLaunchKernelGGL(k_bl_fin, dim3(p->k), dim3(kThreads), 0, s, which, i, j,
                     nblk, p->pt, p->gmcol.as<BlCol>(), p->gmarr.as<double>(),
                     bl_stride(p->irst), p->irst, p->na);
  AINV_HIP_CHECK(hipGetLastError());
}

BlArgs bl_args(const KrylovPlan *p, int i, int j) {
  BlArgs g{};
  g.x = p->v.x;
  g.b = p->v.b;
  g.rt = p->v.q;
  g.y = p->v.p;
  g.z = p->v.z;
  const int64_t vs = (int64_t)p->n * p->K;
  g.r = p->basis.as<double>();
  g.u = g.r + (int64_t)(p->irst + 1) * vs;
  g.vs = vs;
  g.col = p->gmcol.as<BlCol>();
  g.arr = p->gmarr.as<double>();
  g.astr = bl_stride(p->irst);
  g.l = p->irst;
  g.i = i;
  g.j = j;
  return g;
}

void bl_cycle(const KrylovPlan *p, ainv_matrix *a, hipStream_t s) {
  const int32_t n = p->n;
  const int k = p->k, K = p->K, l = p->irst, sm = p->spmv_map;
  const BlArgs g0 = bl_args(p, 0, 0);
  const int64_t vs = g0.vs;
  int g = 0;
  for (int j = 0; j < l; ++j) {  // BiCG part
    bl_fin(p, FB_BETA, 0, j, g, s);
    bl_vec<BL_U>(n, k, K, bl_args(p, 0, j), p->pt, s);
    precond(p, g0.u + j * vs, p->v.z, s);
    g = hll_spmv_dot(a->hll, p->v.z, g0.u + (j + 1) * vs, k, K, g0.rt, 1,
                     p->pt, sm, s);
    bl_fin(p, FB_ALPHA, 0, j, g, s);
    bl_vec<BL_R>(n, k, K, bl_args(p, 0, j), p->pt, s);
    precond(p, g0.r + j * vs, p->v.z, s);
    g = hll_spmv_dot(a->hll, p->v.z, g0.r + (j + 1) * vs, k, K, g0.rt, 1,
                     p->pt, sm, s);
  }
  for (int j = 1; j <= l; ++j) {  // minimal-residual part
    if (j == 1) {
      g = bl_vec<BL_SIG1>(n, k, K, g0, p->pt, s);
      bl_fin(p, FB_SIG, 0, 1, g, s);
      continue;
    }
    g = bl_vec<BL_DOT>(n, k, K, bl_args(p, 1, j), p->pt, s);
    bl_fin(p, FB_TAU, 1, j, g, s);
    for (int i = 1; i < j; ++i) {
      g = bl_vec<BL_MGS>(n, k, K, bl_args(p, i, j), p->pt, s);
      if (i + 1 < j)
        bl_fin(p, FB_TAU, i + 1, j, g, s);
      else
        bl_fin(p, FB_SIG, 0, j, g, s);
    }
  }
  bl_fin(p, FB_GAMMA, 0, 0, 0, s);
  g = bl_vec<BL_FINAL>(n, k, K, g0, p->pt, s);
  bl_fin(p, FB_CONV, 0, 0, g, s);
}

}  // namespace

void bicgstabl_solve(ainv_matrix *a, ainv_solver *m, const double *b,
                     double *x, int k, double eps, int itmax, int l,
                     int32_t *iter, double *err, hipStream_t s, int flags) {
  const int32_t n = a->n;
  const int ll = l <= 0 ? 2 : l;
  if (ll > 16) throw Error(AINV_ERR_ARG, "BICGSTABL: l must be <= 16");
  KrylovPlan *p = get_plan(a, KM_BICGSTABL, m, k, s, ll);
  const int K = p->K;
  const int64_t vs = (int64_t)n * K;
  if (!p->basis.p) {
    p->basis.alloc((size_t)2 * (ll + 1) * vs * sizeof(double));
    // padding columns are read by the row-mapped kernels: keep them 0
    AINV_HIP_CHECK(hipMemsetAsync(p->basis.p, 0, p->basis.bytes, s));
    p->gmcol.alloc(sizeof(BlCol) * kMaxCols);
    p->gmarr.alloc(sizeof(double) * (size_t)bl_stride(ll) * kMaxCols);
  }
  const unsigned gp = (unsigned)((vs + 255) / 256);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, b, p->v.b);
  hipLaunchKernelGGL(k_pack, dim3(gp), dim3(256), 0, s, n, k, K, x, p->v.x);
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<BlCol> init(kMaxCols);
  for (auto &c0 : init) {
    c0 = BlCol{};
    c0.eps = eps;
  }
  AINV_HIP_CHECK(hipMemcpyAsync(p->gmcol.p, init.data(),
                                sizeof(BlCol) * kMaxCols,
                                hipMemcpyHostToDevice, s));
  AINV_HIP_CHECK(hipMemsetAsync(p->na, 0, 3 * sizeof(int32_t), s));
  tune_plan(p, a, m, s);
  const BlArgs g0 = bl_args(p, 0, 0);
  // r0 = b - A x ; rt = r0 ; yhat = 0 ; uhat_0 = 0
  AINV_HIP_CHECK(hipMemcpyAsync(g0.r, p->v.b, (size_t)vs * sizeof(double),
                                hipMemcpyDeviceToDevice, s));
  hll_spmm_axpby(a->hll, -1.0, p->v.x, false, K, 1.0, g0.r, K, k, s,
                 p->spmv_map);
  int g = bl_vec<BL_INIT>(n, k, K, g0, p->pt, s);
  bl_fin(p, FB_INIT, 0, 0, g, s);
  AINV_HIP_CHECK(hipMemcpyAsync(p->pinned, p->na, sizeof(int32_t),
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  const bool eager = (flags & AINV_KRYLOV_EAGER) != 0;
  if (p->pinned[0] > 0 && itmax > 0 && !eager && p->gm_exec.empty()) {
    hipGraph_t gr = nullptr;
    AINV_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      bl_cycle(p, a, s);
    } catch (...) {
      hipGraph_t gx = nullptr;
      (void)hipStreamEndCapture(s, &gx);
      if (gx) (void)hipGraphDestroy(gx);
      throw;
    }
    AINV_HIP_CHECK(hipStreamEndCapture(s, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    AINV_HIP_CHECK(e);
    p->gm_exec.push_back(ex);
  }
  if (p->pinned[0] > 0) {
    // launch cycle c, then look at the count after cycle c - 1
    int c = 0;
    for (int it = 0; it < itmax; it += ll, ++c) {
      if (eager)
        bl_cycle(p, a, s);
      else
        AINV_HIP_CHECK(hipGraphLaunch(p->gm_exec[0], s));
      const int slot = c & 1;
      AINV_HIP_CHECK(hipMemcpyAsync(p->pinned + slot, p->na, sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
      AINV_HIP_CHECK(hipEventRecord(p->ev[slot], s));
      if (c >= 1) {
        AINV_HIP_CHECK(hipEventSynchronize(p->ev[slot ^ 1]));
        if (p->pinned[slot ^ 1] == 0) break;
      }
    }
  }
  // x = x0 + M yhat
  precond(p, g0.y, p->v.z, s);
  bl_vec<BL_XUPD>(n, k, K, g0, p->pt, s);
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)(((int64_t)n * k + 255) / 256)),
                     dim3(256), 0, s, n, k, K, p->v.x, x);
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<BlCol> hc(k);
  AINV_HIP_CHECK(hipMemcpyAsync(hc.data(), p->gmcol.p, sizeof(BlCol) * k,
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    iter[c] = hc[c].brk ? -1 : hc[c].iter;
    err[c] = hc[c].err;
  }
}

namespace {
struct ColMap {
  int cnt;
  int from[kMaxCols];
};

// Column map.from[j] of every workspace vector of one plan to column j of
// the same vector of a plan with fewer slots (same method and solver, so the
// same vector list).
__global__ void k_shrink_cols(int32_t n, int K1, int K2,
                              const double *__restrict__ w1,
                              double *__restrict__ w2, int nvec, int64_t vs1,
                              int64_t vs2, const float *__restrict__ f1,
                              float *__restrict__ f2, int nv32, int64_t fs1,
                              int64_t fs2, ColMap map) {
  // one thread per (row, kept column): neighbouring lanes move the columns
  // of one row, so each load and store instruction covers a few rows
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * map.cnt) return;
  const int64_t row = t / map.cnt;
  const int j = (int)(t - row * map.cnt);
  const int c = map.from[j];
  for (int v = 0; v < nvec; ++v)
    w2[v * vs2 + row * K2 + j] = w1[v * vs1 + row * K1 + c];
  for (int v = 0; v < nv32; ++v)
    f2[v * fs2 + row * K2 + j] = f1[v * fs1 + row * K1 + c];
}

// Column states to the new slots (the rest empty) and the active width the
// next iteration rolls in.
__global__ void k_shrink_state(const Col *__restrict__ c1, Col *__restrict__ c2,
                               int K2, ColMap map, int32_t *__restrict__ na2) {
  const int t = threadIdx.x;
  if (t < K2) {
    Col c = Col{};
    if (t < map.cnt) c = c1[map.from[t]];
    c2[t] = c;
  }
  if (t == 0) {
    int32_t w = 0;
    for (int j = 0; j < map.cnt; ++j)
      if (c1[map.from[j]].active) w = j + 1;
    na2[1] = w;
  }
}
}  // namespace

namespace {
void move_columns(SlotBatch *b, SlotBatch *nb, const std::vector<int> &keep);
}  // namespace

// A batch whose queue has run dry and whose remaining columns fit in half
// of its slots continues on a plan with fewer slots: the columns `keep`
// (their state and every workspace vector) move to slots 0.. of a new batch
// on the same stream, and this batch closes.  The narrower plan's graph runs
// every kernel at the smaller block width.
SlotBatch *batch_shrink(SlotBatch *b, ainv_solver *m,
                        const std::vector<int> &keep) {
  KrylovPlan *p1 = b->p;
  int K2 = 1;
  while (K2 < (int)keep.size()) K2 <<= 1;
  if ((int)keep.size() > kMaxCols || K2 >= b->K)
    throw Error(AINV_ERR_INTERNAL, "shrink to a batch that is not narrower");
  SlotBatch *nb = batch_open(b->a, m, p1->method, K2, b->s);
  try {
    move_columns(b, nb, keep);
  } catch (...) {
    batch_close(nb);  // the caller still owns b
    throw;
  }
  batch_close(b);
  return nb;
}

namespace {
void move_columns(SlotBatch *b, SlotBatch *nb, const std::vector<int> &keep) {
  KrylovPlan *p1 = b->p;
  KrylovPlan *p2 = nb->p;
  const int K2 = nb->K;
  const int32_t n = p1->n;
  const int64_t vs1 = (int64_t)n * p1->K, vs2 = (int64_t)n * p2->K;
  const int nvec = (int)(p1->work.bytes / (sizeof(double) * (size_t)vs1));
  const int nvec2 = (int)(p2->work.bytes / (sizeof(double) * (size_t)vs2));
  const int64_t fs1 = x32_stride(p1), fs2 = x32_stride(p2);
  const int nv32 =
      (p1->work32.p && p2->work32.p)
          ? (int)(p1->work32.bytes / (sizeof(float) * (size_t)fs1))
          : 0;
  if (nvec != nvec2)
    throw Error(AINV_ERR_INTERNAL, "plans of one method differ in vectors");
  ColMap map{};
  map.cnt = (int)keep.size();
  for (int j = 0; j < map.cnt; ++j) map.from[j] = keep[j];
  hipLaunchKernelGGL(k_shrink_cols,
                     dim3((unsigned)(((int64_t)n * map.cnt + 255) / 256)),
                     dim3(256), 0, b->s, n, p1->K, p2->K,
                     p1->work.as<double>(), p2->work.as<double>(), nvec, vs1,
                     vs2, p1->work32.as<float>(), p2->work32.as<float>(), nv32,
                     fs1, fs2, map);
  hipLaunchKernelGGL(k_shrink_state, dim3(1), dim3(kMaxCols), 0, b->s, p1->col,
                     p2->col, K2, map, p2->na);
  AINV_HIP_CHECK(hipGetLastError());
}
}  // namespace

// Block-width autotuner of the served form: for each candidate slot count K
// (8, 16, 32, 64, no wider than the queue needs) one batch is opened (its
// mappings tuned and its iteration captured), K columns of the queue are
// admitted and, after two steps, four replayed steps are timed; the width
// with the least time per column-iteration is kept per (preconditioner,
// method) until the factors change, a wider width only when it is at least
// kWiderGain faster than the narrower ones: a step of twice the width takes
// about twice as long, so every solve's latency grows with the width, and
// the trial's own timings vary by 1-2 %.  The plans the trial made for widths
// above the chosen one are dropped again (each holds the iteration's
// vectors at its width; the narrower ones serve the queue's tail).
constexpr double kWiderGain = 0.03;

int tune_slots(ainv_matrix *a, ainv_solver *m, int mid, const double *b,
               int nrhs, double eps, hipStream_t s) {
  if (m && m->auto_slots[mid] > 0 && m->auto_slots_gen[mid] == m->generation)
    return m->auto_slots[mid];
  std::vector<const KrylovPlan *> before;
  for (auto &q : a->plans) before.push_back(q.get());
  if (m) std::fill(m->slot_ms[mid], m->slot_ms[mid] + 8, 0.0);
  int cap = 1;
  while (cap < nrhs && cap < kMaxCols) cap <<= 1;
  const int64_t n = a->n;
  int best = 0;
  double best_ms = 0.0;
  for (int K = 8; K <= kMaxCols; K <<= 1) {
    if (K > cap && K > 8) break;
    SlotBatch *bt = batch_open(a, m, mid, K, s);
    double ms = 0.0;
    try {
      std::vector<int> slot(K);
      std::vector<const double *> src(K);
      for (int c = 0; c < K; ++c) {
        slot[c] = c;
        src[c] = b + (int64_t)(c % nrhs) * n;
      }
      batch_admit_cols(bt, slot, src, eps);
      for (int r = 0; r < 2; ++r) batch_step(bt);
      AINV_HIP_CHECK(hipStreamSynchronize(s));
      const double t0 = now_seconds();
      for (int r = 0; r < 4; ++r) batch_step(bt);
      AINV_HIP_CHECK(hipStreamSynchronize(s));
      ms = (now_seconds() - t0) * 1e3 / 4.0 / K;
    } catch (...) {
      batch_close(bt);
      throw;
    }
    batch_close(bt);
    if (m) m->slot_ms[mid][log2_pow2(K)] = ms;
    if (best == 0 || ms < best_ms * (1.0 - kWiderGain)) {
      best = K;
      best_ms = ms;
    }
  }
  for (size_t i = a->plans.size(); i-- > 0;) {
    const KrylovPlan *q = a->plans[i].get();
    if (q->k > best && q->holders == 0 &&
        std::find(before.begin(), before.end(), q) == before.end())
      a->plans.erase(a->plans.begin() + i);
  }
  if (m) {
    m->auto_slots[mid] = best;
    m->auto_slots_gen[mid] = m->generation;
  }
  return best;
}

void krylov_stream(const std::string &method, ainv_matrix *a, ainv_solver *m,
                   const double *b, double *x, int nrhs, int slots, double eps,
                   int itmax, int32_t *iter, double *err, double *lat_ms,
                   hipStream_t s) {
  if (nrhs < 0) throw Error(AINV_ERR_ARG, "nrhs must be >= 0");
  if (slots < 0 || slots > kMaxCols)
    throw Error(AINV_ERR_ARG, "slots must be in 0..64 (0: tuned)");
  if (m && m->n != a->n)
    throw Error(AINV_ERR_ARG, "preconditioner and matrix sizes differ");
  const int mid = krylov_method_id(method);
  if (nrhs == 0) return;
  if (slots == 0) slots = tune_slots(a, m, mid, b, nrhs, eps, s);
  const int64_t n = a->n;
  ainv_solver::StreamStats ss;
  const double t_begin = now_seconds();
  SlotBatch *bt = batch_open(a, m, mid, slots, s);
  int K = bt->K;
  std::vector<int> owner(K, -1);
  std::vector<int64_t> first(K, 0);
  std::vector<double> t_adm(K, 0.0);
  std::vector<SlotState> st;
  int next = 0, done = 0;
  int64_t done_iters = 0;
  try {
    while (done < nrhs) {
      // fill every free slot (up to the requested count) from the queue
      std::vector<int> adm;
      std::vector<const double *> src;
      for (int c = 0; c < slots && next < nrhs; ++c) {
        if (owner[c] >= 0) continue;
        owner[c] = next;
        first[c] = batch_next_step(bt);
        t_adm[c] = now_seconds();
        adm.push_back(c);
        src.push_back(b + (int64_t)next * n);
        ++next;
      }
      const double t0 = now_seconds();
      batch_admit_cols(bt, adm, src, eps);
      const double t1 = now_seconds();
      batch_step(bt);
      const double t2 = now_seconds();
      ss.admit += t1 - t0;
      ss.launch += t2 - t1;
      ss.steps++;
      ss.slot_steps += K;
      for (int c = 0; c < K; ++c) ss.busy_slot_steps += owner[c] >= 0;
      // Reading the states one step behind keeps the device busy but leaves
      // a converged column's slot idle for one step: for solves of few
      // iterations (a multilevel preconditioner) that step costs more than
      // the host's turnaround, so wait for the step just launched.
      const bool short_solves = done >= 8 && done_iters < 24 * (int64_t)done;
      const int64_t seen = batch_poll(bt, short_solves, st);
      const double t = now_seconds();
      ss.wait += t - t2;
      if (seen < 0) continue;
      std::vector<int> ext;
      std::vector<double *> dst;
      for (int c = 0; c < K; ++c) {
        if (owner[c] < 0 || seen < first[c]) continue;
        if (st[c].active && st[c].iter < itmax) continue;
        const int j = owner[c];
        ext.push_back(c);
        dst.push_back(x + (int64_t)j * n);
        iter[j] = st[c].iter;
        err[j] = st[c].err;
        if (lat_ms) lat_ms[j] = (t - t_adm[c]) * 1e3;
        owner[c] = -1;
        ++done;
        done_iters += st[c].iter;
      }
      // one launch for the step's finished columns, ordered on the stream
      // before the admissions that reuse their slots and any move to a
      // narrower batch
      batch_extract_cols(bt, ext, dst);
      ss.extract += now_seconds() - t;
      // queue run dry and at most half the slots busy: continue narrower
      if (next >= nrhs && K > 1) {
        std::vector<int> keep;
        for (int c = 0; c < K; ++c)
          if (owner[c] >= 0) keep.push_back(c);
        if (!keep.empty() && (int)keep.size() <= K / 2) {
          ss.shrinks++;
          bt = batch_shrink(bt, m, keep);
          K = bt->K;
          slots = K;
          std::vector<int> ow(K, -1);
          std::vector<int64_t> fi(K, 0);
          std::vector<double> ta(K, 0.0);
          for (size_t j = 0; j < keep.size(); ++j) {
            ow[j] = owner[keep[j]];
            ta[j] = t_adm[keep[j]];
            fi[j] = batch_next_step(bt);
          }
          owner.swap(ow);
          first.swap(fi);
          t_adm.swap(ta);
        }
      }
    }
    AINV_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    batch_close(bt);
    throw;
  }
  batch_close(bt);
  ss.wall = now_seconds() - t_begin;
  if (m) m->last_stream = ss;
}

}  // namespace ainv
