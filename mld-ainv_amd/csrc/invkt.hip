// GPU row-wise sparse inverse of a unit upper triangular matrix, the core of
// the INVK and INVT builds:
//   INVK: src/impl/mld_dsparse_invk.f90:94-131, mld_d_invk_copyin.f90:84-102,
//         mld_d_invk_inv.f90:69-124, mld_d_invk_copyout.f90:62-96
//   INVT: src/impl/mld_dsparse_invt.f90:87-128, mld_d_invt_copyin.f90:82-104,
//         mld_d_invt_inv.f90:69-137, mld_d_invt_copyout.f90:124-286
// Each row i of the inverse is independent, so one wavefront computes one
// row.  The reference walks the row with an index heap: pop the smallest
// pending column k, and if k survives (INVK: row(k) /= 0 and level <= fill;
// INVT: |row(k)| >= thr) subtract row(k) * U(k,:) from the row, adding any
// new column to the heap.  Here the pending set is a sorted array in LDS:
// the pop is a pointer increment, each lane of the wave takes one entry of
// U(k,:), finds its column by binary search, updates it in place or marks it
// new, and the new columns are merged in with one parallel shift.  The
// elimination order and every rounding are those of the reference, so the
// output agrees bit for bit with a sequential restatement.
//
// Copy-out reserves space for the row in a global pool with one atomic add;
// rows are gathered into CSR order afterwards.  A row whose pending set
// exceeds the LDS capacity is recomputed by a second kernel that keeps the
// same arrays in a global scratch slab.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>

#include "ainv_internal.h"

namespace ainv {
namespace {

constexpr int kWave = 64;
constexpr int kCapLds = 512;      // entries per row in LDS (8 KiB)
constexpr int kCapGlobal = 65536; // entries per row in the global tier
constexpr int kGlobalRowsPerLaunch = 256;

enum Status : int32_t {
  ST_OK = 0,
  ST_CAP = 1,   // pending set exceeded the capacity of this tier
  ST_POOL = 2,  // output pool exhausted
  ST_BADU = 3,  // input not strictly upper triangular
};

// classic reference-BLAS dnrm2, sequential (one lane)
__device__ double dnrm2_seq(int64_t n, const double *x) {
  if (n < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int64_t t = 0; t < n; ++t) {
    double v = x[t];
    if (v != 0.0) {
      double ax = fabs(v);
      if (scale < ax) {
        double r = scale / ax;
        ssq = 1.0 + ssq * (r * r);
        scale = ax;
      } else {
        double r = ax / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrt(ssq);
}

// Ordering of candidates at the fill cap: |v| descending, column ascending.
__device__ inline bool before(double va, int32_t ka, double vb, int32_t kb) {
  double fa = fabs(va), fb = fabs(vb);
  if (fa != fb) return fa > fb;
  return ka < kb;
}

struct Shared {
  double thr;
  int32_t flag;     // status of the row (uniform)
  int32_t ncand;
  unsigned long long base;
};

// Body shared by the LDS and global-memory tiers.  key/val/lev point to
// arrays of `cap` entries private to this wavefront; insv to 64 ints.
// lev: INVK fill level; INVT 1 = pending or kept, 0 = dropped.
template <int KIND>
__device__ __forceinline__ void row_inverse(
    int32_t i, int32_t n, const int64_t *__restrict__ uptr,
    const int32_t *__restrict__ uidx, const double *__restrict__ uval,
    int32_t fill, double thresh, int32_t *key, double *val, int32_t *lev,
    int32_t *insv, Shared *sh, int32_t cap, int32_t *__restrict__ out_idx,
    double *__restrict__ out_val, int64_t pool_cap,
    unsigned long long *__restrict__ pool_top, int64_t *__restrict__ row_start,
    int32_t *__restrict__ row_cnt, int32_t *__restrict__ status) {
  const int lane = threadIdx.x;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  const int64_t ub = uptr[i], ue = uptr[i + 1];
  const int32_t m = (int32_t)(ue - ub);
  if (lane == 0) {
    sh->flag = ST_OK;
    if (KIND == AINV_KIND_INVT) {
      double nrmi = dnrm2_seq(m, uval + ub);
      // norm of the row with its unit diagonal (invt.f90:100-104)
      if (nrmi < 1.0)
        nrmi = sqrt(1.0 + nrmi * nrmi);
      else
        nrmi = nrmi * sqrt(1.0 + 1.0 / (nrmi * nrmi));
      sh->thr = thresh * nrmi;
    } else {
      sh->thr = 0.0;
    }
  }
  if (m + 1 > cap) {
    if (lane == 0) {
      status[i] = ST_CAP;
      row_cnt[i] = 0;
    }
    return;
  }
  // copy-in: row(i) = 1, row(k) = -U(i,k), all at level 0 / pending
  if (lane == 0) {
    key[0] = i;
    val[0] = 1.0;
    lev[0] = (KIND == AINV_KIND_INVK) ? 0 : 1;
  }
  int32_t bad = 0;
  for (int32_t t = lane; t < m; t += kWave) {
    int32_t kk = uidx[ub + t];
    bad |= (kk <= i);
    key[t + 1] = kk;
    val[t + 1] = -1.0 * uval[ub + t];
    lev[t + 1] = (KIND == AINV_KIND_INVK) ? 0 : 1;
  }
  if (__ballot(bad)) {
    if (lane == 0) {
      status[i] = ST_BADU;
      row_cnt[i] = 0;
    }
    return;
  }
  __syncthreads();
  const double thr = sh->thr;
  int32_t cnt = m + 1;
  for (int32_t p = 1; p < cnt; ++p) {
    const int32_t k = key[p];
    const double rk = val[p];
    const int32_t lk = lev[p];
    bool expand;
    if (KIND == AINV_KIND_INVK) {
      expand = (rk != 0.0) && (lk <= fill);
    } else {
      expand = !(fabs(rk) < thr);
      __syncthreads();
      if (!expand && lane == 0) {
        val[p] = 0.0;  // dropped (inv.f90:86-92)
        lev[p] = 0;
      }
    }
    if (!expand) continue;
    const int64_t kb = uptr[k], ke = uptr[k + 1];
    for (int64_t c0 = kb; c0 < ke; c0 += kWave) {
      const int64_t ee = c0 + lane;
      const bool has = ee < ke;
      const int32_t j = has ? uidx[ee] : INT_MAX;
      const double uv = has ? uval[ee] : 0.0;
      if (__ballot(has && j <= k)) {
        if (lane == 0) {
          status[i] = ST_BADU;
          row_cnt[i] = 0;
        }
        return;
      }
      // first position in (p, cnt) whose column is >= j
      int32_t lo = p + 1, hi = cnt;
      while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (key[mid] < j)
          lo = mid + 1;
        else
          hi = mid;
      }
      const bool found = has && lo < cnt && key[lo] == j;
      const double t = rk * uv;
      bool ins = false;
      double nv = 0.0;
      if (found) {
        val[lo] = val[lo] - t;
        if (KIND == AINV_KIND_INVK) lev[lo] = min(lev[lo], lk + 1);
      } else if (has) {
        nv = 0.0 - t;
        ins = (KIND == AINV_KIND_INVK) ? true : !(fabs(nv) < thr);
      }
      const uint64_t mask = __ballot(ins);
      const int32_t nins = (int32_t)__popcll(mask);
      if (nins == 0) {
        __syncthreads();  // in-place updates visible to the next step
        continue;
      }
      if (cnt + nins > cap) {
        if (lane == 0) {
          status[i] = ST_CAP;
          row_cnt[i] = 0;
        }
        return;
      }
      const int32_t rank = (int32_t)__popcll(mask & below);
      if (ins) insv[rank] = lo;
      __syncthreads();
      const int32_t first = insv[0];
      // shift the tail up, top chunk first: entry q moves to q + #{ins <= q}
      for (int32_t top = cnt; top > first; top -= kWave) {
        const int32_t q = top - kWave + lane;
        const bool mv = q >= first;
        int32_t kq = 0, lq = 0, s = 0;
        double vq = 0.0;
        if (mv) {
          kq = key[q];
          vq = val[q];
          lq = lev[q];
          int32_t a = 0, b = nins;
          while (a < b) {
            int32_t mid = (a + b) >> 1;
            if (insv[mid] <= q)
              a = mid + 1;
            else
              b = mid;
          }
          s = a;
        }
        __syncthreads();
        if (mv) {
          key[q + s] = kq;
          val[q + s] = vq;
          lev[q + s] = lq;
        }
        __syncthreads();
      }
      if (ins) {
        const int32_t dst = lo + rank;
        key[dst] = j;
        val[dst] = nv;
        lev[dst] = (KIND == AINV_KIND_INVK) ? lk + 1 : 1;
      }
      cnt += nins;
      __syncthreads();
    }
  }
  __syncthreads();
  // ---- copy-out ----
  // candidates: INVK level <= fill (copyout.f90:69-90); INVT kept entries,
  // i.e. the diagonal and every popped entry that passed the threshold
  // (copyout.f90:216-224)
  int32_t ncand = 0;
  for (int32_t q0 = 0; q0 < cnt; q0 += kWave) {
    const int32_t q = q0 + lane;
    bool c = false;
    if (q < cnt)
      c = (KIND == AINV_KIND_INVK) ? (lev[q] <= fill) : (lev[q] == 1);
    ncand += (int32_t)__popcll(__ballot(c));
  }
  int32_t keep_n = ncand;
  const int64_t capk = (int64_t)m + fill;  // INVT: nup + fill_in
  const bool select = (KIND == AINV_KIND_INVT) && ((int64_t)ncand > capk);
  if (select) {
    // keep the capk largest (|v| desc, column asc); mark losers with lev 2
    keep_n = (int32_t)(capk < 0 ? 0 : capk);
    for (int32_t q0 = 0; q0 < cnt; q0 += kWave) {
      const int32_t q = q0 + lane;
      bool lose = false;
      if (q < cnt && lev[q] == 1) {
        const double vq = val[q];
        const int32_t kq = key[q];
        int64_t rnk = 0;
        for (int32_t r = 0; r < cnt; ++r)
          if (lev[r] != 0 && r != q && before(val[r], key[r], vq, kq)) ++rnk;
        lose = rnk >= capk;
      }
      __syncthreads();
      if (lose) lev[q] = 2;
      __syncthreads();
    }
  }
  if (lane == 0) {
    unsigned long long b = atomicAdd(pool_top, (unsigned long long)keep_n);
    sh->base = b;
    if ((int64_t)(b + keep_n) > pool_cap) {
      sh->flag = ST_POOL;
    }
  }
  __syncthreads();
  if (sh->flag != ST_OK) {
    if (lane == 0) {
      status[i] = sh->flag;
      row_cnt[i] = 0;
    }
    return;
  }
  const int64_t base = (int64_t)sh->base;
  int32_t written = 0;
  for (int32_t q0 = 0; q0 < cnt; q0 += kWave) {
    const int32_t q = q0 + lane;
    bool c = false;
    if (q < cnt)
      c = (KIND == AINV_KIND_INVK) ? (lev[q] <= fill) : (lev[q] == 1);
    const uint64_t mk = __ballot(c);
    if (c) {
      const int64_t dst = base + written + (int32_t)__popcll(mk & below);
      out_idx[dst] = key[q];
      out_val[dst] = val[q];
    }
    written += (int32_t)__popcll(mk);
  }
  if (lane == 0) {
    row_start[i] = base;
    row_cnt[i] = keep_n;
    status[i] = ST_OK;
  }
}

template <int KIND>
__global__ __launch_bounds__(kWave) void k_inverse_lds(
    int32_t n, const int64_t *__restrict__ uptr,
    const int32_t *__restrict__ uidx, const double *__restrict__ uval,
    int32_t fill, double thresh, int32_t *__restrict__ out_idx,
    double *__restrict__ out_val, int64_t pool_cap,
    unsigned long long *__restrict__ pool_top, int64_t *__restrict__ row_start,
    int32_t *__restrict__ row_cnt, int32_t *__restrict__ status) {
  __shared__ int32_t key[kCapLds];
  __shared__ double val[kCapLds];
  __shared__ int32_t lev[kCapLds];
  __shared__ int32_t insv[kWave];
  __shared__ Shared sh;
  const int32_t i = blockIdx.x;
  if (i >= n - 1) return;
  row_inverse<KIND>(i, n, uptr, uidx, uval, fill, thresh, key, val, lev, insv,
                    &sh, kCapLds, out_idx, out_val, pool_cap, pool_top,
                    row_start, row_cnt, status);
}

template <int KIND>
__global__ __launch_bounds__(kWave) void k_inverse_global(
    int32_t n, const int32_t *__restrict__ rows, int32_t nrows,
    const int64_t *__restrict__ uptr, const int32_t *__restrict__ uidx,
    const double *__restrict__ uval, int32_t fill, double thresh,
    int32_t *__restrict__ skey, double *__restrict__ sval,
    int32_t *__restrict__ slev, int32_t *__restrict__ out_idx,
    double *__restrict__ out_val, int64_t pool_cap,
    unsigned long long *__restrict__ pool_top, int64_t *__restrict__ row_start,
    int32_t *__restrict__ row_cnt, int32_t *__restrict__ status) {
  __shared__ int32_t insv[kWave];
  __shared__ Shared sh;
  if ((int32_t)blockIdx.x >= nrows) return;
  const int32_t i = rows[blockIdx.x];
  const int64_t off = (int64_t)blockIdx.x * kCapGlobal;
  row_inverse<KIND>(i, n, uptr, uidx, uval, fill, thresh, skey + off,
                    sval + off, slev + off, insv, &sh, kCapGlobal, out_idx,
                    out_val, pool_cap, pool_top, row_start, row_cnt, status);
}

// last row of the inverse is e_n (invk.f90:120-124, invt.f90:117-121)
__global__ void k_last_row(int32_t n, int32_t *__restrict__ out_idx,
                           double *__restrict__ out_val, int64_t pool_cap,
                           unsigned long long *__restrict__ pool_top,
                           int64_t *__restrict__ row_start,
                           int32_t *__restrict__ row_cnt,
                           int32_t *__restrict__ status) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || n < 1) return;
  unsigned long long b = atomicAdd(pool_top, 1ull);
  if ((int64_t)b >= pool_cap) {
    status[n - 1] = ST_POOL;
    row_cnt[n - 1] = 0;
    return;
  }
  out_idx[b] = n - 1;
  out_val[b] = 1.0;
  row_start[n - 1] = (int64_t)b;
  row_cnt[n - 1] = 1;
  status[n - 1] = ST_OK;
}

__global__ void k_gather_rows(int32_t n, const int64_t *__restrict__ ptr,
                              const int64_t *__restrict__ row_start,
                              const int32_t *__restrict__ pidx,
                              const double *__restrict__ pval,
                              int32_t *__restrict__ idx,
                              double *__restrict__ val) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t b = ptr[r], e = ptr[r + 1], s = row_start[r];
  for (int64_t t = b; t < e; ++t) {
    idx[t] = pidx[s + (t - b)];
    val[t] = pval[s + (t - b)];
  }
}

__global__ void k_count64(int32_t n, const int32_t *__restrict__ cnt,
                          int64_t *__restrict__ c64) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) c64[r] = cnt[r];
  if (r == n) c64[r] = 0;
}

}  // namespace

void gpu_sparse_inverse(int kind, const DevCsr &u, int32_t fill, double thresh,
                        DevCsr &z, hipStream_t s, std::string *report) {
  const int32_t n = u.n;
  z = DevCsr();
  z.n = n;
  if (n == 0) {
    z.ptr.alloc(sizeof(int64_t));
    AINV_HIP_CHECK(hipMemsetAsync(z.ptr.p, 0, sizeof(int64_t), s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  DevBuf row_start, row_cnt, status, top;
  row_start.alloc((size_t)n * sizeof(int64_t));
  row_cnt.alloc((size_t)n * sizeof(int32_t));
  status.alloc((size_t)n * sizeof(int32_t));
  top.alloc(sizeof(unsigned long long));
  // pool sized from the input; grown and recomputed if exhausted
  int64_t pool_cap = std::max<int64_t>(4 * u.nnz + 4 * (int64_t)n, 1 << 20);
  DevBuf pidx, pval;
  int lds_overflow_rows = 0, attempts = 0;
  for (;;) {
    pidx.alloc((size_t)pool_cap * sizeof(int32_t));
    pval.alloc((size_t)pool_cap * sizeof(double));
    AINV_HIP_CHECK(hipMemsetAsync(top.p, 0, top.bytes, s));
    if (n > 1) {
      if (kind == AINV_KIND_INVK)
        hipLaunchKernelGGL(k_inverse_lds<AINV_KIND_INVK>, dim3(n - 1),
                           dim3(kWave), 0, s, n, u.ptr.as<int64_t>(),
                           u.idx.as<int32_t>(), u.val.as<double>(), fill,
                           thresh, pidx.as<int32_t>(), pval.as<double>(),
                           pool_cap, top.as<unsigned long long>(),
                           row_start.as<int64_t>(), row_cnt.as<int32_t>(),
                           status.as<int32_t>());
      else
        hipLaunchKernelGGL(k_inverse_lds<AINV_KIND_INVT>, dim3(n - 1),
                           dim3(kWave), 0, s, n, u.ptr.as<int64_t>(),
                           u.idx.as<int32_t>(), u.val.as<double>(), fill,
                           thresh, pidx.as<int32_t>(), pval.as<double>(),
                           pool_cap, top.as<unsigned long long>(),
                           row_start.as<int64_t>(), row_cnt.as<int32_t>(),
                           status.as<int32_t>());
      AINV_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_last_row, dim3(1), dim3(64), 0, s, n,
                       pidx.as<int32_t>(), pval.as<double>(), pool_cap,
                       top.as<unsigned long long>(), row_start.as<int64_t>(),
                       row_cnt.as<int32_t>(), status.as<int32_t>());
    AINV_HIP_CHECK(hipGetLastError());
    std::vector<int32_t> hst(n);
    AINV_HIP_CHECK(hipMemcpyAsync(hst.data(), status.p, status.bytes,
                                  hipMemcpyDeviceToHost, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> over;
    bool pool_full = false;
    for (int32_t r = 0; r < n; ++r) {
      if (hst[r] == ST_BADU)
        throw Error(AINV_ERR_INTERNAL,
                    "row inverse: input is not strictly upper triangular at "
                    "row " + std::to_string(r + 1));
      if (hst[r] == ST_POOL) pool_full = true;
      if (hst[r] == ST_CAP) over.push_back(r);
    }
    // rows whose pending set outgrew LDS: global-memory tier, in batches
    if (!over.empty() && !pool_full) {
      lds_overflow_rows = (int)over.size();
      DevBuf rows, skey, sval, slev;
      rows.alloc(over.size() * sizeof(int32_t));
      AINV_HIP_CHECK(hipMemcpyAsync(rows.p, over.data(), rows.bytes,
                                    hipMemcpyHostToDevice, s));
      const size_t slab = (size_t)kGlobalRowsPerLaunch * kCapGlobal;
      skey.alloc(slab * sizeof(int32_t));
      sval.alloc(slab * sizeof(double));
      slev.alloc(slab * sizeof(int32_t));
      for (size_t b0 = 0; b0 < over.size(); b0 += kGlobalRowsPerLaunch) {
        int32_t nb = (int32_t)std::min<size_t>(kGlobalRowsPerLaunch,
                                               over.size() - b0);
        const int32_t *rp = rows.as<int32_t>() + b0;
        if (kind == AINV_KIND_INVK)
          hipLaunchKernelGGL(k_inverse_global<AINV_KIND_INVK>, dim3(nb),
                             dim3(kWave), 0, s, n, rp, nb, u.ptr.as<int64_t>(),
                             u.idx.as<int32_t>(), u.val.as<double>(), fill,
                             thresh, skey.as<int32_t>(), sval.as<double>(),
                             slev.as<int32_t>(), pidx.as<int32_t>(),
                             pval.as<double>(), pool_cap,
                             top.as<unsigned long long>(),
                             row_start.as<int64_t>(), row_cnt.as<int32_t>(),
                             status.as<int32_t>());
        else
          hipLaunchKernelGGL(k_inverse_global<AINV_KIND_INVT>, dim3(nb),
                             dim3(kWave), 0, s, n, rp, nb, u.ptr.as<int64_t>(),
                             u.idx.as<int32_t>(), u.val.as<double>(), fill,
                             thresh, skey.as<int32_t>(), sval.as<double>(),
                             slev.as<int32_t>(), pidx.as<int32_t>(),
                             pval.as<double>(), pool_cap,
                             top.as<unsigned long long>(),
                             row_start.as<int64_t>(), row_cnt.as<int32_t>(),
                             status.as<int32_t>());
        AINV_HIP_CHECK(hipGetLastError());
      }
      AINV_HIP_CHECK(hipMemcpyAsync(hst.data(), status.p, status.bytes,
                                    hipMemcpyDeviceToHost, s));
      AINV_HIP_CHECK(hipStreamSynchronize(s));
      for (int32_t r : over) {
        if (hst[r] == ST_POOL) pool_full = true;
        if (hst[r] == ST_CAP)
          throw Error(AINV_ERR_CAPACITY,
                      "row inverse: row " + std::to_string(r + 1) +
                          " needs more than " + std::to_string(kCapGlobal) +
                          " pending entries");
        if (hst[r] == ST_BADU)
          throw Error(AINV_ERR_INTERNAL, "row inverse: bad input row");
      }
    }
    if (!pool_full) break;
    unsigned long long used = 0;
    AINV_HIP_CHECK(hipMemcpy(&used, top.p, sizeof(used), hipMemcpyDeviceToHost));
    pool_cap = std::max<int64_t>(2 * pool_cap, (int64_t)used + n);
    if (++attempts > 4)
      throw Error(AINV_ERR_CAPACITY, "row inverse: output pool keeps overflowing");
  }
  // CSR assembly: counts -> scan -> gather per row
  DevBuf c64;
  c64.alloc((size_t)(n + 1) * sizeof(int64_t));
  z.ptr.alloc((size_t)(n + 1) * sizeof(int64_t));
  hipLaunchKernelGGL(k_count64, dim3((n + 256) / 256), dim3(256), 0, s, n,
                     row_cnt.as<int32_t>(), c64.as<int64_t>());
  AINV_HIP_CHECK(hipGetLastError());
  size_t tmp_bytes = 0;
  AINV_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
      nullptr, tmp_bytes, c64.as<int64_t>(), z.ptr.as<int64_t>(), n + 1, s));
  DevBuf tmp;
  tmp.alloc(tmp_bytes ? tmp_bytes : 16);
  AINV_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
      tmp.p, tmp_bytes, c64.as<int64_t>(), z.ptr.as<int64_t>(), n + 1, s));
  AINV_HIP_CHECK(hipMemcpyAsync(&z.nnz, z.ptr.as<int64_t>() + n,
                                sizeof(int64_t), hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  z.idx.alloc((size_t)z.nnz * sizeof(int32_t));
  z.val.alloc((size_t)z.nnz * sizeof(double));
  hipLaunchKernelGGL(k_gather_rows, dim3((n + 255) / 256), dim3(256), 0, s, n,
                     z.ptr.as<int64_t>(), row_start.as<int64_t>(),
                     pidx.as<int32_t>(), pval.as<double>(), z.idx.as<int32_t>(),
                     z.val.as<double>());
  AINV_HIP_CHECK(hipGetLastError());
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  if (report) {
    *report += "{\"rows\": " + std::to_string(n) +
               ", \"nnz\": " + std::to_string(z.nnz) +
               ", \"lds_overflow_rows\": " + std::to_string(lds_overflow_rows) +
               ", \"pool_cap\": " + std::to_string(pool_cap) + "}";
  }
}

}  // namespace ainv
