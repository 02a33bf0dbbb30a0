// Multilevel preconditioner of kind AINV_KIND_ML: the configuration the
// reference's deck tests/pdegen/runs/ppde.inp:10-29 asks MLD2P4 for (ML, AS
// smoother with an AINV-family subdomain solver, SMOOTHED DEC aggregation,
// MULT / POST V-cycle), on one process.  MLD2P4 (>= 2.1.1,
// configure.ac:283-289) is not part of the reference tree; the steps follow
// its published algorithms:
//  - greedy three-phase aggregation (Vanek, Mandel, Brezina, Computing 56
//    (1996) 179-196) over the strength graph |a_ij| > theta sqrt(|a_ii a_jj|);
//  - prolongator P = (I - omega D^-1 A) T, T the tentative (piecewise
//    constant) prolongator, omega = 4 / (3 ||D^-1 A||_inf);
//  - Galerkin coarse matrix R (A P), R = P^T.
//
// Host: aggregation, prolongator and Galerkin products (Gustavson row
// products in stored order over threads, so the hierarchy is the same for
// any thread count), then the factors of every level through the ordinary
// solver build.  Device: every level's A, P and R as FP64 HLL; a cycle is a
// fixed sequence of HLL products on the caller's stream (the residual
// b - A u fused into the product's epilogue), with no allocation, so it is
// captured together with the Krylov iteration.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <system_error>
#include <thread>

#include "objects.h"

namespace ainv {

struct MlLevel {
  int32_t n = 0;   // rows of the level matrix
  int32_t nc = 0;  // aggregates (rows of the next level); 0 on the coarsest
  double theta = 0.0, omega = 0.0;
  HostCsr a;            // level matrix
  HostCsr p, r;         // prolongator (n x nc) and restriction (nc x n)
  std::vector<int32_t> agg;
  Hll ha, hp, hr;       // the same on the device, FP64
  std::unique_ptr<ainv_solver> sv;  // smoother / coarsest-level factors
  int8_t map_a[8], map_p[8], map_r[8];  // tuned mapping per K (-1: none)
  MlLevel() {
    for (int i = 0; i < 8; ++i) map_a[i] = map_p[i] = map_r[i] = -1;
  }
  ~MlLevel() {
    if (sv) sv->stream = nullptr;  // the stream belongs to the ML handle
  }
};

struct MlHier {
  std::vector<std::unique_ptr<MlLevel>> lv;
  int mprec = AINV_PREC_F64;  // storage of the level A, P, R: F64 or F32
  double t_aggr = 0, t_galerkin = 0, t_factors = 0, t_upload = 0;
};

void destroy_ml(MlHier *m) { delete m; }

namespace {

constexpr int kMaxLevels = 16;

int threads_of(const ainv_solver *h) {
  return h->host_threads > 0 ? h->host_threads : build_threads();
}

// f(lo, hi) over nt contiguous row ranges of [0, n), one thread each
void parallel_rows(int32_t n, int nt,
                   const std::function<void(int32_t, int32_t)> &f) {
  nt = std::max(1, std::min<int>(nt, n / 4096 + 1));
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  std::exception_ptr err;
  std::mutex emu;
  for (int t = 0; t < nt; ++t) {
    const int32_t lo = (int32_t)((int64_t)n * t / nt);
    const int32_t hi = (int32_t)((int64_t)n * (t + 1) / nt);
    auto range = [&, lo, hi] {
      try {
        f(lo, hi);
      } catch (...) {
        std::lock_guard<std::mutex> lk(emu);
        if (!err) err = std::current_exception();
      }
    };
    // a range whose thread cannot be started runs on this thread
    try {
      th.emplace_back(range);
    } catch (const std::system_error &) {
      range();
    }
  }
  for (auto &x : th) x.join();
  if (err) std::rethrow_exception(err);
}

std::vector<double> diag_of(const HostCsr &a) {
  std::vector<double> dg(a.n, 0.0);
  for (int32_t i = 0; i < a.n; ++i)
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
      if (a.idx[k] == i) dg[i] = a.val[k];
  return dg;
}

// Aggregates of the rows of a (phase 1: a row whose strong neighbours are
// all free roots an aggregate of itself and them; phase 2: a free row joins
// the phase-1 aggregate of its most strongly coupled neighbour, first in
// column order on ties; phase 3: the remaining rows form aggregates with
// their free strong neighbours).
void aggregate(const HostCsr &a, double theta, int nt,
               std::vector<int32_t> &agg, int32_t &nagg) {
  const int32_t n = a.n;
  const std::vector<double> dg = diag_of(a);
  std::vector<uint8_t> st((size_t)a.nnz());
  parallel_rows(n, nt, [&](int32_t lo, int32_t hi) {
    for (int32_t i = lo; i < hi; ++i)
      for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
        const int32_t j = a.idx[k];
        st[k] = j != i &&
                std::fabs(a.val[k]) > theta * std::sqrt(std::fabs(dg[i] * dg[j]));
      }
  });
  agg.assign(n, -1);
  int32_t na = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool ok = true;
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1] && ok; ++k)
      if (st[k] && agg[a.idx[k]] >= 0) ok = false;
    if (!ok) continue;
    agg[i] = na;
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
      if (st[k]) agg[a.idx[k]] = na;
    ++na;
  }
  const std::vector<int32_t> a1 = agg;
  for (int32_t i = 0; i < n; ++i) {
    if (a1[i] >= 0) continue;
    double bv = -1.0;
    int32_t best = -1;
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
      if (st[k] && a1[a.idx[k]] >= 0 && std::fabs(a.val[k]) > bv) {
        bv = std::fabs(a.val[k]);
        best = a1[a.idx[k]];
      }
    if (best >= 0) agg[i] = best;
  }
  for (int32_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
      if (st[k] && agg[a.idx[k]] < 0) agg[a.idx[k]] = na;
    ++na;
  }
  nagg = na;
}

// C = A B, B with bcols columns: row i accumulates a_ik * B(k, :) over the
// entries of row i of A in stored order, every sum starting from 0.0; the
// pattern is the symbolic product (cancelled entries stay).  Two passes over
// the same row ranges: count the columns of each row, then fill the rows in
// place.
void spgemm(const HostCsr &a, const HostCsr &b, int32_t bcols, HostCsr &c,
            int nt) {
  const int32_t n = a.n;
  c.n = n;
  c.ptr.assign((size_t)n + 1, 0);
  parallel_rows(n, nt, [&](int32_t lo, int32_t hi) {
    std::vector<int32_t> mark(std::max(bcols, 1), -1);
    for (int32_t i = lo; i < hi; ++i) {
      int64_t cnt = 0;
      for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
        const int32_t rr = a.idx[k];
        for (int64_t q = b.ptr[rr]; q < b.ptr[rr + 1]; ++q)
          if (mark[b.idx[q]] != i) {
            mark[b.idx[q]] = i;
            ++cnt;
          }
      }
      c.ptr[i + 1] = cnt;
    }
  });
  for (int32_t i = 0; i < n; ++i) c.ptr[i + 1] += c.ptr[i];
  c.idx.resize(c.ptr[n]);
  c.val.resize(c.ptr[n]);
  parallel_rows(n, nt, [&](int32_t lo, int32_t hi) {
    std::vector<double> acc(std::max(bcols, 1));
    std::vector<int32_t> mark(std::max(bcols, 1), -1);
    for (int32_t i = lo; i < hi; ++i) {
      int32_t *list = c.idx.data() + c.ptr[i];
      int64_t cnt = 0;
      for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
        const double aik = a.val[k];
        const int32_t rr = a.idx[k];
        for (int64_t q = b.ptr[rr]; q < b.ptr[rr + 1]; ++q) {
          const int32_t j = b.idx[q];
          if (mark[j] != i) {
            mark[j] = i;
            list[cnt++] = j;
            acc[j] = 0.0;
          }
          acc[j] = acc[j] + aik * b.val[q];
        }
      }
      std::sort(list, list + cnt);
      double *v = c.val.data() + c.ptr[i];
      for (int64_t t = 0; t < cnt; ++t) v[t] = acc[list[t]];
    }
  });
}

// transpose of an a.n x ncols matrix (rows of the result sorted)
void transpose_rect(const HostCsr &a, int32_t ncols, HostCsr &t) {
  t.n = ncols;
  t.ptr.assign((size_t)ncols + 1, 0);
  for (int64_t k = 0; k < a.nnz(); ++k) ++t.ptr[a.idx[k] + 1];
  for (int32_t j = 0; j < ncols; ++j) t.ptr[j + 1] += t.ptr[j];
  t.idx.resize(a.nnz());
  t.val.resize(a.nnz());
  std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
  for (int32_t i = 0; i < a.n; ++i)
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
      const int64_t q = pos[a.idx[k]]++;
      t.idx[q] = i;
      t.val[q] = a.val[k];
    }
}

// P = T (tentative) or (I - omega D^-1 A) T (smoothed)
void prolongator(const HostCsr &a, const std::vector<int32_t> &agg,
                 int32_t nc, bool smoothed, int nt, HostCsr &p,
                 double &omega) {
  const int32_t n = a.n;
  HostCsr t;
  t.n = n;
  t.ptr.resize((size_t)n + 1);
  for (int32_t i = 0; i <= n; ++i) t.ptr[i] = i;
  t.idx = agg;
  t.val.assign(n, 1.0);
  omega = 0.0;
  if (!smoothed) {
    p = std::move(t);
    return;
  }
  const std::vector<double> dg = diag_of(a);
  double rho = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    if (dg[i] == 0.0)
      throw Error(AINV_ERR_PIVOT,
                  "zero diagonal entry in row " + std::to_string(i) +
                      ": the smoothed prolongator needs D^-1");
    double s = 0.0;
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k)
      s = s + std::fabs(a.val[k]);
    s = s / std::fabs(dg[i]);
    if (s > rho) rho = s;
  }
  const double om = 4.0 / (3.0 * rho);
  omega = om;
  spgemm(a, t, nc, p, nt);
  parallel_rows(n, nt, [&](int32_t lo, int32_t hi) {
    for (int32_t i = lo; i < hi; ++i) {
      const double sc = om / dg[i];
      for (int64_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k) {
        const double tv = (p.idx[k] == agg[i]) ? 1.0 : 0.0;
        p.val[k] = tv - sc * p.val[k];
      }
    }
  });
}

// Rows sorted by length within windows where that saves padding (the
// prolongators' rows hold 1 to ~30 entries: 1.4-1.5x padding in row order).
void upload_hll(const HostCsr &m, int32_t ncols, int prec, Hll &h,
                hipStream_t s) {
  DevCsr d;
  csr_upload(m, d, s);
  hll_from_csr(d, prec, h, s, ncols, kSortAuto);
}

// A matrix whose rows average 16 entries or more also gets its CSR copy,
// for the one-wavefront-per-row mapping.
void upload_hll_wide(const HostCsr &m, int32_t ncols, int prec, Hll &h,
                     hipStream_t s) {
  upload_hll(m, ncols, prec, h, s);
  if (m.n > 0 && m.nnz() >= 16 * (int64_t)m.n) hll_attach_csr(m, h, s);
}

// A, P and R of every level in FP64 (FP64 factors) or FP32 (FP32 / BF16
// factors: the cycle may then run on FP32 vectors)
void upload_levels(MlHier &H, int factor_prec, hipStream_t s) {
  const int mp = factor_prec == AINV_PREC_F64 ? AINV_PREC_F64 : AINV_PREC_F32;
  for (auto &Lp : H.lv) {
    MlLevel &L = *Lp;
    upload_hll_wide(L.a, L.n, mp, L.ha, s);
    if (L.nc > 0) {
      upload_hll_wide(L.p, L.nc, mp, L.hp, s);
      upload_hll_wide(L.r, L.n, mp, L.hr, s);
    }
    for (int i = 0; i < 8; ++i) L.map_a[i] = L.map_p[i] = L.map_r[i] = -1;
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  H.mprec = mp;
}

int64_t hll_bytes(const Hll &h) { return h.bytes(); }

// ---- the cycle --------------------------------------------------------------

int64_t vec_bytes(int32_t n, int64_t K, int64_t elem = sizeof(double)) {
  return ((int64_t)n * K * elem + 255) / 256 * 256;
}

// the cycle's vectors of each level (V: double or float) and the smoother's
// intermediate, carved from one workspace (FP64 sizes cover FP32 ones)
template <typename V> struct Ws {
  V *b[kMaxLevels] = {}, *u[kMaxLevels] = {}, *t[kMaxLevels] = {};
  void *tb = nullptr;
};

template <typename V> Ws<V> carve(const MlHier &H, int64_t K, void *base) {
  Ws<V> w;
  char *p = static_cast<char *>(base);
  for (size_t l = 0; l < H.lv.size(); ++l) {
    const int32_t n = H.lv[l]->n;
    if (l > 0) {
      w.b[l] = reinterpret_cast<V *>(p);
      p += vec_bytes(n, K, sizeof(V));
      w.u[l] = reinterpret_cast<V *>(p);
      p += vec_bytes(n, K, sizeof(V));
    }
    w.t[l] = reinterpret_cast<V *>(p);
    p += vec_bytes(n, K, sizeof(V));
  }
  w.tb = p;
  return w;
}

// mapping of a level product: forced (APPLY_MAP; WIDE where the matrix has
// its CSR copy, COOP elsewhere), else tuned, else split
int pick(const ainv_solver *h, const Hll &m, const int8_t *tuned, int K) {
  if (h->map_mode == AINV_MAP_WIDE)
    return m.cptr.p ? AINV_MAP_WIDE : AINV_MAP_COOP;
  if (h->map_mode >= 0) return h->map_mode;
  const int t = tuned[log2_pow2((int)K)];
  return t < 0 ? AINV_MAP_SPLIT : t;
}

// One cycle on vectors of type V (double, or float with FP32 / BF16 factors
// and FP32 level matrices).
template <typename V> struct Cycle {
  static constexpr bool kF32 = sizeof(V) == sizeof(float);
  const ainv_solver *h;
  const MlHier *H;
  int k;
  int64_t K;
  hipStream_t s;
  const int32_t *kw;
  Ws<V> w;

  // y = M_l x + beta y
  void smooth(int l, const V *x, V *y, double beta) const {
    const ainv_solver *sv = H->lv[l]->sv.get();
    const int mp = tuned_apply_map(sv, (int)K);
    const bool f32t = sv->prec != AINV_PREC_F64;
    hll_spmm_scale(sv->hw, sv->dv.p, x, kF32, K, w.tb, K, k, s, map_w(mp),
                   kw);
    hll_spmm_axpby(sv->hz, 1.0, w.tb, f32t, K, beta, y, K, k, s, map_z(mp),
                   kw);
  }
  // y = M x + beta y for M = R_l or P_l
  void prod(const Hll &m, const int8_t *maps, const V *x, V *y,
            double beta) const {
    hll_spmm_axpby(m, 1.0, x, kF32, K, beta, y, K, k, s, pick(h, m, maps, K),
                   kw);
  }
  // t = b - A_l u
  void resid(int l, const V *b, const V *u, V *t) const {
    const MlLevel &L = *H->lv[l];
    hll_spmm_resid(L.ha, -1.0, u, 1.0, b, t, K, k, s,
                   pick(h, L.ha, L.map_a, K), kw);
  }
  // u = M_l b, then nsw - 1 sweeps u += M_l (b - A_l u)
  void smooth0(int l, int nsw, const V *b, V *u) const {
    smooth(l, b, u, 0.0);
    for (int i = 1; i < nsw; ++i) {
      resid(l, b, u, w.t[l]);
      smooth(l, w.t[l], u, 1.0);
    }
  }
  void run(int l, const V *b, V *u) const {
    const int nl = (int)H->lv.size();
    if (l == nl - 1) {
      smooth0(l, h->ml_csweeps, b, u);
      return;
    }
    const MlLevel &L = *H->lv[l];
    V *bc = w.b[l + 1], *uc = w.u[l + 1];
    if (h->ml_type == AINV_ML_ADD) {
      smooth0(l, h->ml_sweeps, b, u);
      prod(L.hr, L.map_r, b, bc, 0.0);
      run(l + 1, bc, uc);
      prod(L.hp, L.map_p, uc, u, 1.0);
      return;
    }
    const bool pre = h->ml_pos & AINV_ML_PRE, post = h->ml_pos & AINV_ML_POST;
    if (pre) {
      smooth0(l, h->ml_sweeps, b, u);
      resid(l, b, u, w.t[l]);
      prod(L.hr, L.map_r, w.t[l], bc, 0.0);
    } else {
      prod(L.hr, L.map_r, b, bc, 0.0);
    }
    run(l + 1, bc, uc);
    prod(L.hp, L.map_p, uc, u, pre ? 1.0 : 0.0);
    if (post)
      for (int i = 0; i < h->ml_sweeps; ++i) {
        resid(l, b, u, w.t[l]);
        smooth(l, w.t[l], u, 1.0);
      }
  }
};

// the factors have no CSR copy: a forced WIDE runs them cooperatively
int smoother_map(int mode) {
  return mode == AINV_MAP_WIDE ? AINV_MAP_COOP : mode;
}

const MlHier &hier(const ainv_solver *h) {
  if (h->kind != AINV_KIND_ML || !h->ml)
    throw Error(AINV_ERR_STATE, "multilevel preconditioner not built");
  return *h->ml;
}

double time_ms(hipStream_t s, const std::function<void()> &f, int reps) {
  hipEvent_t e0, e1;
  AINV_HIP_CHECK(hipEventCreate(&e0));
  AINV_HIP_CHECK(hipEventCreate(&e1));
  f();
  AINV_HIP_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) f();
  AINV_HIP_CHECK(hipEventRecord(e1, s));
  AINV_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  AINV_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms;
}

// nmaps: kNumApplyMaps, or kNumMaps for a matrix with its CSR copy
int best_map(hipStream_t s, const std::function<void(int)> &f,
             int nmaps = kNumApplyMaps) {
  int best = 0;
  double bt = 0.0;
  for (int map = 0; map < nmaps; ++map) {
    const double t = time_ms(s, [&] { f(map); }, 3);
    if (map == 0 || t < bt) {
      bt = t;
      best = map;
    }
  }
  return best;
}

// y[r, c] = alpha * src[r, c] + beta * y[r, c] (beta == 0: y not read)
__global__ void k_combine(int32_t n, int k, double alpha,
                          const double *__restrict__ src, int64_t lds,
                          double beta, double *__restrict__ y, int64_t ldy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * k) return;
  const int64_t r = t / k;
  const int c = (int)(t % k);
  double v = alpha * src[r * lds + c];
  if (beta != 0.0) v += beta * y[r * ldy + c];
  y[r * ldy + c] = v;
}

}  // namespace

// ---- build -------------------------------------------------------------------

void ml_host_level(const HostCsr &a, double theta, bool smoothed, int threads,
                   std::vector<int32_t> &agg, int32_t &nagg, HostCsr &p,
                   HostCsr &ac, double &omega) {
  const int nt = threads > 0 ? threads : build_threads();
  aggregate(a, theta, nt, agg, nagg);
  prolongator(a, agg, nagg, smoothed, nt, p, omega);
  HostCsr r, ap;
  transpose_rect(p, nagg, r);
  spgemm(a, p, nagg, ap, nt);
  spgemm(r, ap, nagg, ac, nt);
}

void ml_build(ainv_solver *h, const HostCsr &a) {
  std::unique_ptr<MlHier, MlDeleter> H(new MlHier());
  const int nt = threads_of(h);
  if (h->ml_nlev < 1 || h->ml_nlev > kMaxLevels)
    throw Error(AINV_ERR_ARG, "N_PREC_LEVS must be in 1.." +
                                  std::to_string(kMaxLevels));
  HostCsr cur = a;
  double theta = h->ml_thresh;
  for (int l = 0;; ++l) {
    std::unique_ptr<MlLevel> L(new MlLevel());
    L->n = cur.n;
    L->theta = theta;
    bool coarsest = l + 1 >= h->ml_nlev || cur.n <= 1;
    HostCsr next;
    if (!coarsest) {
      double t0 = now_seconds();
      aggregate(cur, theta, nt, L->agg, L->nc);
      H->t_aggr += now_seconds() - t0;
      if (L->nc >= cur.n || L->nc < 1) {
        coarsest = true;  // no reduction: this level is the coarsest
        L->agg.clear();
        L->nc = 0;
      } else {
        t0 = now_seconds();
        prolongator(cur, L->agg, L->nc, h->ml_smoothed != 0, nt, L->p,
                    L->omega);
        transpose_rect(L->p, L->nc, L->r);
        HostCsr ap;
        spgemm(cur, L->p, L->nc, ap, nt);
        spgemm(L->r, ap, L->nc, next, nt);
        H->t_galerkin += now_seconds() - t0;
      }
    }
    L->a = std::move(cur);
    H->lv.push_back(std::move(L));
    if (coarsest) break;
    cur = std::move(next);
    theta *= h->ml_scale;
  }
  const int nl = (int)H->lv.size();
  double t0 = now_seconds();
  // the levels' factors are independent: built concurrently, one host
  // thread and one HIP stream each, the host threads split between them
  for (int l = 0; l < nl; ++l) {
    MlLevel &L = *H->lv[l];
    L.sv.reset(new ainv_solver());
    ainv_solver *sv = L.sv.get();
    sv->kind = l < nl - 1 ? h->ml_sub : h->ml_coarse;
    sv->alg = h->alg;
    sv->fill_in = h->fill_in;
    sv->inv_fill = h->inv_fill;
    sv->thresh = h->thresh;
    sv->inv_thresh = h->inv_thresh;
    sv->prec = h->prec;
    sv->map_mode = smoother_map(h->map_mode);
    sv->host_threads = std::max(1, nt / nl);
    sv->device = h->device;
  }
  {
    std::vector<std::thread> th;
    th.reserve(nl);
    std::exception_ptr err;
    std::mutex emu;
    auto level_build = [&](int l) {
      ainv_solver *sv = H->lv[l]->sv.get();
      // run on the handle's device; a level built on the calling thread
      // leaves that thread's current device as it found it
      int prev = -1;
      try {
        AINV_HIP_CHECK(hipGetDevice(&prev));
        if (prev != h->device) AINV_HIP_CHECK(hipSetDevice(h->device));
        AINV_HIP_CHECK(hipStreamCreate(&sv->stream));
        build_solver(sv, H->lv[l]->a);
      } catch (...) {
        std::lock_guard<std::mutex> lk(emu);
        if (!err) err = std::current_exception();
      }
      // the level keeps the handle's stream from here on
      if (sv->stream && sv->stream != h->stream) {
        (void)hipStreamSynchronize(sv->stream);
        (void)hipStreamDestroy(sv->stream);
      }
      sv->stream = h->stream;
      if (prev >= 0 && prev != h->device) (void)hipSetDevice(prev);
    };
    // a level whose thread cannot be started is built on this thread
    for (int l = 0; l < nl; ++l) {
      try {
        th.emplace_back(level_build, l);
      } catch (const std::system_error &) {
        level_build(l);
      }
    }
    for (auto &x : th) x.join();
    if (err) std::rethrow_exception(err);
  }
  H->t_factors = now_seconds() - t0;
  t0 = now_seconds();
  upload_levels(*H, h->prec, h->stream);
  H->t_upload = now_seconds() - t0;
  h->n = a.n;
  h->ml = std::move(H);
}

// ---- apply -------------------------------------------------------------------

size_t ml_ws_bytes(const ainv_solver *h, int k) {
  const MlHier &H = hier(h);
  size_t b = 0, tb = 0;
  for (size_t l = 0; l < H.lv.size(); ++l) {
    b += (size_t)vec_bytes(H.lv[l]->n, k) * (l > 0 ? 3 : 1);
    tb = std::max(tb, solver_t_bytes(H.lv[l]->sv.get(), k));
  }
  return b + tb;
}

void ml_apply_into(const ainv_solver *h, const double *x, double *y, int k,
                   int64_t K, void *ws, hipStream_t s, const int32_t *kw) {
  const MlHier &H = hier(h);
  Cycle<double> c{h, &H, k, K, s, kw, carve<double>(H, K, ws)};
  c.run(0, x, y);
}

void ml_apply_into32(const ainv_solver *h, const float *x, float *y, int k,
                     int64_t K, void *ws, hipStream_t s, const int32_t *kw) {
  const MlHier &H = hier(h);
  if (H.mprec != AINV_PREC_F32)
    throw Error(AINV_ERR_INTERNAL, "FP32 cycle needs FP32 level matrices");
  Cycle<float> c{h, &H, k, K, s, kw, carve<float>(H, K, ws)};
  c.run(0, x, y);
}

// y = alpha * V-cycle(x) + beta * y for the C ABI: any leading dimensions
// (staged through K-strided buffers when they differ from K)
void ml_apply_abi(ainv_solver *h, double alpha, const double *x, double beta,
                  double *y, char trans, int k, int64_t ldx, int64_t ldy,
                  hipStream_t s) {
  const MlHier &H = hier(h);
  if (trans != 'N' && trans != 'n')
    throw Error(AINV_ERR_ARG,
                "the multilevel preconditioner applies with trans 'N' only");
  if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
  if (ldx < k || ldy < k) throw Error(AINV_ERR_ARG, "ldx and ldy must be >= k");
  if (!x || !y) throw Error(AINV_ERR_ARG, "NULL vector");
  int K = 1;
  while (K < k) K <<= 1;
  const int32_t n = H.lv[0]->n;
  const size_t vb = (size_t)vec_bytes(n, K);
  const size_t need = ml_ws_bytes(h, K) + 2 * vb;
  if (h->tbuf.bytes < need) {
    h->tbuf.alloc(need);
    AINV_HIP_CHECK(hipMemsetAsync(h->tbuf.p, 0, need, s));
  }
  char *base = h->tbuf.as<char>();
  double *xs = reinterpret_cast<double *>(base + ml_ws_bytes(h, K));
  double *ys = reinterpret_cast<double *>(base + ml_ws_bytes(h, K) + vb);
  const double *xin = x;
  if (ldx != K || (reinterpret_cast<uintptr_t>(x) & 15)) {
    AINV_HIP_CHECK(hipMemcpy2DAsync(xs, K * sizeof(double), x,
                                    ldx * sizeof(double), k * sizeof(double),
                                    n, hipMemcpyDeviceToDevice, s));
    xin = xs;
  }
  const bool direct = alpha == 1.0 && beta == 0.0 && ldy == K &&
                      !(reinterpret_cast<uintptr_t>(y) & 15);
  ml_apply_into(h, xin, direct ? y : ys, k, K, base, s);
  if (!direct && n > 0) {
    const int64_t tot = (int64_t)n * k;
    hipLaunchKernelGGL(k_combine, dim3((unsigned)((tot + 255) / 256)),
                       dim3(256), 0, s, n, k, alpha, ys, (int64_t)K, beta, y,
                       ldy);
    AINV_HIP_CHECK(hipGetLastError());
  }
}

template <typename V>
void tune_levels(ainv_solver *h, int K, int k, const V *x0, V *y0, void *ws,
                 hipStream_t s) {
  constexpr bool f32 = sizeof(V) == sizeof(float);
  MlHier &H = *h->ml;
  const int li = log2_pow2(K);
  AINV_HIP_CHECK(hipMemsetAsync(ws, 0, ml_ws_bytes(h, K), s));
  const Ws<V> w = carve<V>(H, K, ws);
  const int nl = (int)H.lv.size();
  for (int l = 0; l < nl; ++l) {
    MlLevel &L = *H.lv[l];
    // x is only read, y and the workspace only written
    const V *x = l > 0 ? w.b[l] : x0;
    V *y = l > 0 ? w.u[l] : y0;
    auto nm = [](const Hll &h) { return h.cptr.p ? kNumMaps : kNumApplyMaps; };
    if (L.map_a[li] < 0)
      L.map_a[li] = (int8_t)best_map(s, [&](int map) {
        hll_spmm_resid(L.ha, -1.0, y, 1.0, x, w.t[l], K, k, s, map);
      }, nm(L.ha));
    if (L.nc > 0) {
      if (L.map_r[li] < 0)
        L.map_r[li] = (int8_t)best_map(s, [&](int map) {
          hll_spmm_axpby(L.hr, 1.0, x, f32, K, 0.0, w.b[l + 1], K, k, s, map);
        }, nm(L.hr));
      if (L.map_p[li] < 0)
        L.map_p[li] = (int8_t)best_map(s, [&](int map) {
          hll_spmm_axpby(L.hp, 1.0, w.u[l + 1], f32, K, 0.0, y, K, k, s, map);
        }, nm(L.hp));
    }
    ainv_solver *sv = L.sv.get();
    if (sv->apply_map[li] < 0) {
      const bool f32t = sv->prec != AINV_PREC_F64;
      const int bw = best_map(s, [&](int map) {
        hll_spmm_scale(sv->hw, sv->dv.p, x, f32, K, w.tb, K, k, s, map);
      });
      const int bz = best_map(s, [&](int map) {
        hll_spmm_axpby(sv->hz, 1.0, w.tb, f32t, K, 0.0, y, K, k, s, map);
      });
      sv->apply_map[li] = (int16_t)map_pair(bw, bz);
    }
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

// One column: split or, for the products with a CSR copy, wide.
template <typename V>
void tune_levels_1(ainv_solver *h, const V *x0, V *y0, void *ws,
                   hipStream_t s) {
  constexpr bool f32 = sizeof(V) == sizeof(float);
  MlHier &H = *h->ml;
  AINV_HIP_CHECK(hipMemsetAsync(ws, 0, ml_ws_bytes(h, 1), s));
  const Ws<V> w = carve<V>(H, 1, ws);
  auto two = [&](const Hll &m, const std::function<void(int)> &f) {
    if (!m.cptr.p) return (int8_t)AINV_MAP_SPLIT;
    const double ts = time_ms(s, [&] { f(AINV_MAP_SPLIT); }, 3);
    const double tw = time_ms(s, [&] { f(AINV_MAP_WIDE); }, 3);
    return (int8_t)(tw < ts ? AINV_MAP_WIDE : AINV_MAP_SPLIT);
  };
  for (size_t l = 0; l < H.lv.size(); ++l) {
    MlLevel &L = *H.lv[l];
    const V *x = l > 0 ? w.b[l] : x0;
    V *y = l > 0 ? w.u[l] : y0;
    if (L.map_a[0] < 0)
      L.map_a[0] = two(L.ha, [&](int map) {
        hll_spmm_resid(L.ha, -1.0, y, 1.0, x, w.t[l], 1, 1, s, map);
      });
    if (L.nc > 0) {
      if (L.map_r[0] < 0)
        L.map_r[0] = two(L.hr, [&](int map) {
          hll_spmm_axpby(L.hr, 1.0, x, f32, 1, 0.0, w.b[l + 1], 1, 1, s, map);
        });
      if (L.map_p[0] < 0)
        L.map_p[0] = two(L.hp, [&](int map) {
          hll_spmm_axpby(L.hp, 1.0, w.u[l + 1], f32, 1, 0.0, y, 1, 1, s, map);
        });
    }
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

// every product of every level has a mapping for block width index li
// (the one-column tuning sets only the level products)
static bool all_tuned(const MlHier &H, int li, bool with_smoothers) {
  for (const auto &L : H.lv) {
    if (L->map_a[li] < 0) return false;
    if (L->nc > 0 && (L->map_p[li] < 0 || L->map_r[li] < 0)) return false;
    if (with_smoothers && L->sv->apply_map[li] < 0) return false;
  }
  return true;
}

void ml_tune(ainv_solver *h, int K, int k, const void *x0, void *y0,
             bool x32, void *ws, hipStream_t s) {
  if (h->map_mode >= 0) return;
  if (all_tuned(*h->ml, log2_pow2(K), K > 1)) return;
  if (K == 1) {
    if (x32)
      tune_levels_1<float>(h, static_cast<const float *>(x0),
                           static_cast<float *>(y0), ws, s);
    else
      tune_levels_1<double>(h, static_cast<const double *>(x0),
                            static_cast<double *>(y0), ws, s);
    return;
  }
  if (k < 2) return;
  if (x32)
    tune_levels<float>(h, K, k, static_cast<const float *>(x0),
                       static_cast<float *>(y0), ws, s);
  else
    tune_levels<double>(h, K, k, static_cast<const double *>(x0),
                        static_cast<double *>(y0), ws, s);
}

// ---- queries -----------------------------------------------------------------

int ml_levels(const ainv_solver *h) { return (int)hier(h).lv.size(); }

const ainv_solver *ml_level_solver(const ainv_solver *h, int l) {
  const MlHier &H = hier(h);
  if (l < 0 || l >= (int)H.lv.size())
    throw Error(AINV_ERR_ARG, "level out of range");
  return H.lv[l]->sv.get();
}

void ml_set_map_mode(ainv_solver *h, int mode) {
  if (!h->ml) return;
  for (auto &L : h->ml->lv) {
    L->sv->map_mode = smoother_map(mode);
    reset_tuning(L->sv.get());
    for (int i = 0; i < 8; ++i) L->map_a[i] = L->map_p[i] = L->map_r[i] = -1;
  }
}

void ml_remold(ainv_solver *h, int prec) {
  for (auto &L : h->ml->lv) solver_remold(L->sv.get(), prec);
  const int mp = prec == AINV_PREC_F64 ? AINV_PREC_F64 : AINV_PREC_F32;
  if (mp != h->ml->mprec) upload_levels(*h->ml, prec, h->stream);
}

int64_t ml_device_bytes(const ainv_solver *h) {
  const MlHier &H = hier(h);
  int64_t b = 0;
  for (const auto &L : H.lv)
    b += hll_bytes(L->ha) + hll_bytes(L->hp) + hll_bytes(L->hr) +
         solver_device_bytes(L->sv.get());
  return b;
}

int64_t ml_nzeros(const ainv_solver *h) {
  const MlHier &H = hier(h);
  int64_t nz = 0;
  for (size_t l = 0; l < H.lv.size(); ++l) {
    const MlLevel &L = *H.lv[l];
    nz += (int64_t)L.n + L.sv->w.nnz + L.sv->z.nnz;
    if (l > 0) nz += L.a.nnz();
    nz += L.p.nnz() + L.r.nnz();
  }
  return nz;
}

std::string ml_report(const ainv_solver *h) {
  const MlHier &H = hier(h);
  std::string s = "{\"levels\": [";
  char buf[512];
  for (size_t l = 0; l < H.lv.size(); ++l) {
    const MlLevel &L = *H.lv[l];
    std::snprintf(buf, sizeof buf,
                  "%s{\"n\": %d, \"nnz\": %lld, \"aggregates\": %d, "
                  "\"nnz_p\": %lld, \"nnz_r\": %lld, "
                  "\"theta\": %.6g, \"omega\": %.17g, \"kind\": %d, "
                  "\"nnz_w\": %lld, \"nnz_z\": %lld, \"value_bytes\": %d, "
                  "\"t_factor\": %.4f, \"t_inverse\": %.4f}",
                  l ? ", " : "", L.n, (long long)L.a.nnz(), L.nc,
                  (long long)L.p.nnz(), (long long)L.r.nnz(), L.theta,
                  L.omega, L.sv->kind, (long long)L.sv->w.nnz,
                  (long long)L.sv->z.nnz, value_bytes(L.sv->prec),
                  L.sv->t_factor, L.sv->t_inverse);
    s += buf;
  }
  // tuned mappings of the level products per block width
  s += "], \"maps\": [";
  static const char *names[kNumMaps] = {"split",  "row",       "coop",
                                        "coop128", "coop_q",   "coop128_q",
                                        "run",     "wide"};
  auto nm = [&](int8_t m) { return m < 0 ? "-" : names[m]; };
  bool first = true;
  for (int i = 0; i < 8; ++i)
    for (size_t l = 0; l < H.lv.size(); ++l) {
      const MlLevel &L = *H.lv[l];
      if (L.map_a[i] < 0 && L.map_p[i] < 0 && L.map_r[i] < 0) continue;
      std::snprintf(buf, sizeof buf,
                    "%s{\"K\": %d, \"level\": %zu, \"a\": \"%s\", "
                    "\"p\": \"%s\", \"r\": \"%s\", \"wide\": %s}",
                    first ? "" : ", ", 1 << i, l, nm(L.map_a[i]),
                    nm(L.map_p[i]), nm(L.map_r[i]),
                    L.ha.cptr.p || L.hp.cptr.p || L.hr.cptr.p ? "true"
                                                              : "false");
      s += buf;
      first = false;
    }
  std::snprintf(buf, sizeof buf,
                "], \"t_aggregation\": %.4f, \"t_galerkin\": %.4f, "
                "\"t_factors\": %.4f, \"t_upload\": %.4f}",
                H.t_aggr, H.t_galerkin, H.t_factors, H.t_upload);
  s += buf;
  return s;
}

void ml_export(const ainv_solver *h, int l, char which, HostCsr &m,
               int32_t &ncols, std::vector<double> &dv,
               std::vector<int32_t> &iv) {
  const MlHier &H = hier(h);
  if (l < 0 || l >= (int)H.lv.size())
    throw Error(AINV_ERR_ARG, "level out of range");
  const MlLevel &L = *H.lv[l];
  switch (which) {
    case 'A':
      m = L.a;
      ncols = L.n;
      return;
    case 'P':
    case 'R':
      if (L.nc == 0)
        throw Error(AINV_ERR_ARG, "the coarsest level has no P or R");
      m = which == 'P' ? L.p : L.r;
      ncols = which == 'P' ? L.nc : L.n;
      return;
    case 'W':
    case 'Z': {
      const ainv_solver *sv = L.sv.get();
      csr_download(which == 'W' ? sv->w : sv->z, m, h->stream);
      ncols = L.n;
      return;
    }
    case 'D':
      dv.resize(L.n);
      if (L.n)
        AINV_HIP_CHECK(hipMemcpy(dv.data(), L.sv->d64.p, L.n * sizeof(double),
                                 hipMemcpyDeviceToHost));
      return;
    case 'G':
      if (L.nc == 0)
        throw Error(AINV_ERR_ARG, "the coarsest level has no aggregates");
      iv = L.agg;
      return;
    default:
      throw Error(AINV_ERR_ARG, "which must be A, P, R, W, Z, D or G");
  }
}

}  // namespace ainv
