// Host-side factorizations of the build phase: ILU(k), ILUT and the sparse
// biconjugation of AINV.  These are sequential row/column recurrences
// (each row or column depends on all earlier ones), so they run on the
// host cores; their outputs are uploaded once and everything the Krylov
// loop touches (inverse factors, apply, vector ops) lives on the GPU.  The
// INVK/INVT row inverses, which ARE independent per row, run on the GPU
// (invkt.hip).
//
// Arithmetic follows the reference operation by operation (one rounding
// per multiply and per add, no contraction: this file is compiled with
// -ffp-contract=off) so that patterns and values agree bit for bit with a
// faithful restatement of:
//   ILU(k):  MLD2P4 mld_iluk_fact as called from
//            src/impl/mld_d_invk_bld.f90:104 (Saad Alg. 10.5, IKJ form)
//   ILUT:    MLD2P4 mld_ilut_fact with maxval scaling,
//            src/impl/mld_d_invt_bld.f90:108 (Saad Alg. 10.6)
//   AINV:    src/impl/mld_d_ainv_bld.F90:115-182,
//            src/impl/mld_dsparse_biconjg_llk.F90:82-364 (LLK),
//            src/impl/mld_dsparse_biconjg_s_llk.F90 (SYM-LLK),
//            src/impl/mld_dsparse_biconjg_s_ft_llk.F90 (STAB-LLK),
//            src/impl/mld_dsparse_biconjg_llk_noth.F90 (LLK-NOTH),
//            src/impl/mld_dsparse_biconjg_mlk.F90 (MLK: same visiting set
//            and arithmetic as LLK through sorted lists),
//            src/impl/mld_d_sparsify.f90:83-143, :152-230
// Ties among equal magnitudes at a fill cap are resolved by "smaller column
// first" (the reference leaves them to PSBLAS heap order).
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <queue>
#include <system_error>
#include <thread>

#include "ainv_internal.h"

namespace ainv {

double now_seconds() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

namespace {

const double kEpsTol = DBL_EPSILON;  // d_epstol

// ascending integer heap with repeats (the psb_i_heap use pattern)
using MinHeap =
    std::priority_queue<int32_t, std::vector<int32_t>, std::greater<int32_t>>;

struct Entry {
  double v;
  int32_t j;
};

// |v| descending, then column ascending
inline bool before(const Entry &a, const Entry &b) {
  double fa = std::fabs(a.v), fb = std::fabs(b.v);
  if (fa != fb) return fa > fb;
  return a.j < b.j;
}

// keep at most `keep` entries of e (largest first), sorted by column
void keep_largest(std::vector<Entry> &e, int64_t keep) {
  if (keep < 0) keep = 0;
  if ((int64_t)e.size() > keep) {
    std::nth_element(e.begin(), e.begin() + keep, e.end(), before);
    e.resize(keep);
  }
  std::sort(e.begin(), e.end(),
            [](const Entry &a, const Entry &b) { return a.j < b.j; });
}

// classic reference BLAS dnrm2
double nrm2(const double *x, int64_t n) {
  if (n < 1) return 0.0;
  if (n == 1) return std::fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int64_t i = 0; i < n; ++i) {
    if (x[i] == 0.0) continue;
    double ax = std::fabs(x[i]);
    if (scale < ax) {
      double r = scale / ax;
      ssq = 1.0 + ssq * (r * r);
      scale = ax;
    } else {
      double r = ax / scale;
      ssq = ssq + r * r;
    }
  }
  return scale * std::sqrt(ssq);
}

double max_abs(const HostCsr &a) {
  double m = 0.0;
  for (double v : a.val) m = std::max(m, std::fabs(v));
  return m;
}

void transpose(const HostCsr &a, HostCsr &t) {
  const int32_t n = a.n;
  t.n = n;
  t.ptr.assign(n + 1, 0);
  t.idx.resize(a.nnz());
  t.val.resize(a.nnz());
  for (int64_t k = 0; k < a.nnz(); ++k) t.ptr[a.idx[k] + 1]++;
  for (int32_t i = 0; i < n; ++i) t.ptr[i + 1] += t.ptr[i];
  std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
  for (int32_t i = 0; i < n; ++i)
    for (int64_t k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
      int64_t q = pos[a.idx[k]]++;
      t.idx[q] = i;
      t.val[q] = a.val[k];
    }
}

}  // namespace

// ---------------------------------------------------------------- ILU(k)
void host_iluk(const HostCsr &a, int32_t fill, HostCsr &l, HostCsr &u,
               std::vector<double> &d) {
  const int32_t n = a.n;
  const int32_t untouched = -(n + 1);
  std::vector<double> row(n, 0.0);
  std::vector<int32_t> lev(n, untouched), seen;
  std::vector<int32_t> ulev;  // level of each stored U entry
  d.assign(n, 0.0);
  l = HostCsr();
  u = HostCsr();
  l.n = u.n = n;
  l.ptr.assign(n + 1, 0);
  u.ptr.assign(n + 1, 0);
  MinHeap heap;
  for (int32_t i = 0; i < n; ++i) {
    for (int64_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
      row[a.idx[e]] = a.val[e];
      lev[a.idx[e]] = 0;
      heap.push(a.idx[e]);
    }
    seen.clear();
    int32_t last = -1;
    while (!heap.empty()) {
      int32_t k = heap.top();
      heap.pop();
      if (k == last) continue;
      last = k;
      seen.push_back(k);
      if (k >= i || row[k] == 0.0 || lev[k] > fill) continue;
      // U is stored scaled by the inverted pivots: the unscaled row(k)
      // is the multiplier of the update, L(i,k) = row(k) * d(k).
      const double rk = row[k];
      const int32_t lk = lev[k];
      for (int64_t jj = u.ptr[k]; jj < u.ptr[k + 1]; ++jj) {
        const int32_t j = u.idx[jj];
        if (lev[j] < 0) {
          heap.push(j);
          lev[j] = -lev[j];
        }
        row[j] = row[j] - rk * u.val[jj];
        lev[j] = std::min(lev[j], lk + ulev[jj] + 1);
      }
    }
    double piv = 0.0;
    const int64_t ubeg = (int64_t)u.idx.size();
    for (int32_t j : seen) {
      if (j < i) {
        if (lev[j] <= fill) {
          l.idx.push_back(j);
          l.val.push_back(row[j] * d[j]);
        }
      } else if (j == i) {
        piv = row[j];
      } else if (lev[j] <= fill) {
        u.idx.push_back(j);
        u.val.push_back(row[j]);
        ulev.push_back(lev[j]);
      }
      row[j] = 0.0;
      lev[j] = untouched;
    }
    if (piv == 0.0)
      throw Error(AINV_ERR_PIVOT,
                  "ILU(k): zero pivot at row " + std::to_string(i + 1));
    d[i] = 1.0 / piv;
    for (int64_t e = ubeg; e < (int64_t)u.val.size(); ++e)
      u.val[e] = u.val[e] * d[i];
    l.ptr[i + 1] = (int64_t)l.idx.size();
    u.ptr[i + 1] = (int64_t)u.idx.size();
  }
}

// ------------------------------------------------------------------ ILUT
void host_ilut(const HostCsr &a, int32_t fill, double thresh, HostCsr &l,
               HostCsr &u, std::vector<double> &d) {
  const int32_t n = a.n;
  const double scale = 1.0 / max_abs(a);
  std::vector<double> row(n, 0.0), rv;
  std::vector<char> flag(n, 0);
  std::vector<int32_t> kept;
  std::vector<Entry> cand;
  d.assign(n, 0.0);
  l = HostCsr();
  u = HostCsr();
  l.n = u.n = n;
  l.ptr.assign(n + 1, 0);
  u.ptr.assign(n + 1, 0);
  MinHeap heap;
  for (int32_t i = 0; i < n; ++i) {
    int32_t nlw = 0, nup = 0;
    rv.clear();
    for (int64_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
      const int32_t k = a.idx[e];
      row[k] = a.val[e] * scale;
      rv.push_back(row[k]);
      heap.push(k);
      flag[k] = 1;
      nlw += (k < i);
      nup += (k > i);
    }
    const double thr = thresh * nrm2(rv.data(), (int64_t)rv.size());
    kept.clear();
    int32_t last = -1;
    while (!heap.empty()) {
      const int32_t k = heap.top();
      heap.pop();
      if (k == last) continue;
      last = k;
      if (k < i) {
        const double rk = row[k];  // multiplier for the scaled U
        row[k] = row[k] * d[k];    // L(i,k), the value the drop rule sees
        if (std::fabs(row[k]) < thr) {
          row[k] = 0.0;
          flag[k] = 0;
          continue;
        }
        for (int64_t jj = u.ptr[k]; jj < u.ptr[k + 1]; ++jj) {
          const int32_t j = u.idx[jj];
          row[j] = row[j] - rk * u.val[jj];
          if (!flag[j]) {
            if (std::fabs(row[j]) < thr) {
              row[j] = 0.0;
            } else {
              heap.push(j);
              flag[j] = 1;
            }
          }
        }
      }
      kept.push_back(k);
      flag[k] = 0;
    }
    cand.clear();
    double piv = 0.0;
    for (int32_t j : kept) {
      if (j < i && std::fabs(row[j]) >= thr) cand.push_back({row[j], j});
      if (j == i) piv = row[j];
    }
    keep_largest(cand, (int64_t)nlw + fill);
    for (const Entry &e : cand) {
      l.idx.push_back(e.j);
      l.val.push_back(e.v);
    }
    if (piv == 0.0)
      throw Error(AINV_ERR_PIVOT,
                  "ILUT: zero pivot at row " + std::to_string(i + 1));
    d[i] = 1.0 / piv;
    cand.clear();
    for (int32_t j : kept)
      if (j > i && std::fabs(row[j]) >= thr) cand.push_back({row[j], j});
    keep_largest(cand, (int64_t)nup + fill);
    for (const Entry &e : cand) {
      u.idx.push_back(e.j);
      u.val.push_back(e.v * d[i]);
    }
    for (int32_t j : kept) row[j] = 0.0;
    l.ptr[i + 1] = (int64_t)l.idx.size();
    u.ptr[i + 1] = (int64_t)u.idx.size();
  }
  for (int32_t i = 0; i < n; ++i) d[i] = d[i] * scale;
}

// ------------------------------------------------------------ AINV (LLK)
namespace {

// Column store of a factor under construction (CSC: column i = vector i).
struct ColStore {
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> idx;
  std::vector<double> val;
};

// Work state of one biconjugation pass (Z or W).
struct PassWork {
  std::vector<double> zval;
  std::vector<char> in_z;    // izkr: index already in the pattern of zval
  std::vector<char> queued;  // izcr: row already queued for a dot product
  MinHeap pat;               // indices of the nonzeros of zval
  MinHeap rows;              // rows (or columns) still to be dotted
  std::vector<Entry> cand;
  // sparse accumulator for the STAB-LLK products
  std::vector<double> acc;
  std::vector<char> acc_on;
  std::vector<int32_t> acc_idx;
};

// Variant switches of the left-looking family.
struct Variant {
  bool no_alpha_test = false;  // LLK-NOTH (_llk_noth.F90:117, :234)
  bool list_sparsify = false;  // MLK: sparsify_list keeps 1 if nzrmax < 1
  bool stabilized = false;     // STAB-LLK (_s_ft_llk.F90)
};

// ww = sum_t v_t M(k_t,:) accumulated per output in entry order, touched
// outputs returned ascending (psb_d_spvspm / psb_d_spmspv,
// mld_d_biconjg_mod.F90:145-296).
void combine_rows(const HostCsr &M, const int32_t *ix, const double *vx,
                  int64_t nx, PassWork &w) {
  w.acc_idx.clear();
  for (int64_t t = 0; t < nx; ++t) {
    const int32_t k = ix[t];
    for (int64_t e = M.ptr[k]; e < M.ptr[k + 1]; ++e) {
      const int32_t c = M.idx[e];
      const double term = vx[t] * M.val[e];
      if (!w.acc_on[c]) {
        w.acc_on[c] = 1;
        w.acc_idx.push_back(c);
        w.acc[c] = term;
      } else {
        w.acc[c] = term + w.acc[c];
      }
    }
  }
  std::sort(w.acc_idx.begin(), w.acc_idx.end());
}

void clear_acc(PassWork &w) {
  for (int32_t c : w.acc_idx) {
    w.acc[c] = 0.0;
    w.acc_on[c] = 0;
  }
  w.acc_idx.clear();
}

// Finished columns of the factor under construction, appended in column
// order by the one thread that builds them.
struct SeqCols {
  ColStore &c;
  void wait(int32_t) const {}
  int64_t size(int32_t j) const { return c.ptr[j + 1] - c.ptr[j]; }
  const int32_t *idx(int32_t j) const { return c.idx.data() + c.ptr[j]; }
  const double *val(int32_t j) const { return c.val.data() + c.ptr[j]; }
  void emit(int32_t, const std::vector<Entry> &e) {
    for (const Entry &x : e) {
      c.idx.push_back(x.j);
      c.val.push_back(x.v);
    }
    c.ptr.push_back((int64_t)c.idx.size());
  }
};

// Columns built concurrently by several threads that claim column indices
// in increasing order: column j lives in its own slot of `cap` entries and
// is published (with its pivot, written before) by a release store of
// ready[j]; a reader of column j < i waits for that flag.  Every column
// below the smallest unfinished one is final and that one's builder only
// waits on smaller columns, so the wait always ends.
struct SlotCols {
  int32_t cap = 0;
  std::vector<int32_t> ix;
  std::vector<double> vx;
  std::vector<int32_t> cnt;
  std::unique_ptr<std::atomic<int32_t>[]> ready;
  std::atomic<bool> *abort = nullptr;
  SlotCols(int32_t n, int32_t c, std::atomic<bool> *ab)
      : cap(c), ix((size_t)n * c), vx((size_t)n * c), cnt(n, 0),
        ready(new std::atomic<int32_t>[n]), abort(ab) {
    for (int32_t j = 0; j < n; ++j) ready[j].store(0, std::memory_order_relaxed);
  }
  void wait(int32_t j) const {
    int spins = 0;
    while (!ready[j].load(std::memory_order_acquire)) {
      if (abort->load(std::memory_order_relaxed))
        throw Error(AINV_ERR_INTERNAL, "biconjugation aborted");
      if (++spins > 64) {
        std::this_thread::yield();
        spins = 0;
      }
    }
  }
  int64_t size(int32_t j) const { return cnt[j]; }
  const int32_t *idx(int32_t j) const { return ix.data() + (size_t)j * cap; }
  const double *val(int32_t j) const { return vx.data() + (size_t)j * cap; }
  void emit(int32_t i, const std::vector<Entry> &e) {
    if ((int64_t)e.size() > cap)
      throw Error(AINV_ERR_INTERNAL, "biconjugation column over capacity");
    for (size_t t = 0; t < e.size(); ++t) {
      ix[(size_t)i * cap + t] = e[t].j;
      vx[(size_t)i * cap + t] = e[t].v;
    }
    cnt[i] = (int32_t)e.size();
    ready[i].store(1, std::memory_order_release);
  }
  void to_colstore(ColStore &c) const {
    const int32_t n = (int32_t)cnt.size();
    c.ptr.assign(1, 0);
    c.idx.clear();
    c.val.clear();
    for (int32_t j = 0; j < n; ++j) {
      c.idx.insert(c.idx.end(), idx(j), idx(j) + cnt[j]);
      c.val.insert(c.val.end(), val(j), val(j) + cnt[j]);
      c.ptr.push_back((int64_t)c.idx.size());
    }
  }
};

// One column i of Z (dot = rows of A, seed/enqueue = columns of A) or of W
// (the same with A^T).  dotm: matrix whose row j is dotted with zval;
// seedm: matrix whose row i / row kr lists the candidates to queue.
// With v.stabilized the dot of step j is (g_j^T M) . zval, g_j being
// column j of the other factor `g` and M = `comb`, and no pivot is formed.
// f gives the finished columns (and receives column i); column j and
// piv[j] are read only after f.wait(j).
template <class Cols>
void bicg_column(int32_t i, const HostCsr &dotm, const HostCsr &seedm,
                 Cols &f, std::vector<double> &piv, int32_t nzrmax,
                 double thresh, const Variant &v, const ColStore *g,
                 const HostCsr *comb, PassWork &w) {
  w.zval[i] = 1.0;
  w.in_z[i] = 1;
  w.pat.push(i);
  for (int64_t e = seedm.ptr[i]; e < seedm.ptr[i + 1]; ++e) {
    const int32_t r = seedm.idx[e];
    if (r < i) {
      w.rows.push(r);
      w.queued[r] = 1;
    }
  }
  int32_t lastj = -1;
  while (!w.rows.empty()) {
    const int32_t j = w.rows.top();
    w.rows.pop();
    if (j <= lastj) continue;
    lastj = j;
    w.queued[j] = 0;
    if (j >= i) continue;
    double s = 0.0;
    if (v.stabilized) {
      const int64_t b = g->ptr[j];
      combine_rows(*comb, g->idx.data() + b, g->val.data() + b,
                   g->ptr[j + 1] - b, w);
      for (int32_t c : w.acc_idx) s = s + w.acc[c] * w.zval[c];
      clear_acc(w);
    } else {
      for (int64_t e = dotm.ptr[j]; e < dotm.ptr[j + 1]; ++e)
        s = s + dotm.val[e] * w.zval[dotm.idx[e]];
    }
    f.wait(j);
    const double alpha = (-s / piv[j]);
    if (!v.no_alpha_test && !(std::fabs(alpha) > thresh)) continue;
    const int32_t *fi = f.idx(j);
    const double *fv = f.val(j);
    const int64_t fn = f.size(j);
    for (int64_t e = 0; e < fn; ++e) {
      const int32_t kr = fi[e];
      w.zval[kr] = w.zval[kr] + alpha * fv[e];
      if (w.in_z[kr]) continue;
      w.pat.push(kr);
      w.in_z[kr] = 1;
      for (int64_t c = seedm.ptr[kr]; c < seedm.ptr[kr + 1]; ++c) {
        const int32_t nj = seedm.idx[c];
        if (!w.queued[nj] && nj > j && nj < i) {
          w.rows.push(nj);
          w.queued[nj] = 1;
        }
      }
    }
  }
  if (!v.stabilized) {
    double s = 0.0;
    for (int64_t e = dotm.ptr[i]; e < dotm.ptr[i + 1]; ++e)
      s = s + dotm.val[e] * w.zval[dotm.idx[e]];
    piv[i] = (std::fabs(s) < kEpsTol) ? 1.0e-3 : s;
  }
  if (v.list_sparsify && nzrmax < 1) nzrmax = 1;  // sparsify.f90:214-222
  // sparsify: diagonal always, then the nzrmax-1 largest with |z| >= thresh
  w.cand.clear();
  double diag = 0.0;
  int32_t last = -1;
  while (!w.pat.empty()) {
    const int32_t k = w.pat.top();
    w.pat.pop();
    if (k == last) continue;
    last = k;
    if (k == i)
      diag = w.zval[k];
    else if (std::fabs(w.zval[k]) >= thresh)
      w.cand.push_back({w.zval[k], k});
    w.zval[k] = 0.0;
    w.in_z[k] = 0;
  }
  const int64_t keep =
      nzrmax >= 1 ? (int64_t)nzrmax - 1 : (int64_t)w.cand.size();
  if ((int64_t)w.cand.size() > keep) {
    std::nth_element(w.cand.begin(), w.cand.begin() + keep, w.cand.end(),
                     before);
    w.cand.resize(keep);
  }
  w.cand.push_back({diag, i});
  std::sort(w.cand.begin(), w.cand.end(),
            [](const Entry &a, const Entry &b) { return a.j < b.j; });
  f.emit(i, w.cand);
}

void init_work(PassWork &w, int32_t n, bool stabilized) {
  w.zval.assign(n, 0.0);
  w.in_z.assign(n, 0);
  w.queued.assign(n, 0);
  if (stabilized) {
    w.acc.assign(n, 0.0);
    w.acc_on.assign(n, 0);
  }
}

// Columns 1..n-1 of one pass on `threads` threads claiming columns in
// increasing order (SlotCols).  Results are bit-identical to the
// sequential loop: each column performs the same operations on the same
// finished columns, whichever thread runs it.
void pipelined_pass(int32_t n, const HostCsr &dotm, const HostCsr &seedm,
                    std::vector<double> &piv, int32_t nzrmax, double thresh,
                    const Variant &v, const ColStore &first, int threads,
                    ColStore &out) {
  const int32_t cap = std::max<int32_t>(nzrmax, 1);
  std::atomic<bool> abort{false};
  SlotCols cols(n, cap, &abort);
  cols.emit(0, {{first.val[0], first.idx[0]}});
  std::atomic<int32_t> next{1};
  std::exception_ptr err;
  std::mutex err_mu;
  auto worker = [&] {
    try {
      PassWork w;
      init_work(w, n, false);
      for (;;) {
        const int32_t i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= n || abort.load(std::memory_order_relaxed)) break;
        bicg_column(i, dotm, seedm, cols, piv, nzrmax, thresh, v, nullptr,
                    nullptr, w);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (!err) err = std::current_exception();
      abort.store(true);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(threads > 1 ? threads - 1 : 0);
  for (int t = 1; t < threads; ++t) {
    try {
      pool.emplace_back(worker);
    } catch (const std::system_error &) {
      break;  // columns are claimed in order by whichever threads run
    }
  }
  worker();
  for (auto &t : pool) t.join();
  if (err) std::rethrow_exception(err);
  cols.to_colstore(out);
}

}  // namespace

int build_threads() {
  if (const char *e = std::getenv("AINV_BUILD_THREADS")) {
    const int t = std::atoi(e);
    if (t >= 1) return std::min(t, 64);
  }
  cpu_set_t set;
  int cpus = 0;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
  if (cpus < 1) cpus = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(cpus, 16));
}

void host_ainv(const HostCsr &a, int alg, int32_t fill_in, double thresh,
               HostCsr &w, HostCsr &z, std::vector<double> &d, int threads,
               AinvBuildInfo *info) {
  // the dispatcher's select (mld_d_biconjg_mod.F90:107-122)
  if (alg < AINV_ALG_LLK || alg > AINV_ALG_MLK)
    throw Error(AINV_ERR_ALG, "Invalid alg " + std::to_string(alg));
  Variant var;
  var.no_alpha_test = (alg == AINV_ALG_LLK_NOTH);
  var.list_sparsify = (alg == AINV_ALG_MLK);
  var.stabilized = (alg == AINV_ALG_S_FT_LLK);
  const int32_t n = a.n;
  // maxval scaling (mld_d_ainv_bld.F90:121-124)
  const double weight = 1.0 / max_abs(a);
  HostCsr as = a;
  for (double &v : as.val) v = v * weight;
  HostCsr at;
  transpose(as, at);
  std::vector<double> p(n, 0.0), q(n, 0.0);
  ColStore zc, wc;
  PassWork pw;
  init_work(pw, n, var.stabilized);
  if (n > 0) {
    for (int64_t e = as.ptr[0]; e < as.ptr[1]; ++e)
      if (as.idx[e] == 0) {
        p[0] = as.val[e];
        break;
      }
    if (std::fabs(p[0]) < kEpsTol) p[0] = 1.0e-3;
    q[0] = p[0];
    zc.idx.push_back(0);
    zc.val.push_back(1.0);
    zc.ptr.push_back(1);
    wc.idx.push_back(0);
    wc.val.push_back(1.0);
    wc.ptr.push_back(1);
  }
  if (threads <= 0) threads = build_threads();
  // bounded columns (a fill cap) can be built by the pipelined passes
  const bool bounded = fill_in >= 1 || var.list_sparsify;
  // The pipelined passes pay off only when a column's dependencies lie
  // well behind it: with a neighbour at distance 1 (stencil orderings)
  // column i needs column i-1 and the pipeline runs one column at a time
  // while paying for the hand-offs.  Median over sampled columns of the
  // distance to the nearest earlier row or column coupled to them.
  if (threads >= 2 && n > 1) {
    std::vector<int32_t> dist;
    // scattered sample (a regular stride can land on one grid boundary)
    for (int64_t t = 1; t <= 1024; ++t) {
      const int32_t i = (int32_t)((t * 2654435761LL) % n);
      if (i == 0) continue;
      int32_t best = i;  // none: as far as it gets
      for (int64_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e)
        if (a.idx[e] < i) best = std::min(best, i - a.idx[e]);
      for (int64_t e = at.ptr[i]; e < at.ptr[i + 1]; ++e)
        if (at.idx[e] < i) best = std::min(best, i - at.idx[e]);
      dist.push_back(best);
    }
    if (!dist.empty()) {
      std::nth_element(dist.begin(), dist.begin() + dist.size() / 2,
                       dist.end());
      if (dist[dist.size() / 2] < 8 * threads) threads = 2;
    }
  }
  const bool two = alg != AINV_ALG_S_LLK;
  bool done = false;
  if (!var.stabilized && n > 1 && threads >= 2 && bounded &&
      (two || threads > 2)) {
    // Without stabilization the Z pass (pivots p) and the W pass (pivots q)
    // share nothing: each runs pipelined on its share of the threads (one
    // thread each: the plain column loop, concurrently).
    const int tz = two ? std::max(1, (threads + 1) / 2) : threads;
    const int tw = std::max(1, threads - tz);
    std::exception_ptr err;
    std::thread wth;
    auto w_pass = [&] {
      try {
        pipelined_pass(n, at, as, q, fill_in, thresh, var, wc, tw, wc);
      } catch (...) {
        err = std::current_exception();
      }
    };
    bool w_here = false;  // the W pass runs after the Z pass on this thread
    if (two) {
      try {
        wth = std::thread(w_pass);
      } catch (const std::system_error &) {
        w_here = true;
      }
    }
    try {
      pipelined_pass(n, as, at, p, fill_in, thresh, var, zc, tz, zc);
    } catch (...) {
      if (wth.joinable()) wth.join();
      throw;
    }
    if (wth.joinable()) wth.join();
    if (w_here) w_pass();
    if (err) std::rethrow_exception(err);
    done = true;
  }
  if (info) {
    info->threads = done ? threads : 1;
    info->pipelined = done;
  }
  SeqCols zs{zc}, ws{wc};
  for (int32_t i = 1; i < n && !done; ++i) {
    // Z: dots with rows of A (STAB: w_j^T A), seeds from column i of A
    bicg_column(i, as, at, zs, p, fill_in, thresh, var, &wc, &as, pw);
    if (alg == AINV_ALG_S_LLK) continue;
    // W: dots with columns of A (STAB: A z_j), seeds from row i of A
    bicg_column(i, at, as, ws, q, fill_in, thresh, var, &zc, &at, pw);
    if (var.stabilized) {
      // q(i) = w_i^T (A z_i), merge dot over sorted indices
      // (mld_dsparse_biconjg_s_ft_llk.F90:377-411)
      const int64_t zb = zc.ptr[i];
      combine_rows(at, zc.idx.data() + zb, zc.val.data() + zb,
                   zc.ptr[i + 1] - zb, pw);
      double t = 0.0;
      size_t u = 0;
      for (int64_t e = wc.ptr[i]; e < wc.ptr[i + 1]; ++e) {
        const int32_t c = wc.idx[e];
        while (u < pw.acc_idx.size() && pw.acc_idx[u] < c) ++u;
        if (u < pw.acc_idx.size() && pw.acc_idx[u] == c)
          t = t + pw.acc[c] * wc.val[e];
      }
      clear_acc(pw);
      q[i] = (std::fabs(t) < kEpsTol) ? 1.0e-3 : t;
      p[i] = q[i];
    }
  }
  if (alg == AINV_ALG_S_LLK) wc = zc;  // W := Z (s_llk.F90:246)
  // Z CSC -> CSR; W CSC holds W^T as CSR (biconjg_mod.F90:131-135)
  HostCsr zcsc;
  zcsc.n = n;
  zcsc.ptr = std::move(zc.ptr);
  zcsc.idx = std::move(zc.idx);
  zcsc.val = std::move(zc.val);
  transpose(zcsc, z);
  w.n = n;
  w.ptr = std::move(wc.ptr);
  w.idx = std::move(wc.idx);
  w.val = std::move(wc.val);
  d.assign(n, 0.0);
  for (int32_t i = 0; i < n; ++i) {
    double v = (std::fabs(p[i]) < kEpsTol) ? 1.0 : 1.0 / p[i];
    d[i] = v * weight;
  }
}

}  // namespace ainv
