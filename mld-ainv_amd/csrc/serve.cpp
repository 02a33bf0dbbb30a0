// Serving loop: solves on many matrices arrive as requests (1..64
// right-hand sides each); each matrix's factors are built on first touch and
// stay resident on the device under an HBM budget with LRU eviction of idle
// matrices.  Worker threads (one HIP stream each, so batches on different
// matrices overlap on the device) take the matrix with the oldest pending
// request and run a slot batch on it: requests are admitted into free
// column slots, one captured Krylov iteration advances every active column,
// and a column that converges frees its slot for the next pending request
// of the same matrix between two iterations (continuous batching).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <system_error>
#include <thread>
#include <unordered_map>

#include "objects.h"

extern "C" int ainv_free(ainv_h h);
extern "C" int ainv_mat_free(ainv_mat_h a);

namespace ainv {
// defined in capi.cpp
HostCsr canonical_csr_public(int32_t n, const int64_t *irp, const int32_t *ja,
                             const double *val, int base);
void build_solver(ainv_solver *h, const HostCsr &a);
ainv_matrix *make_matrix(const HostCsr &a, int device);
void apply_param(ainv_solver *h, const std::string &key,
                 const std::string &val);
int64_t solver_device_bytes(const ainv_solver *h);
}  // namespace ainv

namespace {

using namespace ainv;

// A request's right-hand sides and solution live in device memory: the
// submitting thread uploads b, the worker extracts x on the device and the
// waiting thread copies it out, so no worker blocks on a host copy.
struct Request {
  int64_t id = 0;
  int32_t mat = 0;
  int32_t n = 0;
  int k = 1;
  double eps = 1e-8;
  int itmax = 1000;
  double *dev_b = nullptr;       // n x k, row-interleaved
  double *dev_x = nullptr;       // n x k, written by the worker
  hipEvent_t ready = nullptr;    // recorded after the extraction
  std::vector<int32_t> iters;
  std::vector<double> errs;
  double t_submit = 0, t_start = 0, t_done = 0;
  bool done = false;
  int status = AINV_SUCCESS;
  std::string msg;
};

struct ServedMatrix {
  int32_t id = 0;
  HostCsr a;
  std::vector<std::pair<std::string, std::string>> params;
  int kind = AINV_KIND_INVK;
  int method = KM_BICGSTAB;
  int min_slots = 1;  // SLOTS: fewest column slots a batch opens with
  ainv_matrix *A = nullptr;
  ainv_solver *M = nullptr;
  int64_t bytes = 0;
  uint64_t last_use = 0;
  bool busy = false;  // owned by a worker (served) or a builder (built)
  std::deque<int64_t> pending;
  int builds = 0;
  bool resident() const { return A && M; }
};

}  // namespace

struct ainv_server {
  int device = 0;
  int64_t budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::unordered_map<int32_t, std::unique_ptr<ServedMatrix>> mats;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int32_t next_mat = 1;
  int64_t next_req = 1;
  uint64_t clock = 0;
  int64_t resident = 0;
  bool stop = false;
  int nworkers = 0;  // worker threads started
  int alive = 0;     // workers that may still serve (startup not failed)
  // builder threads: first-touch builds (and rebuilds after eviction) run
  // there, on the host cores and their own streams, while the workers keep
  // serving resident matrices; with no builder alive the workers build
  int nbuilders = 0, builders_alive = 0;
  hipStream_t copy = nullptr;  // uploads and downloads of request data
  // Held while a worker may capture a graph (batch_open) and around every
  // stream-ordered allocation or free: HIP refuses those while any stream
  // of the process is capturing.
  std::mutex hip_mu;
  std::vector<std::thread> workers, builders;
  // statistics
  int64_t builds = 0, evictions = 0, completed = 0, admitted_midflight = 0;
  int64_t steps = 0, batches = 0;
  double build_seconds = 0;
  // overlap of first-touch builds with device work: workers inside a slot
  // batch, the seconds during which at least one was (integrated at every
  // change), and the build seconds that fell inside them
  int in_batch = 0;
  double busy_mark = 0, busy_seconds = 0, build_overlap_seconds = 0;
  // host seconds per phase summed over workers, and request waits
  double open_seconds = 0, admit_seconds = 0, step_seconds = 0,
         extract_seconds = 0, queue_seconds = 0, service_seconds = 0;
  int64_t peak_resident = 0;
};

namespace {

// Seconds so far during which some worker was inside a slot batch (the
// caller holds the lock).
double busy_now_locked(ainv_server *S) {
  const double t = now_seconds();
  if (S->in_batch > 0) S->busy_seconds += t - S->busy_mark;
  S->busy_mark = t;
  return S->busy_seconds;
}

// A worker inside a slot batch from construction to destruction.
struct InBatch {
  ainv_server *S;
  explicit InBatch(ainv_server *s) : S(s) {
    std::unique_lock<std::mutex> lk(S->mu);
    busy_now_locked(S);
    S->in_batch++;
  }
  ~InBatch() {
    std::unique_lock<std::mutex> lk(S->mu);
    busy_now_locked(S);
    S->in_batch--;
  }
};

int64_t plan_bytes(int32_t n, int K, int method) {
  return (int64_t)n * K * (method == KM_CG ? 6 : 10) * 8 + (int64_t)n * K * 8;
}

void evict_locked(ainv_server *S, ServedMatrix *m) {
  // caller holds the lock and m is idle
  if (m->M) ainv_free(m->M);
  if (m->A) ainv_mat_free(m->A);
  m->M = nullptr;
  m->A = nullptr;
  S->resident -= m->bytes;
  m->bytes = 0;
  S->evictions++;
}

// Make m resident, evicting least-recently used idle matrices while the
// budget would be exceeded.  Called with m busy (owned by this worker).
void ensure_resident(ainv_server *S, ServedMatrix *m, int K) {
  if (m->A && m->M) return;
  // estimate: A (value + column) and two factors of about 4x the nonzeros
  const int64_t est = m->a.nnz() * 12 * 9 + plan_bytes(m->a.n, K, m->method);
  {
    std::unique_lock<std::mutex> lk(S->mu);
    while (S->resident + est > S->budget) {
      ServedMatrix *victim = nullptr;
      for (auto &kv : S->mats) {
        ServedMatrix *c = kv.second.get();
        if (c == m || c->busy || !c->A) continue;
        if (!victim || c->last_use < victim->last_use) victim = c;
      }
      if (!victim) break;  // nothing idle to evict: go over budget
      evict_locked(S, victim);
    }
  }
  double busy0;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    busy0 = busy_now_locked(S);
  }
  const double t0 = now_seconds();
  ainv_solver *M = nullptr;
  ainv_matrix *A = nullptr;
  try {
    A = make_matrix(m->a, S->device);
    M = new ainv_solver();
    M->kind = m->kind;
    M->device = S->device;
    // concurrent builds share the host cores
    int builders;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      builders = S->builders_alive > 0 ? S->nbuilders : S->nworkers;
    }
    M->host_threads = std::max(1, build_threads() / std::max(1, builders));
    for (auto &kv : m->params) apply_param(M, kv.first, kv.second);
    build_solver(M, m->a);
  } catch (...) {
    if (M) ainv_free(M);
    if (A) ainv_mat_free(A);
    throw;
  }
  const double dt = now_seconds() - t0;
  std::unique_lock<std::mutex> lk(S->mu);
  m->A = A;
  m->M = M;
  m->bytes = A->hll.bytes() + (int64_t)(A->csr.ptr.bytes + A->csr.idx.bytes +
                                        A->csr.val.bytes) +
             solver_device_bytes(M) + plan_bytes(m->a.n, K, m->method);
  // the multilevel cycle's workspace instead of one apply intermediate
  if (M->kind == AINV_KIND_ML)
    m->bytes += (int64_t)solver_t_bytes(M, K) - (int64_t)m->a.n * K * 8;
  // BiCG builds the transposed layouts of A and of both factors on first use
  if (m->method == KM_BICG)
    m->bytes += A->hll.bytes() + M->hw.bytes() + M->hz.bytes();
  S->resident += m->bytes;
  S->peak_resident = std::max(S->peak_resident, S->resident);
  S->builds++;
  S->build_seconds += dt;
  S->build_overlap_seconds += busy_now_locked(S) - busy0;
  m->builds++;
}

void finish_request(ainv_server *S, Request *r, int status,
                    const std::string &msg) {
  std::unique_lock<std::mutex> lk(S->mu);
  r->status = status;
  r->msg = msg;
  r->t_done = now_seconds();
  r->done = true;
  S->completed++;
  if (r->t_start > 0) {
    S->queue_seconds += r->t_start - r->t_submit;
    S->service_seconds += r->t_done - r->t_start;
  }
  S->cv_done.notify_all();
}

struct SlotUse {
  int64_t req = -1;
  int c0 = 0, kc = 0;
  int64_t first_step = 0;  // first step that runs with this request's columns
};

void serve_matrix(ainv_server *S, ServedMatrix *m, hipStream_t s) {
  int K = 1;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    int cols = 0;
    for (int64_t id : m->pending) cols += S->reqs[id]->k;
    cols = std::min(std::max(cols, m->min_slots), kMaxCols);
    while (K < cols) K <<= 1;
  }
  ensure_resident(S, m, K);
  const int32_t n = m->a.n;
  const double t_open = now_seconds();
  SlotBatch *b = nullptr;
  {
    std::lock_guard<std::mutex> g(S->hip_mu);
    b = batch_open(m->A, m->M, m->method, K, s);
  }
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->open_seconds += now_seconds() - t_open;
  }
  InBatch busy(S);
  double t_admit = 0, t_step = 0, t_extract = 0;
  std::vector<SlotUse> uses;  // admitted requests
  std::vector<int> owner(K, -1);
  std::vector<SlotState> st;
  bool first = true;
  try {
    for (;;) {
      // admit whole requests into free contiguous slot runs, FIFO
      std::vector<Request *> adm;
      std::vector<int> adm_c0;
      {
        std::unique_lock<std::mutex> lk(S->mu);
        while (!m->pending.empty()) {
          Request *r = S->reqs[m->pending.front()].get();
          int c0 = -1;
          for (int c = 0; c + r->k <= K && c0 < 0; ++c) {
            bool free = true;
            for (int j = 0; j < r->k && free; ++j) free = owner[c + j] < 0;
            if (free) c0 = c;
          }
          if (c0 < 0) break;
          m->pending.pop_front();
          for (int j = 0; j < r->k; ++j) owner[c0 + j] = (int)uses.size();
          uses.push_back(SlotUse{r->id, c0, r->k, batch_next_step(b)});
          adm.push_back(r);
          adm_c0.push_back(c0);
          if (!first) S->admitted_midflight++;
        }
      }
      const double t0 = now_seconds();
      for (size_t t = 0; t < adm.size(); ++t) {
        Request *r = adm[t];
        r->t_start = now_seconds();
        batch_admit(b, adm_c0[t], r->k, r->dev_b, r->eps);
      }
      const double t1 = now_seconds();
      t_admit += t1 - t0;
      first = false;
      bool any = false;
      for (int c = 0; c < K; ++c) any |= owner[c] >= 0;
      if (!any) break;
      batch_step(b);
      {
        std::unique_lock<std::mutex> lk(S->mu);
        S->steps++;
      }
      const int64_t seen = batch_poll(b, false, st);
      const double t2 = now_seconds();
      t_step += t2 - t1;
      if (seen >= 0) {
        // complete requests whose columns all stopped (or hit itmax); a
        // snapshot older than a request's admission shows the previous
        // occupants of its slots and says nothing about it
        for (size_t u = 0; u < uses.size(); ++u) {
          SlotUse &su = uses[u];
          if (su.req < 0 || seen < su.first_step) continue;
          Request *r;
          {
            std::unique_lock<std::mutex> lk(S->mu);
            r = S->reqs[su.req].get();
          }
          bool alldone = true;
          for (int j = 0; j < su.kc; ++j) {
            const SlotState &x = st[su.c0 + j];
            if (x.active && x.iter < r->itmax) alldone = false;
          }
          if (!alldone) continue;
          {
            std::lock_guard<std::mutex> g(S->hip_mu);
            AINV_HIP_CHECK(hipMallocAsync(
                (void **)&r->dev_x, (size_t)n * su.kc * sizeof(double), s));
          }
          batch_extract(b, su.c0, su.kc, r->dev_x);
          AINV_HIP_CHECK(
              hipEventCreateWithFlags(&r->ready, hipEventDisableTiming));
          AINV_HIP_CHECK(hipEventRecord(r->ready, s));
          r->iters.resize(su.kc);
          r->errs.resize(su.kc);
          for (int j = 0; j < su.kc; ++j) {
            r->iters[j] = st[su.c0 + j].iter;
            r->errs[j] = st[su.c0 + j].err;
            owner[su.c0 + j] = -1;
          }
          su.req = -1;
          finish_request(S, r, AINV_SUCCESS, "");
        }
      }
      t_extract += now_seconds() - t2;
    }
  } catch (const std::exception &e) {
    batch_close(b);
    std::string msg = e.what();
    for (auto &su : uses)
      if (su.req >= 0) {
        Request *r;
        {
          std::unique_lock<std::mutex> lk(S->mu);
          r = S->reqs[su.req].get();
        }
        finish_request(S, r, AINV_ERR_INTERNAL, msg);
      }
    throw;
  }
  batch_close(b);
  std::unique_lock<std::mutex> lk(S->mu);
  S->batches++;
  S->admit_seconds += t_admit;
  S->step_seconds += t_step;
  S->extract_seconds += t_extract;
}

// A worker that cannot set the device or create its stream leaves; when no
// worker is left, the queued requests fail (their waits return) and
// submissions are refused.
void worker_startup_failed(ainv_server *S) {
  std::vector<Request *> fail;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    if (--S->alive > 0) return;
    for (auto &kv : S->mats)
      while (!kv.second->pending.empty()) {
        fail.push_back(S->reqs[kv.second->pending.front()].get());
        kv.second->pending.pop_front();
      }
  }
  for (Request *r : fail)
    finish_request(S, r, AINV_ERR_HIP, "no server worker could start");
}

// The idle matrix whose oldest pending request is the oldest: among the
// resident ones (which = 1), the non-resident ones (which = -1) or all (0).
// The caller holds the lock.
ServedMatrix *oldest_pending_locked(ainv_server *S, int which) {
  ServedMatrix *m = nullptr;
  double oldest = 0;
  for (auto &kv : S->mats) {
    ServedMatrix *c = kv.second.get();
    if (c->busy || c->pending.empty()) continue;
    if ((which > 0 && !c->resident()) || (which < 0 && c->resident()))
      continue;
    const double t = S->reqs[c->pending.front()]->t_submit;
    if (!m || t < oldest) {
      m = c;
      oldest = t;
    }
  }
  return m;
}

// Fails the requests still queued for m (after its build or batch threw).
void fail_pending(ainv_server *S, ServedMatrix *m, int status,
                  const std::string &msg) {
  std::vector<Request *> fail;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    while (!m->pending.empty()) {
      fail.push_back(S->reqs[m->pending.front()].get());
      m->pending.pop_front();
    }
  }
  for (Request *r : fail) finish_request(S, r, status, msg);
}

// A builder makes resident the non-resident matrix with the oldest pending
// request, then hands it to the workers.
void builder_main(ainv_server *S) {
  if (hipSetDevice(S->device) != hipSuccess) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      --S->builders_alive;  // at zero the workers build themselves
    }
    S->cv_work.notify_all();
    return;
  }
  for (;;) {
    ServedMatrix *m = nullptr;
    int K = 1;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) return;
        m = oldest_pending_locked(S, -1);
        if (m) break;
        S->cv_work.wait(lk);
      }
      m->busy = true;
      int cols = 0;
      for (int64_t id : m->pending) cols += S->reqs[id]->k;
      cols = std::min(std::max(cols, m->min_slots), kMaxCols);
      while (K < cols) K <<= 1;
    }
    try {
      ensure_resident(S, m, K);
    } catch (const std::exception &e) {
      fail_pending(S, m, AINV_ERR_INTERNAL, e.what());
    }
    {
      std::unique_lock<std::mutex> lk(S->mu);
      m->busy = false;
      m->last_use = ++S->clock;
    }
    S->cv_work.notify_all();
  }
}

void worker_main(ainv_server *S) {
  hipStream_t s = nullptr;
  if (hipSetDevice(S->device) != hipSuccess ||
      hipStreamCreate(&s) != hipSuccess) {
    worker_startup_failed(S);
    return;
  }
  for (;;) {
    ServedMatrix *m = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) {
          lk.unlock();
          (void)hipStreamDestroy(s);
          return;
        }
        // resident matrices only, unless no builder is left to build
        m = oldest_pending_locked(S, S->builders_alive > 0 ? 1 : 0);
        if (m) break;
        S->cv_work.wait(lk);
      }
      m->busy = true;
    }
    try {
      serve_matrix(S, m, s);
    } catch (const std::exception &e) {
      fail_pending(S, m, AINV_ERR_INTERNAL, e.what());
    }
    {
      std::unique_lock<std::mutex> lk(S->mu);
      m->busy = false;
      m->last_use = ++S->clock;
    }
    S->cv_work.notify_all();
  }
}

std::vector<std::pair<std::string, std::string>> parse_params(const char *p) {
  std::vector<std::pair<std::string, std::string>> out;
  if (!p) return out;
  std::string s(p);
  size_t i = 0;
  while (i < s.size()) {
    size_t j = s.find(';', i);
    if (j == std::string::npos) j = s.size();
    std::string item = s.substr(i, j - i);
    size_t e = item.find('=');
    if (e != std::string::npos) {
      auto trim = [](std::string t) {
        size_t b = t.find_first_not_of(" \t"), f = t.find_last_not_of(" \t");
        return b == std::string::npos ? std::string() : t.substr(b, f - b + 1);
      };
      out.emplace_back(trim(item.substr(0, e)), trim(item.substr(e + 1)));
    }
    i = j + 1;
  }
  return out;
}

}  // namespace

extern "C" {

int ainv_server_create(int device, int64_t hbm_budget, int workers,
                       ainv_server_h *out) {
  return guard([&] {
    if (!out) throw Error(AINV_ERR_ARG, "NULL output handle");
    if (workers < 1 || workers > 16)
      throw Error(AINV_ERR_ARG, "workers must be 1..16");
    if (hbm_budget <= 0) throw Error(AINV_ERR_ARG, "budget must be positive");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 ||
        device >= count)
      throw Error(AINV_ERR_HIP, "no such HIP device");
    auto *S = new ainv_server();
    S->device = device;
    S->budget = hbm_budget;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) != hipSuccess) {
      if (S->copy) (void)hipStreamDestroy(S->copy);
      delete S;
      throw Error(AINV_ERR_HIP, "cannot create the server's copy stream");
    }
    S->workers.reserve(workers);
    {
      // the workers wait for the lock before they read nworkers / alive
      std::unique_lock<std::mutex> lk(S->mu);
      for (int w = 0; w < workers; ++w) {
        try {
          S->workers.emplace_back(worker_main, S);
        } catch (const std::system_error &) {
          break;  // serve with the workers that started
        }
      }
      S->nworkers = S->alive = (int)S->workers.size();
      // as many builders as workers: the builds' host threads are shared
      S->builders.reserve(S->workers.size());
      for (size_t w = 0; w < S->workers.size(); ++w) {
        try {
          S->builders.emplace_back(builder_main, S);
        } catch (const std::system_error &) {
          break;  // the workers build what no builder takes
        }
      }
      S->nbuilders = S->builders_alive = (int)S->builders.size();
    }
    if (S->workers.empty()) {
      {
        std::unique_lock<std::mutex> lk(S->mu);
        S->stop = true;
      }
      S->cv_work.notify_all();
      for (auto &t : S->builders) t.join();
      (void)hipStreamDestroy(S->copy);
      delete S;
      throw Error(AINV_ERR_INTERNAL, "cannot start a server worker thread");
    }
    *out = S;
  });
}

int ainv_server_add_matrix(ainv_server_h S, int32_t n, const int64_t *irp,
                           const int32_t *ja, const double *val,
                           int index_base, int kind, const char *params,
                           int32_t *mat_id) {
  return guard([&] {
    if (!S || !mat_id) throw Error(AINV_ERR_ARG, "NULL argument");
    if (kind != AINV_KIND_AINV && kind != AINV_KIND_INVK &&
        kind != AINV_KIND_INVT && kind != AINV_KIND_ML)
      throw Error(AINV_ERR_ARG, "unknown solver kind");
    auto m = std::make_unique<ServedMatrix>();
    m->a = canonical_csr_public(n, irp, ja, val, index_base);
    m->kind = kind;
    for (auto &kv : parse_params(params)) {
      std::string k = kv.first;
      for (char &c : k) c = (char)std::toupper((unsigned char)c);
      if (k == "METHOD") {
        std::string v = kv.second;
        for (char &c : v) c = (char)std::toupper((unsigned char)c);
        m->method = krylov_method_id(v);
      } else if (k == "SLOTS") {
        const int v = std::atoi(kv.second.c_str());
        if (v < 1 || v > kMaxCols)
          throw Error(AINV_ERR_ARG, "SLOTS must be 1..64");
        m->min_slots = v;
      } else {
        m->params.emplace_back(k, kv.second);
      }
    }
    if (kind == AINV_KIND_ML && m->method == KM_BICG)
      throw Error(AINV_ERR_ARG,
                  "BICG applies M^T, which the multilevel preconditioner does "
                  "not offer");
    std::unique_lock<std::mutex> lk(S->mu);
    m->id = S->next_mat++;
    *mat_id = m->id;
    S->mats[m->id] = std::move(m);
  });
}

int ainv_server_submit(ainv_server_h S, int32_t mat_id, int k, const double *b,
                       double eps, int itmax, int64_t *req_id) {
  return guard([&] {
    if (!S || !b || !req_id) throw Error(AINV_ERR_ARG, "NULL argument");
    if (k < 1 || k > kMaxCols) throw Error(AINV_ERR_ARG, "k must be 1..64");
    if (itmax < 1) throw Error(AINV_ERR_ARG, "itmax must be >= 1");
    int32_t n = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->alive == 0)
        throw Error(AINV_ERR_HIP, "no server worker is running");
      auto it = S->mats.find(mat_id);
      if (it == S->mats.end()) throw Error(AINV_ERR_ARG, "unknown matrix id");
      n = it->second->a.n;
    }
    auto r = std::make_unique<Request>();
    r->n = n;
    r->mat = mat_id;
    r->k = k;
    r->eps = eps;
    r->itmax = itmax;
    // upload outside the lock: b is the caller's until this call returns
    AINV_HIP_CHECK(hipSetDevice(S->device));
    const size_t bytes = (size_t)n * k * sizeof(double);
    hipEvent_t up = nullptr;
    {
      std::lock_guard<std::mutex> g(S->hip_mu);
      AINV_HIP_CHECK(hipMallocAsync((void **)&r->dev_b, bytes, S->copy));
      AINV_HIP_CHECK(
          hipMemcpyAsync(r->dev_b, b, bytes, hipMemcpyHostToDevice, S->copy));
      AINV_HIP_CHECK(hipEventCreateWithFlags(&up, hipEventDisableTiming));
      AINV_HIP_CHECK(hipEventRecord(up, S->copy));
    }
    const hipError_t e = hipEventSynchronize(up);
    (void)hipEventDestroy(up);
    AINV_HIP_CHECK(e);
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->alive == 0) {  // the last worker failed during the upload
      lk.unlock();
      std::lock_guard<std::mutex> g(S->hip_mu);
      (void)hipFreeAsync(r->dev_b, S->copy);
      throw Error(AINV_ERR_HIP, "no server worker is running");
    }
    auto it = S->mats.find(mat_id);
    r->id = S->next_req++;
    r->t_submit = now_seconds();
    *req_id = r->id;
    it->second->pending.push_back(r->id);
    S->reqs[r->id] = std::move(r);
    S->cv_work.notify_all();  // a worker or, for a matrix to build, a builder
  });
}

namespace {
// Device buffers and event of a request, released on the copy stream after
// everything queued on it (the download) is done.
void release_request(ainv_server *S, Request *r) {
  if (r->dev_b) (void)hipFreeAsync(r->dev_b, S->copy);
  if (r->dev_x) (void)hipFreeAsync(r->dev_x, S->copy);
  if (r->ready) (void)hipEventDestroy(r->ready);
  r->dev_b = r->dev_x = nullptr;
  r->ready = nullptr;
}
}  // namespace

int ainv_server_wait(ainv_server_h S, int64_t req_id, double *x,
                     int32_t *iters, double *errs, double *latency_ms) {
  return guard([&] {
    if (!S) throw Error(AINV_ERR_ARG, "NULL server");
    std::unique_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto it = S->reqs.find(req_id);
      if (it == S->reqs.end()) throw Error(AINV_ERR_ARG, "unknown request id");
      Request *rp = it->second.get();
      S->cv_done.wait(lk, [&] { return rp->done; });
      r = std::move(it->second);
      S->reqs.erase(it);
    }
    AINV_HIP_CHECK(hipSetDevice(S->device));
    if (r->status != AINV_SUCCESS) {
      // a failed batch may have left work on its stream that reads dev_b
      std::lock_guard<std::mutex> g(S->hip_mu);
      (void)hipDeviceSynchronize();
      release_request(S, r.get());
      throw Error(r->status, "request failed: " + r->msg);
    }
    // download outside the server lock, after the worker's extraction; the
    // buffers are freed on the copy stream behind that point in any case
    hipError_t e = hipSuccess;
    hipEvent_t down = nullptr;
    {
      std::lock_guard<std::mutex> g(S->hip_mu);
      if (r->ready) e = hipStreamWaitEvent(S->copy, r->ready, 0);
      if (e == hipSuccess && x && r->dev_x)
        e = hipMemcpyAsync(x, r->dev_x, (size_t)r->n * r->k * sizeof(double),
                           hipMemcpyDeviceToHost, S->copy);
      if (e == hipSuccess)
        e = hipEventCreateWithFlags(&down, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(down, S->copy);
      release_request(S, r.get());
    }
    if (e == hipSuccess) e = hipEventSynchronize(down);
    if (down) (void)hipEventDestroy(down);
    AINV_HIP_CHECK(e);
    if (iters) std::memcpy(iters, r->iters.data(), r->k * sizeof(int32_t));
    if (errs) std::memcpy(errs, r->errs.data(), r->k * sizeof(double));
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  });
}

int ainv_server_stats(ainv_server_h S, char *buf, size_t len) {
  return guard([&] {
    if (!S || !buf || len == 0) throw Error(AINV_ERR_ARG, "NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    int resident_mats = 0;
    for (auto &kv : S->mats) resident_mats += kv.second->A != nullptr;
    std::snprintf(buf, len,
                  "{\"matrices\": %zu, \"resident_matrices\": %d, "
                  "\"resident_bytes\": %lld, \"peak_resident_bytes\": %lld, "
                  "\"budget_bytes\": %lld, \"builds\": %lld, "
                  "\"build_seconds\": %.6f, \"build_overlap_seconds\": %.6f, "
                  "\"batch_busy_seconds\": %.6f, \"evictions\": %lld, "
                  "\"completed\": %lld, \"batches\": %lld, \"steps\": %lld, "
                  "\"admitted_midflight\": %lld, \"workers\": %zu, "
                  "\"builders\": %zu, "
                  "\"open_seconds\": %.6f, \"admit_seconds\": %.6f, "
                  "\"step_seconds\": %.6f, \"extract_seconds\": %.6f, "
                  "\"queue_seconds\": %.6f, \"service_seconds\": %.6f}",
                  S->mats.size(), resident_mats, (long long)S->resident,
                  (long long)S->peak_resident, (long long)S->budget,
                  (long long)S->builds, S->build_seconds,
                  S->build_overlap_seconds, busy_now_locked(S),
                  (long long)S->evictions, (long long)S->completed,
                  (long long)S->batches, (long long)S->steps,
                  (long long)S->admitted_midflight, S->workers.size(),
                  S->builders.size(),
                  S->open_seconds, S->admit_seconds, S->step_seconds,
                  S->extract_seconds, S->queue_seconds, S->service_seconds);
  });
}

int ainv_server_free(ainv_server_h S) {
  return guard([&] {
    if (!S) return;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &t : S->builders) t.join();
    (void)hipSetDevice(S->device);
    (void)hipDeviceSynchronize();  // the workers' streams are done
    {
      std::lock_guard<std::mutex> g(S->hip_mu);
      for (auto &kv : S->reqs) release_request(S, kv.second.get());
    }
    if (S->copy) {
      (void)hipStreamSynchronize(S->copy);
      (void)hipStreamDestroy(S->copy);
    }
    for (auto &kv : S->mats) {
      if (kv.second->M) ainv_free(kv.second->M);
      if (kv.second->A) ainv_mat_free(kv.second->A);
    }
    delete S;
  });
}

}  // extern "C"
