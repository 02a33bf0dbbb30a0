// Solver object (the reference's `sv`) and system matrix, shared by the C
// ABI (capi.cpp) and the Krylov engine (krylov.hip).
#pragma once

#include <memory>
#include <mutex>

#include "ainv_internal.h"

namespace ainv {
// Process-wide counter: every (re)build, re-mold or release of a solver's
// factors takes a new value, so cached Krylov plans (which hold device
// addresses of the factors inside captured graphs) can tell they are stale.
uint64_t next_generation();
struct KrylovPlan;
void destroy_plan(KrylovPlan *p);
struct PlanDeleter {
  void operator()(KrylovPlan *p) const { destroy_plan(p); }
};
// Multilevel hierarchy of a solver of kind AINV_KIND_ML (ml.hip)
struct MlHier;
void destroy_ml(MlHier *m);
struct MlDeleter {
  void operator()(MlHier *m) const { destroy_ml(m); }
};
}  // namespace ainv

struct ainv_solver {
  // ---- parameters (mld_d_ainv_solver.F90:46-67, mld_d_invk_solver.f90:
  // 47-57, mld_d_invt_solver.f90:46-58)
  int kind = AINV_KIND_AINV;
  int alg = AINV_ALG_LLK;
  int32_t fill_in = 0;
  int32_t inv_fill = 0;
  double thresh = 0.0;
  double inv_thresh = 0.0;
  int prec = AINV_PREC_F64;
  // kernel mapping of the apply: -1 autotuned, else AINV_MAP_SPLIT / ROW /
  // COOP / COOP128 for both passes
  int map_mode = -1;
  // slice-group order of the run mapping's passes over W and Z (period and
  // band in groups of 4 slices; 0: chosen by a timed trial at the first
  // Krylov plan of 32 or more columns; period -1: linear), whether that
  // trial has run on the current layout, and its timings (JSON)
  int32_t order_period = 0, order_band = 0;
  // row sort of the factor layouts (Hll::perm): window in rows (a multiple
  // of 64), 0 rows in order, kSortAuto picked from the row lengths
  int32_t slice_sort = ainv::kSortAuto;
  bool order_tuned = false;
  std::string order_report;
  int host_threads = 0;  // host threads of the AINV build (0: the default)
  ainv::AinvBuildInfo bicg;  // how the last AINV build's biconjugation ran
  // multilevel parameters (kind AINV_KIND_ML); the factor parameters above
  // apply to the factors of every level
  int ml_nlev = 3;
  int ml_type = AINV_ML_MULT;
  int ml_pos = AINV_ML_POST;
  int ml_sweeps = 1;
  int ml_csweeps = 15;
  int ml_smoothed = 1;
  double ml_thresh = 0.1;
  double ml_scale = 1.0;
  int ml_sub = AINV_KIND_AINV;
  int ml_coarse = AINV_KIND_AINV;

  // ---- device / stream
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // one call at a time per handle

  // ---- built data
  uint64_t generation = 0;
  bool built = false;
  int32_t n = 0;
  ainv::DevCsr w, z;  // FP64 factors: stored W and Z
  ainv::DevBuf d64;   // FP64 inverted pivots
  ainv::Hll hw, hz;   // apply layout in storage precision
  ainv::DevBuf dv;    // d in storage precision (FP64 or FP32)
  bool have_t = false;
  ainv::Hll hwt, hzt;  // transposed layout for trans = 'T'
  ainv::DevBuf tbuf;   // intermediate t = d o (W x)
  ainv::DevBuf stage_x, stage_y;  // host-pointer staging

  // ---- autotuned kernel mapping per block width K = 2^i (-1: not tuned)
  // (packed pair: first and second pass, map_pair)
  int16_t apply_map[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  double apply_ms[8][2][ainv::kNumMaps] = {};  // [K][pass][map]
  // ---- served block width chosen by a timed trial (krylov_stream with
  // slots = 0): per Krylov method, the slot count (0: not tuned), the
  // factors' generation it was tuned on, and ms per column-iteration of
  // each candidate width K = 2^i
  int auto_slots[8] = {};
  uint64_t auto_slots_gen[8] = {};
  double slot_ms[8][8] = {};  // [method][log2 K]
  // the last krylov_stream on this preconditioner: steps launched, slot
  // steps (the batch width summed over steps) and the slot steps of
  // admitted columns, narrowings, and host seconds spent launching steps,
  // admitting, waiting for column states and extracting solutions
  struct StreamStats {
    int64_t steps = 0, slot_steps = 0, busy_slot_steps = 0, shrinks = 0;
    double wall = 0, launch = 0, admit = 0, wait = 0, extract = 0;
  } last_stream;

  // ---- multilevel hierarchy (kind AINV_KIND_ML)
  std::unique_ptr<ainv::MlHier, ainv::MlDeleter> ml;

  // ---- build report
  double t_factor = 0, t_inverse = 0, t_layout = 0, t_total = 0;
  std::string inv_report;
};

struct ainv_matrix {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t n = 0;
  ainv::DevCsr csr;
  ainv::Hll hll;
  int32_t slice_sort = ainv::kSortAuto;  // row sort of hll and hll_t
  bool have_t = false;  // A^T layout (BiCG), built on first use
  ainv::Hll hll_t;
  int8_t spmv_map[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  double spmv_ms[8][ainv::kNumMaps] = {};
  // Krylov workspaces and captured iteration graphs, most recent first
  std::vector<std::unique_ptr<ainv::KrylovPlan, ainv::PlanDeleter>> plans;
};

namespace ainv {

// y = alpha * M x + beta * y with M the preconditioner of h (device
// pointers, interleaved k columns).  Used by the C ABI and the Krylov loop.
void solver_apply(ainv_solver *h, double alpha, const double *x, double beta,
                  double *y, char trans, int k, int64_t ldx, int64_t ldy,
                  hipStream_t s);
// y = M x ('N', beta = 0, x, t and y with leading dimension ld) with a
// caller-owned intermediate buffer t of n*ld words of the factor's
// intermediate type: no allocation, so it may be captured in a graph.  map:
// packed pair of the two passes' mappings (map_pair).
void solver_apply_into(const ainv_solver *h, const double *x, double *y,
                       int k, int64_t ld, void *t, hipStream_t s,
                       int map = 0, const int32_t *kw = nullptr);
// Transposed layout for 'T' applies (built on first use; call before any
// graph capture) and y = M^T x into a caller buffer t, like
// solver_apply_into.
void solver_ensure_transposed(ainv_solver *h);
void solver_apply_into_t(const ainv_solver *h, const double *x, double *y,
                         int k, int64_t ld, void *t, hipStream_t s, int map,
                         const int32_t *kw = nullptr);
// Index of K = 2^i and the mapping pair the tuner chose for it (split if
// none, the forced mapping for both passes under APPLY_MAP).
int log2_pow2(int K);
// The device address of this GPU's scratch array of the no-allocation
// apply, taken once per device (noalloc.hip); the device must be current.
void *noalloc_setup(int device);
int tuned_apply_map(const ainv_solver *h, int K);
void reset_tuning(ainv_solver *h);
size_t solver_t_bytes(const ainv_solver *h, int k);
// Builds the factors of h (its kind and parameters) for a (capi.cpp).
void build_solver(ainv_solver *h, const HostCsr &a);
int64_t solver_device_bytes(const ainv_solver *h);

// ---- multilevel preconditioner (ml.hip) ------------------------------------
// Builds the smoothed-aggregation hierarchy of h (kind AINV_KIND_ML) for a:
// level matrices, prolongators, restrictions and the factors of each level.
void ml_build(ainv_solver *h, const HostCsr &a);
// Device workspace of one cycle over k columns (leading dimension k).
size_t ml_ws_bytes(const ainv_solver *h, int k);
// y = V-cycle(x): x, y with leading dimension K = pow2 >= k (16-byte
// aligned), ws of ml_ws_bytes(h, K) bytes; no allocation, stream-ordered, so
// it may be captured in a graph.  kw as for the apply kernels.
void ml_apply_into(const ainv_solver *h, const double *x, double *y, int k,
                   int64_t K, void *ws, hipStream_t s,
                   const int32_t *kw = nullptr);
// y = alpha * V-cycle(x) + beta * y for the C ABI (any leading dimensions;
// trans 'N' only), workspace in h->tbuf.
void ml_apply_abi(ainv_solver *h, double alpha, const double *x, double beta,
                  double *y, char trans, int k, int64_t ldx, int64_t ldy,
                  hipStream_t s);
// The same cycle on FP32 vectors (FP32 / BF16 factors: the level matrices
// are then FP32 too).
void ml_apply_into32(const ainv_solver *h, const float *x, float *y, int k,
                     int64_t K, void *ws, hipStream_t s,
                     const int32_t *kw = nullptr);
// Times the kernel mappings of every level's products for block width K and
// keeps the fastest; x0 (read) and y0 (overwritten) are n x K vectors of the
// finest level, FP32 if x32 (the cycle the caller will run), ws the cycle's
// workspace (cleared first).  Before a capture.
void ml_tune(ainv_solver *h, int K, int k, const void *x0, void *y0,
             bool x32, void *ws, hipStream_t s);
// One coarsening step on the host alone: aggregates, prolongator P and
// Galerkin matrix R (A P), on `threads` threads (0: the default count).
void ml_host_level(const HostCsr &a, double theta, bool smoothed, int threads,
                   std::vector<int32_t> &agg, int32_t &nagg, HostCsr &p,
                   HostCsr &ac, double &omega);
// Re-molds the factors of every level to another storage precision.
void ml_remold(ainv_solver *h, int prec);
// A new APPLY_MAP for every level (tuning forgotten).
void ml_set_map_mode(ainv_solver *h, int mode);
// Storage precision change of a built solver's factors (capi.cpp).
void solver_remold(ainv_solver *h, int prec);
int ml_levels(const ainv_solver *h);
int64_t ml_device_bytes(const ainv_solver *h);
int64_t ml_nzeros(const ainv_solver *h);
std::string ml_report(const ainv_solver *h);
// Host copies of level l: 'A', 'P', 'R' (CSR), 'W', 'Z' (factors, CSR),
// 'D' (pivots) and 'G' (aggregates of the rows); ncols of the CSR ones.
void ml_export(const ainv_solver *h, int l, char which, HostCsr &m,
               int32_t &ncols, std::vector<double> &dv,
               std::vector<int32_t> &iv);
const ainv_solver *ml_level_solver(const ainv_solver *h, int l);

// Continuous batching over K column slots of one (matrix, preconditioner):
// admit right-hand sides into free slots between iterations, step one
// captured iteration at a time, read the per-column states one step behind
// the device, extract finished columns.
struct SlotBatch;
struct SlotState {
  bool valid = false;
  int32_t active = 0, iter = 0;
  double err = 0.0;
};
// Krylov methods of the engine (psb_krylov's names)
enum {
  KM_BICGSTAB = 0,
  KM_CG = 1,
  KM_CGS = 2,
  KM_BICG = 3,
  KM_GMRES = 4,
  KM_BICGSTABL = 5
};
int krylov_method_id(const std::string &name);  // throws for others
SlotBatch *batch_open(ainv_matrix *a, ainv_solver *m, int method, int K,
                      hipStream_t s);
void batch_close(SlotBatch *b);
// Drops the matrix's Krylov plans (their captured graphs and workspaces) and
// its tuned product mappings, after a change of its layout; throws
// AINV_ERR_STATE while a slot batch holds a plan.
void matrix_drop_plans(ainv_matrix *a);
// src: device, n x kc interleaved; x0 = 0
void batch_admit(SlotBatch *b, int c0, int kc, const double *src, double eps);
// one right-hand side (a contiguous device n-vector) per listed slot
void batch_admit_cols(SlotBatch *b, const std::vector<int> &slots,
                      const std::vector<const double *> &srcs, double eps);
void batch_step(SlotBatch *b);
// Index of the next step batch_step will launch (0 for the first).
int64_t batch_next_step(const SlotBatch *b);
// Fills out with the states after step `which` and returns `which` (-1 and
// no valid state before the first step).  Columns admitted at or after that
// step still show the states of their slots' previous occupants.
int64_t batch_poll(SlotBatch *b, bool wait_last, std::vector<SlotState> &out);
// dst: device, n x kc interleaved
void batch_extract(SlotBatch *b, int c0, int kc, double *dst);
// slots[j] to the contiguous device n-vector dsts[j], one launch
void batch_extract_cols(SlotBatch *b, const std::vector<int> &slots,
                        const std::vector<double *> &dsts);

// Continuous batching over a device-resident queue: b and x hold nrhs
// n-vectors one after another; `slots` columns run at a time and a slot is
// refilled from the queue as soon as its column stops.  iter, err and
// lat_ms (admission to completion seen by the host) are per right-hand side.
void krylov_stream(const std::string &method, ainv_matrix *a, ainv_solver *m,
                   const double *b, double *x, int nrhs, int slots, double eps,
                   int itmax, int32_t *iter, double *err, double *lat_ms,
                   hipStream_t s);

// flags: AINV_KRYLOV_EAGER disables graph capture (used to check replays)
// irst: restart length of RGMRES (<= 0: 10) and l of BICGSTABL (<= 0: 2),
// ignored by the other methods
void krylov_solve(const std::string &method, ainv_matrix *a, ainv_solver *m,
                  const double *b, double *x, int k, double eps, int itmax,
                  int32_t *iter, double *err, hipStream_t s, int flags,
                  int irst = 0);

}  // namespace ainv
