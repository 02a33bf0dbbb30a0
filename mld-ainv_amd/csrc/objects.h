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
  // Krylov workspaces and captured iteration graphs, most recent first
  std::vector<std::unique_ptr<ainv::KrylovPlan, ainv::PlanDeleter>> plans;
};

namespace ainv {

// y = alpha * M x + beta * y with M the preconditioner of h (device
// pointers, interleaved k columns).  Used by the C ABI and the Krylov loop.
void solver_apply(ainv_solver *h, double alpha, const double *x, double beta,
                  double *y, char trans, int k, int64_t ldx, int64_t ldy,
                  hipStream_t s);
// y = M x ('N', beta = 0, ld = k) with a caller-owned intermediate buffer t
// of n*k words of the factor's intermediate type: no allocation, so it may
// be captured in a graph.
void solver_apply_into(const ainv_solver *h, const double *x, double *y,
                       int k, void *t, hipStream_t s);
size_t solver_t_bytes(const ainv_solver *h, int k);

// flags: AINV_KRYLOV_EAGER disables graph capture (used to check replays)
void krylov_solve(const std::string &method, ainv_matrix *a, ainv_solver *m,
                  const double *b, double *x, int k, double eps, int itmax,
                  int32_t *iter, double *err, hipStream_t s, int flags);

}  // namespace ainv
