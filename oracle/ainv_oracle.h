/*
 * ainv_oracle.h -- CPU restatement of the approximate-inverse preconditioner
 * algorithms of the MLD2P4 AINV plugin (AINV / INVK / INVT), FP64, single
 * thread, reference loop order.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it.  The
 * product library (mld-ainv_amd/) never links, loads or calls it.
 *
 * Conventions: 0-based indices, CSR/CSC with int64 pointers and int32
 * indices.  Every routine cites the reference file:line (1-based Fortran)
 * it restates.  Outputs are malloc'ed and released with or_free().
 */
#ifndef AINV_ORACLE_H
#define AINV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes (mirror the reference's categories) */
#define OR_OK 0
#define OR_ERR_ARG 1      /* invalid argument (psb_err_iarg_invalid_i_) */
#define OR_ERR_ALG 2      /* invalid alg (biconjg_mod.F90:119-122) */
#define OR_ERR_ALLOC 3
#define OR_ERR_INTERNAL 4 /* e.g. non-upper entry inside the row inverse */
#define OR_ERR_PIVOT 5    /* zero pivot in ILU */

/* algorithm ids for biconjugation; names from mld_d_ainv_solver.F90:240-257 */
#define OR_ALG_LLK 0
#define OR_ALG_S_LLK 1
#define OR_ALG_S_FT_LLK 2 /* 'STAB-LLK' */
#define OR_ALG_LLK_NOTH 3
#define OR_ALG_MLK 4

/* A sparse matrix handed out by the oracle (CSR unless stated otherwise). */
typedef struct {
  int32_t n;
  int64_t nnz;
  int64_t *ptr;
  int32_t *idx;
  double *val;
} or_mat;

void or_free(void *p);
void or_mat_free(or_mat *m);

/* sparsify (mld_d_sparsify.f90:73-143), stand-alone form without iheap
 * (:107-116): scans zw[istart..n), keeps idiag plus at most nzrmax-1 entries
 * with |z| >= thresh, largest |z| first (ties: smaller index first), output
 * sorted by index.  zw is zeroed on return.  Returns the count. */
int32_t or_sparsify(int32_t idiag, int32_t nzrmax, double thresh, int32_t n,
                    double *zw, int32_t istart, int32_t *iz, double *valz);

/* Sparse biconjugation (mld_dsparse_biconjg_llk.F90, _s_llk.F90,
 * _s_ft_llk.F90, _llk_noth.F90, _mlk.F90).  Input A CSR (already scaled).
 * Output Z (CSC, n x n, unit upper), W (CSC), p. */
int or_biconjg(int alg, const or_mat *a, int32_t nzrmax, double thresh,
               double *p, or_mat *zcsc, or_mat *wcsc);

/* AINV factor build (mld_d_ainv_bld.F90:115-182, maxval scaling as
 * requested by mld_d_ainv_solver_bld.f90:68-69).  Outputs CSR: w = W^T
 * (unit lower), z = Z (unit upper), d (inverted pivots, unscaled). */
int or_ainv_bld(const or_mat *a, int alg, int32_t fill_in, double thresh,
                or_mat *w, or_mat *z, double *d);

/* ILU(k) (MLD2P4 mld_iluk_fact restated; unit L strictly lower, unit U
 * strictly upper scaled by inverted pivots, d = inverted pivots). */
int or_iluk(const or_mat *a, int32_t fill, or_mat *l, or_mat *u, double *d);
/* ILUT(fill, thresh) with maxval scaling (MLD2P4 mld_ilut_fact restated). */
int or_ilut(const or_mat *a, int32_t fill, double thresh, or_mat *l,
            or_mat *u, double *d);

/* Row-wise sparse inverse of a unit upper U given strictly upper
 * (mld_dsparse_invk.f90:66-131 / mld_dsparse_invt.f90:66-128). Output CSR
 * with the diagonal stored. */
int or_sparse_invk(const or_mat *u, int32_t fill_in, or_mat *z);
int or_sparse_invt(const or_mat *u, int32_t fill_in, double thresh,
                   or_mat *z);

/* INVK / INVT builds (mld_d_invk_bld.f90:104-124, mld_d_invt_bld.f90:
 * 108-136): w = L^{-1} (stored, lower incl. diagonal), z = U^{-1}, d. */
int or_invk_bld(const or_mat *a, int32_t fill1, int32_t fill2, or_mat *w,
                or_mat *z, double *d);
int or_invt_bld(const or_mat *a, int32_t fill, double thresh, int32_t invfill,
                double invthresh, or_mat *w, or_mat *z, double *d);

/* Transpose of a CSR matrix (psb transp + cscnv). */
int or_transpose(const or_mat *a, or_mat *t);

/* Threads for the row-parallel loops (products 'N', dots, vector updates of
 * BiCGSTAB and CG); default 1, the reference's single-rank order. */
void or_set_threads(int t);
/* Compiler and flags this library was built with. */
const char *or_build_info(void);

/* y = alpha*A*x + beta*y (trans 'N') or alpha*A^T*x + beta*y ('T'). */
void or_spmv(char trans, double alpha, const or_mat *a, const double *x,
             double beta, double *y);

/* sv%update_a (mld_d_base_ainv_update_a.f90:53-79) for the shifted
 * operator A + diag(x): ee_i = sum_k W(i,k) * (x_k * Z(k,i)) over the common
 * indices of row i of the stored W and column i of Z, in ascending k
 * (psb_spdot_srtd), then d_i = 1 / (1/d_i + ee_i), in place.  Rows of w
 * and z sorted. */
int or_update_a(const or_mat *w, const or_mat *z, const double *x, double *d);

/* Preconditioner apply (mld_d_base_ainv_solver_apply_vect.f90:112-127):
 * 'N': y = alpha*Z*(d o (W*x)) + beta*y;  'T': y = alpha*W^T*(d o (Z^T x))
 * + beta*y.  Returns OR_ERR_ARG for any other trans (:63-69). */
int or_apply(char trans, double alpha, const or_mat *w, const or_mat *z,
             const double *d, const double *x, double beta, double *y);

/* Preconditioned Krylov solvers with the istop=2 test ||r||/||b|| <= eps,
 * restating PSBLAS psb_dbicgstab / psb_dcg (called from ppde3d.f90:280).
 * The preconditioner is (w, z, d) applied with trans 'N'.  Returns the
 * iteration count in *iter (-1 on breakdown) and the last estimate. */
int or_bicgstab(const or_mat *a, const or_mat *w, const or_mat *z,
                const double *d, const double *b, double *x, double eps,
                int32_t itmax, int32_t *iter, double *err);
/* CGS (Sonneveld), preconditioned as in Barrett et al., Templates (1994),
 * fig. 2.9, with the istop = 2 test after each full step. */
int or_cgs(const or_mat *a, const or_mat *w, const or_mat *z, const double *d,
           const double *b, double *x, double eps, int32_t itmax,
           int32_t *iter, double *err);
/* Restarted GMRES(irst) with right preconditioning (Saad, Iterative Methods
 * for Sparse Linear Systems, 2nd ed., alg. 9.5), the Arnoldi basis by
 * modified Gram-Schmidt (alg. 6.2) and the Hessenberg least-squares problem
 * by Givens rotations (sec. 6.5.3); psb_krylov's 'RGMRES' with restart
 * irst (tests/fileread/runs/dfs.inp:10).  One iteration = one Arnoldi step;
 * a step stops on the rotated residual |rs_{i+1}| <= eps ||b||; at the end
 * of a cycle x += M (V y) and, unless converged, the true residual starts
 * the next cycle (and is tested with the same istop = 2 rule). */
int or_gmres(const or_mat *a, const or_mat *w, const or_mat *z,
             const double *d, const double *b, double *x, double eps,
             int32_t itmax, int32_t irst, int32_t *iter, double *err);
/* BiCGStab(l) (Sleijpen and Fokkema, ETNA 1 (1993) 11-32, alg. 3.1) with
 * right preconditioning (operator A M, x = x0 + M yhat at exit); psb_krylov's
 * 'BICGSTABL' with l = irst (tests/pdegen/runs/mld3d.inp:1, :10).  A cycle
 * is l BiCG steps then the minimal-residual step by modified Gram-Schmidt;
 * it counts l iterations and the istop = 2 test follows each cycle. */
int or_bicgstabl(const or_mat *a, const or_mat *w, const or_mat *z,
                 const double *d, const double *b, double *x, double eps,
                 int32_t itmax, int32_t l, int32_t *iter, double *err);
/* BiCG, preconditioned as in Barrett et al., Templates (1994), fig. 2.8
 * (M^T = the trans 'T' apply), with the istop = 2 test. */
int or_bicg(const or_mat *a, const or_mat *w, const or_mat *z, const double *d,
            const double *b, double *x, double eps, int32_t itmax,
            int32_t *iter, double *err);
int or_cg(const or_mat *a, const or_mat *w, const or_mat *z, const double *d,
          const double *b, double *x, double eps, int32_t itmax,
          int32_t *iter, double *err);


/* ---- multilevel (smoothed aggregation; MLD2P4's published algorithm, the
 * framework itself being absent from the reference tree) ---------------- */
/* rectangular transpose: a has a->n rows and ncols columns */
int or_transpose_rect(const or_mat *a, int32_t ncols, or_mat *t);
/* greedy three-phase aggregation; agg[n] receives the aggregate of each
 * row, *nagg their count */
int or_aggregate(const or_mat *a, double theta, int32_t *agg, int32_t *nagg);
/* C = A B, B with bcols columns (Gustavson, stored order) */
int or_spgemm(const or_mat *a, const or_mat *b, int32_t bcols, or_mat *c);
/* one coarsening step: aggregates, prolongator P (n x nagg; smoothed or
 * tentative), R = P^T, Ac = R (A P); omega of the smoothing step */
int or_ml_level(const or_mat *a, double theta, int smoothed, int32_t *agg,
                int32_t *nagg, or_mat *p, or_mat *r, or_mat *ac,
                double *omega);
#define OR_ML_MULT 0
#define OR_ML_ADD 1
#define OR_ML_PRE 1
#define OR_ML_POST 2
#define OR_ML_TWOSIDE 3
typedef struct {
  int32_t nlev;
  const or_mat *a; /* [nlev] level matrices */
  const or_mat *p; /* [nlev - 1] prolongators */
  const or_mat *r; /* [nlev - 1] restrictions */
  const or_mat *w; /* [nlev] smoother / coarse factors: stored W */
  const or_mat *z; /* [nlev] Z */
  const double *const *d; /* [nlev] pivots */
  int32_t type;    /* OR_ML_MULT or OR_ML_ADD */
  int32_t pos;     /* OR_ML_PRE, OR_ML_POST or OR_ML_TWOSIDE (MULT) */
  int32_t sweeps;  /* smoother sweeps per visit */
  int32_t csweeps; /* coarsest-level sweeps */
} or_ml;
/* y = V-cycle(x) */
int or_ml_apply(const or_ml *ml, const double *x, double *y);
/* Krylov method (BICGSTAB, CG, CGS, RGMRES, BICGSTABL) preconditioned by
 * the cycle; irst as for or_gmres / or_bicgstabl */
int or_ml_krylov(const char *method, const or_mat *a, const or_ml *ml,
                 const double *b, double *x, double eps, int32_t itmax,
                 int32_t irst, int32_t *iter, double *err);

#ifdef __cplusplus
}
#endif
#endif
