/*
 * mld_ainv_hip.h -- C ABI of libmldainv_hip.so, the MI355X-native
 * implementation of the MLD2P4 approximate-inverse sub-solvers
 * (mld_d_ainv_solver, mld_d_invk_solver, mld_d_invt_solver).
 *
 * One handle = one solver object (the reference's `sv`).  Every entry point
 * returns an int info code (0 = success, AINV_ERR_* otherwise, mirroring the
 * PSBLAS error categories); ainv_last_error() returns the message of the
 * last failure on the calling thread.  Matrices cross the boundary as CSR
 * with int64 row pointers, int32 column indices and FP64 values, 0- or
 * 1-based (index_base).  Vectors handed to the apply / Krylov calls are
 * device pointers unless AINV_HOST_PTRS is set.  A block of k right-hand
 * sides is stored row-interleaved: entry (i, c) at x[i*ldx + c], ldx >= k.
 *
 * Reference interfaces replaced (files of sfilippone/mld-ainv):
 *   type-bound procedures of mld_d_base_ainv_solver_type
 *     src/mld_d_base_ainv_mod.F90:45-68 and its extensions
 *     src/mld_d_ainv_solver.F90:46-67, src/mld_d_invk_solver.f90:47-57,
 *     src/mld_d_invt_solver.f90:46-58.
 */
#ifndef MLD_AINV_HIP_H
#define MLD_AINV_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- info codes ---------------------------------------------------- */
#define AINV_SUCCESS 0
#define AINV_ERR_ARG 1      /* invalid argument (psb_err_iarg_invalid_i_) */
#define AINV_ERR_ALG 2      /* invalid algorithm variant */
#define AINV_ERR_ALLOC 3    /* allocation failure (psb_err_alloc_request_) */
#define AINV_ERR_INTERNAL 4 /* internal error (psb_err_internal_error_) */
#define AINV_ERR_PIVOT 5    /* zero pivot in the incomplete factorization */
#define AINV_ERR_STATE 6    /* call not valid in the object's state */
#define AINV_ERR_HIP 7      /* HIP runtime failure or no usable device */
#define AINV_ERR_CAPACITY 8 /* a working set exceeded the kernel capacity */

/* ---- solver kinds -------------------------------------------------- */
#define AINV_KIND_AINV 1 /* mld_d_ainv_solver_type (biconjugation) */
#define AINV_KIND_INVK 2 /* mld_d_invk_solver_type (inverse of ILU(k)) */
#define AINV_KIND_INVT 3 /* mld_d_invt_solver_type (inverse of ILUT) */
#define AINV_KIND_ML 4   /* multilevel preconditioner whose smoothers and
                            coarsest-level solver are AINV-family factors */

/* ---- AINV algorithm variants, in the order of the reference's ids
 * (src/mld_base_ainv_mod.F90:40-60; the absolute values there are offsets
 * from MLD2P4 constants, these are this library's own) ---------------- */
#define AINV_ALG_LLK 1      /* 'LLK'      left-looking */
#define AINV_ALG_S_LLK 2    /* 'SYM-LLK'  symmetric left-looking, W = Z */
#define AINV_ALG_S_FT_LLK 3 /* 'STAB-LLK' stabilized left-looking */
#define AINV_ALG_LLK_NOTH 4 /* left-looking without the alpha drop */
#define AINV_ALG_MLK 5      /* 'MLK'      left-looking, list merge */

/* ---- factor storage precision (the MI355X counterpart of the
 * reference's amold/vmold re-mold, mld_d_ainv_solver_bld.f90:71-82) ---- */
#define AINV_PREC_F64 0  /* FP64 values, FP64 d, FP64 intermediate */
#define AINV_PREC_F32 1  /* FP32 values and d, FP64 accumulation */
#define AINV_PREC_BF16 2 /* BF16 values, FP32 d, FP32 accumulation */

/* ---- multilevel preconditioner (kind AINV_KIND_ML) ----------------- */
/* The reference's solvers are MLD2P4 smoothers; its deck
 * tests/pdegen/runs/ppde.inp:10-29 runs them inside MLD2P4's multilevel
 * preconditioner (prec%init('ML')), a framework outside the reference tree
 * (MLD2P4 >= 2.1.1, configure.ac:283-289).  Kind AINV_KIND_ML is that
 * configuration on one process: smoothed aggregation (greedy three-phase
 * aggregation over the strength graph |a_ij| > theta sqrt(|a_ii a_jj|),
 * prolongator P = (I - omega D^-1 A) T with omega = 4 / (3 ||D^-1 A||_inf),
 * Galerkin R A P with R = P^T), AINV-family factors as the smoother of each
 * level (AS / block Jacobi with one block) and as the coarsest-level solver,
 * and a V-cycle.  Keys of the setters for this kind:
 *   ML_TYPE          'MULT' | 'ADD'                  (default MULT)
 *   SMOOTHER_POS     'PRE' | 'POST' | 'TWOSIDE'      (MULT; default POST)
 *   SMOOTHER_SWEEPS  sweeps of the smoother per visit (default 1)
 *   N_PREC_LEVS      levels, coarsest included       (default 3)
 *   AGGR_KIND        'SMOOTHED' | 'NONSMOOTHED'      (default SMOOTHED)
 *   AGGR_ALG         'DEC'                           (the only one)
 *   AGGR_THRESH      theta of the finest level       (default 0.10)
 *   AGGR_SCALE       theta of level l+1 = theta of level l * AGGR_SCALE
 *                    (default 1)
 *   SUB_SOLVE        'AINV' | 'INVK' | 'INVT': smoother factors (AINV)
 *   COARSE_SUBSOLVE  'AINV' | 'INVK' | 'INVT': coarsest-level factors (AINV)
 *   COARSE_SWEEPS    sweeps u += M (b - A u) of the coarsest level (15)
 * and the factor keys below (SUB_FILLIN, AINV_ALG, SUB_ILUTHRS, INV_FILLIN,
 * INV_THRESH, FACTOR_PREC, APPLY_MAP), which hold for the factors of every
 * level.  Coarsening stops before N_PREC_LEVS when a step no longer reduces
 * the size.  Applies only with trans 'N'; Krylov methods other than BICG. */
#define AINV_ML_MULT 0
#define AINV_ML_ADD 1
#define AINV_ML_PRE 1
#define AINV_ML_POST 2
#define AINV_ML_TWOSIDE 3

/* ---- flags --------------------------------------------------------- */
#define AINV_HOST_PTRS 1 /* x/y (or b/x) are host pointers */
#define AINV_KRYLOV_EAGER 2 /* Krylov: launch kernels one by one instead of
                               replaying the captured iteration graph */

typedef struct ainv_solver *ainv_h;
typedef struct ainv_matrix *ainv_mat_h;

/* Last error message of the calling thread ("" if none). */
const char *ainv_last_error(void);

/* Library / device set-up: selects the HIP device used by handles created
 * afterwards on this thread.  Returns AINV_ERR_HIP if no device is usable. */
int ainv_init(int device);

/* Create / destroy.  create also applies the defaults (d_*_solver_default,
 * mld_d_ainv_solver.F90:206-220, mld_d_invk_solver.f90:243-256,
 * mld_d_invt_solver.f90:258-272). free releases everything
 * (mld_d_base_ainv_solver_free.f90). */
int ainv_create(int kind, ainv_h *h);
int ainv_free(ainv_h h);
int ainv_default(ainv_h h);
int ainv_get_kind(ainv_h h, int *kind);

/* Setters keyed by case-insensitive strings (sv%cseti / sv%csetr /
 * sv%csetc: mld_d_ainv_solver_cseti.f90:55-62, _csetr.f90:55-60,
 * _csetc.f90:56-61, mld_d_invk_solver_cseti.f90:55-62,
 * mld_d_invt_solver_cseti.f90:55-62, mld_d_invt_solver_csetr.f90:55-62).
 * Keys: SUB_FILLIN, AINV_ALG, SUB_ILUTHRS, INV_FILLIN, INV_THRESH, plus
 * FACTOR_PREC ('F64'|'F32'|'BF16' or AINV_PREC_*) and APPLY_MAP
 * ('AUTO'|'SPLIT'|'ROW'|'COOP'|'COOP128'|'COOP_Q'|'COOP128_Q'|'RUN' or
 * -1/0/1/2/3/4/5/6: kernel mapping of block applies with a power-of-two
 * leading dimension, the _Q ones with four waves per 64-row slice, RUN the
 * one whose row groups hand gathered x rows on along runs of consecutive
 * diagonals and gather only the rows they do not hold; AUTO times each for
 * each of the two passes once per block width and keeps the fastest;
 * multilevel handles also take 'WIDE' or 7: one wavefront per row for the
 * level products that keep a CSR copy, COOP for the rest), ORDER_PERIOD and
 * ORDER_BAND (integers: the order in which the RUN mapping's workgroups take
 * their groups of 4 slices within each XCD's range -- strips of ORDER_BAND
 * groups across planes of ORDER_PERIOD groups; ORDER_PERIOD 0, the default,
 * lets the Krylov tuner time linear and strip orders once per layout,
 * -1 keeps linear order; results are the same in every order), SLICE_SORT
 * (integer: the row sort of the factor layouts as in
 * ainv_mat_set_slice_sort -- -1, the default, picks the window from the row
 * lengths, 0 keeps rows in order, else a window of that many rows, a
 * multiple of 64; set on a built handle it rebuilds the layouts; every
 * apply's result is the same in each).  Unknown
 * keys are passed to the base solver, which ignores them (info = 0). */
int ainv_cseti(ainv_h h, const char *what, int val);
int ainv_csetr(ainv_h h, const char *what, double val);
int ainv_csetc(ainv_h h, const char *what, const char *val);
/* Getters for the same keys (integer and real). */
int ainv_geti(ainv_h h, const char *what, int *val);
int ainv_getr(ainv_h h, const char *what, double *val);

/* sv%check: coerce illegal parameters to their defaults
 * (mld_d_ainv_solver_check.f90:53-56, mld_d_invk_solver_check.f90,
 * mld_d_invt_solver_check.f90). */
int ainv_check(ainv_h h);

/* sv%descr: text description into buf (NUL-terminated, truncated to len)
 * (mld_d_ainv_solver_descr.f90:63-66, mld_d_invk_solver_descr.f90:63-65,
 * mld_d_invt_solver_descr.f90:63-67). */
int ainv_descr(ainv_h h, char *buf, size_t len);

/* AINV variant name <-> id (d_ainv_stringval / d_ainv_algname,
 * mld_d_ainv_solver.F90:232-292). stringval returns -1 if unknown. */
int ainv_stringval(const char *name);
int ainv_algname(int alg, char *buf, size_t len);

/* sv%build (mld_d_ainv_solver_bld.f90:68-82 -> mld_d_ainv_bld.F90,
 * mld_d_invk_solver_bld.f90 -> mld_d_invk_bld.f90, mld_d_invt_solver_bld
 * .f90 -> mld_d_invt_bld.f90).  The n x n local block in host CSR; the
 * factors end up resident on the device in the storage precision set by
 * FACTOR_PREC.  Calls check first. */
int ainv_build_csr(ainv_h h, int32_t n, const int64_t *irp, const int32_t *ja,
                   const double *val, int index_base);

/* sv%cnv: re-mold the resident factors to another storage precision
 * (mld_d_base_ainv_solver_cnv.f90:56-60). */
int ainv_cnv(ainv_h h, int prec);

/* sv%apply_v / sv%apply_a (mld_d_base_ainv_solver_apply_vect.f90:110-143,
 * mld_d_base_ainv_solver_apply.f90:100-121):
 *   trans 'N': y = alpha * Z * (d o (W * x)) + beta * y
 *   trans 'T'/'C': y = alpha * W^T * (d o (Z^T * x)) + beta * y
 * for k right-hand sides (1 <= k <= 64) at once.  beta == 0 overwrites y
 * without reading it.  Asynchronous on `stream` (NULL = the handle's
 * stream) unless AINV_HOST_PTRS is set. */
int ainv_apply(ainv_h h, double alpha, const double *x, double beta, double *y,
               char trans, int k, int64_t ldx, int64_t ldy, int flags,
               void *stream);

/* Accessors (mld_d_base_ainv_mod.F90:179-214). */
int ainv_get_nzeros(ainv_h h, int64_t *nz);  /* n + nnz(W) + nnz(Z) */
int ainv_sizeof(ainv_h h, int64_t *bytes);   /* resident bytes */
int ainv_get_wrksz(ainv_h h, int *wrksz);    /* = 2 work vectors */
int ainv_get_n(ainv_h h, int32_t *n);

/* clone / clone_settings / clear_data (mld_d_ainv_solver_clone.f90:57-82,
 * _clone_settings.f90, mld_d_base_ainv_solver_clear_data.f90:52-62). */
int ainv_clone(ainv_h h, ainv_h *out);
int ainv_clone_settings(ainv_h h, ainv_h out);
int ainv_clear_data(ainv_h h);

/* Factor export for parity checks and dumps (the data sv%dump writes,
 * mld_d_base_ainv_solver_dmp.f90:77-106).  which: 'W' (stored W, i.e. W^T
 * for AINV and L^{-1} for INVK/INVT), 'Z', or 'D'.  Call with NULL arrays
 * to get n and nnz, then with arrays of that size (0-based, FP64). */
int ainv_export(ainv_h h, char which, int32_t *n, int64_t *nnz, int64_t *irp,
                int32_t *ja, double *val);

/* sv%dump (mld_d_base_ainv_solver_dmp.f90:56-106): with solver != 0 writes
 * the stored W, d and Z as MatrixMarket files named
 * <prefix>_p000_l<level>_{wmat,diag,zmat}.mtx (prefix NULL: "dump_ainv_d";
 * rank 000: one process per library instance), W and Z as 1-based
 * coordinate files, d as a dense array, values with 17 significant digits;
 * head (may be NULL) becomes a comment line.  solver == 0 writes nothing, as
 * in the reference. */
int ainv_dump(ainv_h h, const char *prefix, int level, int solver,
              const char *head);

/* Factor exchange: make h apply factors built elsewhere (for instance the
 * files of a reference sv%dump): stored W and Z as n x n CSR (index_base 0
 * or 1) and the inverted pivots d.  No counterpart in the reference, whose
 * factors only come from its own build; the kind and build parameters of h
 * are left as they are.  The factors are stored in h's FACTOR_PREC. */
int ainv_import_factors(ainv_h h, int32_t n, const int64_t *w_irp,
                        const int32_t *w_ja, const double *w_val,
                        const int64_t *z_irp, const int32_t *z_ja,
                        const double *z_val, const double *d, int index_base);

/* Multilevel hierarchy of a built AINV_KIND_ML handle (no counterpart in
 * the reference, whose hierarchy lives in MLD2P4): the number of levels, and
 * level l's (0 = finest) 'A' (level matrix), 'P' (prolongator, nrows x
 * ncols), 'R' (restriction P^T), 'W', 'Z' (factors) as 0-based CSR, 'D'
 * (pivots, nnz = nrows values) or 'G' (aggregate of each row, nnz = nrows
 * entries in ja).  Call with NULL arrays to get the sizes first. */
int ainv_ml_levels(ainv_h h, int *nlev);
int ainv_ml_export(ainv_h h, int level, char which, int32_t *nrows,
                   int32_t *ncols, int64_t *nnz, int64_t *irp, int32_t *ja,
                   double *val);

/* One coarsening step of the multilevel build on the host alone (no device
 * involved; no counterpart in the reference): aggregates agg[n] and their
 * count, the prolongator P (n x nagg, SMOOTHED if smoothed != 0, else the
 * tentative one) and the Galerkin matrix R (A P) (nagg x nagg), as 0-based
 * CSR with row pointers of n+1 (P) and nagg+1 (R A P) entries and at most
 * `cap` entries each (larger: AINV_ERR_ARG with the sizes in *p_nnz,
 * *c_nnz), and omega of the smoothing step; on `threads` host threads (0:
 * the default count), bit-identical results for any count. */
int ainv_ml_host_level(int32_t n, const int64_t *irp, const int32_t *ja,
                       const double *val, int index_base, double theta,
                       int smoothed, int threads, int64_t cap, int32_t *agg,
                       int32_t *nagg, int64_t *p_irp, int32_t *p_ja,
                       double *p_val, int64_t *p_nnz, int64_t *c_irp,
                       int32_t *c_ja, double *c_val, int64_t *c_nnz,
                       double *omega);

/* MatrixMarket readers (the input path of tests/fileread/df_sample.f90:
 * 156-175): a square coordinate file (real, integer or pattern; general,
 * symmetric or skew-symmetric) as 0-based CSR with sorted columns and
 * duplicates summed; a dense single-column array as a vector.  Call with
 * NULL arrays to get the sizes first. */
int ainv_mm_read(const char *path, int32_t *n, int64_t *nnz, int64_t *irp,
                 int32_t *ja, double *val);
int ainv_mm_read_vector(const char *path, int64_t *n, double *val);

/* The AINV factorization alone, on the host (no device involved):
 * mld_d_ainv_bld (src/impl/mld_d_ainv_bld.F90:35-194) with the
 * biconjugation variant alg (AINV_ALG_*), fill cap fill_in and drop
 * threshold thresh taken as given (no check coercion), on `threads` host
 * threads (0: AINV_BUILD_THREADS, else the CPUs available, at most 16;
 * bit-identical results for any count).  Outputs (0-based CSR, FP64): the
 * stored W (W^T, as the reference keeps it) and Z with row pointers of n+1
 * entries and at most `cap` entries each (larger: AINV_ERR_ARG with the
 * sizes in *w_nnz, *z_nnz), and the n entries of d. */
int ainv_host_ainv_bld(int32_t n, const int64_t *irp, const int32_t *ja,
                       const double *val, int index_base, int alg,
                       int32_t fill_in, double thresh, int threads,
                       int64_t cap, int64_t *w_irp, int32_t *w_ja,
                       double *w_val, int64_t *w_nnz, int64_t *z_irp,
                       int32_t *z_ja, double *z_val, int64_t *z_nnz,
                       double *d);

/* The incomplete factorization of the INVK / INVT builds alone, on the
 * host (no device involved): ILU(fill) when kind = AINV_KIND_INVK (the
 * mld_iluk_fact call of src/impl/mld_d_invk_bld.f90:104-105), ILUT(fill,
 * thresh) with maxval scaling when kind = AINV_KIND_INVT (the
 * mld_ilut_fact call of src/impl/mld_d_invt_bld.f90:108-109).  Outputs
 * (0-based CSR, FP64): L strictly lower (unit diagonal implied), U strictly
 * upper (unit diagonal implied, rows scaled by the inverted pivots), with
 * row pointers of n+1 entries and at most `cap` entries each (larger:
 * AINV_ERR_ARG with the sizes in *l_nnz, *u_nnz), and the n inverted
 * pivots d, so that A = L diag(1/d) (I + U).  A zero pivot returns
 * AINV_ERR_PIVOT. */
int ainv_host_ilu_fact(int32_t n, const int64_t *irp, const int32_t *ja,
                       const double *val, int index_base, int kind,
                       int32_t fill, double thresh, int64_t cap,
                       int64_t *l_irp, int32_t *l_ja, double *l_val,
                       int64_t *l_nnz, int64_t *u_irp, int32_t *u_ja,
                       double *u_val, int64_t *u_nnz, double *d);

/* sv%update_a (mld_d_base_ainv_update_a.f90:53-79): adapt the built
 * preconditioner to the shifted operator A + diag(x) by updating the
 * inverted pivots, d_i = 1 / (1/d_i + sum_k W(i,k) x_k Z(k,i)) (stored W by
 * rows, Z by columns); W and Z are kept.  x has nx >= n entries (fewer:
 * AINV_ERR_ARG, the reference's "Wrong size in update"); device pointer, or
 * host pointer with flags = AINV_HOST_PTRS. */
int ainv_update_a(ainv_h h, const double *x, int64_t nx, int flags);

/* The apply of ainv_apply without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as ainv_apply with device
 * pointers; the intermediate t of the two passes lives in a 64 MiB device
 * array the library holds once per GPU until the process ends (its address
 * taken at ainv_init or at the first call on a device), so n x K x 4 bytes
 * (FP32 / BF16 factors; x 8 with FP64 factors; K = k rounded up to a power
 * of two when ldx = ldy = K, else k) must fit in it (larger:
 * AINV_ERR_CAPACITY).  Refused (AINV_ERR_ARG / AINV_ERR_STATE) where
 * ainv_apply would allocate: host pointers, the multilevel kind, and a
 * 'T' / 'C' apply before ainv_apply has built the transposed layout.
 * Every call on one GPU shares its array: calls that overlap on a GPU
 * (other threads or streams) race and must be serialized by the caller. */
int ainv_apply_noalloc(ainv_h h, double alpha, const double *x, double beta,
                       double *y, char trans, int k, int64_t ldx, int64_t ldy,
                       void *stream);

/* Timing and layout report of the last build (JSON text). */
int ainv_report(ainv_h h, char *buf, size_t len);

/* ---- system matrix and Krylov driver (the PSBLAS psb_krylov call the
 * reference drivers make, tests/pdegen/ppde3d.f90:280) ----------------- */
int ainv_mat_create(int32_t n, const int64_t *irp, const int32_t *ja,
                    const double *val, int index_base, ainv_mat_h *a);
int ainv_mat_free(ainv_mat_h a);
/* Row sort of A's sliced layout: rows sorted by length (longest first,
 * ties in row order) within windows of `window` rows before they are cut
 * into 64-row slices, so that a slice pads little (SELL-C-sigma); window
 * -1 (the default at ainv_mat_create) picks the window from the row
 * lengths -- none when the rows in order pad by less than a quarter or the
 * sort saves less than a fifth, else the smallest of 256, 1024 and 4096
 * rows within 5 % of the best of them -- 0 keeps the rows in order, else a
 * multiple of 64.  Products are the same in every layout; dot products
 * over rows are summed in another order.  Rebuilds the layout and drops the
 * matrix's Krylov plans and tuned mappings (AINV_ERR_STATE while a served
 * stream holds one).  No counterpart in the reference, whose matrices stay
 * in PSBLAS CSR. */
int ainv_mat_set_slice_sort(ainv_mat_h a, int window);
/* Layout of A (JSON text): {"n", "slice_sort", "a": {"nnz", "padded"
 * (stored entries), "unsorted_padded" (those of the rows in order),
 * "sort_window", "slices", "max_width", "width_mean", "width_p50",
 * "width_p99" (slice widths in entries per row slot: the load balance of
 * the slices' waves)}, "a_t": the same for A^T once BiCG has built it}.
 * AINV_ERR_ARG when len is too small. */
int ainv_mat_report(ainv_mat_h a, char *buf, size_t len);
/* The row sort on the host, without a device: for the CSR row pointers irp
 * (n + 1 entries) and window as in ainv_mat_set_slice_sort, perm (may be
 * NULL; 64 * ceil(n / 64) entries) gets the stored order -- stored row i
 * holds row perm[i], n past the last row -- padded the stored entries of
 * that layout, unsorted_padded those of the rows in order and window_used
 * the window applied (0: none).  No counterpart in the reference. */
int ainv_slice_sort_plan(int32_t n, const int64_t *irp, int window,
                         int32_t *perm, int64_t *padded,
                         int64_t *unsorted_padded, int *window_used);
/* y = alpha*A*x + beta*y for k interleaved columns (device pointers). */
int ainv_mat_spmm(ainv_mat_h a, double alpha, const double *x, double beta,
                  double *y, int k, int64_t ldx, int64_t ldy, void *stream);

/* method: "BICGSTAB", "CG", "CGS", "BICG", "RGMRES" or "BICGSTABL"
 * (psb_krylov's names, the methods of tests/pdegen/runs/mld3d.inp:1; CGS
 * and BiCG in the preconditioned forms of Barrett et al., Templates, figs.
 * 2.9 and 2.8; BiCG applies A^T and M^T through transposed layouts built on
 * its first call; RGMRES is right-preconditioned GMRES restarted every irst
 * steps, one iteration per Arnoldi step, its in-cycle test on the rotated
 * residual; BICGSTABL is right-preconditioned BiCGStab(l) of Sleijpen and
 * Fokkema with l = irst, counting l iterations per cycle and testing after
 * each cycle; irst as psb_krylov's argument,
 * tests/fileread/df_sample.f90:368-369, <= 0 meaning 10 for RGMRES and 2
 * for BICGSTABL); istop = 2
 * (||r||_2 <= eps ||b||_2); x holds
 * x0 on entry.  k interleaved right-hand sides, each with its own
 * recurrence and stopping test; iter[c] and err[c] (host arrays of k) get
 * the iteration count and final ||r_c||/||b_c|| of column c.  prec may be
 * NULL (no preconditioner). */
int ainv_krylov(const char *method, ainv_mat_h a, ainv_h prec, const double *b,
                double *x, int k, double eps, int itmax, int istop, int irst,
                int32_t *iter, double *err, int flags, void *stream);

/* Served form of the same solve (no counterpart in the reference, whose
 * Krylov driver takes one right-hand side at a time): a device-resident
 * queue of nrhs right-hand sides b (n-vectors one after another) is solved
 * through `slots` (1..64) column slots of one captured Krylov
 * iteration (BICGSTAB, CG, CGS or BICG); a slot is refilled from the queue
 * between two iterations as soon as its column stops (continuous batching).
 * slots = 0 chooses the width: the first such call on a preconditioner
 * times six iterations at 8, 16, 32 and 64 slots (no wider than the queue
 * rounded up to a power of two) on the queue's first columns and keeps the
 * width with the least time per column-iteration for that preconditioner
 * and method until its factors change (ainv_report "served_slots"; with
 * prec NULL the trial runs on every call).  x (device, nrhs
 * n-vectors) gets the solutions, x0 = 0.  iter, err, lat_ms are host arrays
 * of nrhs: iterations, final ||r||/||b|| and milliseconds from admission to
 * completion (lat_ms may be NULL).  istop = 2 semantics as above. */
int ainv_krylov_stream(const char *method, ainv_mat_h a, ainv_h prec,
                       const double *b, double *x, int nrhs, int slots,
                       double eps, int itmax, int32_t *iter, double *err,
                       double *lat_ms, void *stream);

/* ---- synthetic matrices of the benchmark configurations (host CSR).
 * Call with NULL arrays to get n and nnz, then with allocated arrays. ---- */
/* 3-D convection-diffusion -div(grad u) + b.grad u, b = (1,1,1)/sqrt(3),
 * 7-point centred differences on an N^3 interior grid; hconv 0 uses
 * h = 1/(N+1), hconv 1 uses the reference generator's h = 1/(N+2)
 * (tests/pdegen/genpde_mod.f90:141, :330-389). */
int ainv_gen_pde3d(int32_t N, int hconv, int32_t *n, int64_t *nnz,
                   int64_t *irp, int32_t *ja, double *val);
/* 2-D anisotropic diffusion -(eps u_xx + u_yy), 5-point, h = 1/(N+1). */
int ainv_gen_aniso2d(int32_t N, double eps, int32_t *n, int64_t *nnz,
                     int64_t *irp, int32_t *ja, double *val);
/* Random irregular nonsymmetric matrix: off-diagonal counts
 * round(LogNormal(ln(median), sigma)) capped at cap, columns uniform in
 * [i-band, i+band] plus nlong uniform long-range columns, N(0,1) values,
 * diagonal = dom * row absolute sum. */
int ainv_gen_random(int32_t nrows, uint64_t seed, double median, double sigma,
                    int32_t cap, int32_t band, int32_t nlong, double dom,
                    int32_t *n, int64_t *nnz, int64_t *irp, int32_t *ja,
                    double *val);
/* n*k N(0,1) values from the library RNG (xoshiro256** seeded by
 * SplitMix64), so CPU and GPU paths see byte-identical inputs. */
int ainv_gen_normal(uint64_t seed, int64_t count, double *out);

/* ---- serving loop: requests on many matrices, continuous batching ------
 * A server keeps each registered matrix's factors resident on the device
 * after first touch, under an HBM budget with LRU eviction of idle matrices.
 * Worker threads (one HIP stream each) batch the pending right-hand sides of
 * one resident matrix into up to 64 column slots; a converged column's slot
 * is refilled by the next pending request between two Krylov iterations.
 * First-touch builds and rebuilds after an eviction run on builder threads
 * (as many as workers), so a build never holds a worker.
 * No counterpart in the reference (its drivers solve one system at a time);
 * serving is the deployment form of the apply + Krylov path. */
typedef struct ainv_server *ainv_server_h;
int ainv_server_create(int device, int64_t hbm_budget_bytes, int workers,
                       ainv_server_h *srv);
/* kind: AINV_KIND_*; params: "KEY=VALUE;..." with the keys of
 * ainv_cseti/csetr/csetc plus METHOD=BICGSTAB|CG|CGS|BICG and SLOTS=s (a
 * batch on this matrix opens with at least s column slots, rounded up to a
 * power of two; default: as many as the pending columns).  Host CSR is
 * copied. */
int ainv_server_add_matrix(ainv_server_h srv, int32_t n, const int64_t *irp,
                           const int32_t *ja, const double *val,
                           int index_base, int kind, const char *params,
                           int32_t *mat_id);
/* b: host, n x k interleaved (1 <= k <= 64), x0 = 0, istop = 2; copied to
 * device memory before the call returns (b may be reused at once). */
int ainv_server_submit(ainv_server_h srv, int32_t mat_id, int k,
                       const double *b, double eps, int itmax,
                       int64_t *req_id);
/* Blocks until the request completes and copies its solution from device
 * memory into x (host n x k; may be NULL); latency from submit to
 * completion on the device.  Releases the request. */
int ainv_server_wait(ainv_server_h srv, int64_t req_id, double *x,
                     int32_t *iters, double *errs, double *latency_ms);
/* Counters as JSON: builds, evictions, completions, batches, steps, host
 * seconds per phase; build_seconds (first-touch builds, summed over
 * workers), batch_busy_seconds (wall seconds with at least one worker inside
 * a slot batch) and build_overlap_seconds (the build seconds that fell
 * inside those: builds on one worker hidden behind other workers' device
 * work). */
int ainv_server_stats(ainv_server_h srv, char *json, size_t len);
int ainv_server_free(ainv_server_h srv);

/* ---- device plumbing for hosts without a GPU array library (Python via
 * ctypes, Fortran via iso_c_binding).  No counterpart in the reference,
 * where PSBLAS vectors own their storage. -------------------------------- */
#define AINV_H2D 1
#define AINV_D2H 2
#define AINV_D2D 3
int ainv_dev_alloc(size_t bytes, void **p);
int ainv_dev_free(void *p);
int ainv_memcpy(void *dst, const void *src, size_t bytes, int kind);
int ainv_memset(void *p, int value, size_t bytes);
int ainv_device_sync(void);
/* Runs `reps` applies of h on device vectors and returns the average
 * duration (ms) of each of its two kernels, timed with HIP events on the
 * stream they are launched on: ms_first = d o (W x) (or Z^T for 'T'),
 * ms_second = alpha Z t + beta y.  flags (FP32 / BF16 factors only):
 * AINV_PROFILE_X32, x holds FP32 values (the input the Krylov engine's
 * BiCGSTAB and CG give their preconditioner); AINV_PROFILE_Y32, y gets FP32
 * values (their preconditioned vectors).  A multilevel handle times whole
 * cycles instead (ms_first, ms_second = 0; k a power of two; X32 and Y32
 * together for the FP32 cycle). */
#define AINV_PROFILE_X32 1
#define AINV_PROFILE_Y32 2
int ainv_apply_profile(ainv_h h, const void *x, double *y, int k, int reps,
                       int flags, double *ms_first, double *ms_second);
/* Same for the system matrix product y = A x. */
int ainv_mat_profile(ainv_mat_h a, const double *x, double *y, int k, int reps,
                     double *ms);
/* The Krylov step's fused system-matrix product in one mapping (map:
 * AINV_MAP_SPLIT, _ROW, _COOP, _COOP128 or _RUN as numbered for APPLY_MAP):
 * y = A x for k columns interleaved with leading dimension K (the power of
 * two >= k; x FP32 when x_is_f32), and per column c the dots
 * dots[c] = (w_c, y_c) (np = 1, BiCGSTAB's (r~, v)) or dots[c] = (y_c, y_c),
 * dots[k + c] = (y_c, w_c) (np = 2, its (t, t) and (t, s)), summed over the
 * blocks in the order the Krylov step sums them (host array of np * k).
 * x, y, w device pointers; ms (may be NULL): average device time of ten
 * more launches.  No counterpart in the reference (psb_spmm and psb_gedot
 * there, tests/pdegen/ppde3d.f90:280 through psb_krylov); for checking the
 * mappings against each other. */
int ainv_mat_spmv_dot(ainv_mat_h a, const void *x, int x_is_f32, double *y,
                      int k, const double *w, int np, int map, double *dots,
                      double *ms);
/* Average device time (ms) of `reps` Krylov solves from x0 = 0 with the
 * last solve's iteration counts and errors (timed with HIP events). */
int ainv_krylov_profile(const char *method, ainv_mat_h a, ainv_h prec,
                        const double *b, double *x, int k, double eps,
                        int itmax, int reps, int32_t *iter, double *err,
                        double *ms);
/* STREAM-style copy of `bytes` (multiple of 16) from src to dst, `reps`
 * times; *ms = average duration of one copy kernel (HIP events).  The
 * achievable-bandwidth reference and the calibration kernel for hardware
 * traffic counters. */
int ainv_stream_copy(void *dst, const void *src, size_t bytes, int reps,
                     double *ms);
/* Layout / traffic facts of the resident factors: n, nnz(W), nnz(Z),
 * padded entries of each HLL, value bytes. */
int ainv_layout(ainv_h h, int64_t *n, int64_t *nnz_w, int64_t *nnz_z,
                int64_t *pad_w, int64_t *pad_z, int *value_bytes);

#ifdef __cplusplus
}
#endif
#endif
