"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the CPU
baseline of the benchmark use this module as the checker.  The product
package never imports it.

Matrices cross this boundary as scipy.sparse CSR (0-based, int64 indptr,
int32 indices, float64 data, rows sorted by column).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

ALG = {"LLK": 0, "SYM-LLK": 1, "STAB-LLK": 2, "LLK-NOTH": 3, "MLK": 4}


class OrMat(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64),
                ("ptr", C.POINTER(C.c_int64)), ("idx", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double))]


class OrMl(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("a", C.POINTER(OrMat)),
                ("p", C.POINTER(OrMat)), ("r", C.POINTER(OrMat)),
                ("w", C.POINTER(OrMat)), ("z", C.POINTER(OrMat)),
                ("d", C.POINTER(C.POINTER(C.c_double))),
                ("type", C.c_int32), ("pos", C.c_int32),
                ("sweeps", C.c_int32), ("csweeps", C.c_int32)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.POINTER
        L.or_mat_free.argtypes = [P(OrMat)]
        for name, args in {
            "or_sparsify": [C.c_int32, C.c_int32, C.c_double, C.c_int32,
                            P(C.c_double), C.c_int32, P(C.c_int32), P(C.c_double)],
            "or_biconjg": [C.c_int, P(OrMat), C.c_int32, C.c_double,
                           P(C.c_double), P(OrMat), P(OrMat)],
            "or_ainv_bld": [P(OrMat), C.c_int, C.c_int32, C.c_double,
                            P(OrMat), P(OrMat), P(C.c_double)],
            "or_iluk": [P(OrMat), C.c_int32, P(OrMat), P(OrMat), P(C.c_double)],
            "or_ilut": [P(OrMat), C.c_int32, C.c_double, P(OrMat), P(OrMat),
                        P(C.c_double)],
            "or_sparse_invk": [P(OrMat), C.c_int32, P(OrMat)],
            "or_sparse_invt": [P(OrMat), C.c_int32, C.c_double, P(OrMat)],
            "or_invk_bld": [P(OrMat), C.c_int32, C.c_int32, P(OrMat), P(OrMat),
                            P(C.c_double)],
            "or_invt_bld": [P(OrMat), C.c_int32, C.c_double, C.c_int32,
                            C.c_double, P(OrMat), P(OrMat), P(C.c_double)],
            "or_update_a": [P(OrMat), P(OrMat), P(C.c_double),
                            P(C.c_double)],
            "or_apply": [C.c_char, C.c_double, P(OrMat), P(OrMat),
                         P(C.c_double), P(C.c_double), C.c_double, P(C.c_double)],
            "or_bicgstab": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                            P(C.c_double), P(C.c_double), C.c_double, C.c_int32,
                            P(C.c_int32), P(C.c_double)],
            "or_bicg": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                        P(C.c_double), P(C.c_double), C.c_double, C.c_int32,
                        P(C.c_int32), P(C.c_double)],
            "or_cgs": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                       P(C.c_double), P(C.c_double), C.c_double, C.c_int32,
                       P(C.c_int32), P(C.c_double)],
            "or_gmres": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                         P(C.c_double), P(C.c_double), C.c_double, C.c_int32,
                         C.c_int32, P(C.c_int32), P(C.c_double)],
            "or_bicgstabl": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                             P(C.c_double), P(C.c_double), C.c_double,
                             C.c_int32, C.c_int32, P(C.c_int32),
                             P(C.c_double)],
            "or_cg": [P(OrMat), P(OrMat), P(OrMat), P(C.c_double),
                      P(C.c_double), P(C.c_double), C.c_double, C.c_int32,
                      P(C.c_int32), P(C.c_double)],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int32 if name == "or_sparsify" else C.c_int
        L.or_aggregate.argtypes = [P(OrMat), C.c_double, P(C.c_int32),
                                   P(C.c_int32)]
        L.or_spgemm.argtypes = [P(OrMat), P(OrMat), C.c_int32, P(OrMat)]
        L.or_transpose_rect.argtypes = [P(OrMat), C.c_int32, P(OrMat)]
        L.or_ml_level.argtypes = [P(OrMat), C.c_double, C.c_int,
                                  P(C.c_int32), P(C.c_int32), P(OrMat),
                                  P(OrMat), P(OrMat), P(C.c_double)]
        L.or_ml_apply.argtypes = [P(OrMl), P(C.c_double), P(C.c_double)]
        L.or_ml_krylov.argtypes = [C.c_char_p, P(OrMat), P(OrMl),
                                   P(C.c_double), P(C.c_double), C.c_double,
                                   C.c_int32, C.c_int32, P(C.c_int32),
                                   P(C.c_double)]
        for f in ("or_aggregate", "or_spgemm", "or_transpose_rect",
                  "or_ml_level", "or_ml_apply", "or_ml_krylov"):
            getattr(L, f).restype = C.c_int
        L.or_spmv.argtypes = [C.c_char, C.c_double, P(OrMat), P(C.c_double),
                              C.c_double, P(C.c_double)]
        L.or_spmv.restype = None
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = None
        L.or_build_info.argtypes = []
        L.or_build_info.restype = C.c_char_p
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class _In:
    """Keeps the numpy buffers of an input matrix alive with its OrMat."""

    def __init__(self, a):
        a = sp.csr_matrix(a)
        a.sort_indices()
        self.ptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(a.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(a.data, dtype=np.float64)
        self.m = OrMat(a.shape[0], len(self.val),
                       self.ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                       _ip(self.idx), _dp(self.val))

    @property
    def ref(self):
        return C.byref(self.m)


def _out_to_csr(m):
    n, nnz = m.n, m.nnz
    ptr = np.ctypeslib.as_array(m.ptr, shape=(n + 1,)).copy()
    idx = np.ctypeslib.as_array(m.idx, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().or_mat_free(C.byref(m))
    return sp.csr_matrix((val, idx, ptr), shape=(n, n))


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with code {rc}")


def sparsify(zw, idiag, nzrmax, thresh, istart=0):
    zw = np.array(zw, dtype=np.float64)
    n = len(zw)
    iz = np.zeros(n + 1, dtype=np.int32)
    vz = np.zeros(n + 1, dtype=np.float64)
    nz = lib().or_sparsify(idiag, nzrmax, thresh, n, _dp(zw), istart, _ip(iz),
                           _dp(vz))
    return iz[:nz].copy(), vz[:nz].copy()


def biconjg(a, alg="LLK", nzrmax=8, thresh=0.1):
    """Returns (p, Z, W) with Z, W as CSC-column matrices (scipy CSC)."""
    A = _In(a)
    n = A.m.n
    p = np.zeros(n, dtype=np.float64)
    z, w = OrMat(), OrMat()
    _check(lib().or_biconjg(ALG[alg], A.ref, nzrmax, thresh, _dp(p),
                            C.byref(z), C.byref(w)), "biconjg")
    # the CSC arrays of Z are the CSR arrays of Z^T
    return p, _out_to_csr(z).T.tocsc(), _out_to_csr(w).T.tocsc()


def ainv_bld(a, alg="LLK", fill_in=8, thresh=0.1):
    """Returns (W_stored, Z, d) as built by mld_d_ainv_bld."""
    A = _In(a)
    d = np.zeros(A.m.n, dtype=np.float64)
    w, z = OrMat(), OrMat()
    _check(lib().or_ainv_bld(A.ref, ALG[alg], fill_in, thresh, C.byref(w),
                             C.byref(z), _dp(d)), "ainv_bld")
    return _out_to_csr(w), _out_to_csr(z), d


def iluk(a, fill):
    A = _In(a)
    d = np.zeros(A.m.n, dtype=np.float64)
    l, u = OrMat(), OrMat()
    _check(lib().or_iluk(A.ref, fill, C.byref(l), C.byref(u), _dp(d)), "iluk")
    return _out_to_csr(l), _out_to_csr(u), d


def ilut(a, fill, thresh):
    A = _In(a)
    d = np.zeros(A.m.n, dtype=np.float64)
    l, u = OrMat(), OrMat()
    _check(lib().or_ilut(A.ref, fill, thresh, C.byref(l), C.byref(u), _dp(d)),
           "ilut")
    return _out_to_csr(l), _out_to_csr(u), d


def sparse_invk(u, fill):
    U = _In(u)
    z = OrMat()
    _check(lib().or_sparse_invk(U.ref, fill, C.byref(z)), "sparse_invk")
    return _out_to_csr(z)


def sparse_invt(u, fill, thresh):
    U = _In(u)
    z = OrMat()
    _check(lib().or_sparse_invt(U.ref, fill, thresh, C.byref(z)), "sparse_invt")
    return _out_to_csr(z)


def invk_bld(a, fill1, fill2):
    A = _In(a)
    d = np.zeros(A.m.n, dtype=np.float64)
    w, z = OrMat(), OrMat()
    _check(lib().or_invk_bld(A.ref, fill1, fill2, C.byref(w), C.byref(z),
                             _dp(d)), "invk_bld")
    return _out_to_csr(w), _out_to_csr(z), d


def invt_bld(a, fill, thresh, invfill, invthresh):
    A = _In(a)
    d = np.zeros(A.m.n, dtype=np.float64)
    w, z = OrMat(), OrMat()
    _check(lib().or_invt_bld(A.ref, fill, thresh, invfill, invthresh,
                             C.byref(w), C.byref(z), _dp(d)), "invt_bld")
    return _out_to_csr(w), _out_to_csr(z), d


def update_a(w, z, d, x):
    """sv%update_a: inverted pivots for A + diag(x) with the same W, Z."""
    W, Z = _In(w), _In(z)
    d = np.array(d, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    _check(lib().or_update_a(W.ref, Z.ref, _dp(x), _dp(d)), "update_a")
    return d


def apply(w, z, d, x, alpha=1.0, beta=0.0, y=None, trans="N"):
    W, Z = _In(w), _In(z)
    d = np.ascontiguousarray(d, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x) if y is None else np.array(y, dtype=np.float64)
    _check(lib().or_apply(trans.encode(), alpha, W.ref, Z.ref, _dp(d), _dp(x),
                          beta, _dp(y)), "apply")
    return y


def spmv(a, x, alpha=1.0, beta=0.0, y=None, trans="N"):
    A = _In(a)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.m.n) if y is None else np.array(y, dtype=np.float64)
    lib().or_spmv(trans.encode(), alpha, A.ref, _dp(x), beta, _dp(y))
    return y


def set_threads(t):
    """Threads of the row-parallel loops (1 = sequential, the default)."""
    lib().or_set_threads(int(t))


def build_info():
    return lib().or_build_info().decode()


def wrap(a):
    """A matrix converted once for repeated calls (krylov, apply)."""
    return a if isinstance(a, _In) else _In(a)


def krylov(method, a, w, z, d, b, eps, itmax=1000, x0=None, irst=10):
    """Returns (x, iterations, last ||r||/||b||).  irst: restart of
    RGMRES, l of BICGSTABL."""
    A, W, Z = wrap(a), wrap(w), wrap(z)
    d = np.ascontiguousarray(d, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int32(0)
    err = C.c_double(0.0)
    meth = method.upper()
    if meth in ("RGMRES", "BICGSTABL"):
        f = lib().or_gmres if meth == "RGMRES" else lib().or_bicgstabl
        _check(f(A.ref, W.ref, Z.ref, _dp(d), _dp(b), _dp(x), eps, itmax,
                 irst, C.byref(it), C.byref(err)), method)
        return x, it.value, err.value
    f = {"BICGSTAB": lib().or_bicgstab, "CG": lib().or_cg,
         "CGS": lib().or_cgs, "BICG": lib().or_bicg}[meth]
    _check(f(A.ref, W.ref, Z.ref, _dp(d), _dp(b), _dp(x), eps, itmax,
             C.byref(it), C.byref(err)), method)
    return x, it.value, err.value


# ---- multilevel (smoothed aggregation) -----------------------------------
ML_TYPE = {"MULT": 0, "ADD": 1}
ML_POS = {"PRE": 1, "POST": 2, "TWOSIDE": 3}


def _out_rect(m, ncols):
    n, nnz = m.n, m.nnz
    ptr = np.ctypeslib.as_array(m.ptr, shape=(n + 1,)).copy()
    idx = np.ctypeslib.as_array(m.idx, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().or_mat_free(C.byref(m))
    return sp.csr_matrix((val, idx, ptr), shape=(n, ncols))


class _InRect(_In):
    """A rectangular CSR input (the OrMat carries only the row count)."""

    def __init__(self, a):
        a = sp.csr_matrix(a)
        a.sort_indices()
        self.ptr = np.ascontiguousarray(a.indptr, dtype=np.int64)
        self.idx = np.ascontiguousarray(a.indices, dtype=np.int32)
        self.val = np.ascontiguousarray(a.data, dtype=np.float64)
        self.m = OrMat(a.shape[0], len(self.val),
                       self.ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                       _ip(self.idx), _dp(self.val))


def aggregate(a, theta):
    """Returns (agg, nagg): the aggregate of each row and their count."""
    A = wrap(a)
    agg = np.zeros(A.m.n, dtype=np.int32)
    na = C.c_int32(0)
    _check(lib().or_aggregate(A.ref, theta, _ip(agg), C.byref(na)),
           "aggregate")
    return agg, na.value


def spgemm(a, b):
    """A @ B in the oracle's fixed summation order (explicit zeros kept)."""
    A, B = _InRect(a), _InRect(b)
    c = OrMat()
    _check(lib().or_spgemm(A.ref, B.ref, b.shape[1], C.byref(c)), "spgemm")
    return _out_rect(c, b.shape[1])


def ml_level(a, theta, smoothed=True):
    """One coarsening step: (agg, P, R, Ac, omega)."""
    A = wrap(a)
    n = A.m.n
    agg = np.zeros(n, dtype=np.int32)
    na = C.c_int32(0)
    p, r, ac = OrMat(), OrMat(), OrMat()
    om = C.c_double(0.0)
    _check(lib().or_ml_level(A.ref, theta, int(bool(smoothed)), _ip(agg),
                             C.byref(na), C.byref(p), C.byref(r), C.byref(ac),
                             C.byref(om)), "ml_level")
    nc = na.value
    return (agg, _out_rect(p, nc), _out_rect(r, n), _out_to_csr(ac),
            om.value)


def factor_bld(a, kind, fill_in=0, inv_fill=0, thresh=0.0, inv_thresh=0.0,
               alg="LLK"):
    """(W, Z, d) of one AINV-family factorization with the library's
    parameter names and checks (AINV: fill >= 1)."""
    kind = kind.upper()
    if kind == "AINV":
        return ainv_bld(a, alg, max(fill_in, 1), max(thresh, 0.0))
    if kind == "INVK":
        return invk_bld(a, max(fill_in, 0), max(inv_fill, 0))
    if kind == "INVT":
        return invt_bld(a, max(fill_in, 0), max(thresh, 0.0),
                        max(inv_fill, 0), max(inv_thresh, 0.0))
    raise ValueError(kind)


class MlHierarchy:
    """Levels of a smoothed-aggregation hierarchy with AINV-family smoothers
    (sub) and coarsest-level factors (coarse).  Coarsening stops early when
    a step no longer reduces the size."""

    def __init__(self, a, nlev=3, theta=0.1, smoothed=True, sub=None,
                 coarse=None, ml_type="MULT", pos="POST", sweeps=1,
                 csweeps=1, scale=1.0):
        """theta of level l + 1 = theta of level l * scale."""
        sub = dict(sub or {"kind": "INVK"})
        coarse = dict(coarse or sub)
        self.A, self.P, self.R, self.agg, self.omega = [sp.csr_matrix(a)], [], [], [], []
        th = float(theta)
        while len(self.A) < nlev and self.A[-1].shape[0] > 1:
            agg, P, R, Ac, om = ml_level(self.A[-1], th, smoothed)
            th = th * scale
            if Ac.shape[0] >= self.A[-1].shape[0] or Ac.shape[0] == 0:
                break
            self.agg.append(agg)
            self.P.append(P)
            self.R.append(R)
            self.A.append(Ac)
            self.omega.append(om)
        L = len(self.A)
        self.F = [factor_bld(self.A[l], **(sub if l < L - 1 else coarse))
                  for l in range(L)]
        self.type, self.pos = ML_TYPE[ml_type.upper()], ML_POS[pos.upper()]
        self.sweeps, self.csweeps = sweeps, csweeps
        self._pack()

    @classmethod
    def from_levels(cls, A, P, R, F, ml_type="MULT", pos="POST", sweeps=1,
                    csweeps=1):
        """A hierarchy given level by level (scipy CSR A_l, P_l, R_l and
        factors F_l = (W, Z, d)), for instance exported from the library."""
        self = cls.__new__(cls)
        self.A = [sp.csr_matrix(x) for x in A]
        self.P = [sp.csr_matrix(x) for x in P]
        self.R = [sp.csr_matrix(x) for x in R]
        self.F = [(sp.csr_matrix(w), sp.csr_matrix(z), np.asarray(d))
                  for w, z, d in F]
        self.agg, self.omega = [], []
        self.type, self.pos = ML_TYPE[ml_type.upper()], ML_POS[pos.upper()]
        self.sweeps, self.csweeps = sweeps, csweeps
        self._pack()
        return self

    def _pack(self):
        L = len(self.A)
        self._keep = [[_In(m) for m in self.A], [_InRect(m) for m in self.P],
                      [_InRect(m) for m in self.R],
                      [_In(f[0]) for f in self.F], [_In(f[1]) for f in self.F]]
        self._d = [np.ascontiguousarray(f[2], dtype=np.float64) for f in self.F]

        def arr(objs):
            t = (OrMat * max(len(objs), 1))()
            for i, o in enumerate(objs):
                t[i] = o.m
            return t
        self._arrs = [arr(k) for k in self._keep]
        dp = (C.POINTER(C.c_double) * L)(*[_dp(d) for d in self._d])
        self._dp = dp
        self.m = OrMl(L, self._arrs[0], self._arrs[1], self._arrs[2],
                      self._arrs[3], self._arrs[4], dp, self.type, self.pos,
                      self.sweeps, self.csweeps)

    @property
    def nlev(self):
        return len(self.A)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().or_ml_apply(C.byref(self.m), _dp(x), _dp(y)), "ml_apply")
        return y


def ml_krylov(method, a, ml, b, eps, itmax=1000, x0=None, irst=10):
    """Krylov method preconditioned by the V-cycle of ml (MlHierarchy):
    (x, iterations, last ||r||/||b||)."""
    A = wrap(a)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int32(0)
    err = C.c_double(0.0)
    _check(lib().or_ml_krylov(method.upper().encode(), A.ref,
                              C.byref(ml.m), _dp(b), _dp(x), eps, itmax, irst,
                              C.byref(it), C.byref(err)), method)
    return x, it.value, err.value
