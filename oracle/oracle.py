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
