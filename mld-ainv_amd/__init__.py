"""MI355X-native AINV / INVK / INVT preconditioners: Python host mirror.

The objects here mirror the MLD2P4 solver objects of the reference plugin
(``mld_d_ainv_solver_type``, ``mld_d_invk_solver_type``,
``mld_d_invt_solver_type``; src/mld_d_ainv_solver.F90:46-67,
src/mld_d_invk_solver.f90:47-57, src/mld_d_invt_solver.f90:46-58) with the
same procedure names (build, apply, cseti/csetr/csetc/set, check, descr,
clone, clone_settings, clear_data, free, sizeof, get_nzeros, get_wrksz,
cnv, dump) and the same error behaviour: every call that the reference
reports through ``info /= 0`` raises :class:`AinvError` carrying that code.

All computation runs in ``lib/libmldainv_hip.so`` (HIP kernels for gfx950);
this module only marshals arguments.  If the library is missing it raises
at import of the first object, it never computes anything itself.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmldainv_hip.so")

# info codes (include/mld_ainv_hip.h)
SUCCESS, ERR_ARG, ERR_ALG, ERR_ALLOC, ERR_INTERNAL, ERR_PIVOT, ERR_STATE, \
    ERR_HIP, ERR_CAPACITY = range(9)
KIND_AINV, KIND_INVK, KIND_INVT, KIND_ML = 1, 2, 3, 4
ALG_LLK, ALG_S_LLK, ALG_S_FT_LLK, ALG_LLK_NOTH, ALG_MLK = 1, 2, 3, 4, 5
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
HOST_PTRS = 1
H2D, D2H, D2D = 1, 2, 3


class AinvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"info={code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads the HIP library; fails loudly if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AinvError(ERR_STATE, f"{LIB_PATH} is not built; run "
                        "__graft_entry__.build() or make -C mld-ainv_amd/csrc")
    L = C.CDLL(LIB_PATH)
    P, vp = C.POINTER, C.c_void_p
    i64p, i32p, dp = P(C.c_int64), P(C.c_int32), P(C.c_double)
    sig = {
        "ainv_init": [C.c_int],
        "ainv_create": [C.c_int, P(vp)],
        "ainv_free": [vp],
        "ainv_default": [vp],
        "ainv_get_kind": [vp, P(C.c_int)],
        "ainv_cseti": [vp, C.c_char_p, C.c_int],
        "ainv_csetr": [vp, C.c_char_p, C.c_double],
        "ainv_csetc": [vp, C.c_char_p, C.c_char_p],
        "ainv_geti": [vp, C.c_char_p, P(C.c_int)],
        "ainv_getr": [vp, C.c_char_p, dp],
        "ainv_check": [vp],
        "ainv_descr": [vp, C.c_char_p, C.c_size_t],
        "ainv_algname": [C.c_int, C.c_char_p, C.c_size_t],
        "ainv_build_csr": [vp, C.c_int32, i64p, i32p, dp, C.c_int],
        "ainv_cnv": [vp, C.c_int],
        "ainv_apply": [vp, C.c_double, vp, C.c_double, vp, C.c_char, C.c_int,
                       C.c_int64, C.c_int64, C.c_int, vp],
        "ainv_apply_noalloc": [vp, C.c_double, vp, C.c_double, vp, C.c_char,
                               C.c_int, C.c_int64, C.c_int64, vp],
        "ainv_get_nzeros": [vp, i64p],
        "ainv_sizeof": [vp, i64p],
        "ainv_get_wrksz": [vp, P(C.c_int)],
        "ainv_get_n": [vp, i32p],
        "ainv_clone": [vp, P(vp)],
        "ainv_clone_settings": [vp, vp],
        "ainv_clear_data": [vp],
        "ainv_export": [vp, C.c_char, i32p, i64p, i64p, i32p, dp],
        "ainv_ml_levels": [vp, P(C.c_int)],
        "ainv_ml_host_level": [C.c_int32, i64p, i32p, dp, C.c_int, C.c_double,
                               C.c_int, C.c_int, C.c_int64, i32p, i32p, i64p,
                               i32p, dp, P(C.c_int64), i64p, i32p, dp,
                               P(C.c_int64), dp],
        "ainv_ml_export": [vp, C.c_int, C.c_char, i32p, i32p, i64p, i64p,
                           i32p, dp],
        "ainv_report": [vp, C.c_char_p, C.c_size_t],
        "ainv_mat_create": [C.c_int32, i64p, i32p, dp, C.c_int, P(vp)],
        "ainv_mat_free": [vp],
        "ainv_mat_set_slice_sort": [vp, C.c_int],
        "ainv_mat_report": [vp, C.c_char_p, C.c_size_t],
        "ainv_slice_sort_plan": [C.c_int32, i64p, C.c_int, P(C.c_int32),
                                 i64p, i64p, P(C.c_int)],
        "ainv_mat_spmm": [vp, C.c_double, vp, C.c_double, vp, C.c_int,
                          C.c_int64, C.c_int64, vp],
        "ainv_dump": [vp, C.c_char_p, C.c_int, C.c_int, C.c_char_p],
        "ainv_update_a": [vp, vp, C.c_int64, C.c_int],
        "ainv_import_factors": [vp, C.c_int32, i64p, i32p, dp, i64p, i32p, dp,
                                dp, C.c_int],
        "ainv_mm_read": [C.c_char_p, P(C.c_int32), P(C.c_int64), i64p, i32p,
                         dp],
        "ainv_mm_read_vector": [C.c_char_p, P(C.c_int64), dp],
        "ainv_host_ainv_bld": [C.c_int32, i64p, i32p, dp, C.c_int, C.c_int,
                               C.c_int32, C.c_double, C.c_int, C.c_int64,
                               i64p, i32p, dp, P(C.c_int64), i64p, i32p, dp,
                               P(C.c_int64), dp],
        "ainv_host_ilu_fact": [C.c_int32, i64p, i32p, dp, C.c_int, C.c_int,
                               C.c_int32, C.c_double, C.c_int64, i64p, i32p,
                               dp, P(C.c_int64), i64p, i32p, dp,
                               P(C.c_int64), dp],
        "ainv_krylov_stream": [C.c_char_p, vp, vp, vp, vp, C.c_int, C.c_int,
                               C.c_double, C.c_int, vp, vp, vp, vp],
        "ainv_krylov": [C.c_char_p, vp, vp, vp, vp, C.c_int, C.c_double,
                        C.c_int, C.c_int, C.c_int, i32p, dp, C.c_int, vp],
        "ainv_gen_pde3d": [C.c_int32, C.c_int, i32p, i64p, i64p, i32p, dp],
        "ainv_gen_aniso2d": [C.c_int32, C.c_double, i32p, i64p, i64p, i32p, dp],
        "ainv_gen_random": [C.c_int32, C.c_uint64, C.c_double, C.c_double,
                            C.c_int32, C.c_int32, C.c_int32, C.c_double, i32p,
                            i64p, i64p, i32p, dp],
        "ainv_gen_normal": [C.c_uint64, C.c_int64, dp],
        "ainv_dev_alloc": [C.c_size_t, P(vp)],
        "ainv_dev_free": [vp],
        "ainv_memcpy": [vp, vp, C.c_size_t, C.c_int],
        "ainv_memset": [vp, C.c_int, C.c_size_t],
        "ainv_device_sync": [],
        "ainv_apply_profile": [vp, vp, vp, C.c_int, C.c_int, C.c_int, dp, dp],
        "ainv_mat_profile": [vp, vp, vp, C.c_int, C.c_int, dp],
        "ainv_mat_spmv_dot": [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                              C.c_int, dp, dp],
        "ainv_krylov_profile": [C.c_char_p, vp, vp, vp, vp, C.c_int,
                                C.c_double, C.c_int, C.c_int, i32p, dp, dp],
        "ainv_layout": [vp, i64p, i64p, i64p, i64p, i64p, P(C.c_int)],
        "ainv_stream_copy": [vp, vp, C.c_size_t, C.c_int, dp],
        "ainv_server_create": [C.c_int, C.c_int64, C.c_int, P(vp)],
        "ainv_server_add_matrix": [vp, C.c_int32, i64p, i32p, dp, C.c_int,
                                   C.c_int, C.c_char_p, i32p],
        "ainv_server_submit": [vp, C.c_int32, C.c_int, dp, C.c_double, C.c_int,
                               i64p],
        "ainv_server_wait": [vp, C.c_int64, dp, i32p, dp, dp],
        "ainv_server_stats": [vp, C.c_char_p, C.c_size_t],
        "ainv_server_free": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.ainv_last_error.argtypes = []
    L.ainv_last_error.restype = C.c_char_p
    L.ainv_stringval.argtypes = [C.c_char_p]
    L.ainv_stringval.restype = C.c_int
    _lib = L
    return L


def _ck(rc):
    if rc != SUCCESS:
        raise AinvError(rc, lib().ainv_last_error().decode(errors="replace"))


def _b(s):
    return s.encode() if isinstance(s, str) else s


_initialized = False


def init(device=0):
    """Selects the HIP device (raises ERR_HIP when no GPU is usable)."""
    global _initialized
    _ck(lib().ainv_init(device))
    _initialized = True


def _ensure_init():
    if not _initialized:
        init(0)


# ----------------------------------------------------------------- memory
class DeviceArray:
    """FP64 array in device memory (n rows x k interleaved columns)."""

    def __init__(self, shape, dtype=np.float64):
        _ensure_init()
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ck(lib().ainv_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _ck(lib().ainv_memcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes,
                                  H2D))
        return d

    def to_numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _ck(lib().ainv_memcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  self.nbytes, D2H))
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _ck(lib().ainv_memcpy(self.ptr, a.ctypes.data_as(C.c_void_p),
                              self.nbytes, H2D))

    def zero(self):
        _ck(lib().ainv_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            _ck(lib().ainv_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.ainv_dev_free(self.ptr)
        except Exception:
            pass


def synchronize():
    _ck(lib().ainv_device_sync())


def stream_copy_gbs(nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth of a plain copy kernel (read + write bytes
    over the kernel time), GB/s, and the per-copy time in ms."""
    src = DeviceArray((nbytes // 8,))
    dst = DeviceArray((nbytes // 8,))
    src.zero()
    ms = C.c_double()
    _ck(lib().ainv_stream_copy(dst.ptr, src.ptr, nbytes, reps, C.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def runtime_copy_gbs(nbytes=1 << 30, reps=10):
    """A second reference for the achievable bandwidth: the HIP runtime's
    own device-to-device copy (hipMemcpy) of nbytes, timed on the host over
    reps copies after one warm-up, with a device synchronization on both
    sides (a device-to-device hipMemcpy may return before the copy ends);
    read + write bytes over the time, GB/s."""
    import time
    src = DeviceArray((nbytes // 8,))
    dst = DeviceArray((nbytes // 8,))
    src.zero()
    _ck(lib().ainv_memcpy(dst.ptr, src.ptr, nbytes, 3))
    _ck(lib().ainv_device_sync())
    t0 = time.perf_counter()
    for _ in range(reps):
        _ck(lib().ainv_memcpy(dst.ptr, src.ptr, nbytes, 3))
    _ck(lib().ainv_device_sync())
    dt = (time.perf_counter() - t0) / reps
    src.free()
    dst.free()
    return 2.0 * nbytes / dt / 1e9


# ------------------------------------------------------------------- CSR in
def _csr_args(a):
    """(n, irp int64, ja int32, val f64) from scipy CSR or a tuple."""
    if isinstance(a, tuple):
        n, irp, ja, val = a
    else:
        a = a.tocsr()
        n, irp, ja, val = a.shape[0], a.indptr, a.indices, a.data
    irp = np.ascontiguousarray(irp, dtype=np.int64)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return int(n), irp, ja, val


def _ptrs(irp, ja, val):
    return (irp.ctypes.data_as(C.POINTER(C.c_int64)),
            ja.ctypes.data_as(C.POINTER(C.c_int32)),
            val.ctypes.data_as(C.POINTER(C.c_double)))


# ----------------------------------------------------------- solver objects
class _BaseAinvSolver:
    """mld_d_base_ainv_solver_type (src/mld_d_base_ainv_mod.F90:45-68)."""

    KIND = None

    def __init__(self, _handle=None):
        if _handle is None:
            h = C.c_void_p()
            _ck(lib().ainv_create(self.KIND, C.byref(h)))
            _handle = h
        self._h = _handle

    # -- parameters -------------------------------------------------------
    def cseti(self, what, val):
        _ck(lib().ainv_cseti(self._h, _b(what), int(val)))

    def csetr(self, what, val):
        _ck(lib().ainv_csetr(self._h, _b(what), float(val)))

    def csetc(self, what, val):
        _ck(lib().ainv_csetc(self._h, _b(what), _b(val)))

    def set(self, what, val):
        """Generic set dispatching on the value's type, as MLD2P4 does."""
        if isinstance(val, str):
            self.csetc(what, val)
        elif isinstance(val, (int, np.integer)) and not isinstance(val, bool):
            self.cseti(what, val)
        else:
            self.csetr(what, val)

    def geti(self, what):
        v = C.c_int()
        _ck(lib().ainv_geti(self._h, _b(what), C.byref(v)))
        return v.value

    def getr(self, what):
        v = C.c_double()
        _ck(lib().ainv_getr(self._h, _b(what), C.byref(v)))
        return v.value

    def default(self):
        _ck(lib().ainv_default(self._h))

    def check(self):
        _ck(lib().ainv_check(self._h))

    def descr(self):
        buf = C.create_string_buffer(2048)
        _ck(lib().ainv_descr(self._h, buf, len(buf)))
        return buf.value.decode()

    # -- build / apply ----------------------------------------------------
    def build(self, a, index_base=0):
        _ensure_init()
        n, irp, ja, val = _csr_args(a)
        _ck(lib().ainv_build_csr(self._h, n, *_ptrs(irp, ja, val), index_base))

    def cnv(self, prec):
        if isinstance(prec, str):
            prec = {"F64": PREC_F64, "F32": PREC_F32, "BF16": PREC_BF16}[
                prec.upper()]
        _ck(lib().ainv_cnv(self._h, int(prec)))

    def apply(self, x, y=None, alpha=1.0, beta=0.0, trans="N"):
        """y = alpha * prec(x) + beta * y.  x, y: numpy (host, n or n x k)
        or DeviceArray (device, interleaved n x k).  Returns y."""
        n = self.n
        if isinstance(x, DeviceArray):
            k = 1 if len(x.shape) == 1 else x.shape[1]
            if y is None:
                y = DeviceArray(x.shape)
            _ck(lib().ainv_apply(self._h, alpha, x.ptr, beta, y.ptr,
                                 _b(trans[0]), k, k, k, 0, None))
            synchronize()
            return y
        x = np.ascontiguousarray(x, dtype=np.float64)
        k = 1 if x.ndim == 1 else x.shape[1]
        if n and x.shape[0] != n:  # an unbuilt solver reports ERR_STATE
            raise AinvError(ERR_ARG, "x has the wrong number of rows")
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(
            y, dtype=np.float64)
        _ck(lib().ainv_apply(self._h, alpha, x.ctypes.data_as(C.c_void_p),
                             beta, y.ctypes.data_as(C.c_void_p),
                             _b(trans[0]), k, k, k, HOST_PTRS, None))
        return y

    # -- accessors --------------------------------------------------------
    @property
    def n(self):
        v = C.c_int32()
        _ck(lib().ainv_get_n(self._h, C.byref(v)))
        return v.value

    def get_nzeros(self):
        v = C.c_int64()
        _ck(lib().ainv_get_nzeros(self._h, C.byref(v)))
        return v.value

    def sizeof(self):
        v = C.c_int64()
        _ck(lib().ainv_sizeof(self._h, C.byref(v)))
        return v.value

    def get_wrksz(self):
        v = C.c_int()
        _ck(lib().ainv_get_wrksz(self._h, C.byref(v)))
        return v.value

    def layout(self):
        vals = [C.c_int64() for _ in range(5)]
        vb = C.c_int()
        _ck(lib().ainv_layout(self._h, *[C.byref(v) for v in vals],
                              C.byref(vb)))
        n, nw, nz, pw, pz = (v.value for v in vals)
        return dict(n=n, nnz_w=nw, nnz_z=nz, pad_w=pw, pad_z=pz,
                    value_bytes=vb.value)

    def report(self):
        import json
        buf = C.create_string_buffer(65536)
        _ck(lib().ainv_report(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    # -- lifecycle ----------------------------------------------------------
    def clone(self):
        h = C.c_void_p()
        _ck(lib().ainv_clone(self._h, C.byref(h)))
        return type(self)(_handle=h)

    def clone_settings(self, other):
        _ck(lib().ainv_clone_settings(self._h, other._h))

    def clear_data(self):
        _ck(lib().ainv_clear_data(self._h))

    def free(self):
        if self._h:
            _ck(lib().ainv_free(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib is not None:
                _lib.ainv_free(self._h)
        except Exception:
            pass

    # -- factors ------------------------------------------------------------
    def export(self, which):
        """Returns the stored factor 'W' or 'Z' as scipy CSR, or 'D' as a
        numpy vector (FP64, as built)."""
        import scipy.sparse as sp
        n = C.c_int32()
        nnz = C.c_int64()
        _ck(lib().ainv_export(self._h, _b(which), C.byref(n), C.byref(nnz),
                              None, None, None))
        if which.upper() == "D":
            d = np.empty(n.value)
            _ck(lib().ainv_export(self._h, b"D", None, None, None, None,
                                  d.ctypes.data_as(C.POINTER(C.c_double))))
            return d
        irp = np.empty(n.value + 1, dtype=np.int64)
        ja = np.empty(max(nnz.value, 1), dtype=np.int32)
        val = np.empty(max(nnz.value, 1), dtype=np.float64)
        _ck(lib().ainv_export(self._h, _b(which), None, None,
                              *_ptrs(irp, ja, val)))
        return sp.csr_matrix((val[:nnz.value], ja[:nnz.value], irp),
                             shape=(n.value, n.value))

    def dump(self, level=1, prefix=None, head=None, solver=True):
        """sv%dump (mld_d_base_ainv_solver_dmp.f90:56-106): stored W, d and
        Z as MatrixMarket files <prefix>_p000_l<level>_{wmat,diag,zmat}.mtx
        (prefix default "dump_ainv_d"); nothing unless solver is true.
        Returns the three file names."""
        _ck(lib().ainv_dump(self._h, _b(prefix) if prefix else None, level,
                            1 if solver else 0, _b(head) if head else None))
        base = f"{prefix or 'dump_ainv_d'}_p000_l{level:03d}_"
        return [base + s + ".mtx" for s in ("wmat", "diag", "zmat")]

    def update_a(self, x):
        """sv%update_a: inverted pivots for the shifted operator A + diag(x)
        (mld_d_base_ainv_update_a.f90:53-79); x: numpy (host) or
        DeviceArray."""
        if isinstance(x, DeviceArray):
            _ck(lib().ainv_update_a(self._h, x.ptr, x.shape[0], 0))
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            _ck(lib().ainv_update_a(self._h, x.ctypes.data_as(C.c_void_p),
                                    x.shape[0], HOST_PTRS))

    def import_factors(self, w, z, d):
        """Applies factors built elsewhere (stored W and Z as scipy CSR,
        inverted pivots d), e.g. read back from a dump."""
        _ensure_init()
        n, wi, wj, wv = _csr_args(w)
        n2, zi, zj, zv = _csr_args(z)
        d = np.ascontiguousarray(d, dtype=np.float64)
        if n2 != n or d.shape != (n,):
            raise AinvError(ERR_ARG, "W, Z and d sizes differ")
        _ck(lib().ainv_import_factors(self._h, n, *_ptrs(wi, wj, wv),
                                      *_ptrs(zi, zj, zv),
                                      d.ctypes.data_as(C.POINTER(C.c_double)),
                                      0))


class MldDAinvSolver(_BaseAinvSolver):
    """mld_d_ainv_solver_type: AINV by sparse biconjugation."""
    KIND = KIND_AINV

    @staticmethod
    def stringval(name):
        return lib().ainv_stringval(_b(name))

    @staticmethod
    def algname(alg):
        buf = C.create_string_buffer(64)
        _ck(lib().ainv_algname(int(alg), buf, len(buf)))
        return buf.value.decode()


class MldDInvkSolver(_BaseAinvSolver):
    """mld_d_invk_solver_type: inverse of ILU(k) factors with fill level."""
    KIND = KIND_INVK


class MldDInvtSolver(_BaseAinvSolver):
    """mld_d_invt_solver_type: inverse of ILUT factors with threshold."""
    KIND = KIND_INVT


class MldMlPrec(_BaseAinvSolver):
    """Multilevel preconditioner (prec%init('ML') of MLD2P4 as the deck
    tests/pdegen/runs/ppde.inp:10-29 configures it): smoothed aggregation,
    AINV-family factors as smoother and coarsest-level solver, V-cycle.
    Keys besides the factor keys: ML_TYPE, SMOOTHER_POS, SMOOTHER_SWEEPS,
    N_PREC_LEVS, AGGR_KIND, AGGR_ALG, AGGR_THRESH, AGGR_SCALE, SUB_SOLVE,
    COARSE_SUBSOLVE, COARSE_SWEEPS."""
    KIND = KIND_ML

    def levels(self):
        n = C.c_int()
        _ck(lib().ainv_ml_levels(self._h, C.byref(n)))
        return n.value

    def level(self, level, which):
        """Level data: 'A', 'P', 'R', 'W', 'Z' as scipy CSR, 'D' (pivots)
        as a float64 vector, 'G' (aggregates) as an int32 vector."""
        import scipy.sparse as sp
        w = which.upper()
        nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        _ck(lib().ainv_ml_export(self._h, level, _b(w), C.byref(nr),
                                 C.byref(nc), C.byref(nnz), None, None, None))
        if w == "D":
            d = np.empty(max(nnz.value, 1))
            _ck(lib().ainv_ml_export(self._h, level, b"D", None, None, None,
                                     None, None,
                                     d.ctypes.data_as(C.POINTER(C.c_double))))
            return d[:nnz.value]
        if w == "G":
            g = np.empty(max(nnz.value, 1), dtype=np.int32)
            _ck(lib().ainv_ml_export(self._h, level, b"G", None, None, None,
                                     None,
                                     g.ctypes.data_as(C.POINTER(C.c_int32)),
                                     None))
            return g[:nnz.value]
        irp = np.empty(nr.value + 1, dtype=np.int64)
        ja = np.empty(max(nnz.value, 1), dtype=np.int32)
        val = np.empty(max(nnz.value, 1), dtype=np.float64)
        _ck(lib().ainv_ml_export(self._h, level, _b(w), None, None, None,
                                 *_ptrs(irp, ja, val)))
        return sp.csr_matrix((val[:nnz.value], ja[:nnz.value], irp),
                             shape=(nr.value, nc.value))


SOLVERS = {"AINV": MldDAinvSolver, "INVK": MldDInvkSolver,
           "INVT": MldDInvtSolver, "ML": MldMlPrec}


# ---------------------------------------------------- system matrix, Krylov
class SystemMatrix:
    """The operator A of the Krylov solve, resident on the device (FP64)."""

    def __init__(self, a, index_base=0):
        _ensure_init()
        n, irp, ja, val = _csr_args(a)
        h = C.c_void_p()
        _ck(lib().ainv_mat_create(n, *_ptrs(irp, ja, val), index_base,
                                  C.byref(h)))
        self._h = h
        self.n = n

    def spmm(self, x, y, alpha=1.0, beta=0.0):
        k = 1 if len(x.shape) == 1 else x.shape[1]
        _ck(lib().ainv_mat_spmm(self._h, alpha, x.ptr, beta, y.ptr, k, k, k,
                                None))
        synchronize()
        return y

    def set_slice_sort(self, window):
        """Row sort of A's sliced layout: -1 picks the window from the row
        lengths (the default), 0 keeps rows in order, else a window of that
        many rows (a multiple of 64).  Drops the matrix's Krylov plans."""
        _ck(lib().ainv_mat_set_slice_sort(self._h, int(window)))

    def report(self):
        """Layout of A: stored entries with and without the row sort, the
        sort window and the slice widths (JSON as a dict)."""
        import json
        buf = C.create_string_buffer(4096)
        _ck(lib().ainv_mat_report(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def spmv_dot(self, x, w, np_=1, mapping="COOP", timed=False):
        """The Krylov step's fused product in one mapping: y = A x for the
        columns of x (numpy (n, K), K a power of two; FP32 or FP64) and the
        per-column dots (w, y) (np_ = 1) or (y, y), (y, w) (np_ = 2).
        Returns (y, dots[np_, K], ms or None)."""
        maps = {"SPLIT": 0, "ROW": 1, "COOP": 2, "COOP128": 3, "RUN": 6}
        x = np.ascontiguousarray(x)
        n, K = x.shape
        f32 = x.dtype == np.float32
        xd = DeviceArray.from_numpy(x if f32 else x.astype(np.float64))
        wd = DeviceArray.from_numpy(np.ascontiguousarray(w, dtype=np.float64))
        yd = DeviceArray((n, K))
        yd.zero()
        dots = np.zeros(np_ * K, dtype=np.float64)
        ms = C.c_double(0.0)
        _ck(lib().ainv_mat_spmv_dot(
            self._h, xd.ptr, int(f32), yd.ptr, K, wd.ptr, np_,
            maps[mapping.upper()], dots.ctypes.data_as(C.POINTER(C.c_double)),
            C.byref(ms) if timed else None))
        y = yd.to_numpy()
        for d in (xd, wd, yd):
            d.free()
        return y, dots.reshape(np_, K), (ms.value if timed else None)

    def free(self):
        if self._h:
            _ck(lib().ainv_mat_free(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib is not None:
                _lib.ainv_mat_free(self._h)
        except Exception:
            pass


KRYLOV_EAGER = 2


def krylov(method, a, prec, b, x=None, eps=1e-8, itmax=1000, istop=2,
           eager=False, irst=0):
    """psb_krylov(method, a, prec, b, x, eps, ...) for k interleaved
    right-hand sides.  a: SystemMatrix; prec: solver object or None;
    b, x: numpy (host) or DeviceArray; irst: restart of RGMRES / l of
    BICGSTABL (0: 10 / 2).  eager=True
    launches the kernels of each iteration one by one instead of replaying
    the captured graphs.  Returns (x, iters[k], errs[k])."""
    host = not isinstance(b, DeviceArray)
    if host:
        b = np.ascontiguousarray(b, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(
            x, dtype=np.float64).copy()
        bp, xp, flags = (b.ctypes.data_as(C.c_void_p),
                         x.ctypes.data_as(C.c_void_p), HOST_PTRS)
    else:
        k = 1 if len(b.shape) == 1 else b.shape[1]
        if x is None:
            x = DeviceArray(b.shape)
            x.zero()
        bp, xp, flags = b.ptr, x.ptr, 0
    if eager:
        flags |= KRYLOV_EAGER
    iters = np.zeros(k, dtype=np.int32)
    errs = np.zeros(k, dtype=np.float64)
    _ck(lib().ainv_krylov(_b(method), a._h, prec._h if prec else None, bp, xp,
                          k, eps, itmax, istop, irst,
                          iters.ctypes.data_as(C.POINTER(C.c_int32)),
                          errs.ctypes.data_as(C.POINTER(C.c_double)), flags,
                          None))
    if not host:
        synchronize()
    return x, iters, errs


def krylov_stream(method, a, prec, b, x=None, slots=0, eps=1e-8,
                  itmax=1000):
    """Served form: solve a device-resident queue of right-hand sides
    (DeviceArray b of shape (nrhs, n), one right-hand side per row) through
    `slots` continuously refilled column slots (0: the width among 8, 16,
    32, 64 with the least time per column-iteration in a trial on the
    queue, kept per preconditioner and method; report()["served_slots"]).
    Returns (x, iters[nrhs], errs[nrhs], latency_ms[nrhs]) with x a
    DeviceArray like b."""
    if not isinstance(b, DeviceArray) or len(b.shape) != 2:
        raise AinvError(ERR_ARG, "krylov_stream takes a (nrhs, n) DeviceArray")
    nrhs = b.shape[0]
    if x is None:
        x = DeviceArray(b.shape)
    iters = np.zeros(nrhs, dtype=np.int32)
    errs = np.zeros(nrhs, dtype=np.float64)
    lat = np.zeros(nrhs, dtype=np.float64)
    _ck(lib().ainv_krylov_stream(
        _b(method), a._h, prec._h if prec else None, b.ptr, x.ptr, nrhs,
        slots, eps, itmax, iters.ctypes.data_as(C.c_void_p),
        errs.ctypes.data_as(C.c_void_p), lat.ctypes.data_as(C.c_void_p),
        None))
    synchronize()
    return x, iters, errs, lat


def mm_read(path):
    """MatrixMarket coordinate file -> scipy CSR (square; duplicates summed;
    symmetric / skew-symmetric storage expanded)."""
    import scipy.sparse as sp
    n = C.c_int32()
    nnz = C.c_int64()
    _ck(lib().ainv_mm_read(_b(path), C.byref(n), C.byref(nnz), None, None,
                           None))
    irp = np.empty(n.value + 1, dtype=np.int64)
    ja = np.empty(nnz.value, dtype=np.int32)
    val = np.empty(nnz.value, dtype=np.float64)
    _ck(lib().ainv_mm_read(_b(path), None, None, *_ptrs(irp, ja, val)))
    return sp.csr_matrix((val, ja, irp), shape=(n.value, n.value))


def mm_read_vector(path):
    """MatrixMarket dense single-column array -> numpy vector."""
    n = C.c_int64()
    _ck(lib().ainv_mm_read_vector(_b(path), C.byref(n), None))
    v = np.empty(n.value, dtype=np.float64)
    _ck(lib().ainv_mm_read_vector(_b(path), None,
                                  v.ctypes.data_as(C.POINTER(C.c_double))))
    return v


# AINV_ALG_* of include/mld_ainv_hip.h by the reference's names
ALGS = {"LLK": 1, "SYM-LLK": 2, "STAB-LLK": 3, "LLK-NOTH": 4, "MLK": 5}


def host_ainv_bld(a, alg="LLK", fill_in=8, thresh=0.1, threads=0):
    """mld_d_ainv_bld on the host (ainv_host_ainv_bld): returns
    (W_stored, Z, d) as scipy CSR / numpy.  threads = 0: the default."""
    import scipy.sparse as sp
    a = sp.csr_matrix(a)
    a.sort_indices()
    n = a.shape[0]
    irp = np.ascontiguousarray(a.indptr, dtype=np.int64)
    ja = np.ascontiguousarray(a.indices, dtype=np.int32)
    val = np.ascontiguousarray(a.data, dtype=np.float64)
    cap = int(n) * max(int(fill_in), 1) + 1 if fill_in >= 1 else \
        int(n) * int(n)
    outs = []
    for _ in range(2):
        outs.append((np.empty(n + 1, dtype=np.int64),
                     np.empty(cap, dtype=np.int32),
                     np.empty(cap, dtype=np.float64)))
    wn, zn = C.c_int64(), C.c_int64()
    d = np.empty(n, dtype=np.float64)
    (wi, wj, wv), (zi, zj, zv) = outs
    _ck(lib().ainv_host_ainv_bld(
        n, *_ptrs(irp, ja, val), 0, ALGS[alg], int(fill_in), float(thresh),
        int(threads), cap, *_ptrs(wi, wj, wv), C.byref(wn),
        *_ptrs(zi, zj, zv), C.byref(zn),
        d.ctypes.data_as(C.POINTER(C.c_double))))
    w = sp.csr_matrix((wv[:wn.value], wj[:wn.value], wi), shape=(n, n))
    z = sp.csr_matrix((zv[:zn.value], zj[:zn.value], zi), shape=(n, n))
    return w, z, d


def host_ilu_fact(a, kind="ILUT", fill=1, thresh=0.0):
    """The INVK / INVT builds' incomplete factorization on the host
    (ainv_host_ilu_fact): kind "ILUK" (mld_iluk_fact) or "ILUT"
    (mld_ilut_fact, maxval scaling).  Returns (L, U, d) as scipy CSR /
    numpy: L strictly lower, U strictly upper (unit diagonals implied),
    d the inverted pivots."""
    import scipy.sparse as sp
    n, irp, ja, val = _csr_args(a)
    kid = {"ILUK": KIND_INVK, "ILUT": KIND_INVT}[kind.upper()]
    cap = max(4 * len(val), 1024)
    while True:
        outs = [(np.empty(n + 1, dtype=np.int64), np.empty(cap, dtype=np.int32),
                 np.empty(cap, dtype=np.float64)) for _ in range(2)]
        ln, un = C.c_int64(), C.c_int64()
        d = np.empty(n, dtype=np.float64)
        (li, lj, lv), (ui, uj, uv) = outs
        rc = lib().ainv_host_ilu_fact(
            n, *_ptrs(irp, ja, val), 0, kid, int(fill), float(thresh), cap,
            *_ptrs(li, lj, lv), C.byref(ln), *_ptrs(ui, uj, uv), C.byref(un),
            d.ctypes.data_as(C.POINTER(C.c_double)))
        if rc == ERR_ARG and max(ln.value, un.value) > cap:
            cap = max(ln.value, un.value)
            continue
        _ck(rc)
        break
    L = sp.csr_matrix((lv[:ln.value], lj[:ln.value], li), shape=(n, n))
    U = sp.csr_matrix((uv[:un.value], uj[:un.value], ui), shape=(n, n))
    return L, U, d


# --------------------------------------------------------------- generators
def ml_host_level(a, theta=0.1, smoothed=True, threads=0):
    """One coarsening step of the multilevel build on the host (no device):
    (agg, P, R A P, omega)."""
    import scipy.sparse as sp
    n, irp, ja, val = _csr_args(a)
    cap = max(16 * len(val), 1024)
    while True:
        agg = np.empty(max(n, 1), dtype=np.int32)
        nagg = C.c_int32()
        p_irp = np.empty(n + 1, dtype=np.int64)
        c_irp = np.empty(n + 1, dtype=np.int64)
        p_ja = np.empty(cap, dtype=np.int32)
        c_ja = np.empty(cap, dtype=np.int32)
        p_val = np.empty(cap)
        c_val = np.empty(cap)
        pn, cn, om = C.c_int64(), C.c_int64(), C.c_double()
        rc = lib().ainv_ml_host_level(
            n, *_ptrs(irp, ja, val), 0, theta, int(bool(smoothed)),
            threads, cap, agg.ctypes.data_as(C.POINTER(C.c_int32)),
            C.byref(nagg), *_ptrs(p_irp, p_ja, p_val), C.byref(pn),
            *_ptrs(c_irp, c_ja, c_val), C.byref(cn), C.byref(om))
        if rc == ERR_ARG and max(pn.value, cn.value) > cap:
            cap = max(pn.value, cn.value)
            continue
        _ck(rc)
        break
    nc = nagg.value
    P = sp.csr_matrix((p_val[:pn.value], p_ja[:pn.value], p_irp),
                      shape=(n, nc))
    Ac = sp.csr_matrix((c_val[:cn.value], c_ja[:cn.value], c_irp[:nc + 1]),
                       shape=(nc, nc))
    return agg[:n], P, Ac, om.value


def _gen(fn, *args):
    import scipy.sparse as sp
    n = C.c_int32()
    nnz = C.c_int64()
    _ck(fn(*args, C.byref(n), C.byref(nnz), None, None, None))
    irp = np.empty(n.value + 1, dtype=np.int64)
    ja = np.empty(nnz.value, dtype=np.int32)
    val = np.empty(nnz.value, dtype=np.float64)
    _ck(fn(*args, None, None, *_ptrs(irp, ja, val)))
    return sp.csr_matrix((val, ja, irp), shape=(n.value, n.value))


def gen_pde3d(N, hconv=0):
    """3-D convection-diffusion, 7-point (tests/pdegen/genpde_mod.f90)."""
    return _gen(lib().ainv_gen_pde3d, int(N), int(hconv))


def gen_aniso2d(N, eps=1e-3):
    return _gen(lib().ainv_gen_aniso2d, int(N), float(eps))


def gen_random(n, seed=40, median=12.0, sigma=1.0, cap=400, band=5000,
               nlong=2, dom=1.1):
    return _gen(lib().ainv_gen_random, int(n), int(seed), float(median),
                float(sigma), int(cap), int(band), int(nlong), float(dom))


def slice_sort_plan(indptr, window=-1):
    """The row sort of the sliced layouts computed on the host (no device):
    for CSR row pointers `indptr` and a window as in
    SystemMatrix.set_slice_sort, returns (perm, padded, unsorted_padded,
    window_used): stored row i holds row perm[i] (n past the last row)."""
    irp = np.ascontiguousarray(indptr, dtype=np.int64)
    n = len(irp) - 1
    perm = np.empty(64 * ((n + 63) // 64), dtype=np.int32)
    pad, upad, used = C.c_int64(0), C.c_int64(0), C.c_int(0)
    _ck(lib().ainv_slice_sort_plan(
        n, irp.ctypes.data_as(C.POINTER(C.c_int64)), int(window),
        perm.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(pad),
        C.byref(upad), C.byref(used)))
    return perm, pad.value, upad.value, used.value


def gen_normal(seed, shape):
    count = int(np.prod(shape))
    out = np.empty(count, dtype=np.float64)
    _ck(lib().ainv_gen_normal(int(seed), count,
                              out.ctypes.data_as(C.POINTER(C.c_double))))
    return out.reshape(shape)


# ----------------------------------------------------------------- serving
class Server:
    """Serving loop (ainv_server_*): matrices registered once, factors built
    on first touch and kept resident under an HBM budget with LRU eviction,
    requests of 1..64 right-hand sides batched per matrix with continuous
    admission into freed column slots."""

    KINDS = {"AINV": KIND_AINV, "INVK": KIND_INVK, "INVT": KIND_INVT,
             "ML": KIND_ML}

    def __init__(self, hbm_budget_bytes=64 << 30, workers=2, device=0):
        _ensure_init()
        h = C.c_void_p()
        _ck(lib().ainv_server_create(device, int(hbm_budget_bytes), workers,
                                     C.byref(h)))
        self._h = h
        self._n = {}

    def add_matrix(self, a, kind="INVK", **params):
        """params: solver keys (SUB_FILLIN=1, FACTOR_PREC='F32', ...) and
        METHOD='BICGSTAB'|'CG'|'CGS'|'BICG', SLOTS=s (fewest column slots of
        a batch).  Returns the matrix id."""
        n, irp, ja, val = _csr_args(a)
        spec = ";".join(f"{k}={v}" for k, v in params.items())
        mid = C.c_int32()
        _ck(lib().ainv_server_add_matrix(self._h, n, *_ptrs(irp, ja, val), 0,
                                         self.KINDS[kind], spec.encode(),
                                         C.byref(mid)))
        self._n[mid.value] = n
        return mid.value

    def submit(self, mat_id, b, eps=1e-8, itmax=1000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        rid = C.c_int64()
        _ck(lib().ainv_server_submit(self._h, mat_id, k,
                                     b.ctypes.data_as(C.POINTER(C.c_double)),
                                     eps, itmax, C.byref(rid)))
        return (rid.value, mat_id, k, b.ndim)

    def wait(self, ticket):
        rid, mat_id, k, nd = ticket
        n = self._n[mat_id]
        x = np.empty((n, k))
        it = np.empty(k, dtype=np.int32)
        err = np.empty(k)
        lat = C.c_double()
        _ck(lib().ainv_server_wait(self._h, rid,
                                   x.ctypes.data_as(C.POINTER(C.c_double)),
                                   it.ctypes.data_as(C.POINTER(C.c_int32)),
                                   err.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(lat)))
        return (x[:, 0] if nd == 1 else x), it, err, lat.value

    def stats(self):
        import json
        buf = C.create_string_buffer(4096)
        _ck(lib().ainv_server_stats(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def free(self):
        if self._h:
            _ck(lib().ainv_server_free(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _lib is not None:
                _lib.ainv_server_free(self._h)
        except Exception:
            pass
