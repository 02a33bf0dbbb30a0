"""The no-allocation apply (ainv_apply_noalloc) through the C ABI against
ainv_apply: the same kernels in the same mappings, so the results agree
bit for bit.  Two calls in a row, the second on a smaller problem (the
shared scratch array keeps the first call's contents), and the refusals:
a size beyond the scratch array, host pointers' kind (multilevel), and a
'T' apply before the transposed layout exists."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build(gpu, kind, a, prec, **params):
    sv = gpu.SOLVERS[kind]()
    for key, v in params.items():
        sv.set(key, v)
    sv.set("FACTOR_PREC", prec)
    sv.build(a)
    return sv


def noalloc(gpu, sv, x, y, alpha, beta, trans="N"):
    """x, y: DeviceArray (n x k interleaved)."""
    k = 1 if len(x.shape) == 1 else x.shape[1]
    return gpu.lib().ainv_apply_noalloc(sv._h, alpha, x.ptr, beta, y.ptr,
                                        trans.encode(), k, k, k, None)


@pytest.mark.parametrize("prec", ["F64", "F32", "BF16"])
def test_noalloc_matches_apply(gpu, prec):
    cases = [(gpu.gen_pde3d(20), "INVK", dict(SUB_FILLIN=1, INV_FILLIN=1), 32),
             (gpu.gen_pde3d(9), "AINV", dict(AINV_ALG="LLK", SUB_FILLIN=8,
                                             SUB_ILUTHRS=0.1), 5),
             # factors with rows sorted by length within 256-row windows
             (gpu.gen_random(600, seed=5, band=200), "AINV",
              dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=5e-2,
                   SLICE_SORT=256), 3)]
    for a, kind, params, k in cases:  # the second call is the smaller one
        n = a.shape[0]
        sv = build(gpu, kind, a, prec, **params)
        if "SLICE_SORT" in params:
            assert sv.report()["z"]["sort_window"] == 256
        x = gpu.gen_normal(31, (n, k))
        y0 = gpu.gen_normal(32, (n, k))
        dx = gpu.DeviceArray.from_numpy(x)
        for trans in ("N", "T"):
            ref = sv.apply(dx, y=gpu.DeviceArray.from_numpy(y0), alpha=-0.5,
                           beta=2.0, trans=trans).to_numpy()
            dy = gpu.DeviceArray.from_numpy(y0)
            assert noalloc(gpu, sv, dx, dy, -0.5, 2.0, trans) == 0, \
                gpu.lib().ainv_last_error()
            gpu.synchronize()
            assert np.array_equal(dy.to_numpy(), ref), (kind, prec, trans)
        # beta = 0 overwrites y without reading it
        dy = gpu.DeviceArray.from_numpy(np.full((n, k), np.nan))
        assert noalloc(gpu, sv, dx, dy, 1.0, 0.0) == 0
        gpu.synchronize()
        assert np.array_equal(dy.to_numpy(), sv.apply(x))


def test_noalloc_refusals(gpu):
    # an intermediate beyond the 64 MiB scratch array: 262,144 rows x 64
    # FP64 columns = 128 MiB
    a = gpu.gen_pde3d(64)
    n = a.shape[0]
    sv = build(gpu, "INVK", a, "F64", SUB_FILLIN=0, INV_FILLIN=0)
    x = gpu.DeviceArray((n, 64))
    y = gpu.DeviceArray((n, 64))
    assert noalloc(gpu, sv, x, y, 1.0, 0.0) == gpu.ERR_CAPACITY
    assert "scratch" in gpu.lib().ainv_last_error().decode()
    # 'T' before the transposed layout exists (ainv_apply builds it)
    small = build(gpu, "INVK", gpu.gen_pde3d(8), "F64", SUB_FILLIN=1,
                  INV_FILLIN=1)
    xs = gpu.DeviceArray.from_numpy(np.ones(512))
    ys = gpu.DeviceArray((512,))
    assert noalloc(gpu, small, xs, ys, 1.0, 0.0, "T") == gpu.ERR_STATE
    # the multilevel kind
    ml = gpu.MldMlPrec()
    ml.set("SUB_SOLVE", "INVK")
    ml.set("COARSE_SUBSOLVE", "INVK")
    ml.build(gpu.gen_pde3d(8))
    assert noalloc(gpu, ml, xs, ys, 1.0, 0.0) == gpu.ERR_ARG
