"""GPU parity of the multilevel preconditioner (kind AINV_KIND_ML) against
the oracle's restatement, through the C ABI.

Bars: the hierarchy (aggregates, P, R, level matrices) and every level's
factors bit-exact (host products in the oracle's summation order, factor
builds as for the one-level solvers); the cycle within 1e-11 relative of the
oracle's for every cycle type and 1-64 columns; Krylov iteration counts
within +-1 of the oracle's; graph replays bit-identical to eager launches.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def same_csr(a, b):
    a, b = a.tocsr(), b.tocsr()
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr)
            and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


SUBS = {
    "INVK": (dict(SUB_FILLIN=1, INV_FILLIN=1),
             {"kind": "INVK", "fill_in": 1, "inv_fill": 1}),
    "INVT": (dict(SUB_FILLIN=4, SUB_ILUTHRS=1e-2, INV_FILLIN=4,
                  INV_THRESH=1e-2),
             {"kind": "INVT", "fill_in": 4, "thresh": 1e-2, "inv_fill": 4,
              "inv_thresh": 1e-2}),
    "AINV": (dict(AINV_ALG="LLK", SUB_FILLIN=6, SUB_ILUTHRS=5e-2),
             {"kind": "AINV", "alg": "LLK", "fill_in": 6, "thresh": 5e-2}),
}


def build_ml(gpu, O, a, sub="INVK", coarse=None, nlev=3, theta=0.05,
             ml_type="MULT", pos="POST", sweeps=1, csweeps=4, scale=1.0,
             smoothed=True):
    coarse = coarse or sub
    pv = gpu.MldMlPrec()
    for k, v in SUBS[sub][0].items():
        pv.set(k, v)
    pv.set("SUB_SOLVE", sub)
    pv.set("COARSE_SUBSOLVE", coarse)
    pv.set("N_PREC_LEVS", nlev)
    pv.set("AGGR_THRESH", theta)
    pv.set("AGGR_SCALE", scale)
    pv.set("AGGR_KIND", "SMOOTHED" if smoothed else "NONSMOOTHED")
    pv.set("ML_TYPE", ml_type)
    pv.set("SMOOTHER_POS", pos)
    pv.set("SMOOTHER_SWEEPS", sweeps)
    pv.set("COARSE_SWEEPS", csweeps)
    pv.build(a)
    ref = O.MlHierarchy(a, nlev=nlev, theta=theta, smoothed=smoothed,
                        sub=SUBS[sub][1], coarse=SUBS[coarse][1],
                        ml_type=ml_type, pos=pos, sweeps=sweeps,
                        csweeps=csweeps, scale=scale)
    return pv, ref


def rand_nonsym(gpu, n, seed):
    return gpu.gen_random(n, seed=seed, median=8.0, band=400)


@pytest.mark.parametrize("sub,coarse", [("INVK", "INVK"), ("INVT", "INVK"),
                                        ("AINV", "AINV")])
@pytest.mark.parametrize("mat", ["pde3d", "random"])
def test_hierarchy_bit_exact(gpu, O, sub, coarse, mat):
    a = gpu.gen_pde3d(14, 1) if mat == "pde3d" else rand_nonsym(gpu, 5000, 3)
    print("hierarchy", sub, coarse, mat, flush=True)
    pv, ref = build_ml(gpu, O, a, sub, coarse, scale=0.5)
    assert pv.levels() == ref.nlev >= 2
    for l in range(ref.nlev):
        assert same_csr(pv.level(l, "A"), ref.A[l]), l
        w, z, d = ref.F[l]
        assert same_csr(pv.level(l, "W"), w), l
        assert same_csr(pv.level(l, "Z"), z), l
        assert np.array_equal(pv.level(l, "D"), d), l
        if l < ref.nlev - 1:
            assert np.array_equal(pv.level(l, "G"), ref.agg[l]), l
            assert same_csr(pv.level(l, "P"), ref.P[l]), l
            assert same_csr(pv.level(l, "R"), ref.R[l]), l


@pytest.mark.parametrize("ml_type,pos,sweeps", [
    ("MULT", "POST", 1), ("MULT", "PRE", 2), ("MULT", "TWOSIDE", 1),
    ("ADD", "POST", 2)])
@pytest.mark.parametrize("smoothed", [True, False])
def test_cycle_matches_oracle(gpu, O, ml_type, pos, sweeps, smoothed):
    a = gpu.gen_pde3d(12, 1)
    n = a.shape[0]
    print("cycle", ml_type, pos, sweeps, smoothed, flush=True)
    pv, ref = build_ml(gpu, O, a, "INVK", ml_type=ml_type, pos=pos,
                       sweeps=sweeps, csweeps=3, smoothed=smoothed)
    X = gpu.gen_normal(5, (n, 64))
    for k in (1, 2, 5, 8, 32, 33, 64):
        Y = pv.apply(np.ascontiguousarray(X[:, :k]))
        Y = Y.reshape(n, k)
        for j in range(0, k, max(1, k // 4)):
            yr = ref.apply(X[:, j])
            err = np.linalg.norm(Y[:, j] - yr) / np.linalg.norm(yr)
            assert err < 1e-11, (k, j, err)


@pytest.mark.parametrize("mapping", ["ROW", "COOP", "COOP128", "COOP_Q",
                                     "RUN"])
def test_cycle_every_forced_mapping(gpu, O, mapping):
    """Every level product (the square level matrices and factors, and the
    rectangular P, whose x has fewer rows than P, and R, whose x has more)
    in each block mapping: the cycle matches the oracle's.  A mapping that
    sized its reads of x by the product's row count would read past x
    (P) or miss rows of it (R)."""
    a = gpu.gen_pde3d(16, 1)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "INVK", csweeps=3)
    pv.set("APPLY_MAP", mapping)
    X = gpu.gen_normal(15, (n, 32))
    for k in (4, 16, 32):
        Y = pv.apply(np.ascontiguousarray(X[:, :k])).reshape(n, k)
        for j in (0, k - 1):
            yr = ref.apply(X[:, j])
            err = np.linalg.norm(Y[:, j] - yr) / np.linalg.norm(yr)
            assert err < 1e-11, (mapping, k, j, err)


def test_cycle_alpha_beta_and_device_arrays(gpu, O):
    a = gpu.gen_pde3d(10)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "INVK")
    X = gpu.gen_normal(6, (n, 3))
    Y0 = gpu.gen_normal(7, (n, 3))
    Y = pv.apply(X, y=Y0.copy(), alpha=-2.0, beta=0.5)
    for j in range(3):
        want = -2.0 * ref.apply(X[:, j]) + 0.5 * Y0[:, j]
        assert np.linalg.norm(Y[:, j] - want) <= 1e-11 * np.linalg.norm(want)
    dx = gpu.DeviceArray.from_numpy(np.ascontiguousarray(X))
    dy = pv.apply(dx).to_numpy()
    assert np.allclose(dy, pv.apply(X), rtol=1e-14, atol=0)
    with pytest.raises(gpu.AinvError) as e:
        pv.apply(X, trans="T")
    assert e.value.code == gpu.ERR_ARG


@pytest.mark.parametrize("method", ["BICGSTAB", "CGS", "RGMRES", "BICGSTABL",
                                    "CG"])
@pytest.mark.parametrize("k", [1, 6])
def test_krylov_counts_match_oracle(gpu, O, method, k):
    if method == "CG":
        a = gpu.gen_aniso2d(40, 1e-2)
        sub, pos = "AINV", "TWOSIDE"
    else:
        a = gpu.gen_pde3d(14, 1)
        sub, pos = "INVK", "POST"
    n = a.shape[0]
    print("krylov", method, k, flush=True)
    pv, ref = build_ml(gpu, O, a, sub, pos=pos)
    B = gpu.gen_normal(11, (n, k))
    A = gpu.SystemMatrix(a)
    irst = 2 if method == "BICGSTABL" else 10
    x, it, err = gpu.krylov(method, A, pv, B, eps=1e-8, itmax=300, irst=irst)
    x = x.reshape(n, k)
    for j in range(k):
        xr, itr, errr = O.ml_krylov(method, a, ref, B[:, j], 1e-8, 300,
                                    irst=irst)
        step = 2 if method == "BICGSTABL" else 1
        assert abs(int(it[j]) - itr) <= step, (j, it[j], itr)
        assert err[j] <= 1e-8
        r = B[:, j] - a @ x[:, j]
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(B[:, j])


def test_bicg_refused(gpu, O):
    a = gpu.gen_pde3d(8)
    pv, _ = build_ml(gpu, O, a, "INVK", nlev=2)
    A = gpu.SystemMatrix(a)
    with pytest.raises(gpu.AinvError) as e:
        gpu.krylov("BICG", A, pv, a @ np.ones(a.shape[0]))
    assert "M^T" in str(e.value)


@pytest.mark.parametrize("method", ["BICGSTAB", "RGMRES"])
def test_graph_replays_match_eager(gpu, O, method):
    a = gpu.gen_pde3d(12, 1)
    n = a.shape[0]
    pv, _ = build_ml(gpu, O, a, "INVK")
    B = gpu.gen_normal(21, (n, 3))
    A = gpu.SystemMatrix(a)
    x_e, it_e, err_e = gpu.krylov(method, A, pv, B, eps=1e-9, eager=True)
    for rep in range(3):
        x_g, it_g, err_g = gpu.krylov(method, A, pv, B, eps=1e-9)
        assert np.array_equal(it_g, it_e), (rep, it_g, it_e)
        assert np.array_equal(x_g, x_e), rep
        assert np.array_equal(err_g, err_e), rep
        pv.apply(B[:, 0])  # eager work between replays
    # re-molded factors invalidate the captured graphs
    pv.cnv("F32")
    x32, it32, _ = gpu.krylov(method, A, pv, B, eps=1e-9)
    x32e, it32e, _ = gpu.krylov(method, A, pv, B, eps=1e-9, eager=True)
    assert np.array_equal(it32, it32e) and np.array_equal(x32, x32e)


def test_reduced_precision_factors(gpu, O):
    a = gpu.gen_pde3d(12, 1)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "INVK")
    X = gpu.gen_normal(3, (n, 4))
    y64 = pv.apply(X)
    A = gpu.SystemMatrix(a)
    _, it64, _ = gpu.krylov("BICGSTAB", A, pv, X, eps=1e-8)
    for prec, tol in (("F32", 2e-6), ("BF16", 3e-2)):
        pv.cnv(prec)
        y = pv.apply(X)
        assert np.linalg.norm(y - y64) <= tol * np.linalg.norm(y64), prec
        _, it, err = gpu.krylov("BICGSTAB", A, pv, X, eps=1e-8)
        assert np.all(err <= 1e-8) and np.all(it <= it64 + 3), (prec, it)
    # back to FP64 factors: the same cycle up to the summation order of the
    # kernel mappings the solves above tuned
    pv.cnv("F64")
    assert np.linalg.norm(pv.apply(X) - y64) <= 1e-13 * np.linalg.norm(y64)


def test_served_form(gpu, O):
    a = gpu.gen_pde3d(12, 1)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "INVK")
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(31, (10, n))
    x, it, err, lat = gpu.krylov_stream(
        "BICGSTAB", A, pv, gpu.DeviceArray.from_numpy(B), slots=4, eps=1e-8)
    X = x.to_numpy()
    for j in range(10):
        _, itr, _ = O.ml_krylov("BICGSTAB", a, ref, B[j], 1e-8, 1000)
        assert abs(int(it[j]) - itr) <= 1, (j, it[j], itr)
        r = B[j] - a @ X[j]
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(B[j])


def test_accessors_and_dump(gpu, O, tmp_path):
    a = gpu.gen_pde3d(10)
    pv, ref = build_ml(gpu, O, a, "INVK", nlev=3)
    assert pv.geti("N_PREC_LEVS") == 3 and pv.getr("AGGR_THRESH") == 0.05
    assert pv.geti("SUB_SOLVE") == gpu.KIND_INVK
    text = pv.descr()
    assert "Multilevel" in text and "INVK" in text
    rep = pv.report()
    assert len(rep["multilevel"]["levels"]) == ref.nlev
    nz = sum(ref.A[l].shape[0] + ref.F[l][0].nnz + ref.F[l][1].nnz
             for l in range(ref.nlev))
    nz += sum(ref.A[l].nnz for l in range(1, ref.nlev))
    nz += sum(p.nnz + r.nnz for p, r in zip(ref.P, ref.R))
    assert pv.get_nzeros() == nz
    files = []
    prefix = str(tmp_path / "ml")
    pv.dump(level=1, prefix=prefix)
    for l in range(ref.nlev):
        f = f"{prefix}_p000_l{l + 1:03d}_wmat.mtx"
        assert os.path.exists(f)
        files.append(f)
    w1 = gpu.mm_read(files[1])
    assert same_csr(w1, ref.F[1][0])
    with pytest.raises(gpu.AinvError) as e:
        pv.export("W")
    assert e.value.code == gpu.ERR_STATE
    with pytest.raises(gpu.AinvError):
        pv.update_a(np.ones(a.shape[0]))
    with pytest.raises(gpu.AinvError):
        pv.clone()
    pv.clear_data()
    assert pv.levels() == 0


def test_hierarchy_without_reduction_is_one_level(gpu, O):
    a = gpu.gen_pde3d(8)
    pv, ref = build_ml(gpu, O, a, "INVK", theta=10.0, csweeps=2)
    assert pv.levels() == ref.nlev == 1
    x = gpu.gen_normal(2, (a.shape[0],))
    yr = ref.apply(x)
    assert np.linalg.norm(pv.apply(x) - yr) <= 1e-12 * np.linalg.norm(yr)


@pytest.mark.parametrize("prec,tol", [("F32", 2e-6), ("BF16", 1e-5)])
def test_fp32_cycle(gpu, O, prec, tol):
    """The cycle on FP32 vectors (what BiCGSTAB and CG give FP32 / BF16
    factors), left in y by the profiling entry point, against the FP64
    cycle on the same factors and FP32 level matrices."""
    import ctypes as C
    a = gpu.gen_pde3d(12, 1)
    n, k = a.shape[0], 8
    pv, _ = build_ml(gpu, O, a, "INVK")
    pv.cnv(prec)
    X = gpu.gen_normal(4, (n, k)).astype(np.float32)
    y64 = pv.apply(X.astype(np.float64))
    dx = gpu.DeviceArray.from_numpy(X)
    dy = gpu.DeviceArray((n, k), dtype=np.float32)
    ms1, ms2 = C.c_double(), C.c_double()
    gpu._ck(gpu.lib().ainv_apply_profile(pv._h, dx.ptr, dy.ptr, k, 2, 3,
                                         C.byref(ms1), C.byref(ms2)))
    y32 = dy.to_numpy().astype(np.float64)
    assert ms1.value > 0 and ms2.value == 0
    assert np.linalg.norm(y32 - y64) <= tol * np.linalg.norm(y64)
    # and the served BiCGSTAB on that cycle converges like the FP64 one
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(12, (n, 4))
    _, it32, err32 = gpu.krylov("BICGSTAB", A, pv, B, eps=1e-8)
    pv.cnv("F64")
    _, it64, _ = gpu.krylov("BICGSTAB", A, pv, B, eps=1e-8)
    assert np.all(err32 <= 1e-8) and np.all(it32 <= it64 + 2), (it32, it64)


def test_cg_on_the_fp32_cycle(gpu, O):
    """CG hands FP32 / BF16 factors the FP32 copy of r and keeps z FP32: the
    multilevel cycle then runs on FP32 vectors; counts close to the FP64
    cycle's, graph replays bit-identical.  (POST: with these undamped
    smoothers the TWOSIDE cycle, though symmetric, needs some 340 CG
    iterations on this matrix in FP64, too close to failure to compare.)"""
    a = gpu.gen_aniso2d(48, 1e-2)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "AINV", pos="POST")
    pv.set("AINV_ALG", "SYM-LLK")
    pv.build(a)
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(14, (n, 4))
    _, it64, err64 = gpu.krylov("CG", A, pv, B, eps=1e-8, itmax=400)
    pv.cnv("F32")
    x32, it32, err32 = gpu.krylov("CG", A, pv, B, eps=1e-8, itmax=400)
    x32e, it32e, _ = gpu.krylov("CG", A, pv, B, eps=1e-8, itmax=400,
                                eager=True)
    assert np.all(err64 <= 1e-8) and np.all(err32 <= 1e-8)
    assert np.all(np.abs(it32.astype(int) - it64.astype(int)) <= 2), \
        (it32, it64)
    assert np.array_equal(it32, it32e) and np.array_equal(x32, x32e)
    for j in range(4):
        r = B[:, j] - a @ x32[:, j]
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(B[:, j])


@pytest.mark.parametrize("k", [1, 8, 32])
def test_wide_mapping_parity(gpu, O, k):
    """APPLY_MAP=WIDE (multilevel only) runs every level product that has a
    CSR copy (rows of 16 entries or more on average: the anisotropic
    problem's Galerkin levels) one wavefront per row, the rest
    cooperatively; the cycle and the solve match the oracle."""
    a = gpu.gen_aniso2d(64, 1e-3)
    n = a.shape[0]
    pv, ref = build_ml(gpu, O, a, "AINV", nlev=5, theta=0.1, csweeps=2)
    # SYM-LLK 8 / 5e-2 (21 iterations here; LLK 6 needs ~150, too erratic a
    # convergence to compare counts)
    sub = {"kind": "AINV", "alg": "SYM-LLK", "fill_in": 8, "thresh": 5e-2}
    pv.set("AINV_ALG", "SYM-LLK")
    pv.set("SUB_FILLIN", 8)
    pv.build(a)
    ref = O.MlHierarchy(a, nlev=5, theta=0.1, sub=sub, csweeps=2)
    pv.set("APPLY_MAP", "WIDE")
    assert pv.geti("APPLY_MAP") == 7
    X = gpu.gen_normal(23, (n, k))
    Y = pv.apply(X)
    for j in range(0, k, 3):
        yr = ref.apply(X[:, j])
        assert np.linalg.norm(Y[:, j] - yr) <= 1e-11 * np.linalg.norm(yr), j
    assert any(L["n"] > 0 for L in pv.report()["multilevel"]["levels"])
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(24, (n, k))
    x, it, err = gpu.krylov("BICGSTAB", A, pv, B, eps=1e-8, itmax=300)
    for j in range(0, k, 5):
        _, itr, _ = O.ml_krylov("BICGSTAB", a, ref, B[:, j], 1e-8, 300)
        assert abs(int(it[j]) - itr) <= 1, (j, it[j], itr)
    # the one-level kinds refuse it (their factors have no CSR copy)
    sv = gpu.MldDInvkSolver()
    with pytest.raises(gpu.AinvError):
        sv.set("APPLY_MAP", "WIDE")


def test_tuner_report_lists_level_mappings(gpu, O):
    a = gpu.gen_aniso2d(64, 1e-3)
    n = a.shape[0]
    pv, _ = build_ml(gpu, O, a, "AINV", nlev=5, theta=0.1, csweeps=2)
    A = gpu.SystemMatrix(a)
    gpu.krylov("BICGSTAB", A, pv, gpu.gen_normal(25, (n, 16)), eps=1e-6,
               itmax=50)
    maps = [e for e in pv.report()["multilevel"]["maps"] if e["K"] == 16]
    assert len(maps) == pv.levels()
    names = {"split", "row", "coop", "coop128", "coop_q", "coop128_q",
             "run", "wide"}
    for e in maps:
        assert e["a"] in names
    assert any(e["wide"] for e in maps)
