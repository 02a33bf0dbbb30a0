"""GPU parity of the HIP library against the CPU oracle, through the C ABI.

Bars: factor patterns and values bit-exact (the build kernels reproduce the
reference's operation order with no contraction); FP64 apply within 1e-12
relative; FP32 / BF16 factor storage within the stated tolerances of the FP64
apply; Krylov iteration counts within +-1 of the oracle at the same stopping
tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def same_csr(a, b):
    a, b = a.tocsr(), b.tocsr()
    a.sort_indices()
    b.sort_indices()
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr)
            and np.array_equal(a.indices, b.indices)
            and np.array_equal(a.data, b.data))


def build(mld, kind, a, **params):
    sv = mld.SOLVERS[kind]()
    for k, v in params.items():
        sv.set(k, v)
    sv.build(a)
    return sv


def check_factors(sv, ref):
    w_ref, z_ref, d_ref = ref
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    assert same_csr(w, w_ref), "W differs from the oracle"
    assert same_csr(z, z_ref), "Z differs from the oracle"
    assert np.array_equal(d, d_ref), "d differs from the oracle"


@pytest.mark.parametrize("fill1,fill2", [(0, 0), (0, 1), (1, 1), (1, 2), (2, 1)])
def test_invk_factors_bit_exact(gpu, O, fill1, fill2):
    a = gpu.gen_pde3d(9)
    sv = build(gpu, "INVK", a, SUB_FILLIN=fill1, INV_FILLIN=fill2)
    check_factors(sv, O.invk_bld(a, fill1, fill2))


@pytest.mark.parametrize("fill,thr,ifill,ithr", [
    (10, 1e-2, 10, 1e-2), (0, 1e-1, 0, 1e-1), (3, 1e-3, 2, 1e-2),
    (10, 0.0, 5, 0.0)])
def test_invt_factors_bit_exact(gpu, O, fill, thr, ifill, ithr):
    a = gpu.gen_pde3d(9)
    sv = build(gpu, "INVT", a, SUB_FILLIN=fill, SUB_ILUTHRS=thr,
               INV_FILLIN=ifill, INV_THRESH=ithr)
    check_factors(sv, O.invt_bld(a, fill, thr, ifill, ithr))


def test_invt_random_matrix_bit_exact(gpu, O):
    a = gpu.gen_random(3000, seed=11, median=6.0, band=200)
    sv = build(gpu, "INVT", a, SUB_FILLIN=5, SUB_ILUTHRS=1e-2, INV_FILLIN=4,
               INV_THRESH=5e-2)
    check_factors(sv, O.invt_bld(a, 5, 1e-2, 4, 5e-2))


def test_invk_global_tier_exact_inverse(gpu, O):
    """Rows of the exact inverse exceed the LDS capacity (512 pending
    entries) and go through the global-memory tier."""
    a = gpu.gen_pde3d(10)
    n = a.shape[0]
    sv = build(gpu, "INVK", a, SUB_FILLIN=0, INV_FILLIN=n)
    rep = sv.report()
    assert rep["inverse"]["upper"]["lds_overflow_rows"] > 0
    check_factors(sv, O.invk_bld(a, 0, n))


@pytest.mark.parametrize("alg,thr,fill", [
    ("LLK", 0.1, 8), ("LLK", 0.0, 4), ("LLK", 0.05, 0), ("STAB-LLK", 0.1, 8),
    ("LLK-NOTH", 0.1, 6), ("MLK", 0.1, 8)])
def test_ainv_llk_bit_exact(gpu, O, alg, thr, fill):
    a = gpu.gen_pde3d(8)
    sv = build(gpu, "AINV", a, AINV_ALG=alg, SUB_ILUTHRS=thr, SUB_FILLIN=fill)
    # build calls check: fill 0 is coerced to 1 as the reference does
    check_factors(sv, O.ainv_bld(a, alg, max(fill, 1), thr))


def test_ainv_sym_llk_bit_exact(gpu, O):
    a = gpu.gen_aniso2d(30, 1e-3)
    sv = build(gpu, "AINV", a, AINV_ALG="SYM-LLK", SUB_ILUTHRS=5e-2,
               SUB_FILLIN=8)
    check_factors(sv, O.ainv_bld(a, "SYM-LLK", 8, 5e-2))


@pytest.mark.parametrize("alg", ["LLK", "LLK-NOTH", "MLK", "SYM-LLK"])
def test_ainv_pipelined_build_bit_exact(gpu, O, monkeypatch, alg):
    """The biconjugation pipelined over several host threads (taken when a
    column's dependencies lie far back, as in this random banded matrix,
    whose median distance to the nearest earlier coupling is 96 columns)
    gives the oracle's factors bit for bit, and its applies on the GPU
    match the oracle's."""
    monkeypatch.setenv("AINV_BUILD_THREADS", "4")
    a = gpu.gen_random(6000, seed=23, median=8.0, band=2500)
    sv = build(gpu, "AINV", a, AINV_ALG=alg, SUB_ILUTHRS=5e-2, SUB_FILLIN=8)
    bc = sv.report()["biconjugation"]
    assert bc["pipelined"] and bc["threads"] == 4, bc
    w, z, d = O.ainv_bld(a, alg, 8, 5e-2)
    check_factors(sv, (w, z, d))
    x = gpu.gen_normal(6, (a.shape[0],))
    for trans in ("N", "T"):
        y = sv.apply(x, trans=trans)
        ref = O.apply(w, z, d, x, trans=trans)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    # one thread: the column loop, the same factors
    monkeypatch.setenv("AINV_BUILD_THREADS", "1")
    sv1 = build(gpu, "AINV", a, AINV_ALG=alg, SUB_ILUTHRS=5e-2, SUB_FILLIN=8)
    assert not sv1.report()["biconjugation"]["pipelined"]
    check_factors(sv1, (w, z, d))


def test_tiny_and_diagonal_matrices(gpu, O):
    for a in [sp.csr_matrix(np.array([[4.0]])),
              sp.csr_matrix(np.diag([2.0, -3.0, 5.0])),
              sp.csr_matrix(np.array([[4.0, 1.0], [-1.0, 3.0]]))]:
        for kind, ref in [("INVK", lambda m: O.invk_bld(m, 1, 1)),
                          ("INVT", lambda m: O.invt_bld(m, 2, 0.0, 2, 0.0))]:
            sv = build(gpu, kind, a, SUB_FILLIN=1, INV_FILLIN=1) \
                if kind == "INVK" else build(gpu, kind, a, SUB_FILLIN=2,
                                             INV_FILLIN=2)
            check_factors(sv, ref(a))
            x = np.arange(1.0, a.shape[0] + 1)
            assert np.allclose(sv.apply(x), np.linalg.solve(a.toarray(), x),
                               rtol=1e-12)


@pytest.mark.parametrize("kind", ["INVK", "INVT", "AINV"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_apply_fp64_matches_oracle(gpu, O, kind, trans):
    a = gpu.gen_pde3d(11)
    params = {"INVK": dict(SUB_FILLIN=1, INV_FILLIN=1),
              "INVT": dict(SUB_FILLIN=10, SUB_ILUTHRS=1e-2, INV_FILLIN=10,
                           INV_THRESH=1e-2),
              "AINV": dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=0.1)}
    sv = build(gpu, kind, a, **params[kind])
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    n = a.shape[0]
    x = gpu.gen_normal(3, (n,))
    y0 = gpu.gen_normal(4, (n,))
    for alpha, beta in [(1.0, 0.0), (-2.5, 0.75)]:
        y = sv.apply(x, y=y0.copy(), alpha=alpha, beta=beta, trans=trans)
        ref = O.apply(w, z, d, x, alpha=alpha, beta=beta, y=y0, trans=trans)
        rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        assert rel < 1e-12, rel


@pytest.mark.parametrize("k", [2, 3, 16, 33, 64])
def test_block_apply_matches_columns(gpu, O, k):
    a = gpu.gen_pde3d(10)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    n = a.shape[0]
    x = gpu.gen_normal(5, (n, k))
    y = sv.apply(x)
    for c in range(k):
        ref = O.apply(w, z, d, x[:, c])
        assert np.linalg.norm(y[:, c] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec,tol", [("F64", 1e-12), ("F32", 2e-6),
                                      ("BF16", 2e-2)])
@pytest.mark.parametrize("k", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("mapping", ["ROW", "COOP", "COOP128", "RUN"])
def test_row_mapping_apply(gpu, O, prec, tol, k, mapping):
    """The row-per-lane and cooperative kernels (leading dimension a
    power of two) against the oracle column by column, alpha / beta
    included, and against the split mapping."""
    a = gpu.gen_pde3d(10)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC=prec,
               APPLY_MAP=mapping)
    assert sv.geti("APPLY_MAP") == {"ROW": 1, "COOP": 2, "COOP128": 3,
                                    "RUN": 6}[mapping]
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    n = a.shape[0]
    x = gpu.gen_normal(7, (n, k))
    y0 = gpu.gen_normal(8, (n, k))
    y = sv.apply(x, y=y0.copy(), alpha=-1.5, beta=0.5)
    for c in range(k):
        ref = O.apply(w, z, d, x[:, c], alpha=-1.5, beta=0.5, y=y0[:, c])
        rel = np.linalg.norm(y[:, c] - ref) / np.linalg.norm(ref)
        assert rel < tol, (c, rel)
    sv.set("APPLY_MAP", "SPLIT")
    ys = sv.apply(x, y=y0.copy(), alpha=-1.5, beta=0.5)
    assert np.linalg.norm(ys - y) <= 4 * tol * np.linalg.norm(ys)


@pytest.mark.parametrize("prec", ["F64", "F32", "BF16"])
@pytest.mark.parametrize("k", [1, 8, 32, 64])
def test_run_mapping_bit_identical(gpu, O, prec, k):
    """RUN (row groups of consecutive rows hand gathered x rows on along
    runs of consecutive diagonals and gather only the others) on the 7-point
    stencil's INVK factors, whose rows hold runs of 2-3 diagonals: the same
    products summed in the same order as the cooperative mapping, so the
    results agree bit for bit (both transposes, alpha / beta), and with the
    oracle to the storage precision.  Rows near the grid faces break the
    runs (held columns differ), so both branches of the hand-on run."""
    a = gpu.gen_pde3d(24)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC=prec,
               APPLY_MAP="RUN")
    n = a.shape[0]
    x = gpu.gen_normal(9, (n, k))
    y0 = gpu.gen_normal(10, (n, k))
    if k == 1:
        x, y0 = x[:, 0].copy(), y0[:, 0].copy()
    outs = {}
    for mapping in ("RUN", "COOP"):
        sv.set("APPLY_MAP", mapping)
        outs[mapping] = [sv.apply(x, y=y0.copy(), alpha=0.75, beta=-0.5,
                                  trans=t) for t in ("N", "T")]
    for yr, yc in zip(outs["RUN"], outs["COOP"]):
        assert np.array_equal(yr, yc)
    if k > 1:
        # every mapping sums a row's entries in entry order: a column alone
        # (the split mapping at k = 1) equals the same column in the block
        for c in (0, k - 1):
            y1 = sv.apply(np.ascontiguousarray(x[:, c]),
                          y=np.ascontiguousarray(y0[:, c]), alpha=0.75,
                          beta=-0.5)
            assert np.array_equal(y1, outs["COOP"][0][:, c]), c
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    tol = {"F64": 1e-12, "F32": 2e-6, "BF16": 2e-2}[prec]
    X = x.reshape(n, -1)
    Y0 = y0.reshape(n, -1)
    Y = outs["RUN"][0].reshape(n, -1)
    YT = outs["RUN"][1].reshape(n, -1)
    for c in {0, X.shape[1] - 1}:
        ref = O.apply(w, z, d, X[:, c], alpha=0.75, beta=-0.5, y=Y0[:, c])
        assert np.linalg.norm(Y[:, c] - ref) <= tol * np.linalg.norm(ref)
        ref = O.apply(w, z, d, X[:, c], alpha=0.75, beta=-0.5, y=Y0[:, c],
                      trans="T")
        assert np.linalg.norm(YT[:, c] - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("prec", ["F64", "F32"])
@pytest.mark.parametrize("k", [8, 32, 64])
def test_run_slice_order_bit_identical(gpu, O, prec, k):
    """The run mapping's slice-group order (strips of ORDER_BAND groups of
    4 slices taken across ORDER_PERIOD-group planes of each XCD's range)
    changes only the order in which workgroups take their slices: the
    applies are the linear order's bits for periods and bands that do and
    do not divide the ranges (64 FP32 / 32 FP64 columns run as two column
    chunks), and agree with the oracle to the storage precision."""
    a = gpu.gen_pde3d(40)  # 1,000 slices: 250 groups, 31-32 per XCD range
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC=prec,
               APPLY_MAP="RUN", ORDER_PERIOD=-1)
    n = a.shape[0]
    x = gpu.gen_normal(31, (n, k))
    y0 = gpu.gen_normal(32, (n, k))
    ylin = sv.apply(x, y=y0.copy(), alpha=0.75, beta=-0.5)
    assert sv.report()["slice_order"]["w"]["period"] == 0
    for period, band in [(25, 2), (7, 3), (10, 10), (3, 1), (1000, 7)]:
        sv.set("ORDER_PERIOD", period)
        sv.set("ORDER_BAND", band)
        so = sv.report()["slice_order"]
        assert so["w"] == {"period": period, "band": min(band, period)}
        assert so["z"] == so["w"]
        y = sv.apply(x, y=y0.copy(), alpha=0.75, beta=-0.5)
        assert np.array_equal(y, ylin), (period, band)
    sv.set("ORDER_PERIOD", -1)
    assert sv.report()["slice_order"]["w"]["period"] == 0
    assert np.array_equal(sv.apply(x, y=y0.copy(), alpha=0.75, beta=-0.5),
                          ylin)
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    tol = {"F64": 1e-12, "F32": 2e-6}[prec]
    for c in (0, k - 1):
        ref = O.apply(w, z, d, x[:, c], alpha=0.75, beta=-0.5, y=y0[:, c])
        assert np.linalg.norm(ylin[:, c] - ref) <= tol * np.linalg.norm(ref)


def test_run_slice_order_trial(gpu):
    """ORDER_PERIOD 0 (the default): the first Krylov plan of 32 or more
    columns that runs the run mapping times linear order against strip
    orders over planes of R / 2 and R groups (R the factor's median reach)
    and keeps one, reported with the trial; the solve is the same bits as
    with linear order."""
    a = gpu.gen_pde3d(40)
    n = a.shape[0]
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(33, (n, 32))
    out = {}
    for period in (-1, 0):
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1,
                   FACTOR_PREC="F32", APPLY_MAP="RUN", ORDER_PERIOD=period)
        x, it, err = gpu.krylov("BICGSTAB", A, sv, B, eps=1e-8)
        out[period] = (x, it, sv.report()["slice_order"])
    assert np.array_equal(out[0][0], out[-1][0])
    assert np.array_equal(out[0][1], out[-1][1])
    assert out[-1][2]["trial"] is None
    tr = out[0][2]["trial"]
    assert tr["K"] == 32
    for f in ("w", "z"):
        reach = tr[f]["reach_groups"]
        # 40^3 INVK(1,1): the far cluster lies 2 x 1,600 rows back
        assert 10 <= reach <= 15, reach
        ms = tr[f]["ms"]
        assert ms[0][:2] == [0, 0] and all(m[2] > 0 for m in ms)
        kept = [out[0][2][f]["period"], out[0][2][f]["band"]]
        assert kept == [0, 0] or kept in [m[:2] for m in ms[1:]]
    A.free()


@pytest.mark.parametrize("prec", ["F64", "F32", "BF16"])
@pytest.mark.parametrize("k", [1, 8, 32, 64])
def test_slice_sort_factor_applies_bit_identical(gpu, O, prec, k):
    """Factor layouts whose rows are sorted by length within windows
    (SLICE_SORT) store each row's entries in the same order and write each
    row where it belongs: every mapping's apply, both transposes, is the
    rows-in-order layout's bits.  The automatic window is the host plan's
    for the exported factor, and the sort cuts the random matrix's padding;
    the applies agree with the oracle to the storage precision."""
    a = gpu.gen_random(5000, seed=17, median=12.0, band=1500)
    sv = build(gpu, "AINV", a, AINV_ALG="LLK", SUB_FILLIN=8,
               SUB_ILUTHRS=5e-2, FACTOR_PREC=prec, SLICE_SORT=0)
    n = a.shape[0]
    x = gpu.gen_normal(41, (n, k))
    y0 = gpu.gen_normal(42, (n, k))
    if k == 1:
        x, y0 = x[:, 0].copy(), y0[:, 0].copy()
    maps = ["SPLIT"] if k == 1 else ["SPLIT", "ROW", "COOP", "COOP128",
                                     "COOP_Q", "COOP128_Q", "RUN"]

    def applies():
        out = {}
        for mapping in maps:
            sv.set("APPLY_MAP", mapping)
            out[mapping] = [sv.apply(x, y=y0.copy(), alpha=0.75, beta=-0.5,
                                     trans=t) for t in ("N", "T")]
        return out

    rep = sv.report()
    assert rep["w"]["sort_window"] == 0
    assert rep["w"]["padded"] == rep["w"]["unsorted_padded"]
    ref = applies()
    for window in (256, -1):
        sv.set("SLICE_SORT", window)
        rep = sv.report()
        for f in ("w", "z"):
            fac = sv.export(f.upper())
            _, pad, upad, used = gpu.slice_sort_plan(fac.indptr, window)
            lay = rep[f]
            assert lay["sort_window"] == used
            assert lay["padded"] == pad and lay["unsorted_padded"] == upad
            assert lay["max_width"] >= lay["width_p99"] >= lay["width_p50"]
        if window == 256:
            assert rep["z"]["padded"] < rep["z"]["unsorted_padded"]
        out = applies()
        for mapping in maps:
            for t in (0, 1):
                assert np.array_equal(out[mapping][t], ref[mapping][t]), \
                    (window, mapping, t)
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    tol = {"F64": 1e-12, "F32": 2e-6, "BF16": 2e-2}[prec]
    X, Y0 = x.reshape(n, -1), y0.reshape(n, -1)
    for t in (0, 1):
        Y = ref[maps[0]][t].reshape(n, -1)
        c = X.shape[1] - 1
        r = O.apply(w, z, d, X[:, c], alpha=0.75, beta=-0.5, y=Y0[:, c],
                    trans="NT"[t])
        assert np.linalg.norm(Y[:, c] - r) <= tol * np.linalg.norm(r)


@pytest.mark.parametrize("xprec", ["F64", "F32"])
@pytest.mark.parametrize("k", [1, 8, 64])
def test_slice_sort_system_matrix(gpu, xprec, k):
    """The system matrix's layout sorts rows by length within windows by
    default (the random matrix's rows in order pad ~6x): the products of
    every mapping of the fused product + dots, and the plain product, are
    the rows-in-order layout's bits; the dots, summed over rows in another
    order, agree to rounding; BiCG (which adds the A^T layout) takes the
    same iterations and reaches the same solution to the tolerance."""
    a = gpu.gen_random(3000, seed=5)
    n = a.shape[0]
    A = gpu.SystemMatrix(a)
    rep = A.report()
    _, pad, upad, used = gpu.slice_sort_plan(a.indptr, -1)
    assert rep["slice_sort"] == -1 and rep["a"]["sort_window"] == used > 0
    assert rep["a"]["padded"] == pad < upad == rep["a"]["unsorted_padded"]
    assert pad < 1.5 * a.nnz
    K = max(k, 2)
    x = gpu.gen_normal(43, (n, K))
    if xprec == "F32":
        x = x.astype(np.float32)
    w = gpu.gen_normal(44, (n, K))
    maps = ("SPLIT", "ROW", "COOP", "COOP128", "RUN")

    def products():
        out = {m: [A.spmv_dot(x, w, np_=p, mapping=m) for p in (1, 2)]
               for m in maps}
        X = gpu.DeviceArray.from_numpy(np.ascontiguousarray(
            x[:, :k].astype(np.float64)))
        Y = gpu.DeviceArray((n, k) if k > 1 else (n,))
        Y.zero()
        A.spmm(X, Y)
        out["spmm"] = Y.to_numpy()
        X.free()
        Y.free()
        return out

    srt = products()
    A.set_slice_sort(0)
    rep = A.report()
    assert rep["slice_sort"] == 0 and rep["a"]["sort_window"] == 0
    assert rep["a"]["padded"] == upad
    ord_ = products()
    assert np.array_equal(srt["spmm"], ord_["spmm"])
    y_ref = a @ x[:, :k].astype(np.float64)
    assert np.linalg.norm(srt["spmm"].reshape(n, -1) - y_ref) <= \
        1e-14 * np.linalg.norm(y_ref)
    nw = np.linalg.norm(w, axis=0)
    for m in maps:
        for p, ((ys, ds, _), (yo, do, _)) in enumerate(zip(srt[m], ord_[m])):
            assert np.array_equal(ys, yo), m
            ny = np.linalg.norm(ys, axis=0)
            # dots to rounding of the sums' magnitudes: |y| |w|, |y|^2
            scale = (np.stack([ny * ny, ny * nw]) if p else
                     (ny * nw)[None, :])
            assert np.all(np.abs(ds - do) <= 1e-13 * scale), m
    # BiCG on both layouts (A^T is built on first use, in the same sort)
    sv = build(gpu, "AINV", a, AINV_ALG="LLK", SUB_FILLIN=8,
               SUB_ILUTHRS=5e-2)
    b = gpu.gen_normal(45, (n, k)) if k > 1 else gpu.gen_normal(45, (n,))
    xo, ito, _ = gpu.krylov("BICG", A, sv, b, eps=1e-8)
    A.set_slice_sort(-1)
    xs, its, _ = gpu.krylov("BICG", A, sv, b, eps=1e-8)
    assert A.report()["a_t"]["sort_window"] > 0
    assert np.all(np.abs(np.asarray(its) - np.asarray(ito)) <= 1)
    assert np.linalg.norm(xs - xo) <= 1e-6 * np.linalg.norm(xo)
    A.free()


@pytest.mark.parametrize("kind", ["pde3d", "aniso2d", "random"])
@pytest.mark.parametrize("xprec", ["F64", "F32"])
@pytest.mark.parametrize("k", [2, 8, 16, 32, 64])
def test_fused_product_mappings(gpu, kind, xprec, k):
    """The Krylov step's fused system-matrix product v = A f and its dots
    in every mapping: the products agree bit for bit across mappings (each
    sums a row's entries in entry order), including the run mapping that
    hands x rows on along the stencils' runs of diagonals (the random
    matrix has almost none, so its rows gather everything), and with the
    FP64 CSR product to rounding; the dots (BiCGSTAB's (r~, v), and (t, t),
    (t, s)) agree with numpy's to rounding."""
    a = {"pde3d": lambda: gpu.gen_pde3d(14),
         "aniso2d": lambda: gpu.gen_aniso2d(50, 1e-3),
         "random": lambda: gpu.gen_random(3000, seed=5)}[kind]()
    n = a.shape[0]
    A = gpu.SystemMatrix(a)
    x = gpu.gen_normal(21, (n, k))
    if xprec == "F32":
        x = x.astype(np.float32)
    w = gpu.gen_normal(22, (n, k))
    y_ref = a @ x.astype(np.float64)
    for np_ in (1, 2):
        ys = {}
        for mapping in ("SPLIT", "ROW", "COOP", "COOP128", "RUN"):
            y, dots, _ = A.spmv_dot(x, w, np_=np_, mapping=mapping)
            ys[mapping] = y
            assert np.linalg.norm(y - y_ref) <= 1e-14 * np.linalg.norm(y_ref)
            ny = np.linalg.norm(y, axis=0)
            nw = np.linalg.norm(w, axis=0)
            if np_ == 1:
                ref = np.einsum("ij,ij->j", w, y)[None, :]
                scale = (nw * ny)[None, :]
            else:
                ref = np.stack([np.einsum("ij,ij->j", y, y),
                                np.einsum("ij,ij->j", y, w)])
                scale = np.stack([ny * ny, ny * nw])
            assert np.all(np.abs(dots - ref) <= 1e-13 * scale), mapping
        for mapping, y in ys.items():
            assert np.array_equal(y, ys["SPLIT"]), mapping
    A.free()


@pytest.mark.parametrize("method,k,mapping", [
    ("BICGSTAB", 8, "ROW"), ("BICGSTAB", 5, "ROW"), ("CG", 16, "ROW"),
    ("BICGSTAB", 8, "COOP"), ("BICGSTAB", 3, "COOP"), ("CG", 32, "COOP"),
    ("BICGSTAB", 40, "COOP"), ("BICGSTAB", 64, "COOP128"),
    ("CG", 32, "COOP128"), ("BICGSTAB", 32, "RUN"), ("CG", 16, "RUN")])
def test_row_mapping_krylov(gpu, O, method, k, mapping):
    """Krylov solves with the row-mapped apply and system-matrix product
    (a padded leading dimension when k is not a power of two): iteration
    counts within +-1 of the oracle per column."""
    if method == "CG":
        a = gpu.gen_aniso2d(40, 1e-2)
        sv = build(gpu, "AINV", a, AINV_ALG="SYM-LLK", SUB_FILLIN=8,
                   SUB_ILUTHRS=5e-2, APPLY_MAP=mapping)
        w, z, d = O.ainv_bld(a, "SYM-LLK", 8, 5e-2)
    else:
        a = gpu.gen_pde3d(12)
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1,
                   APPLY_MAP=mapping)
        w, z, d = O.invk_bld(a, 1, 1)
    n = a.shape[0]
    B = gpu.gen_normal(11, (n, k))
    A = gpu.SystemMatrix(a)
    X, it, err = gpu.krylov(method, A, sv, B, eps=1e-8)
    for c in range(k):
        _, it_ref, _ = O.krylov(method, a, w, z, d, B[:, c], 1e-8)
        assert abs(int(it[c]) - it_ref) <= 1, (c, it[c], it_ref)
        assert np.linalg.norm(B[:, c] - a @ X[:, c]) <= \
            1e-7 * np.linalg.norm(B[:, c])


def test_autotuner_report(gpu):
    """AUTO mapping: a block solve tunes the apply once per block width
    and reports both timings and the choice."""
    a = gpu.gen_pde3d(16)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(12, (a.shape[0], 8))
    gpu.krylov("BICGSTAB", A, sv, B, eps=1e-6)
    rep = sv.report()
    assert rep["apply_map"] == "auto"
    (t,) = [e for e in rep["tuning"] if e["K"] == 8]
    for pass_, key in (("w_ms", "map_w"), ("z_ms", "map_z")):
        ms = t[pass_]
        assert set(ms) == {"split", "row", "coop", "coop128", "coop_q",
                           "coop128_q", "run"}
        assert min(ms.values()) > 0
        assert t[key] == min(ms, key=ms.get)


@pytest.mark.parametrize("prec,tol", [("F32", 2e-6), ("BF16", 2e-2)])
def test_reduced_precision_factors(gpu, O, prec, tol):
    """FP32 / BF16 factor storage: error against the FP64 apply bounded by
    the storage precision (the factor values are the only rounded input)."""
    a = gpu.gen_pde3d(12)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC=prec)
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    x = gpu.gen_normal(6, (a.shape[0], 4))
    y = sv.apply(x)
    for c in range(4):
        ref = O.apply(w, z, d, x[:, c])
        rel = np.linalg.norm(y[:, c] - ref) / np.linalg.norm(ref)
        assert rel < tol, (prec, rel)
    sv.cnv("F64")
    y64 = sv.apply(x[:, 0])
    ref = O.apply(w, z, d, x[:, 0])
    assert np.linalg.norm(y64 - ref) < 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("method,prec", [("BICGSTAB", "F32"),
                                         ("BICGSTAB", "BF16"), ("CG", "F32"),
                                         ("CG", "BF16")])
def test_reduced_precision_krylov(gpu, O, method, prec):
    """With FP32 / BF16 factors the Krylov engine hands the preconditioner
    FP32 copies of its inputs and keeps the preconditioned vectors in FP32
    (recurrences, dots and residuals stay FP64): iteration counts stay close
    to the FP64-factor solve's (within 2 for FP32, 25 % + 2 for BF16), every
    column reaches the tolerance in true residual, and the oracle's FP64
    solve with the FP32 factors' values agrees on the count within 2."""
    if method == "CG":
        a = gpu.gen_aniso2d(48, 1e-2)
        sv = build(gpu, "AINV", a, AINV_ALG="SYM-LLK", SUB_FILLIN=6,
                   SUB_ILUTHRS=5e-2)
    else:
        a = gpu.gen_pde3d(12)
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    n = a.shape[0]
    B = gpu.gen_normal(29, (n, 5))
    A = gpu.SystemMatrix(a)
    _, it64, _ = gpu.krylov(method, A, sv, B, eps=1e-8)
    sv.cnv(prec)
    X, it, err = gpu.krylov(method, A, sv, B, eps=1e-8)
    for c in range(5):
        if prec == "F32":
            assert abs(int(it[c]) - int(it64[c])) <= 2, (c, it, it64)
        else:
            assert it[c] <= 1.25 * it64[c] + 2, (c, it, it64)
        assert err[c] <= 1e-8
        assert np.linalg.norm(B[:, c] - a @ X[:, c]) <= \
            1e-7 * np.linalg.norm(B[:, c])
    if prec == "F32":
        # the factors as stored (FP32 values) in an FP64 oracle solve
        w = sv.export("W")
        z = sv.export("Z")
        d = sv.export("D")
        _, it_ref, _ = O.krylov(method, a, w, z, d, B[:, 0], 1e-8)
        assert abs(int(it[0]) - it_ref) <= 2, (it[0], it_ref)


def test_krylov_bicgstab_iterations(gpu, O):
    a = gpu.gen_pde3d(14)
    n = a.shape[0]
    b = a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    for kind, params, ref in [
            ("INVK", dict(SUB_FILLIN=1, INV_FILLIN=1),
             lambda: O.invk_bld(a, 1, 1)),
            ("AINV", dict(SUB_FILLIN=8, SUB_ILUTHRS=0.1),
             lambda: O.ainv_bld(a, "LLK", 8, 0.1))]:
        sv = build(gpu, kind, a, **params)
        x, it, err = gpu.krylov("BICGSTAB", A, sv, b, eps=1e-8)
        w, z, d = ref()
        _, it_ref, _ = O.krylov("BICGSTAB", a, w, z, d, b, 1e-8)
        assert abs(int(it[0]) - it_ref) <= 1, (kind, it, it_ref)
        true = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
        assert err[0] <= 1e-8 and true < 1e-7


def test_krylov_cg_block(gpu, O):
    a = gpu.gen_aniso2d(40, 1e-3)
    n = a.shape[0]
    k = 4
    B = gpu.gen_normal(3, (n, k))
    A = gpu.SystemMatrix(a)
    sv = build(gpu, "AINV", a, AINV_ALG="SYM-LLK", SUB_FILLIN=8,
               SUB_ILUTHRS=5e-2)
    X, it, err = gpu.krylov("CG", A, sv, B, eps=1e-8)
    w, z, d = O.ainv_bld(a, "SYM-LLK", 8, 5e-2)
    for c in range(k):
        _, it_ref, _ = O.krylov("CG", a, w, z, d, B[:, c], 1e-8)
        assert abs(int(it[c]) - it_ref) <= 1, (c, it[c], it_ref)
        assert np.linalg.norm(B[:, c] - a @ X[:, c]) <= \
            1e-7 * np.linalg.norm(B[:, c])


@pytest.mark.parametrize("k", [1, 5])
def test_krylov_cgs(gpu, O, k):
    """CGS on the GPU: per-column iteration counts within +-1 of the
    oracle's CGS (itself checked against scipy's cgs), true residuals."""
    a = gpu.gen_pde3d(12)
    n = a.shape[0]
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    w, z, d = O.invk_bld(a, 1, 1)
    B = gpu.gen_normal(15, (n, k)) if k > 1 else a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    X, it, err = gpu.krylov("CGS", A, sv, B, eps=1e-8)
    Bc = B.reshape(n, -1)
    Xc = X.reshape(n, -1)
    for c in range(k):
        _, it_ref, _ = O.krylov("CGS", a, w, z, d, Bc[:, c], 1e-8)
        assert abs(int(it[c]) - it_ref) <= 1, (c, it[c], it_ref)
        assert np.linalg.norm(Bc[:, c] - a @ Xc[:, c]) <= \
            1e-7 * np.linalg.norm(Bc[:, c])


@pytest.mark.parametrize("kind,k", [("INVK", 1), ("INVK", 5), ("AINV", 3),
                                    ("NONE", 2)])
def test_krylov_bicg(gpu, O, kind, k):
    """BiCG on the GPU (A^T and M^T through the transposed layouts):
    per-column iteration counts within +-1 of the oracle's BiCG (itself
    checked against scipy's bicg), true residuals."""
    a = gpu.gen_pde3d(12)
    n = a.shape[0]
    if kind == "INVK":
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
        w, z, d = O.invk_bld(a, 1, 1)
    elif kind == "AINV":
        sv = build(gpu, "AINV", a, AINV_ALG="LLK", SUB_FILLIN=8,
                   SUB_ILUTHRS=0.1)
        w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    else:
        sv = None
        w = z = sp.identity(n, format="csr")
        d = np.ones(n)
    B = gpu.gen_normal(17, (n, k)) if k > 1 else a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    X, it, err = gpu.krylov("BICG", A, sv, B, eps=1e-8, itmax=400)
    Bc = B.reshape(n, -1)
    Xc = X.reshape(n, -1)
    for c in range(k):
        _, it_ref, _ = O.krylov("BICG", a, w, z, d, Bc[:, c], 1e-8, 400)
        assert abs(int(it[c]) - it_ref) <= 1, (c, it[c], it_ref)
        assert np.linalg.norm(Bc[:, c] - a @ Xc[:, c]) <= \
            1e-7 * np.linalg.norm(Bc[:, c])


@pytest.mark.parametrize("kind,k,irst", [("INVK", 1, 10), ("INVK", 5, 2),
                                         ("AINV", 3, 30), ("NONE", 2, 5),
                                         ("INVK", 33, 8)])
def test_krylov_rgmres(gpu, O, kind, k, irst):
    """Restarted GMRES on the GPU: per-column iteration counts within +-1 of
    the oracle's RGMRES (itself pinned by the least-squares minimum over the
    Krylov space and by scipy's gmres), true residuals."""
    a = gpu.gen_pde3d(12)
    n = a.shape[0]
    if kind == "INVK":
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
        w, z, d = O.invk_bld(a, 1, 1)
    elif kind == "AINV":
        sv = build(gpu, "AINV", a, AINV_ALG="LLK", SUB_FILLIN=8,
                   SUB_ILUTHRS=0.1)
        w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    else:
        sv = None
        w = z = sp.identity(n, format="csr")
        d = np.ones(n)
    B = gpu.gen_normal(19, (n, k)) if k > 1 else a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    X, it, err = gpu.krylov("RGMRES", A, sv, B, eps=1e-8, itmax=600,
                            irst=irst)
    Bc = B.reshape(n, -1)
    Xc = X.reshape(n, -1)
    for c in range(k):
        _, it_ref, _ = O.krylov("RGMRES", a, w, z, d, Bc[:, c], 1e-8, 600,
                                irst=irst)
        assert abs(int(it[c]) - it_ref) <= 1, (c, it[c], it_ref)
        assert err[c] <= 1e-8
        assert np.linalg.norm(Bc[:, c] - a @ Xc[:, c]) <= \
            1e-7 * np.linalg.norm(Bc[:, c])


@pytest.mark.parametrize("kind,k,l", [("INVK", 1, 1), ("INVK", 5, 2),
                                      ("AINV", 3, 4), ("NONE", 2, 2),
                                      ("INVK", 17, 3)])
def test_krylov_bicgstabl(gpu, O, kind, k, l):
    """BiCGStab(l) on the GPU: per-column iteration counts within one cycle
    (+-l) of the oracle's BiCGStab(l) (itself pinned by BiCGStab(1) =
    BiCGStab and a least-squares restatement), true residuals."""
    a = gpu.gen_pde3d(12)
    n = a.shape[0]
    if kind == "INVK":
        sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
        w, z, d = O.invk_bld(a, 1, 1)
    elif kind == "AINV":
        sv = build(gpu, "AINV", a, AINV_ALG="LLK", SUB_FILLIN=8,
                   SUB_ILUTHRS=0.1)
        w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    else:
        sv = None
        w = z = sp.identity(n, format="csr")
        d = np.ones(n)
    B = gpu.gen_normal(23, (n, k)) if k > 1 else a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    X, it, err = gpu.krylov("BICGSTABL", A, sv, B, eps=1e-8, itmax=800,
                            irst=l)
    Bc = B.reshape(n, -1)
    Xc = X.reshape(n, -1)
    for c in range(k):
        _, it_ref, _ = O.krylov("BICGSTABL", a, w, z, d, Bc[:, c], 1e-8, 800,
                                irst=l)
        assert it[c] > 0 and it[c] % l == 0
        assert abs(int(it[c]) - it_ref) <= l, (c, it[c], it_ref)
        assert err[c] <= 1e-8
        assert np.linalg.norm(Bc[:, c] - a @ Xc[:, c]) <= \
            1e-7 * np.linalg.norm(Bc[:, c])


def test_rgmres_itmax_and_served_form(gpu, O):
    """RGMRES stops at itmax mid-cycle with x updated from the steps taken
    (the oracle's x to rounding); the served form refuses it."""
    a = gpu.gen_pde3d(10)
    n = a.shape[0]
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    w, z, d = O.invk_bld(a, 1, 1)
    b = a @ np.ones(n)
    A = gpu.SystemMatrix(a)
    x, it, err = gpu.krylov("RGMRES", A, sv, b, eps=1e-14, itmax=7, irst=5)
    xr, itr, errr = O.krylov("RGMRES", a, w, z, d, b, 1e-14, 7, irst=5)
    assert int(it[0]) == itr == 7
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
    assert err[0] == pytest.approx(errr, rel=1e-6)
    with pytest.raises(gpu.AinvError) as e:
        gpu.krylov_stream("RGMRES", A, sv,
                          gpu.DeviceArray.from_numpy(np.stack([b, b])),
                          slots=2)
    assert "RGMRES" in str(e.value)


@pytest.mark.parametrize("method", ["BICGSTAB", "CG", "CGS", "BICG",
                                    "RGMRES", "BICGSTABL"])
def test_graph_replays_match_eager(gpu, method):
    """The captured iteration graph, replayed solve after solve with eager
    work and synchronizations in between, gives bit-identical iterates and
    counts to launching the same kernels one by one."""
    a = gpu.gen_aniso2d(48, 1e-2) if method == "CG" else gpu.gen_pde3d(12)
    n = a.shape[0]
    B = gpu.gen_normal(21, (n, 3))
    A = gpu.SystemMatrix(a)
    kind = dict(AINV_ALG="SYM-LLK", SUB_FILLIN=6, SUB_ILUTHRS=5e-2) \
        if method == "CG" else dict(SUB_FILLIN=1, INV_FILLIN=1)
    sv = build(gpu, "AINV" if method == "CG" else "INVK", a, **kind)
    x_e, it_e, err_e = gpu.krylov(method, A, sv, B, eps=1e-9, eager=True)
    for rep in range(4):
        x_g, it_g, err_g = gpu.krylov(method, A, sv, B, eps=1e-9)
        assert np.array_equal(it_g, it_e), (rep, it_g, it_e)
        assert np.array_equal(x_g, x_e), rep
        assert np.array_equal(err_g, err_e), rep
        sv.apply(B[:, 0])  # eager work between replays
    # a rebuilt preconditioner invalidates the captured graph
    sv.cnv("F32")
    x32, it32, _ = gpu.krylov(method, A, sv, B, eps=1e-9)
    x32e, it32e, _ = gpu.krylov(method, A, sv, B, eps=1e-9, eager=True)
    assert np.array_equal(it32, it32e) and np.array_equal(x32, x32e)


def test_error_paths(gpu):
    sv = gpu.MldDInvkSolver()
    x = np.ones(10)
    with pytest.raises(gpu.AinvError) as e:
        sv.apply(x)
    assert e.value.code == gpu.ERR_STATE
    a = gpu.gen_pde3d(4)
    sv = build(gpu, "INVK", a, SUB_FILLIN=1, INV_FILLIN=1)
    with pytest.raises(gpu.AinvError) as e:
        sv.apply(np.ones(a.shape[0]), trans="X")
    assert e.value.code == gpu.ERR_ARG
    bad = gpu.MldDAinvSolver()
    bad.set("AINV_ALG", "NO-SUCH-ALG")
    with pytest.raises(gpu.AinvError) as e:
        bad.build(a)
    assert e.value.code == gpu.ERR_ALG


def test_clone_clear_and_accessors(gpu, O):
    a = gpu.gen_pde3d(8)
    sv = build(gpu, "INVT", a, SUB_FILLIN=4, SUB_ILUTHRS=1e-2, INV_FILLIN=4,
               INV_THRESH=1e-2)
    c = sv.clone()
    x = gpu.gen_normal(9, (a.shape[0],))
    assert np.array_equal(sv.apply(x), c.apply(x))
    w, z, d = O.invt_bld(a, 4, 1e-2, 4, 1e-2)
    assert sv.get_nzeros() == a.shape[0] + w.nnz + z.nnz
    assert sv.get_wrksz() == 2
    assert sv.sizeof() > 0
    sv.clear_data()
    with pytest.raises(gpu.AinvError):
        sv.apply(x)
    c.free()


def test_dump_and_import_roundtrip(gpu, O, tmp_path):
    """sv%dump writes the reference's three files (names, MatrixMarket
    coordinate W / Z, dense d, the head line); reading them back gives the
    resident factors exactly, solver=False writes nothing, and the factors
    imported into another object apply like the original and the oracle."""
    import os
    a = gpu.gen_pde3d(9)
    sv = build(gpu, "INVT", a, SUB_FILLIN=5, SUB_ILUTHRS=1e-2, INV_FILLIN=5,
               INV_THRESH=1e-2)
    files = sv.dump(level=2, prefix=str(tmp_path / "run"), head="invt test")
    assert [os.path.basename(f) for f in files] == [
        "run_p000_l002_wmat.mtx", "run_p000_l002_diag.mtx",
        "run_p000_l002_zmat.mtx"]
    with open(files[0]) as f:
        assert f.readline().startswith("%%MatrixMarket matrix coordinate")
        assert f.readline().strip() == "% invt test"
    w, d, z = (gpu.mm_read(files[0]), gpu.mm_read_vector(files[1]),
               gpu.mm_read(files[2]))
    assert same_csr(w, sv.export("W")) and same_csr(z, sv.export("Z"))
    assert np.array_equal(d, sv.export("D"))
    sv.dump(level=3, prefix=str(tmp_path / "none"), solver=False)
    assert not list(tmp_path.glob("none*"))
    sv2 = gpu.MldDAinvSolver()
    sv2.import_factors(w, z, d)
    x = gpu.gen_normal(9, (a.shape[0],))
    assert np.array_equal(sv2.apply(x), sv.apply(x))
    ref = O.apply(w, z, d, x)
    assert np.linalg.norm(sv2.apply(x) - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.array_equal(sv2.apply(x, trans="T"), sv.apply(x, trans="T"))


@pytest.mark.parametrize("kind", ["INVK", "AINV"])
def test_update_a_matches_oracle(gpu, O, kind):
    """sv%update_a on the GPU: the updated pivots equal the oracle's bit for
    bit, the factors stay, the apply uses the new pivots (also through a
    captured Krylov graph, recaptured after the update), and a short x is
    rejected."""
    a = gpu.gen_pde3d(10)
    n = a.shape[0]
    params = (dict(SUB_FILLIN=1, INV_FILLIN=1) if kind == "INVK" else
              dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=0.1))
    sv = build(gpu, kind, a, **params)
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    x = gpu.gen_normal(12, (n,)) * 0.1 + 0.5
    A = gpu.SystemMatrix(a)
    b = a @ np.ones(n)
    gpu.krylov("BICGSTAB", A, sv, b, eps=1e-8)      # capture a graph first
    sv.update_a(x)
    d_ref = O.update_a(w, z, d, x)
    assert np.array_equal(sv.export("D"), d_ref)
    assert same_csr(sv.export("W"), w) and same_csr(sv.export("Z"), z)
    v = gpu.gen_normal(13, (n,))
    ref = O.apply(w, z, d_ref, v)
    assert np.linalg.norm(sv.apply(v) - ref) <= 1e-12 * np.linalg.norm(ref)
    # the shifted system solved with the updated preconditioner
    ash = (a + sp.diags(x)).tocsr()
    Ash = gpu.SystemMatrix(ash)
    bsh = ash @ np.ones(n)
    xs, it, err = gpu.krylov("BICGSTAB", Ash, sv, bsh, eps=1e-8)
    _, it_ref, _ = O.krylov("BICGSTAB", ash, w, z, d_ref, bsh, 1e-8)
    assert abs(int(it[0]) - it_ref) <= 1
    assert np.linalg.norm(bsh - ash @ xs) <= 1e-7 * np.linalg.norm(bsh)
    # the graph captured before the update is recaptured with the new d
    _, it2, _ = gpu.krylov("BICGSTAB", A, sv, b, eps=1e-8)
    _, it2_ref, _ = O.krylov("BICGSTAB", a, w, z, d_ref, b, 1e-8)
    assert abs(int(it2[0]) - it2_ref) <= 1
    with pytest.raises(gpu.AinvError):
        sv.update_a(x[:-1])
