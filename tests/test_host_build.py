"""Host AINV factorization through the C ABI (ainv_host_ainv_bld, no GPU):
bit-identical to the oracle's mld_d_ainv_bld restatement for every
biconjugation variant and for any number of host threads (the pipelined
passes claim columns in order and wait for the finished columns they
read)."""
import numpy as np
import pytest
import scipy.sparse as sp


def same(x, y):
    x, y = sp.csr_matrix(x), sp.csr_matrix(y)
    return (np.array_equal(x.indptr, y.indptr)
            and np.array_equal(x.indices, y.indices)
            and np.array_equal(x.data, y.data))


@pytest.mark.parametrize("alg,fill,thresh", [("LLK", 8, 0.1),
                                             ("LLK", 4, 0.0),
                                             ("SYM-LLK", 8, 5e-2),
                                             ("LLK-NOTH", 6, 0.1),
                                             ("MLK", 8, 0.1),
                                             ("STAB-LLK", 8, 0.1),
                                             ("LLK", 0, 0.1)])
@pytest.mark.parametrize("threads", [1, 2, 3, 8])
def test_host_ainv_matches_oracle(mld, O, alg, fill, thresh, threads):
    a = mld.gen_pde3d(14, 1) if alg != "SYM-LLK" else mld.gen_aniso2d(40)
    if alg == "LLK" and fill == 0:
        a = mld.gen_pde3d(7, 1)  # no fill cap: keeps whole columns
    w, z, d = mld.host_ainv_bld(a, alg, fill, thresh, threads)
    wr, zr, dr = O.ainv_bld(a, alg, fill, thresh)
    assert same(w, wr) and same(z, zr)
    assert np.array_equal(d, dr)


@pytest.mark.parametrize("threads", [2, 5, 8])
def test_host_ainv_random_matrix_threads(mld, O, threads):
    """A random banded matrix (columns within +-5000 plus long-range ones,
    so the nearest coupled column is hundreds of columns back and the
    pipelined passes run): the passes run far ahead of the columns they
    wait for."""
    a = mld.gen_random(20000, seed=41, median=8.0, band=5000, nlong=2)
    w1, z1, d1 = mld.host_ainv_bld(a, "LLK", 8, 5e-2, 1)
    wt, zt, dt = mld.host_ainv_bld(a, "LLK", 8, 5e-2, threads)
    assert same(w1, wt) and same(z1, zt) and np.array_equal(d1, dt)
    wr, zr, dr = O.ainv_bld(a, "LLK", 8, 5e-2)
    assert same(w1, wr) and same(z1, zr) and np.array_equal(d1, dr)


def same_bits(x, y):
    x, y = sp.csr_matrix(x), sp.csr_matrix(y)
    return (x.shape == y.shape and np.array_equal(x.indptr, y.indptr)
            and np.array_equal(x.indices, y.indices)
            and np.array_equal(x.data.view(np.int64), y.data.view(np.int64)))


@pytest.mark.parametrize("mat", ["pde3d", "aniso2d", "random"])
@pytest.mark.parametrize("theta,smoothed", [(0.0, True), (0.1, True),
                                            (0.05, False)])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_host_ml_level_matches_oracle(mld, O, mat, theta, smoothed, threads):
    """One coarsening step of the multilevel build through the C ABI
    (ainv_ml_host_level, no GPU): aggregates, prolongator, Galerkin matrix
    and omega bit-identical to the oracle for any host thread count."""
    a = {"pde3d": lambda: mld.gen_pde3d(20, 1),
         "aniso2d": lambda: mld.gen_aniso2d(64, 1e-3),
         "random": lambda: mld.gen_random(20000, seed=3, band=3000)}[mat]()
    agg, P, Ac, om = mld.ml_host_level(a, theta, smoothed, threads)
    ragg, rP, rR, rAc, rom = O.ml_level(a, theta, smoothed)
    assert np.array_equal(agg, ragg)
    assert same_bits(P, rP) and same_bits(Ac, rAc)
    assert om == rom


def _digests():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "fullsize_hashes.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["random1e6_llk_8_5e-2",
                                  "aniso2048_symllk_8_5e-2"])
def test_host_ainv_fullsize_matches_oracle_digests(mld, name):
    """The library's host biconjugation at the benchmark sizes (the random
    n = 10^6 matrix, pipelined, and the 2048^2 anisotropic Laplacian)
    bit for bit equal to the oracle's factors, through the digests of
    tests/golden/make_fullsize_hashes.py."""
    import refcheck as R
    ref = _digests()[name]
    if ref["matrix"] == "random":
        a = mld.gen_random(ref["size"], seed=40, median=12.0, sigma=1.0,
                           cap=400, band=5000, nlong=2, dom=1.1)
    else:
        a = mld.gen_aniso2d(ref["size"], 1e-3)
    assert R.csr_digest(a) == ref["a"]
    p = ref["params"]
    w, z, d = mld.host_ainv_bld(a, p["alg"], p["fill_in"], p["thresh"], 0)
    got = R.factor_digests(w, z, d)
    for key in ("w_nnz", "z_nnz", "w_pattern", "z_pattern", "w", "z", "d"):
        assert got[key] == ref[key], key


def test_host_ilu_fullsize_digest_inputs(mld):
    """The INVK / INVT digests at 160^3 are of factors of this generator's
    matrix: the generator at full size against the recorded digest (the
    GPU tests compare the factors)."""
    import refcheck as R
    ref = _digests()["pde3d160_invk_1_1"]
    a = mld.gen_pde3d(160)
    assert a.nnz == ref["a_nnz"] and R.csr_digest(a) == ref["a"]
