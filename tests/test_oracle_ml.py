"""Pins of the oracle's multilevel restatement (smoothed aggregation with
AINV-family smoothers) against closed forms, invariants and independent
library computations.  MLD2P4, where the reference's multilevel framework
lives, is absent from the reference tree; what pins the restatement is
listed per test:

* aggregation: closed form on the 1D Laplacian, singletons when every
  coupling is weak, the plus-shaped first aggregate of a 2D grid, and
  partition / phase-1 structure invariants on nonsymmetric matrices;
* sparse product: exact equality with scipy on integer-valued rectangular
  matrices (any summation order gives the same integers);
* prolongator: omega = 2/3 on the 1D Laplacian, constants preserved when A
  has zero row sums, equality with a dense numpy evaluation;
* Galerkin matrix: dense R A P, symmetry, zero row sums carried to the
  coarse level;
* cycle: with an exact coarsest-level solve, M A P = P (POST),
  R A M A = R A (PRE) and R A (M - S) = R (ADD), which a swapped R / P, a
  missing correction or a wrong smoothing order breaks; dense evaluation of
  the recursion for every cycle type and sweep count;
* Krylov: iteration counts of scipy's methods with the cycle as M.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle as O  # noqa: E402


def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)],
                    [-1, 0, 1], format="csr")


def grid2d(m, conv=0.0):
    """5-point operator on an m x m grid, optional first-order convection."""
    i1 = lap1d(m)
    a = sp.kron(sp.eye(m), i1) + sp.kron(i1, sp.eye(m))
    if conv:
        c = sp.diags([-np.ones(m - 1), np.ones(m)], [-1, 0])
        a = a + conv * sp.kron(sp.eye(m), c)
    return sp.csr_matrix(a)


def graph_laplacian(m, seed):
    """Zero-row-sum Laplacian of a random weighted graph on an m x m grid."""
    rng = np.random.default_rng(seed)
    n = m * m
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in (i + 1, i + m):
            if (j == i + 1 and (i % m) == m - 1) or j >= n:
                continue
            w = rng.uniform(0.5, 2.0)
            rows += [i, j]
            cols += [j, i]
            vals += [-w, -w]
    a = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return sp.csr_matrix(a - sp.diags(np.asarray(a.sum(axis=1)).ravel()))


def rand_nonsym(n, seed, dens=0.01):
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=dens, random_state=rng, format="csr")
    a = a - a.T * rng.uniform(0.2, 0.8)
    d = np.asarray(abs(a).sum(axis=1)).ravel() + 1.0
    return sp.csr_matrix(a + sp.diags(d))


# ---------------------------------------------------------------- aggregation
@pytest.mark.parametrize("n", [2, 3, 11, 12, 13, 14, 30])
def test_aggregate_1d_closed_form(n):
    # roots at 0, 3, 6, ...: {0,1}, {2,3,4}, {5,6,7}, ...; a last node left
    # alone (n = 0 mod 3) joins its neighbour's aggregate in phase 2
    agg, na = O.aggregate(lap1d(n), 0.0)
    want = (np.arange(n) + 1) // 3
    if n % 3 == 0:
        want[-1] -= 1
    assert np.array_equal(agg, want)
    assert na == want.max() + 1


def test_aggregate_all_weak_gives_singletons():
    a = grid2d(9)
    agg, na = O.aggregate(a, 10.0)
    assert na == a.shape[0] and np.array_equal(agg, np.arange(a.shape[0]))


def test_aggregate_2d_first_aggregates():
    m = 8
    agg, _ = O.aggregate(grid2d(m), 0.0)
    # node 0 roots {0, 1, m}; node 2 is the next free node whose neighbours
    # (1, 3, 2 + m) include an aggregated one, so the next root is 3:
    # {2, 3, 4, 3 + m}
    assert set(np.flatnonzero(agg == 0)) >= {0, 1, m}
    assert agg[0] == agg[1] == agg[m] == 0
    assert agg[3] == agg[4] == agg[3 + m] == 1


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("theta", [0.0, 0.08, 0.25])
def test_aggregate_invariants(seed, theta):
    a = rand_nonsym(600, seed)
    agg, na = O.aggregate(a, theta)
    assert agg.min() == 0 and agg.max() == na - 1
    assert np.all(np.bincount(agg, minlength=na) > 0)
    d = a.diagonal()
    # strength graph, independently
    rows = np.repeat(np.arange(a.shape[0]), np.diff(a.indptr))
    strong = (np.abs(a.data) > theta * np.sqrt(np.abs(d[rows] * d[a.indices]))
              ) & (rows != a.indices)
    S = sp.csr_matrix((np.ones(strong.sum()), (rows[strong],
                                               a.indices[strong])),
                      shape=a.shape)
    # every aggregate contains a node whose strong neighbourhood it covers
    # or is connected through strong links to a member (phases 2 and 3)
    for c in range(na):
        mem = set(np.flatnonzero(agg == c))
        rooted = any(set(S[i].indices) <= mem for i in mem)
        linked = len(mem) == 1 or all(
            set(S[i].indices) & mem or any(i in set(S[j].indices) for j in mem)
            for i in mem)
        assert rooted or linked, (c, mem)
    # a node with a strong neighbour is never left a singleton unless all
    # its strong neighbours are taken by phase-3 aggregates before it
    assert na < a.shape[0] or theta > 0


# ---------------------------------------------------------------- products
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spgemm_exact_on_integers(seed):
    rng = np.random.default_rng(seed)
    a = sp.random(300, 200, density=0.03, random_state=rng, format="csr")
    b = sp.random(200, 170, density=0.04, random_state=rng, format="csr")
    a.data = rng.integers(-5, 6, a.nnz).astype(float)
    b.data = rng.integers(-5, 6, b.nnz).astype(float)
    c = O.spgemm(a, b)
    assert c.shape == (300, 170)
    assert np.array_equal(c.toarray(), (a @ b).toarray())
    # sorted columns, pattern = symbolic product (cancelled entries kept)
    for i in range(c.shape[0]):
        cols = c.indices[c.indptr[i]:c.indptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
    pa = a.copy()
    pa.data[:] = 1
    pb = b.copy()
    pb.data[:] = 1
    assert c.nnz == (pa @ pb).nnz


def test_prolongator_1d_omega():
    _, _, _, _, om = O.ml_level(lap1d(20), 0.0)
    assert om == pytest.approx(2.0 / 3.0, rel=1e-15)


@pytest.mark.parametrize("seed", [4, 5])
def test_prolongator_preserves_constants(seed):
    a = graph_laplacian(16, seed)
    agg, P, R, Ac, om = O.ml_level(a, 0.05)
    ones_c = np.ones(P.shape[1])
    assert np.allclose(P @ ones_c, 1.0, atol=1e-13)
    # and the Galerkin matrix keeps zero row sums
    assert np.allclose(Ac @ ones_c, 0.0, atol=1e-12)


@pytest.mark.parametrize("smoothed", [True, False])
@pytest.mark.parametrize("theta", [0.0, 0.1])
def test_prolongator_and_galerkin_dense(smoothed, theta):
    a = sp.csr_matrix(grid2d(12, conv=0.7) + rand_nonsym(144, 9, 0.02) * 0.1)
    agg, P, R, Ac, om = O.ml_level(a, theta, smoothed)
    n, nc = P.shape
    A = a.toarray()
    T = np.zeros((n, nc))
    T[np.arange(n), agg] = 1.0
    if smoothed:
        DA = A / np.diag(A)[:, None]
        assert om == pytest.approx(4.0 / (3.0 * np.abs(DA).sum(axis=1).max()),
                                   rel=1e-14)
        Pd = T - om * DA @ T
    else:
        assert om == 0.0
        Pd = T
    assert np.allclose(P.toarray(), Pd, rtol=0, atol=1e-14)
    assert np.array_equal(R.toarray(), P.toarray().T)
    G = Pd.T @ A @ Pd
    assert np.allclose(Ac.toarray(), G, rtol=0, atol=1e-12 * np.abs(G).max())


def test_galerkin_symmetric():
    a = grid2d(14)
    _, _, _, Ac, _ = O.ml_level(a, 0.1)
    D = Ac.toarray()
    assert np.allclose(D, D.T, rtol=0, atol=1e-14 * np.abs(D).max())


# ---------------------------------------------------------------- cycle
EXACT = {"kind": "INVK", "fill_in": 10 ** 6, "inv_fill": 10 ** 6}


def exact_pair(a):
    """factors whose apply is A^{-1} up to rounding"""
    w, z, d = O.invk_bld(a, 10 ** 6, 10 ** 6)
    S = (z @ sp.diags(d) @ w).toarray()
    assert np.allclose(S @ a.toarray(), np.eye(a.shape[0]), atol=1e-9)


@pytest.mark.parametrize("sub", ["INVK", "AINV", "INVT"])
def test_cycle_post_exact_coarse_keeps_range_of_p(sub):
    a = grid2d(10, conv=0.5)
    exact_pair(a @ sp.eye(a.shape[0]))
    H = O.MlHierarchy(a, nlev=2, theta=0.0,
                      sub={"kind": sub, "fill_in": 1, "inv_fill": 1,
                           "thresh": 0.01, "inv_thresh": 0.01},
                      coarse=EXACT, pos="POST")
    P = H.P[0].toarray()
    for j in range(0, P.shape[1], 3):
        got = H.apply(a @ P[:, j])
        assert np.allclose(got, P[:, j], rtol=0, atol=1e-9)


def test_cycle_pre_exact_coarse():
    a = grid2d(10, conv=0.5)
    H = O.MlHierarchy(a, nlev=2, theta=0.0, sub={"kind": "INVK"},
                      coarse=EXACT, pos="PRE")
    R, A = H.R[0].toarray(), a.toarray()
    rng = np.random.default_rng(3)
    for _ in range(4):
        v = rng.standard_normal(a.shape[0])
        lhs = R @ (A @ H.apply(A @ v))
        assert np.allclose(lhs, R @ (A @ v), rtol=0,
                           atol=1e-9 * np.abs(lhs).max())


def test_cycle_add_exact_coarse():
    a = grid2d(10, conv=0.5)
    H = O.MlHierarchy(a, nlev=2, theta=0.0, sub={"kind": "INVK"},
                      coarse=EXACT, ml_type="ADD")
    w, z, d = H.F[0]
    R, A = H.R[0].toarray(), a.toarray()
    rng = np.random.default_rng(4)
    x = rng.standard_normal(a.shape[0])
    corr = H.apply(x) - O.apply(w, z, d, x)
    assert np.allclose(R @ (A @ corr), R @ x, rtol=0, atol=1e-9)


def dense_cycle(H, x):
    """The V-cycle evaluated with dense matrices, level by level."""
    A = [m.toarray() for m in H.A]
    P = [m.toarray() for m in H.P]
    S = [(f[1] @ sp.diags(f[2]) @ f[0]).toarray() for f in H.F]
    L = len(A)

    def smooth0(l, nsw, b):
        u = S[l] @ b
        for _ in range(nsw - 1):
            u = u + S[l] @ (b - A[l] @ u)
        return u

    def run(l, b):
        if l == L - 1:
            return smooth0(l, H.csweeps, b)
        if H.type == 1:
            return smooth0(l, H.sweeps, b) + P[l] @ run(l + 1, P[l].T @ b)
        pre, post = H.pos & 1, H.pos & 2
        u = smooth0(l, H.sweeps, b) if pre else np.zeros_like(b)
        u = u + P[l] @ run(l + 1, P[l].T @ (b - A[l] @ u))
        if post:
            for _ in range(H.sweeps):
                u = u + S[l] @ (b - A[l] @ u)
        return u

    return run(0, x)


@pytest.mark.parametrize("ml_type,pos", [("MULT", "POST"), ("MULT", "PRE"),
                                         ("MULT", "TWOSIDE"), ("ADD", "POST")])
@pytest.mark.parametrize("sweeps,csweeps", [(1, 1), (2, 3)])
@pytest.mark.parametrize("nlev", [2, 3])
def test_cycle_dense(ml_type, pos, sweeps, csweeps, nlev):
    a = grid2d(16, conv=0.3)
    H = O.MlHierarchy(a, nlev=nlev, theta=0.05,
                      sub={"kind": "INVK", "fill_in": 1, "inv_fill": 1},
                      ml_type=ml_type, pos=pos, sweeps=sweeps,
                      csweeps=csweeps)
    assert H.nlev == nlev
    x = np.random.default_rng(nlev).standard_normal(a.shape[0])
    y = H.apply(x)
    yd = dense_cycle(H, x)
    assert np.allclose(y, yd, rtol=0, atol=1e-12 * np.abs(yd).max())


def test_hierarchy_stops_without_reduction():
    a = grid2d(8)
    H = O.MlHierarchy(a, nlev=5, theta=10.0)  # all weak: no coarsening
    assert H.nlev == 1
    x = np.ones(a.shape[0])
    w, z, d = H.F[0]
    # a one-level "cycle" is the coarsest-level smoother with its sweeps
    assert np.allclose(H.apply(x), O.apply(w, z, d, x))


# ---------------------------------------------------------------- Krylov
@pytest.mark.parametrize("method,sfn", [("BICGSTAB", spla.bicgstab),
                                        ("RGMRES", spla.gmres),
                                        ("CG", spla.cg)])
def test_ml_krylov_matches_scipy(method, sfn):
    a = grid2d(24) if method == "CG" else grid2d(24, conv=0.4)
    n = a.shape[0]
    pos = "TWOSIDE" if method == "CG" else "POST"
    sub = {"kind": "AINV", "alg": "SYM-LLK", "fill_in": 4} \
        if method == "CG" else {"kind": "INVK", "fill_in": 1, "inv_fill": 1}
    H = O.MlHierarchy(a, nlev=3, theta=0.05, sub=sub, pos=pos, csweeps=4)
    b = a @ np.ones(n)
    x, it, err = O.ml_krylov(method, a, H, b, 1e-8, itmax=200, irst=30)
    assert err <= 1e-8
    assert np.linalg.norm(b - a @ x) / np.linalg.norm(b) < 1e-6
    M = spla.LinearOperator((n, n), matvec=H.apply)
    count = [0]

    def cb(_):
        count[0] += 1

    kw = dict(restart=30, callback_type="pr_norm") if method == "RGMRES" \
        else {}
    sfn(a, b, rtol=1e-8, atol=0, M=M, maxiter=200, callback=cb, **kw)
    assert abs(count[0] - it) <= 1, (it, count[0])


def test_ml_krylov_bicg_refused():
    a = grid2d(8)
    H = O.MlHierarchy(a, nlev=2, theta=0.0)
    with pytest.raises(O.OracleError):
        O.ml_krylov("BICG", a, H, np.ones(a.shape[0]), 1e-6)


def test_hierarchy_from_levels_is_the_same_cycle():
    """A hierarchy rebuilt from its levels (how the benchmark's CPU
    baseline takes the library's exported hierarchy) applies the same
    cycle bit for bit."""
    a = grid2d(20, conv=0.3)
    H = O.MlHierarchy(a, nlev=3, theta=0.0,
                      sub={"kind": "INVK", "fill_in": 0, "inv_fill": 1},
                      csweeps=2)
    H2 = O.MlHierarchy.from_levels(H.A, H.P, H.R, H.F, csweeps=2)
    x = np.random.default_rng(5).standard_normal(a.shape[0])
    assert np.array_equal(H.apply(x), H2.apply(x))
