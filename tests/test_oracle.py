"""Pins the CPU oracle (CPU only, no GPU).

The reference ships no golden vectors (its tests are sample drivers that
print timings), so every oracle function is pinned here by what the
mathematics fixes -- exact inverses and factorizations when nothing is
dropped, biconjugation invariants, graph-distance patterns -- and by
independent dense restatements (tests/refcheck.py) on small inputs, plus the
committed fixtures of tests/golden/.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import refcheck as R

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rand_sparse(n, density, seed, dom=1.5, sym=False):
    rng = np.random.default_rng(seed)
    m = sp.random(n, n, density=density, random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr()
    if sym:
        m = m + m.T
    d = np.asarray(abs(m).sum(axis=1)).ravel() * dom + 1.0
    m = (m + sp.diags(d)).tocsr()
    m.sort_indices()
    return m


def as_dict_rows(m):
    return R.rows_of(m)


# ------------------------------------------------------------- generator
def test_generator_matches_independent_numpy(mld):
    for N in (3, 6):
        a = mld.gen_pde3d(N)
        b = R.pde3d_numpy(N)
        assert (a != b).nnz == 0
        assert a.nnz == 7 * N ** 3 - 6 * N ** 2
    # genpde's own spacing h = 1/(N+2)
    a = mld.gen_pde3d(4, hconv=1)
    assert np.allclose(a.diagonal(), 6.0 * 36.0)


def test_aniso_and_random_generators(mld):
    a = mld.gen_aniso2d(16, 1e-3)
    n = 256
    assert a.nnz == 5 * n - 4 * 16
    assert abs(a - a.T).max() == 0.0
    assert np.all(np.linalg.eigvalsh(a.toarray()) > 0)
    r = mld.gen_random(5000, seed=3, median=12.0, sigma=1.0, band=100)
    off = r - sp.diags(r.diagonal())
    rowabs = np.asarray(abs(off).sum(axis=1)).ravel()
    assert np.allclose(r.diagonal(), np.where(rowabs > 0, 1.1 * rowabs, 1.0))
    lens = np.diff(r.indptr) - 1
    assert 10 < np.median(lens - 2) < 14  # lognormal median 12 (+2 long)
    r2 = mld.gen_random(5000, seed=3, median=12.0, sigma=1.0, band=100)
    assert (r != r2).nnz == 0  # deterministic


# ------------------------------------------------------------- sparsify
def test_sparsify_brute_force(O):
    rng = np.random.default_rng(0)
    for trial in range(50):
        n = int(rng.integers(5, 40))
        z = rng.standard_normal(n) * (rng.random(n) < 0.6)
        idiag = int(rng.integers(0, n))
        nzrmax = int(rng.integers(1, 8))
        thr = float(rng.choice([0.0, 0.1, 0.5]))
        iz, vz = O.sparsify(z, idiag, nzrmax, thr)
        cand = [j for j in range(n) if j != idiag and abs(z[j]) >= thr]
        cand.sort(key=lambda j: -abs(z[j]))
        keep = sorted(cand[:nzrmax - 1] + [idiag])
        assert list(iz) == keep
        assert np.array_equal(vz, z[keep])
    # nzrmax = 0 never hits the exit test: every candidate is kept
    z = np.array([0.5, 0.0, 2.0, -0.1])
    iz, _ = O.sparsify(z, 1, 0, 0.2)
    assert list(iz) == [0, 1, 2]


# ------------------------------------------------------------- biconjugation
@pytest.mark.parametrize("alg", ["LLK", "SYM-LLK", "STAB-LLK", "LLK-NOTH",
                                 "MLK"])
def test_biconjg_no_drop_is_exact_inverse(O, alg):
    a = rand_sparse(60, 0.05, 1, sym=True) if alg == "SYM-LLK" else \
        R.pde3d_numpy(4)
    n = a.shape[0]
    w, z, d = O.ainv_bld(a, alg, n, 0.0)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    assert np.abs(M @ a.toarray() - np.eye(n)).max() < 1e-10
    assert np.allclose(sp.triu(z).toarray(), z.toarray())
    assert np.all(z.diagonal() == 1.0)  # zval(i) is never updated
    # invariant of the biconjugation: W^T A Z diagonal, = weight * diag(p)
    p, zc, wc = O.biconjg(a / abs(a).max(), alg, n, 0.0)
    G = wc.toarray().T @ (a / abs(a).max()).toarray() @ zc.toarray()
    assert np.abs(G - np.diag(np.diag(G))).max() < 1e-10
    assert np.allclose(np.diag(G), p)


@pytest.mark.parametrize("seed,nzrmax,thr", [(0, 3, 0.1), (1, 5, 0.05),
                                             (2, 2, 0.0), (3, 0, 0.2)])
def test_biconjg_matches_dense_restatement(O, seed, nzrmax, thr):
    a = rand_sparse(50, 0.06, seed)
    a = a / abs(a).max()
    p, zc, wc = O.biconjg(a, "LLK", nzrmax, thr)
    Z, W, pd = R.dense_llk(a, nzrmax, thr)
    assert np.array_equal(zc.toarray(), Z)
    assert np.array_equal(wc.toarray(), W)
    assert np.array_equal(p, pd)
    # pattern including stored zeros
    zc = zc.tocsc()
    for i in range(a.shape[0]):
        col = set(zc.indices[zc.indptr[i]:zc.indptr[i + 1]])
        assert i in col


@pytest.mark.parametrize("seed,nzrmax,thr", [(5, 3, 0.1), (6, 6, 0.02)])
def test_llk_noth_matches_dense_restatement(O, seed, nzrmax, thr):
    a = rand_sparse(50, 0.06, seed)
    a = a / abs(a).max()
    p, zc, wc = O.biconjg(a, "LLK-NOTH", nzrmax, thr)
    Z, W, pd = R.dense_llk(a, nzrmax, thr, noth=True)
    assert np.array_equal(zc.toarray(), Z)
    assert np.array_equal(wc.toarray(), W)
    assert np.array_equal(p, pd)


def test_mlk_is_llk_except_for_nzrmax_zero(O):
    """MLK visits the same rows in the same order with sorted lists instead
    of heaps (mld_dsparse_biconjg_mlk.F90:116-440), so it reproduces LLK;
    its sparsify_list keeps max(1, min(count, nzrmax)) entries
    (mld_d_sparsify.f90:214-222), only the diagonal when nzrmax = 0."""
    a = rand_sparse(60, 0.06, 8)
    a = a / abs(a).max()
    p1, z1, w1 = O.biconjg(a, "LLK", 4, 0.05)
    p2, z2, w2 = O.biconjg(a, "MLK", 4, 0.05)
    assert (z1 != z2).nnz == 0 and (w1 != w2).nnz == 0
    assert np.array_equal(p1, p2)
    _, z0, w0 = O.biconjg(a, "MLK", 0, 0.05)
    assert z0.nnz == a.shape[0] and w0.nnz == a.shape[0]
    assert np.array_equal(z0.toarray(), np.eye(a.shape[0]))


def test_stab_llk_pivots_are_w_a_z(O):
    """STAB-LLK forms p_i = q_i = w_i^T A z_i after both columns
    (mld_dsparse_biconjg_s_ft_llk.F90:377-411): with dropping, the pivots
    equal the diagonal of W^T A Z computed densely."""
    a = rand_sparse(50, 0.06, 12)
    a = a / abs(a).max()
    p, zc, wc = O.biconjg(a, "STAB-LLK", 5, 0.05)
    G = wc.toarray().T @ a.toarray() @ zc.toarray()
    d = np.diag(G)
    assert np.allclose(p[1:], np.where(np.abs(d[1:]) < 2.2e-16, 1e-3, d[1:]),
                       rtol=1e-13, atol=0)


def test_sym_llk_matches_dense_restatement(O):
    a = rand_sparse(50, 0.06, 9, sym=True)
    a = a / abs(a).max()
    p, zc, wc = O.biconjg(a, "SYM-LLK", 4, 0.05)
    Z, W, pd = R.dense_llk(a, 4, 0.05, sym=True)
    assert np.array_equal(zc.toarray(), Z)
    assert np.array_equal(wc.toarray(), Z)
    assert np.array_equal(p, pd)


def test_ainv_bld_scaling_and_pivots(O):
    """d = weight / p with weight = 1/max|a| (mld_d_ainv_bld.F90:121-173):
    scaling A by any c > 0 leaves W, Z unchanged and scales d by 1/c."""
    a = R.pde3d_numpy(4)
    w1, z1, d1 = O.ainv_bld(a, "LLK", 6, 0.05)
    w2, z2, d2 = O.ainv_bld(a * 8.0, "LLK", 6, 0.05)
    assert (w1 != w2).nnz == 0 and (z1 != z2).nnz == 0
    assert np.allclose(d1, 8.0 * d2, rtol=1e-15)


# ------------------------------------------------------------- ILU
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_iluk_full_fill_is_lu(O, seed):
    a = rand_sparse(40, 0.08, seed)
    n = a.shape[0]
    l, u, d = O.iluk(a, n)
    L = l.toarray() + np.eye(n)
    U = u.toarray() + np.eye(n)
    assert np.abs(L @ np.diag(1 / d) @ U - a.toarray()).max() < 1e-12
    # pivots of textbook Gaussian elimination without pivoting
    G = a.toarray().copy()
    for k in range(n):
        G[k + 1:, k] /= G[k, k]
        G[k + 1:, k + 1:] -= np.outer(G[k + 1:, k], G[k, k + 1:])
    assert np.allclose(1 / d, np.diag(G), rtol=1e-12)


@pytest.mark.parametrize("fill", [0, 1, 2, 3])
def test_iluk_pattern_matches_kij_levels(O, fill):
    a = rand_sparse(45, 0.05, 10 + fill)
    l, u, d = O.iluk(a, fill)
    keep = R.dense_iluk_levels(a, fill)
    n = a.shape[0]
    pat = np.zeros((n, n), dtype=bool)
    pat[np.nonzero((l + u).toarray())] = True
    lu = (l + u).tocoo()
    pat[:] = False
    pat[lu.row, lu.col] = True
    np.fill_diagonal(keep, False)
    assert np.array_equal(pat, keep)


def test_ilut_no_drop_is_lu_and_drops_shrink(O):
    a = rand_sparse(40, 0.1, 4)
    n = a.shape[0]
    l, u, d = O.ilut(a, n, 0.0)
    L = l.toarray() + np.eye(n)
    U = u.toarray() + np.eye(n)
    assert np.abs(L @ np.diag(1 / d) @ U - a.toarray()).max() < 1e-12
    l2, u2, d2 = O.ilut(a, 1, 0.05)
    assert l2.nnz + u2.nnz < l.nnz + u.nnz
    # fill cap: at most (original count) + fill per row and triangle
    al = sp.tril(a, -1).tocsr()
    au = sp.triu(a, 1).tocsr()
    assert np.all(np.diff(l2.indptr) <= np.diff(al.indptr) + 1)
    assert np.all(np.diff(u2.indptr) <= np.diff(au.indptr) + 1)


# ------------------------------------------------------------- INVK / INVT
@pytest.mark.parametrize("fill", [0, 1, 2])
def test_invk_matches_dense_scan_and_bfs(O, fill):
    a = rand_sparse(60, 0.05, 20 + fill)
    l, u, d = O.iluk(a, 1)
    z = O.sparse_invk(u, fill)
    assert R.rows_of(z) == R.dense_row_inverse_invk(u, fill)
    pat = [sorted(r) for r in R.rows_of(z)]
    assert pat == R.invk_pattern_bfs(u, fill)


def test_invk_full_fill_is_exact_inverse(O):
    a = R.pde3d_numpy(4)
    n = a.shape[0]
    l, u, d = O.iluk(a, 1)
    z = O.sparse_invk(u, n)
    U = u.toarray() + np.eye(n)
    ref = sla.solve_triangular(U, np.eye(n), unit_diagonal=True)
    assert np.abs(z.toarray() - ref).max() < 1e-12
    w, z, d = O.invk_bld(a, n, n)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    assert np.abs(M - np.linalg.inv(a.toarray())).max() < 1e-12 * \
        np.abs(np.linalg.inv(a.toarray())).max() * 100


@pytest.mark.parametrize("fill,thr", [(0, 0.1), (2, 0.05), (5, 0.01),
                                      (1, 0.0)])
def test_invt_matches_dense_scan(O, fill, thr):
    a = rand_sparse(60, 0.06, 30 + fill)
    l, u, d = O.ilut(a, 3, 1e-3)
    z = O.sparse_invt(u, fill, thr)
    assert R.rows_of(z) == R.dense_row_inverse_invt(u, fill, thr)


def test_invt_no_drop_is_exact_inverse(O):
    a = R.pde3d_numpy(4)
    n = a.shape[0]
    w, z, d = O.invt_bld(a, n, 0.0, n, 0.0)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    inv = np.linalg.inv(a.toarray())
    assert np.abs(M - inv).max() < 1e-10 * np.abs(inv).max()


def test_invt_diagonal_is_not_protected(O):
    """The copy-out keeps the nup+fill largest entries with no exemption for
    the unit diagonal (invt_copyout.f90:240-248)."""
    u = sp.csr_matrix(np.array([[0, 3.0, 0], [0, 0, 2.0], [0, 0, 0]]))
    z = O.sparse_invt(u, 0, 0.0)
    row0 = R.rows_of(z)[0]
    # row 0 of U^{-1} = [1, -3, 6]; cap = nup + 0 = 1 keeps only the 6
    assert row0 == {2: 6.0}


# ------------------------------------------------------------- apply
@pytest.mark.parametrize("trans", ["N", "T"])
def test_apply_closed_form(O, trans):
    a = R.pde3d_numpy(4)
    w, z, d = O.invk_bld(a, 1, 1)
    x = np.random.default_rng(1).standard_normal(a.shape[0])
    y0 = np.random.default_rng(2).standard_normal(a.shape[0])
    M = z.toarray() @ np.diag(d) @ w.toarray()
    if trans == "T":
        M = M.T
    y = O.apply(w, z, d, x, alpha=-1.5, beta=0.25, y=y0, trans=trans)
    assert np.allclose(y, -1.5 * M @ x + 0.25 * y0, rtol=1e-13, atol=0)
    with pytest.raises(O.OracleError):
        O.apply(w, z, d, x, trans="X")


# ------------------------------------------------------------- update_a
def test_update_a_matrix_form(O):
    """Arbitrary sparse W, Z: d_new = 1 / (1/d + diag(W diag(x) Z)) (the
    stored W by rows, Z by columns; a transposed operand fails this)."""
    rng = np.random.default_rng(4)
    n = 30
    w = sp.random(n, n, density=0.2, random_state=1, format="csr") + sp.eye(n)
    z = sp.random(n, n, density=0.2, random_state=2, format="csr") + sp.eye(n)
    d = rng.uniform(0.5, 2.0, n)
    x = rng.standard_normal(n)
    got = O.update_a(w, z, d, x)
    ee = np.diag(w.toarray() @ np.diag(x) @ z.toarray())
    assert np.allclose(got, 1.0 / (1.0 / d + ee), rtol=1e-13, atol=0)
    assert not np.allclose(
        got, 1.0 / (1.0 / d + np.diag(w.toarray().T @ np.diag(x) @ z.toarray())))
    # x = 0 leaves d as 1/(1/d)
    assert np.allclose(O.update_a(w, z, d, np.zeros(n)), d, rtol=1e-15)


@pytest.mark.parametrize("kind", ["INVK", "AINV"])
def test_update_a_exact_factors_give_shifted_pivots(O, kind):
    """With exact factors (no dropping) W A Z = diag(1/d); the update must
    give the pivots of the same W, Z against A + diag(x):
    d_new = 1 / diag(W (A + X) Z)."""
    a = R.pde3d_numpy(3)
    n = a.shape[0]
    if kind == "INVK":
        w, z, d = O.invk_bld(a, n, n)
    else:
        w, z, d = O.ainv_bld(a, "LLK", n, 0.0)
    W, Z = w.toarray(), z.toarray()
    assert np.allclose(W @ a.toarray() @ Z, np.diag(1.0 / d), atol=1e-10)
    x = np.random.default_rng(5).uniform(0.1, 1.0, n)
    got = O.update_a(w, z, d, x)
    ref = 1.0 / np.diag(W @ (a.toarray() + np.diag(x)) @ Z)
    assert np.allclose(got, ref, rtol=1e-10, atol=0)


# ------------------------------------------------------------- Krylov
def _scipy_iters(method, a, M, b, rtol):
    count = [0]

    def cb(_):
        count[0] += 1
    f = {"BICGSTAB": spla.bicgstab, "CG": spla.cg, "CGS": spla.cgs,
         "BICG": spla.bicg}[method]
    x, info = f(a, b, rtol=rtol, atol=0.0, M=M, maxiter=2000, callback=cb)
    assert info == 0
    return count[0], x


@pytest.mark.parametrize("method,kind", [("BICGSTAB", "INVK"),
                                         ("BICGSTAB", "AINV"),
                                         ("CG", "SYM"), ("CGS", "INVK"),
                                         ("CGS", "AINV"), ("BICG", "INVK"),
                                         ("BICG", "AINV")])
def test_krylov_against_scipy(O, method, kind):
    if kind == "SYM":
        a = sp.csr_matrix(R.pde3d_numpy(6))
        a = ((a + a.T) * 0.5).tocsr()
        a.sort_indices()
        w, z, d = O.ainv_bld(a, "SYM-LLK", 8, 0.05)
    elif kind == "INVK":
        a = R.pde3d_numpy(7)
        w, z, d = O.invk_bld(a, 1, 1)
    else:
        a = R.pde3d_numpy(7)
        w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    n = a.shape[0]
    b = a @ np.ones(n)
    M = spla.LinearOperator((n, n), matvec=lambda v: O.apply(w, z, d, v),
                            rmatvec=lambda v: O.apply(w, z, d, v, trans="T"))
    it_sc, x_sc = _scipy_iters(method, a, M, b, 1e-8)
    x, it, err = O.krylov(method, a, w, z, d, b, 1e-8)
    assert abs(it - it_sc) <= 1, (it, it_sc)
    assert err <= 1e-8
    assert np.linalg.norm(b - a @ x) <= 1e-7 * np.linalg.norm(b)


@pytest.mark.parametrize("m", [5, 30])
def test_gmres_unpreconditioned_against_scipy(O, m):
    """With M = I right and left preconditioning coincide: iteration counts
    of scipy's restarted gmres (inner steps, 'pr_norm' callback)."""
    a = R.pde3d_numpy(7)
    n = a.shape[0]
    w = z = sp.identity(n, format="csr")
    d = np.ones(n)
    b = a @ np.ones(n)
    count = [0]

    def cb(_):
        count[0] += 1
    x_sc, info = spla.gmres(a, b, rtol=1e-8, atol=0.0, restart=m,
                            maxiter=500, callback=cb, callback_type="pr_norm")
    assert info == 0
    x, it, err = O.krylov("RGMRES", a, w, z, d, b, 1e-8, 2000, irst=m)
    assert abs(it - count[0]) <= 1, (it, count[0])
    assert np.linalg.norm(b - a @ x) <= 1e-7 * np.linalg.norm(b)


@pytest.mark.parametrize("kind", ["INVK", "AINV"])
def test_gmres_minimal_residual(O, kind):
    """Before a restart, step i of right-preconditioned GMRES reaches the
    least-squares minimum of ||b - A M u|| over the Krylov space
    K_i(A M, b) (x0 = 0), computed here with numpy from the explicit
    Krylov basis."""
    a = R.pde3d_numpy(5)
    n = a.shape[0]
    if kind == "INVK":
        w, z, d = O.invk_bld(a, 1, 1)
    else:
        w, z, d = O.ainv_bld(a, "LLK", 4, 0.1)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    AM = a.toarray() @ M
    b = np.random.default_rng(8).standard_normal(n)
    basis = [b / np.linalg.norm(b)]
    for i in range(1, 7):
        x, it, err = O.krylov("RGMRES", a, w, z, d, b, 1e-30, i, irst=50)
        assert it == i
        K = np.stack(basis, axis=1)
        c, *_ = np.linalg.lstsq(AM @ K, b, rcond=None)
        rmin = np.linalg.norm(b - AM @ K @ c)
        assert abs(err * np.linalg.norm(b) - rmin) <= 1e-7 * rmin, (i, err)
        assert np.linalg.norm(b - a @ x) == pytest.approx(rmin, rel=1e-7)
        nxt = AM @ basis[-1]
        for q in basis:  # orthonormal basis keeps the lstsq well posed
            nxt = nxt - (q @ nxt) * q
        basis.append(nxt / np.linalg.norm(nxt))


def test_gmres_finite_termination_and_restart(O):
    """Without restart GMRES solves an n x n system in at most n steps; a
    short restart still converges on a convection-diffusion matrix."""
    rng = np.random.default_rng(9)
    n = 9
    a = sp.csr_matrix(rng.standard_normal((n, n)) + 4 * np.eye(n))
    i1 = sp.identity(n, format="csr")
    b = rng.standard_normal(n)
    x, it, err = O.krylov("RGMRES", a, i1, i1, np.ones(n), b, 1e-12, 100,
                          irst=n)
    assert it <= n and err <= 1e-12
    assert np.linalg.norm(b - a @ x) <= 1e-11 * np.linalg.norm(b)
    a = R.pde3d_numpy(6)
    w, z, d = O.invk_bld(a, 1, 1)
    b = a @ np.ones(a.shape[0])
    x2, it2, err2 = O.krylov("RGMRES", a, w, z, d, b, 1e-8, 2000, irst=2)
    x30, it30, _ = O.krylov("RGMRES", a, w, z, d, b, 1e-8, 2000, irst=30)
    assert err2 <= 1e-8 and it2 >= it30
    for xx in (x2, x30):
        assert np.linalg.norm(b - a @ xx) <= 1e-7 * np.linalg.norm(b)


def _bicgstabl_lstsq(A, M, b, l, cycles):
    """BiCGStab(l) on B = A M with the minimal-residual step solved by
    least squares (min ||r_0 - [r_1 .. r_l] c||), not by the Gram-Schmidt
    gamma / gamma' / gamma'' recurrences; returns (x, ||r||) per cycle."""
    B = A @ M
    r = [b.copy()] + [np.zeros_like(b) for _ in range(l)]
    u = [np.zeros_like(b) for _ in range(l + 1)]
    rt = b.copy()
    y = np.zeros_like(b)
    rho0, alpha, omega = 1.0, 0.0, 1.0
    out = []
    for _ in range(cycles):
        rho0 = -omega * rho0
        for j in range(l):
            rho1 = r[j] @ rt
            beta = alpha * rho1 / rho0
            rho0 = rho1
            for i in range(j + 1):
                u[i] = r[i] - beta * u[i]
            u[j + 1] = B @ u[j]
            alpha = rho0 / (u[j + 1] @ rt)
            for i in range(j + 1):
                r[i] = r[i] - alpha * u[i + 1]
            r[j + 1] = B @ r[j]
            y = y + alpha * u[0]
        R = np.stack(r[1:], axis=1)
        c, *_ = np.linalg.lstsq(R, r[0], rcond=None)
        y = y + sum(c[j - 1] * r[j - 1] for j in range(1, l + 1))
        u[0] = u[0] - sum(c[j - 1] * u[j] for j in range(1, l + 1))
        r[0] = r[0] - R @ c
        omega = c[l - 1]
        out.append((M @ y, np.linalg.norm(r[0])))
    return out


@pytest.mark.parametrize("l", [2, 4])
def test_bicgstabl_against_least_squares_form(O, l):
    """The oracle's Gram-Schmidt minimal-residual step equals the least
    squares one of an independent dense restatement, cycle by cycle."""
    a = R.pde3d_numpy(5)
    n = a.shape[0]
    w, z, d = O.invk_bld(a, 1, 1)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    b = np.random.default_rng(10).standard_normal(n)
    ref = _bicgstabl_lstsq(a.toarray(), M, b, l, 3)
    for cyc in range(1, 4):
        x, it, err = O.krylov("BICGSTABL", a, w, z, d, b, 1e-30, cyc * l,
                              irst=l)
        xr, rn = ref[cyc - 1]
        assert it == cyc * l
        assert err * np.linalg.norm(b) == pytest.approx(rn, rel=1e-7)
        assert np.linalg.norm(x - xr) <= 1e-7 * np.linalg.norm(xr)


def test_bicgstabl_one_is_bicgstab_and_converges(O):
    """BiCGStab(1) is BiCGStab (same iterations +-1, same solution to the
    tolerance); l = 2, 4, 8 converge to true residuals at the tolerance."""
    a = R.pde3d_numpy(7)
    n = a.shape[0]
    for w, z, d in (O.invk_bld(a, 1, 1), O.ainv_bld(a, "LLK", 8, 0.1)):
        b = a @ np.ones(n)
        x1, it1, _ = O.krylov("BICGSTABL", a, w, z, d, b, 1e-9, irst=1)
        xs, its, _ = O.krylov("BICGSTAB", a, w, z, d, b, 1e-9)
        assert abs(it1 - its) <= 1, (it1, its)
        for l in (1, 2, 4, 8):
            x, it, err = O.krylov("BICGSTABL", a, w, z, d, b, 1e-9, irst=l)
            assert it > 0 and it % l == 0 and err <= 1e-9
            assert np.linalg.norm(b - a @ x) <= 1e-8 * np.linalg.norm(b)


@pytest.mark.parametrize("method", ["BICGSTAB", "CG"])
def test_threaded_loops_same_solve(O, method):
    """The row-parallel loops (CPU baseline on all host cores) change only
    the summation order of the dots: same iterations +-1, same solution to
    the tolerance, and one thread reproduces the sequential result bit for
    bit."""
    a = R.pde3d_numpy(9)
    if method == "CG":
        a = sp.csr_matrix(((a + a.T) * 0.5))
        a.sort_indices()
        w, z, d = O.ainv_bld(a, "SYM-LLK", 8, 0.05)
    else:
        w, z, d = O.invk_bld(a, 1, 1)
    b = a @ np.ones(a.shape[0])
    x1, it1, _ = O.krylov(method, a, w, z, d, b, 1e-9)
    try:
        O.set_threads(4)
        x4, it4, e4 = O.krylov(method, a, w, z, d, b, 1e-9)
        O.set_threads(1)
        x1b, it1b, _ = O.krylov(method, a, w, z, d, b, 1e-9)
    finally:
        O.set_threads(1)
    assert abs(it4 - it1) <= 1 and e4 <= 1e-9
    assert np.linalg.norm(b - a @ x4) <= 1e-8 * np.linalg.norm(b)
    assert it1b == it1 and np.array_equal(x1b, x1)


# ------------------------------------------------------------- fixtures
def test_golden_fixtures(O):
    """Committed vectors (tests/golden/make_golden.py): exact inverses from
    numpy and iteration counts from scipy at the config-C1 settings on a
    small grid."""
    g = np.load(os.path.join(GOLDEN, "golden_small.npz"))
    a = sp.csr_matrix((g["a_val"], g["a_ja"], g["a_irp"]),
                      shape=(int(g["n"]), int(g["n"])))
    n = a.shape[0]
    w, z, d = O.ainv_bld(a, "LLK", n, 0.0)
    M = z.toarray() @ np.diag(d) @ w.toarray()
    assert np.abs(M - g["ainv"]).max() < 1e-10 * np.abs(g["ainv"]).max()
    w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    _, it, _ = O.krylov("BICGSTAB", a, w, z, d, np.ones(n), 1e-9)
    assert abs(it - int(g["scipy_bicgstab_iters_ainv"])) <= 1
    w, z, d = O.invk_bld(a, 1, 1)
    _, it, _ = O.krylov("BICGSTAB", a, w, z, d, a @ np.ones(n), 1e-8)
    assert abs(it - int(g["scipy_bicgstab_iters_invk"])) <= 1


# ------------------------------------------------- ILUT against Saad 10.6
def ilut_case_matrix(n, density, seed, mag=250.0):
    """Nonsymmetric, rows of very different magnitude (so tau_i differs from
    row to row and the pivots from 1 after maxval scaling), max|a| far from
    1, diagonally dominant by a random factor in [1.05, 3]."""
    rng = np.random.default_rng(seed)
    m = sp.random(n, n, density=density, random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k)).tocsr()
    m = m - sp.diags(m.diagonal())
    rowabs = np.asarray(abs(m).sum(axis=1)).ravel()
    dom = rng.uniform(1.05, 3.0, n)
    m = (m + sp.diags(dom * rowabs + 0.1)).tocsr()
    rs = 10.0 ** rng.uniform(-1.0, 1.0, n)
    m = (sp.diags(rs * mag) @ m).tocsr()
    m.sort_indices()
    return m


def ilut_rows(l, u):
    return R.rows_of(l), R.rows_of(u)


# (fill, thresh) binding: the threshold alone (fill never caps), the caps
# alone (thresh 0 drops nothing), and both
ILUT_CASES = [("thresh", 10 ** 6, 0.05), ("caps", 0, 0.0), ("caps", 1, 0.0),
              ("both", 1, 0.02), ("both", 2, 0.01)]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("what,fill,thresh", ILUT_CASES)
def test_ilut_matches_dense_saad(O, seed, what, fill, thresh):
    a = ilut_case_matrix(48, 0.12, 300 + seed)
    fill = min(fill, a.shape[0])
    l, u, d = O.ilut(a, fill, thresh)
    Lr, Ur, dr, st = R.dense_ilut(a, fill, thresh)
    assert ilut_rows(l, u) == (Lr, Ur)
    assert np.array_equal(d, dr)
    # the case binds what it claims to
    if what in ("thresh", "both"):
        assert st["thresh_drops"] > 0
    if what in ("caps", "both"):
        assert st["cap_cuts"] > 0
    if what == "thresh":
        assert st["cap_cuts"] == 0
    if what == "caps":
        assert st["thresh_drops"] == 0


@pytest.mark.parametrize("mutate", ["keep_smallest", "unscaled_norm",
                                    "drop_before_scale"])
def test_ilut_restatement_detects_planted_mistakes(O, mutate):
    """Each plausible mistake, planted in the dense restatement, changes
    the factors on at least one of the cases above, so agreement with the
    oracle rules it out there."""
    differs = False
    for seed in (0, 1, 2):
        a = ilut_case_matrix(48, 0.12, 300 + seed)
        for _, fill, thresh in ILUT_CASES:
            l, u, d = O.ilut(a, fill, thresh)
            try:
                Lr, Ur, dr, _ = R.dense_ilut(a, fill, thresh, mutate=mutate)
            except ZeroDivisionError:
                differs = True
                continue
            if ilut_rows(l, u) != (Lr, Ur) or not np.array_equal(d, dr):
                differs = True
    assert differs


def test_ilut_tie_goes_to_smaller_column(O):
    """Equal magnitudes at the cap: the smaller column stays (the tie rule
    the oracle fixes for the PSBLAS heap)."""
    # row 0 has two equal upper entries under a cap of nup + 1 = 3: nothing
    # to cut; row 1 (no upper entry of its own) gains equal fill at columns
    # 2 and 3 from row 0 and keeps nup + fill = 1 of them
    a = sp.csr_matrix(np.array([[4.0, 0.0, 1.0, 1.0],
                                [1.0, 4.0, 0.0, 0.0],
                                [0.0, 0.0, 4.0, 0.0],
                                [0.0, 0.0, 0.0, 4.0]]))
    l, u, d = O.ilut(a, 1, 0.0)
    Lr, Ur, dr, st = R.dense_ilut(a, 1, 0.0)
    assert ilut_rows(l, u) == (Lr, Ur)
    assert st["cap_cuts"] == 1
    assert list(Ur[1].keys()) == [2]


@pytest.mark.parametrize("what,fill,thresh", ILUT_CASES)
def test_library_host_ilut_matches_saad_and_oracle(mld, O, what, fill, thresh):
    """The library's own ILUT (ainv_host_ilu_fact, the factorization the
    INVT build runs) against the dense restatement, bit for bit."""
    a = ilut_case_matrix(60, 0.1, 310)
    l, u, d = mld.host_ilu_fact(a, "ILUT", fill, thresh)
    Lr, Ur, dr, _ = R.dense_ilut(a, fill, thresh)
    assert ilut_rows(l, u) == (Lr, Ur)
    assert np.array_equal(d, dr)
    lo, uo, do = O.ilut(a, fill, thresh)
    assert (l != lo).nnz == 0 and (u != uo).nnz == 0


@pytest.mark.parametrize("fill", [0, 1, 2])
def test_library_host_iluk_matches_oracle_and_levels(mld, O, fill):
    a = rand_sparse(50, 0.06, 40 + fill)
    l, u, d = mld.host_ilu_fact(a, "ILUK", fill)
    lo, uo, do = O.iluk(a, fill)
    assert (l != lo).nnz == 0 and (u != uo).nnz == 0
    assert np.array_equal(d, do)
    keep = R.dense_iluk_levels(a, fill)
    np.fill_diagonal(keep, False)
    lu = (l + u).tocoo()
    pat = np.zeros_like(keep)
    pat[lu.row, lu.col] = True
    assert np.array_equal(pat, keep)


def test_library_host_ilu_zero_pivot_and_args(mld):
    a = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 1.0]]))
    with pytest.raises(mld.AinvError) as e:
        mld.host_ilu_fact(a, "ILUT", 1, 0.0)
    assert e.value.code == mld.ERR_PIVOT
    b = rand_sparse(10, 0.2, 1)
    with pytest.raises(mld.AinvError):
        mld.host_ilu_fact(b, "ILUT", -1, 0.0)
