"""The serving stream's recipe at reduced sizes on the GPU: a pool of 24
matrices (12 pde3d grids with N <= 30 and INVK(1,1); 12 random
nonsymmetric matrices with n log-uniform in [1e3, 3e4] and AINV-LLK 5e-2),
factors built on first touch, BiCGSTAB 1e-8, requests of 1-64 columns
arriving as a Poisson stream.

* FP32 factors (the recipe's): every result reaches ||r|| <= 1e-8 ||b|| and
  the solution of the same right-hand sides solved alone through the Krylov
  call (||A (x_served - x_alone)|| <= 2e-7 ||b||).  Iteration counts are
  held to 25 % (at least 4): with FP32 factors the iteration rounds the
  FP64 vectors to the FP32 copies the preconditioner reads, so the 1e-16
  differences of dot-product trees of different block widths can cross an
  FP32 rounding boundary (a 6e-8 step), and BiCGSTAB amplifies them.
  Measured on 12 columns of the pool's pde3d 30^3 matrix: with FP64 factors
  a block, each column alone, the served form and the oracle all take the
  same 29-31 iterations; with FP32 factors the three GPU forms alike spread
  over 27-34 (no systematic shift), and 33 served against 27 alone was
  seen for one column.
* FP64 factors: the same stream against the oracle's BiCGSTAB on a sample
  of columns, iterations +-1 and true residuals."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FACTORS = {"INVK": dict(SUB_FILLIN=1, INV_FILLIN=1),
           "AINV": dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=5e-2)}


def c5_pool(gpu):
    rng = np.random.default_rng(5)
    pool = []
    for _ in range(12):
        N = int(rng.integers(6, 31))
        pool.append((gpu.gen_pde3d(N), "INVK"))
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
        pool.append((gpu.gen_random(n, seed=100 + i, median=12.0, sigma=1.0,
                                    cap=400, band=5000, nlong=2, dom=1.1),
                     "AINV"))
    return pool


def stream(gpu, pool, prec, nreq, seed):
    """Poisson arrivals (mean gap 10 ms) of nreq requests of U{1..64}
    columns over the pool; returns [(matrix, b, (x, it, err, lat))]."""
    srv = gpu.Server(hbm_budget_bytes=16 << 30, workers=4)
    ids = [srv.add_matrix(a, kind, FACTOR_PREC=prec, METHOD="BICGSTAB",
                          **FACTORS[kind]) for a, kind in pool]
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(nreq):
        j = int(rng.integers(0, len(pool)))
        k = int(rng.integers(1, 65))
        reqs.append((j, gpu.gen_normal(seed * 1000 + r,
                                       (pool[j][0].shape[0], k))))
    gaps = rng.exponential(0.01, nreq)
    tickets = [None] * nreq
    results = [None] * nreq

    def waiter():
        for i in range(nreq):
            while tickets[i] is None:
                time.sleep(0.001)
            results[i] = srv.wait(tickets[i])

    th = threading.Thread(target=waiter)
    th.start()
    for i, (j, b) in enumerate(reqs):
        time.sleep(gaps[i])
        tickets[i] = srv.submit(ids[j], b, eps=1e-8, itmax=1000)
    th.join()
    st = srv.stats()
    assert st["completed"] == nreq and st["builds"] <= len(pool)
    srv.free()
    return [(j, b, res) for (j, b), res in zip(reqs, results)]


def test_c5_stream_fp32_against_alone(gpu):
    pool = c5_pool(gpu)
    out = stream(gpu, pool, "F32", 48, 7)
    solvers = {}
    for i, (j, b, (x, it, err, lat)) in enumerate(out):
        a, kind = pool[j]
        assert np.all(err <= 1e-8) and lat > 0
        if j not in solvers:
            sv = gpu.SOLVERS[kind]()
            for key, v in FACTORS[kind].items():
                sv.set(key, v)
            sv.set("FACTOR_PREC", "F32")
            sv.build(a)
            solvers[j] = (sv, gpu.SystemMatrix(a))
        sv, A = solvers[j]
        xa, ita, _ = gpu.krylov("BICGSTAB", A, sv, b, eps=1e-8, itmax=1000)
        xa = xa.reshape(x.shape)
        for c in range(b.shape[1]):
            i1, i2 = int(it[c]), int(ita[c])
            assert abs(i1 - i2) <= max(4, (25 * max(i1, i2) + 99) // 100), \
                (i, j, c, i1, i2)
            nb = np.linalg.norm(b[:, c])
            assert np.linalg.norm(b[:, c] - a @ x[:, c]) <= 1e-7 * nb, (i, c)
            assert np.linalg.norm(a @ (x[:, c] - xa[:, c])) <= 2e-7 * nb, \
                (i, c)
    for sv, A in solvers.values():
        sv.free()


def test_c5_stream_fp64_against_oracle(gpu, O):
    pool = c5_pool(gpu)
    out = stream(gpu, pool, "F64", 24, 11)
    facs = {}
    for i, (j, b, (x, it, err, lat)) in enumerate(out):
        a, kind = pool[j]
        assert np.all(err <= 1e-8)
        if j not in facs:
            facs[j] = (O.invk_bld(a, 1, 1) if kind == "INVK" else
                       O.ainv_bld(a, "LLK", 8, 5e-2))
        w, z, d = facs[j]
        for c in sorted({0, b.shape[1] // 2, b.shape[1] - 1}):
            xo, ito, _ = O.krylov("BICGSTAB", a, w, z, d, b[:, c], 1e-8,
                                  itmax=1000)
            assert abs(int(it[c]) - ito) <= 1, (i, j, c, it[c], ito)
            nb = np.linalg.norm(b[:, c])
            assert np.linalg.norm(b[:, c] - a @ x[:, c]) <= 1e-7 * nb
            assert np.linalg.norm(a @ (x[:, c] - xo)) <= 2e-7 * nb
