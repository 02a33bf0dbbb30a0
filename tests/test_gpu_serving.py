"""Serving loop on the GPU: each result is the same alone and in the mix
(within the Krylov tolerance; iteration counts within +-1), requests that
arrive while others run are admitted into freed column slots mid-batch, and
an HBM budget smaller than the pool forces LRU evictions without changing
results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def alone(mld, a, kind, params, b, eps, method="BICGSTAB"):
    sv = mld.SOLVERS[kind]()
    for k, v in params.items():
        sv.set(k, v)
    sv.build(a)
    x, it, err = mld.krylov(method, mld.SystemMatrix(a), sv, b, eps=eps)
    return x, it


def pool(mld):
    mats = [
        (mld.gen_pde3d(12), "INVK", dict(SUB_FILLIN=1, INV_FILLIN=1,
                                          FACTOR_PREC="F32")),
        (mld.gen_pde3d(9), "INVT", dict(SUB_FILLIN=5, SUB_ILUTHRS=1e-2,
                                         INV_FILLIN=5, INV_THRESH=1e-2)),
        (mld.gen_random(4000, seed=8, median=8.0, band=300), "AINV",
         dict(SUB_FILLIN=8, SUB_ILUTHRS=5e-2, FACTOR_PREC="F32")),
    ]
    return mats


def test_alone_vs_mix_and_midflight_admission(gpu):
    mats = pool(gpu)
    srv = gpu.Server(hbm_budget_bytes=8 << 30, workers=2)
    ids = [srv.add_matrix(a, kind, **p) for a, kind, p in mats]
    # first touch: the builder threads build all three; the requests below
    # then arrive while the workers' batches run on resident matrices
    for m, (a, kind, p) in enumerate(mats):
        b = gpu.gen_normal(90 + m, (a.shape[0], 1))
        x, it, err, _ = srv.wait(srv.submit(ids[m], b, eps=1e-9))
        assert np.all(err <= 1e-9)
    st = srv.stats()
    assert st["builds"] == 3 and st["builders"] == 2
    rng = np.random.default_rng(0)
    tickets = []
    for r in range(30):
        m = r % 3
        n = mats[m][0].shape[0]
        k = int(rng.integers(1, 9))
        b = gpu.gen_normal(100 + r, (n, k))
        tickets.append((m, b, srv.submit(ids[m], b, eps=1e-9)))
    for m, b, t in tickets:
        x, it, err, lat = srv.wait(t)
        a, kind, p = mats[m]
        assert np.all(err <= 1e-9)
        xa, ita = alone(gpu, a, kind, p, b, 1e-9)
        for c in range(b.shape[1]):
            assert abs(int(it[c]) - int(ita[c])) <= 1, (m, c, it, ita)
            tr = np.linalg.norm(b[:, c] - a @ x[:, c]) / np.linalg.norm(b[:, c])
            assert tr < 1e-7, (m, c, tr)
            # both iterates have relative residual <= 1e-9; the pool's
            # condition numbers are below 1.1e3, so they agree to ~1e-6
            rel = np.linalg.norm(x[:, c] - xa[:, c]) / np.linalg.norm(xa[:, c])
            assert rel < 2e-5, (m, c, rel)
        assert lat > 0
    st = srv.stats()
    assert st["completed"] == 33
    assert st["admitted_midflight"] >= 1
    assert st["builds"] == 3 and st["evictions"] == 0
    srv.free()


def test_lru_eviction_under_budget(gpu):
    mats = pool(gpu)[:2]
    srv = gpu.Server(hbm_budget_bytes=1, workers=1)  # room for none: evict
    ids = [srv.add_matrix(a, kind, **p) for a, kind, p in mats]
    for rnd in range(3):
        for m in range(2):
            n = mats[m][0].shape[0]
            b = gpu.gen_normal(7 + rnd, (n, 2))
            x, it, err, _ = srv.wait(srv.submit(ids[m], b, eps=1e-8))
            assert np.all(err <= 1e-8)
    st = srv.stats()
    assert st["evictions"] >= 4 and st["builds"] >= 5
    srv.free()


@pytest.mark.parametrize("method,slots,nrhs", [("BICGSTAB", 8, 21),
                                               ("BICGSTAB", 5, 3),
                                               ("CG", 16, 40),
                                               ("CGS", 8, 12),
                                               ("BICG", 6, 10)])
def test_krylov_stream_matches_single_solves(gpu, method, slots, nrhs):
    """The served form (a device queue through continuously refilled slots)
    gives each right-hand side the iterations (+-1) and solution of solving
    it alone; latencies are positive."""
    if method == "CG":
        a = gpu.gen_aniso2d(40, 1e-2)
        params = dict(AINV_ALG="SYM-LLK", SUB_FILLIN=8, SUB_ILUTHRS=5e-2,
                      FACTOR_PREC="F32")
        kind = "AINV"
    else:
        a = gpu.gen_pde3d(12)
        params = dict(SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC="F32")
        kind = "INVK"
    n = a.shape[0]
    sv = gpu.SOLVERS[kind]()
    for k, v in params.items():
        sv.set(k, v)
    sv.build(a)
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(31, (nrhs, n))
    X, it, err, lat = gpu.krylov_stream(method, A, sv,
                                        gpu.DeviceArray.from_numpy(B),
                                        slots=slots, eps=1e-9)
    Xh = X.to_numpy()
    assert np.all(err <= 1e-9) and np.all(lat > 0)
    for j in range(nrhs):
        xs, its, _ = gpu.krylov(method, A, sv, B[j], eps=1e-9)
        assert abs(int(it[j]) - int(its[0])) <= 1, (j, it[j], its)
        tr = np.linalg.norm(B[j] - a @ Xh[j]) / np.linalg.norm(B[j])
        assert tr < 1e-8, (j, tr)
        assert np.linalg.norm(Xh[j] - xs) <= 1e-6 * np.linalg.norm(xs)


@pytest.mark.parametrize("method", ["BICGSTAB", "CG", "CGS", "BICG"])
def test_krylov_stream_zero_rhs_and_slot_reuse(gpu, method):
    """A zero right-hand side stops at admission with x = 0 (iteration 0),
    also in a slot whose previous occupant left its iterate behind; every
    other column's iterations (+-1: the queue's tail narrows at other
    steps, and narrower plans sum the dots in another order) and solution
    match those of the same queue without the zero column."""
    if method == "CG":
        a = gpu.gen_aniso2d(30, 1e-2)
        kind, params = "AINV", dict(AINV_ALG="SYM-LLK", SUB_FILLIN=8,
                                    SUB_ILUTHRS=5e-2, FACTOR_PREC="F32")
    else:
        a = gpu.gen_pde3d(10)
        kind, params = "INVK", dict(SUB_FILLIN=1, INV_FILLIN=1,
                                    FACTOR_PREC="F32")
    n = a.shape[0]
    sv = gpu.SOLVERS[kind]()
    for k, v in params.items():
        sv.set(k, v)
    sv.build(a)
    A = gpu.SystemMatrix(a)
    B = gpu.gen_normal(77, (12, n))
    Bz = np.insert(B, 9, 0.0, axis=0)  # admitted after 9 others: a reused slot
    X, it, err, _ = gpu.krylov_stream(method, A, sv,
                                      gpu.DeviceArray.from_numpy(B), slots=4,
                                      eps=1e-9)
    Xz, itz, errz, _ = gpu.krylov_stream(method, A, sv,
                                         gpu.DeviceArray.from_numpy(Bz),
                                         slots=4, eps=1e-9)
    Xz = Xz.to_numpy()
    assert itz[9] == 0 and errz[9] == 0.0
    assert np.all(Xz[9] == 0.0)
    keep = [j for j in range(13) if j != 9]
    assert np.all(np.abs(itz[keep] - it) <= 1), (itz, it)
    assert np.all(errz[keep] <= 1e-9)
    Xh = X.to_numpy()
    for j, jz in enumerate(keep):
        assert np.linalg.norm(Xz[jz] - Xh[j]) <= 1e-6 * np.linalg.norm(Xh[j])


def test_krylov_stream_tuned_width(gpu):
    """slots = 0: the first call times 8, 16, 32 and 64 slots (no wider than
    the queue rounded up to a power of two), keeps the narrowest width that
    no wider one beats by 3 % per column-iteration, reports it and every
    width's time, keeps the choice for later calls and tunes again after
    the factors change; the results are those of solving alone."""
    a = gpu.gen_pde3d(12)
    n = a.shape[0]
    sv = gpu.SOLVERS["INVK"]()
    for k, v in dict(SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC="F32").items():
        sv.set(k, v)
    sv.build(a)
    A = gpu.SystemMatrix(a)
    assert sv.report()["served_slots"] == []
    B = gpu.gen_normal(41, (40, n))
    dB = gpu.DeviceArray.from_numpy(B)
    X, it, err, lat = gpu.krylov_stream("BICGSTAB", A, sv, dB, slots=0,
                                        eps=1e-9)
    ss = sv.report()["served_slots"]
    assert len(ss) == 1 and ss[0]["method"] == "BICGSTAB"
    assert ss[0]["slots"] in (8, 16, 32, 64)
    t = ss[0]["ms_per_col_iter"]
    assert sorted(t, key=int) == ["8", "16", "32", "64"]
    assert all(v > 0 for v in t.values())
    best = 8
    for w in (16, 32, 64):
        if t[str(w)] < t[str(best)] * (1 - 0.03):
            best = w
    assert ss[0]["slots"] == best
    Xh = X.to_numpy()
    assert np.all(err <= 1e-9) and np.all(lat > 0)
    for j in (0, 17, 39):
        xs, its, _ = gpu.krylov("BICGSTAB", A, sv, B[j], eps=1e-9)
        assert abs(int(it[j]) - int(its[0])) <= 1, (j, it[j], its)
        assert np.linalg.norm(Xh[j] - xs) <= 1e-6 * np.linalg.norm(xs)
    # kept: a second call runs no trial (the timings are unchanged)
    gpu.krylov_stream("BICGSTAB", A, sv, dB, slots=0, eps=1e-9)
    assert sv.report()["served_slots"] == ss
    # a short queue on another method tries only 8 slots
    X5, it5, err5, _ = gpu.krylov_stream(
        "BICG", A, sv, gpu.DeviceArray.from_numpy(B[:5]), slots=0, eps=1e-9)
    assert np.all(err5 <= 1e-9)
    e5 = [e for e in sv.report()["served_slots"] if e["method"] == "BICG"]
    assert e5[0]["slots"] == 8 and list(e5[0]["ms_per_col_iter"]) == ["8"]
    # new factors: no width kept until the next call tunes again
    sv.cnv("F64")
    assert sv.report()["served_slots"] == []
    X2, it2, err2, _ = gpu.krylov_stream("BICGSTAB", A, sv, dB, slots=0,
                                         eps=1e-9)
    assert np.all(err2 <= 1e-9)
    assert sv.report()["served_slots"][0]["slots"] in (8, 16, 32, 64)


def test_failed_build_and_unwaited_requests(gpu):
    """A matrix whose build fails fails its requests with the build's error
    and leaves the server serving the others; requests still queued or
    unwaited when the server is freed release their device buffers."""
    good = gpu.gen_pde3d(8)
    # a zero diagonal entry: ILU(0) meets a zero pivot
    bad = good.tolil()
    bad[0, 0] = 0.0
    bad = bad.tocsr()
    bad.eliminate_zeros()
    srv = gpu.Server(hbm_budget_bytes=4 << 30, workers=2)
    ig = srv.add_matrix(good, "INVK", SUB_FILLIN=0, INV_FILLIN=0)
    ib = srv.add_matrix(bad, "INVK", SUB_FILLIN=0, INV_FILLIN=0)
    n = good.shape[0]
    tb = srv.submit(ib, gpu.gen_normal(3, (n, 2)), eps=1e-8)
    tg = srv.submit(ig, gpu.gen_normal(4, (n, 3)), eps=1e-8)
    with pytest.raises(gpu.AinvError) as e:
        srv.wait(tb)
    assert "pivot" in str(e.value).lower()
    x, it, err, _ = srv.wait(tg)
    assert np.all(err <= 1e-8)
    for r in range(6):  # left unwaited
        srv.submit(ig, gpu.gen_normal(10 + r, (n, 4)), eps=1e-8)
    srv.free()
    # a new server on the same device works afterwards
    srv2 = gpu.Server(hbm_budget_bytes=4 << 30, workers=1)
    i2 = srv2.add_matrix(good, "INVK", SUB_FILLIN=1, INV_FILLIN=1)
    x, it, err, _ = srv2.wait(srv2.submit(i2, gpu.gen_normal(5, (n, 1)),
                                          eps=1e-8))
    assert np.all(err <= 1e-8)
    srv2.free()


def test_multilevel_matrix_in_the_pool(gpu):
    """A matrix served with the multilevel preconditioner (kind ML, FP32
    level factors, so the served BiCGSTAB runs the FP32 cycle) beside a
    one-level one: results as the single solves give them; a real key given
    an integer-looking value still reaches its setter; BiCG is refused when
    the matrix is registered."""
    a_ml = gpu.gen_pde3d(14, 1)
    p_ml = dict(SUB_SOLVE="INVK", COARSE_SUBSOLVE="INVK", SUB_FILLIN=0,
                INV_FILLIN=1, N_PREC_LEVS=3, AGGR_THRESH=0,
                COARSE_SWEEPS=2, FACTOR_PREC="F32")
    a_1 = gpu.gen_pde3d(10)
    p_1 = dict(SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC="F32")
    srv = gpu.Server(hbm_budget_bytes=8 << 30, workers=2)
    mid = srv.add_matrix(a_ml, "ML", **p_ml)
    oid = srv.add_matrix(a_1, "INVK", **p_1)
    with pytest.raises(gpu.AinvError) as e:
        srv.add_matrix(a_ml, "ML", METHOD="BICG", **p_ml)
    assert "M^T" in str(e.value)
    tickets = []
    for r in range(8):
        a, m = (a_ml, mid) if r % 2 == 0 else (a_1, oid)
        b = gpu.gen_normal(300 + r, (a.shape[0], 1 + r % 5))
        tickets.append((a, r % 2, b, srv.submit(m, b, eps=1e-9)))
    # the reference object gets the threshold as a real (an integer goes to
    # cseti, which ignores real keys as the reference's setters do; the
    # server's "KEY=value" parsing sends real keys to csetr)
    ref = gpu.MldMlPrec()
    for k, v in p_ml.items():
        ref.set(k, float(v) if k == "AGGR_THRESH" else v)
    ref.build(a_ml)
    assert ref.getr("AGGR_THRESH") == 0.0
    assert ref.report()["multilevel"]["levels"][0]["theta"] == 0.0
    for a, one, b, t in tickets:
        x, it, err, lat = srv.wait(t)
        assert np.all(err <= 1e-9)
        if one:
            _, ita = alone(gpu, a, "INVK", p_1, b, 1e-9)
        else:
            _, ita, _ = gpu.krylov("BICGSTAB", gpu.SystemMatrix(a), ref, b,
                                   eps=1e-9)
        for c in range(b.shape[1]):
            assert abs(int(it[c]) - int(ita[c])) <= 1, (one, c, it, ita)
            tr = np.linalg.norm(b[:, c] - a @ x[:, c]) / np.linalg.norm(b[:, c])
            assert tr < 1e-7, (one, c, tr)
    st = srv.stats()
    assert st["completed"] == 8
    srv.free()


def test_first_touch_build_does_not_block_serving(gpu):
    """A first-touch build runs on a builder thread: with a single worker,
    requests on a resident matrix submitted after a request on an unbuilt
    one are served while it builds (the build overlaps their batches and
    they finish first), and both matrices' results are right."""
    big = gpu.gen_random(100000, seed=12, median=12.0, band=5000)
    small = gpu.gen_pde3d(16)
    srv = gpu.Server(hbm_budget_bytes=8 << 30, workers=1)
    id_big = srv.add_matrix(big, "AINV", AINV_ALG="LLK", SUB_FILLIN=8,
                            SUB_ILUTHRS=5e-2, FACTOR_PREC="F32")
    id_small = srv.add_matrix(small, "INVK", SUB_FILLIN=1, INV_FILLIN=1,
                              FACTOR_PREC="F32")
    n = small.shape[0]
    srv.wait(srv.submit(id_small, gpu.gen_normal(60, (n, 1)), eps=1e-9))
    b_big = gpu.gen_normal(61, (big.shape[0], 4))
    t_big = srv.submit(id_big, b_big, eps=1e-8)
    tickets = [srv.submit(id_small, gpu.gen_normal(62 + r, (n, 8)), eps=1e-10)
               for r in range(8)]
    lat_small = []
    for t in tickets:
        x, it, err, lat = srv.wait(t)
        assert np.all(err <= 1e-10)
        lat_small.append(lat)
    x, it, err, lat_big = srv.wait(t_big)
    assert np.all(err <= 1e-8)
    for c in range(4):
        tr = (np.linalg.norm(b_big[:, c] - big @ x[:, c]) /
              np.linalg.norm(b_big[:, c]))
        assert tr < 1e-7
    st = srv.stats()
    assert st["builds"] == 2 and st["builders"] == 1
    # the small matrix's first request was done while the big one built
    assert lat_small[0] < lat_big, (lat_small, lat_big)
    assert st["build_overlap_seconds"] > 0
    srv.free()
