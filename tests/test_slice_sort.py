"""Row sort of the sliced (HLL) layouts on the host: the stored order, its
padding and the window the library picks (no device needed)."""
import numpy as np
import pytest

SLICE = 64


def _padding(lens, order):
    """Stored entries of 64-row slices holding the rows in `order`."""
    n = len(lens)
    ls = np.zeros(SLICE * ((n + SLICE - 1) // SLICE), dtype=np.int64)
    ls[:n] = lens[order]
    return int(ls.reshape(-1, SLICE).max(axis=1).sum() * SLICE)


def _window_order(lens, window):
    """Rows sorted by length (longest first, ties in row order) within
    consecutive windows of `window` rows."""
    order = np.arange(len(lens))
    for b in range(0, len(lens), window):
        seg = order[b:b + window]
        order[b:b + window] = seg[np.lexsort((seg, -lens[seg]))]
    return order


def _weight(lens, rows):
    """Stored entries of the 64-row slices of a group of rows."""
    return sum(int(lens[rows[i:i + SLICE]].max()) * SLICE
               for i in range(0, len(rows), SLICE))


def _check_perm(lens, perm, window):
    n = len(lens)
    assert len(perm) == SLICE * ((n + SLICE - 1) // SLICE)
    assert (perm[n:] == n).all()
    p = perm[:n]
    assert sorted(p.tolist()) == list(range(n))
    if window == 0:
        assert (p == np.arange(n)).all()
        return
    # the windows' sorted rows cut into groups of 4 slices (256 rows); the
    # full groups of each eighth of the groups stored heaviest first, the
    # last, partial group last
    ref = _window_order(lens, window)
    GR = 4 * SLICE
    ns = (n + SLICE - 1) // SLICE
    ng = (ns + 3) // 4
    nfull = n // GR
    for x in range(8):
        b, e = ng * x // 8, min(ng * (x + 1) // 8, nfull)
        if e <= b:
            continue
        stored = [tuple(p[g * GR:(g + 1) * GR]) for g in range(b, e)]
        refg = [tuple(ref[g * GR:(g + 1) * GR]) for g in range(b, e)]
        assert sorted(stored) == sorted(refg)
        w = [_weight(lens, np.array(g)) for g in stored]
        assert all(w[i] >= w[i + 1] for i in range(len(w) - 1))
        # equal weights keep the windows' order
        for i in range(len(w) - 1):
            if w[i] == w[i + 1]:
                assert refg.index(stored[i]) < refg.index(stored[i + 1])
    assert (p[nfull * GR:] == ref[nfull * GR:]).all()


@pytest.mark.parametrize("n,window", [(1, 64), (63, 64), (64, 64),
                                      (1000, 256), (5000, 1024), (777, 4096),
                                      (3000, 0), (20480, 256), (70000, 4096),
                                      (33000, 64)])
def test_plan_is_a_windowed_length_sort(mld, n, window):
    rng = np.random.default_rng(n + window)
    lens = np.minimum(np.round(rng.lognormal(np.log(8), 1.0, n)), 300)
    lens = lens.astype(np.int64)
    irp = np.concatenate([[0], np.cumsum(lens)])
    perm, pad, upad, used = mld.slice_sort_plan(irp, window)
    assert used == window
    _check_perm(lens, perm, window)
    assert pad == _padding(lens, perm[:n])
    assert upad == _padding(lens, np.arange(n))
    assert pad <= upad


def test_auto_window_rule(mld):
    # lognormal rows as in the random nonsymmetric matrix: the rows in order
    # pad ~6x; the pick is the smallest window within 5 % of the best of
    # 256, 1024, 4096
    rng = np.random.default_rng(7)
    n = 200000
    lens = np.minimum(np.round(rng.lognormal(np.log(12), 1.0, n)), 400)
    lens = lens.astype(np.int64)
    irp = np.concatenate([[0], np.cumsum(lens)])
    perm, pad, upad, used = mld.slice_sort_plan(irp, -1)
    assert upad > 5 * lens.sum()
    cand = {w: mld.slice_sort_plan(irp, w)[1] for w in (256, 1024, 4096)}
    best = min(cand.values())
    assert used in cand and cand[used] == pad
    assert pad * 100 <= best * 105
    assert all(cand[w] * 100 > best * 105 for w in cand if w < used)
    _check_perm(lens, perm, used)


def test_auto_keeps_regular_rows_in_order(mld):
    # a 7-point stencil pads ~1.5 % in row order: no sort
    a = mld.gen_pde3d(20)
    perm, pad, upad, used = mld.slice_sort_plan(a.indptr, -1)
    assert used == 0 and pad == upad
    assert (perm[:a.shape[0]] == np.arange(a.shape[0])).all()
    # rows padding > 25 % whose sort saves < 20 %: one slice (the sort
    # cannot do better than the rows in order) ...
    lens = np.ones(SLICE, dtype=np.int64)
    lens[5] = 40
    irp = np.concatenate([[0], np.cumsum(lens)])
    _, pad, upad, used = mld.slice_sort_plan(irp, -1)
    assert upad > 1.25 * lens.sum() and used == 0 and pad == upad
    # ... and with a second slice of rows of 5: sorted, the long row takes
    # 63 of them into its slice and both widths stay (40, 5)
    lens = np.concatenate([lens, np.full(SLICE, 5)])
    irp = np.concatenate([[0], np.cumsum(lens)])
    assert mld.slice_sort_plan(irp, 128)[1] == (40 + 5) * SLICE
    assert mld.slice_sort_plan(irp, -1)[3] == 0


def test_plan_argument_errors(mld):
    irp = np.array([0, 1, 3], dtype=np.int64)
    for w in (-2, 1, 100):
        with pytest.raises(mld.AinvError) as e:
            mld.slice_sort_plan(irp, w)
        assert e.value.code == mld.ERR_ARG
    with pytest.raises(mld.AinvError):
        mld.slice_sort_plan(np.array([0, 2, 1], dtype=np.int64), 64)
