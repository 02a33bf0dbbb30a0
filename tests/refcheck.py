"""Independent small-case formulations used to pin the oracle (CPU only).

Each routine restates a reference algorithm in a structurally different
form -- dense work vectors scanned in full instead of heaps, KIJ loop order
instead of IKJ, graph distances instead of level propagation -- so that a
dropped term, a wrong index or a wrong visiting rule in the oracle shows up
as a mismatch.  Pure Python, for matrices of at most a few hundred rows.
"""
import numpy as np
import scipy.sparse as sp


def pde3d_numpy(N, h=None):
    """7-point convection-diffusion (tests/pdegen/genpde_mod.f90:330-389,
    a = 1, b = 1/sqrt(3), c = 0), written independently of the library."""
    h = 1.0 / (N + 1) if h is None else h
    b = 1.0 / np.sqrt(3.0)
    rows, cols, vals = [], [], []
    idx = lambda x, y, z: (x * N + y) * N + z  # noqa: E731
    for x in range(N):
        for y in range(N):
            for z in range(N):
                i = idx(x, y, z)
                for dx, dy, dz, s in [(-1, 0, 0, -1), (0, -1, 0, -1),
                                      (0, 0, -1, -1), (0, 0, 1, 1),
                                      (0, 1, 0, 1), (1, 0, 0, 1)]:
                    X, Y, Z = x + dx, y + dy, z + dz
                    if 0 <= X < N and 0 <= Y < N and 0 <= Z < N:
                        rows.append(i)
                        cols.append(idx(X, Y, Z))
                        vals.append(-1.0 / (h * h) + s * b / (2 * h))
                rows.append(i)
                cols.append(i)
                vals.append(6.0 / (h * h))
    a = sp.csr_matrix((vals, (rows, cols)), shape=(N ** 3, N ** 3))
    a.sort_indices()
    return a


def key(v, j):
    return (-abs(v), j)


def dense_llk(a, nzrmax, thresh, sym=False, noth=False):
    """Left-looking biconjugation with dense work vectors, visiting every
    earlier column j (the sparse code visits only those whose row overlaps
    the current pattern; for the others the dot product is exactly 0 and
    |alpha| = 0 is never > thresh; with noth (LLK-NOTH, no alpha test) the
    zero update only touches entries that sparsify then drops, so thresh
    must be > 0 for the two to agree).  Input a already scaled.  Returns dense
    Z, W (columns) and p."""
    a = sp.csr_matrix(a)
    a.sort_indices()
    at = sp.csr_matrix(a.T)
    at.sort_indices()
    n = a.shape[0]
    eps = np.finfo(float).eps

    def rowdot(m, j, z):
        s = 0.0
        for t in range(m.indptr[j], m.indptr[j + 1]):
            s = s + m.data[t] * z[m.indices[t]]
        return s

    def sparsify(i, z, touched):
        cand = [(z[k], k) for k in sorted(touched) if k != i
                and abs(z[k]) >= thresh]
        keep = nzrmax - 1 if nzrmax >= 1 else len(cand)
        cand.sort(key=lambda e: key(*e))
        cand = cand[:keep] + [(z[i], i)]
        col = {k: v for v, k in cand}
        return col

    def one_pass(i, dotm, cols, piv):
        z = np.zeros(n)
        z[i] = 1.0
        touched = {i}
        for j in range(i):
            s = rowdot(dotm, j, z)
            alpha = -s / piv[j]
            if noth or abs(alpha) > thresh:
                for k, v in sorted(cols[j].items()):
                    z[k] = z[k] + alpha * v
                    touched.add(k)
        pi = rowdot(dotm, i, z)
        piv[i] = 1e-3 if abs(pi) < eps else pi
        return sparsify(i, z, touched)

    p = np.zeros(n)
    q = np.zeros(n)
    p[0] = a[0, 0] if a[0, 0] != 0 else 0.0
    if abs(p[0]) < eps:
        p[0] = 1e-3
    q[0] = p[0]
    zc = [{0: 1.0}]
    wc = [{0: 1.0}]
    for i in range(1, n):
        zc.append(one_pass(i, a, zc, p))
        if not sym:
            wc.append(one_pass(i, at, wc, q))
    if sym:
        wc = zc
    Z = np.zeros((n, n))
    W = np.zeros((n, n))
    for i in range(n):
        for k, v in zc[i].items():
            Z[k, i] = v
        for k, v in wc[i].items():
            W[k, i] = v
    return Z, W, p


def dense_row_inverse_invk(u, fill):
    """INVK row inverse with a dense row and a full ascending scan over
    k = i+1..n-1 instead of a heap (mld_d_invk_inv.f90:69-124).  Returns
    one {column: value} dict per row (explicit zeros kept)."""
    u = sp.csr_matrix(u)
    n = u.shape[0]
    rows = []
    INF = n + 1
    for i in range(n - 1):
        row = np.zeros(n)
        lev = np.full(n, -INF)
        for t in range(u.indptr[i], u.indptr[i + 1]):
            row[u.indices[t]] = -1.0 * u.data[t]
            lev[u.indices[t]] = 0
        row[i] = 1.0
        lev[i] = 0
        for k in range(i + 1, n):
            if lev[k] < 0:
                continue  # never touched: never on the heap
            if row[k] != 0.0 and lev[k] <= fill:
                for t in range(u.indptr[k], u.indptr[k + 1]):
                    j = u.indices[t]
                    if lev[j] < 0:
                        lev[j] = INF
                    row[j] = row[j] - row[k] * u.data[t]
                    lev[j] = min(lev[j], lev[k] + 1)
        rows.append({j: row[j] for j in range(i, n) if 0 <= lev[j] <= fill})
    rows.append({n - 1: 1.0})
    return rows


def invk_pattern_bfs(u, fill):
    """Pattern of row i of INVK's inverse as graph distance: j is kept iff a
    path i -> j in graph(U) of at most fill+1 edges exists (levels grow by
    one per edge, uplevs = 0), barring exact cancellation."""
    u = sp.csr_matrix(u)
    n = u.shape[0]
    pat = []
    for i in range(n - 1):
        dist = {i: 0}
        frontier = [i]
        for d in range(1, fill + 2):
            nxt = []
            for k in frontier:
                for t in range(u.indptr[k], u.indptr[k + 1]):
                    j = u.indices[t]
                    if j not in dist:
                        dist[j] = d
                        nxt.append(j)
            frontier = nxt
        pat.append(sorted(dist))
    pat.append([n - 1])
    return pat


def dnrm2(x):
    """Reference-BLAS dnrm2 (scaled sum of squares, the published
    algorithm; LAPACK 3.9 and earlier)."""
    x = list(x)
    if len(x) == 0:
        return 0.0
    if len(x) == 1:
        return abs(x[0])
    scale, ssq = 0.0, 1.0
    for v in x:
        if v != 0.0:
            av = abs(v)
            if scale < av:
                ssq = 1.0 + ssq * ((scale / av) * (scale / av))
                scale = av
            else:
                ssq = ssq + (av / scale) * (av / scale)
    return scale * np.sqrt(ssq)


def dense_row_inverse_invt(u, fill, thresh):
    """INVT row inverse with a dense row and a full scan over k, flags as in
    mld_d_invt_inv.f90:69-137 and the copy-out of invt_copyout.f90:201-286
    (top nup+fill by |value|, ties to the smaller column).  Returns one
    {column: value} dict per row."""
    u = sp.csr_matrix(u)
    n = u.shape[0]
    rows = []
    for i in range(n - 1):
        vals = u.data[u.indptr[i]:u.indptr[i + 1]]
        nup = len(vals)
        nrm = dnrm2(vals)
        nrm = np.sqrt(1.0 + nrm * nrm) if nrm < 1.0 else \
            nrm * np.sqrt(1.0 + 1.0 / (nrm * nrm))
        thr = thresh * nrm
        row = np.zeros(n)
        flag = np.zeros(n, dtype=bool)
        for t in range(u.indptr[i], u.indptr[i + 1]):
            row[u.indices[t]] = -1.0 * u.data[t]
            flag[u.indices[t]] = True
        row[i] = 1.0
        kept = [i]
        for k in range(i + 1, n):
            if not flag[k]:
                continue
            if abs(row[k]) < thr:
                row[k] = 0.0
                continue
            for t in range(u.indptr[k], u.indptr[k + 1]):
                j = u.indices[t]
                row[j] = row[j] - row[k] * u.data[t]
                if not flag[j]:
                    if abs(row[j]) < thr:
                        row[j] = 0.0
                    else:
                        flag[j] = True
            kept.append(k)
        cand = [(row[j], j) for j in kept if j == i or abs(row[j]) >= thr]
        cap = nup + fill
        if len(cand) > cap:
            cand = sorted(cand, key=lambda e: key(*e))[:cap]
        rows.append({j: v for v, j in cand})
    rows.append({n - 1: 1.0})
    return rows


def rows_of(m):
    """CSR -> list of {column: value} dicts."""
    m = sp.csr_matrix(m)
    return [{int(m.indices[t]): float(m.data[t])
             for t in range(m.indptr[i], m.indptr[i + 1])}
            for i in range(m.shape[0])]


def dense_iluk_levels(a, fill):
    """ILU(k) fill levels by the KIJ symbolic recurrence on a dense level
    matrix (entries above `fill` do not take part)."""
    n = a.shape[0]
    A = sp.csr_matrix(a).toarray()
    INF = 10 ** 9
    lev = np.where(A != 0, 0, INF)
    for k in range(n):
        for i in range(k + 1, n):
            if lev[i, k] > fill:
                continue
            for j in range(k + 1, n):
                if lev[k, j] <= fill:
                    lev[i, j] = min(lev[i, j], lev[i, k] + lev[k, j] + 1)
    return lev <= fill


def dense_ilut(a, fill, thresh, mutate=None):
    """ILUT(p, tau) of Saad, Iterative Methods for Sparse Linear Systems,
    2nd ed., Alg. 10.6, restated with a dense work row scanned in full
    (no heap), dense pattern flags and numpy selection, with the
    conventions the INVT build relies on (mld_d_invt_bld.f90:108-109 calls
    mld_ilut_fact with iscale = maxval):

    * A is scaled by 1 / max|a_ij| first; tau_i = tau * ||scaled row i||_2
      (reference-BLAS dnrm2 over the stored entries in stored order);
    * elimination (lines 3-8): for k < i in ascending order, if k is in the
      row's pattern, the multiplier L(i,k) = w_k * d_k (d_k the inverted
      pivot of row k) is dropped when |L(i,k)| < tau_i; otherwise
      w_j -= w_k * U(k,j) over the stored unit-upper row k, and a position
      j outside the pattern whose new value is below tau_i is set to 0 and
      stays outside (entries of A's own pattern are only tested at the
      end);
    * dropping at copy-out (line 10): entries below tau_i go; L keeps the
      nlw + p and U the nup + p largest |values| (nlw, nup the row's own
      counts in A; ties to the smaller column); the pivot always stays;
    * U is stored unit upper (scaled by the row's inverted pivot), d holds
      the inverted pivots times the scale factor (A = L diag(1/d) U).

    `mutate` plants one plausible mistake for the tests that check this
    restatement's power: "keep_smallest", "unscaled_norm" (tau_i from the
    unscaled row), "drop_before_scale" (the drop test on w_k instead of
    w_k d_k).  Returns (L rows, U rows, d, stats) with rows as
    {column: value} dicts and stats counting threshold and cap drops."""
    a = sp.csr_matrix(a)
    n = a.shape[0]
    scale = 1.0 / np.abs(a.data).max()
    Lr, Ur = [], []
    d = np.zeros(n)
    stats = {"thresh_drops": 0, "cap_cuts": 0}
    for i in range(n):
        lo, hi = a.indptr[i], a.indptr[i + 1]
        cols = a.indices[lo:hi]
        vals = [float(v) * scale for v in a.data[lo:hi]]
        w = np.zeros(n)
        pat = np.zeros(n, dtype=bool)
        for c, v in zip(cols, vals):
            w[c] = v
            pat[c] = True
        nlw = int(np.sum(cols < i))
        nup = int(np.sum(cols > i))
        nv = dnrm2(a.data[lo:hi]) if mutate == "unscaled_norm" else \
            dnrm2(vals)
        tau = thresh * nv
        for k in range(i):
            if not pat[k]:
                continue
            wk = w[k]
            lik = wk * d[k]
            test = wk if mutate == "drop_before_scale" else lik
            if abs(test) < tau:
                w[k] = 0.0
                pat[k] = False
                stats["thresh_drops"] += 1
                continue
            w[k] = lik
            for j, ukj in Ur[k].items():
                w[j] = w[j] - wk * ukj
                if not pat[j]:
                    if abs(w[j]) < tau:
                        w[j] = 0.0
                        stats["thresh_drops"] += 1
                    else:
                        pat[j] = True
        piv = w[i] if pat[i] else 0.0
        if piv == 0.0:
            raise ZeroDivisionError(f"zero pivot at row {i}")

        def keep(js, cap):
            js = np.asarray(js, dtype=np.int64)
            big = js[np.abs(w[js]) >= tau] if len(js) else js
            stats["thresh_drops"] += len(js) - len(big)
            if len(big) > cap:
                stats["cap_cuts"] += len(big) - cap
                mag = np.abs(w[big])
                order = np.lexsort((big, mag if mutate == "keep_smallest"
                                    else -mag))
                big = big[order[:max(cap, 0)]]
            return np.sort(big)

        lj = keep(np.nonzero(pat[:i])[0], nlw + fill)
        Lr.append({int(j): float(w[j]) for j in lj})
        d[i] = 1.0 / piv
        uj = keep(np.nonzero(pat[i + 1:])[0] + i + 1, nup + fill)
        Ur.append({int(j): float(w[j] * d[i]) for j in uj})
    return Lr, Ur, d * scale, stats


def csr_digest(m, values=True):
    """SHA-256 of a CSR matrix's row pointers (int64), column indices
    (int32) and, with values, its FP64 values, little-endian, columns
    sorted within rows: equal digests mean bit-identical matrices."""
    import hashlib
    m = sp.csr_matrix(m)
    m.sort_indices()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(m.indptr, dtype="<i8").tobytes())
    h.update(np.ascontiguousarray(m.indices, dtype="<i4").tobytes())
    if values:
        h.update(np.ascontiguousarray(m.data, dtype="<f8").tobytes())
    return h.hexdigest()


def vec_digest(v):
    import hashlib
    return hashlib.sha256(
        np.ascontiguousarray(v, dtype="<f8").tobytes()).hexdigest()


def factor_digests(w, z, d):
    """Digests of stored W, Z (patterns and full) and d, plus the entry
    counts, as committed for the full-size factor checks."""
    return {"w_nnz": int(w.nnz), "z_nnz": int(z.nnz),
            "w_pattern": csr_digest(w, values=False),
            "z_pattern": csr_digest(z, values=False),
            "w": csr_digest(w), "z": csr_digest(z), "d": vec_digest(d)}
