"""Writes tests/golden/fullsize_hashes.json: SHA-256 digests of the
oracle's factors on the full-size benchmark matrices, for the GPU tests
that compare the library's factors with them bit for bit
(tests/test_gpu_fullsize.py).

* pde3d 160^3 (n = 4,096,000), INVK with ILU(1) and inverse fill 1;
* pde3d 160^3, INVT with ILUT fill 10, drop 1e-2, inverse fill 10,
  inverse drop 1e-2;
* aniso2d 2048^2, eps = 1e-3, SYM-LLK AINV, fill 8, drop 5e-2;
* random nonsymmetric n = 10^6 (seed 40, lognormal row lengths of median
  12 and sigma 1, band 5000, 2 long-range columns), AINV-LLK fill 8,
  drop 5e-2.

The matrices come from the library's generators (pinned against an
independent numpy generator at small sizes in tests/test_oracle.py); their
digests are recorded too, so the GPU test also pins the generators at full
size.  Needs the built library and oracle; about two minutes on one core.
Run from the repository root:
    python tests/golden/make_fullsize_hashes.py
"""
import importlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..", "..")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402
import refcheck as R  # noqa: E402

mld = importlib.import_module("mld-ainv_amd")

CASES = {
    "pde3d160_invk_1_1": ("pde3d", 160, "INVK",
                          dict(fill1=1, fill2=1)),
    "pde3d160_invt_10_1e-2_10_1e-2": ("pde3d", 160, "INVT",
                                      dict(fill=10, thresh=1e-2, invfill=10,
                                           invthresh=1e-2)),
    "aniso2048_symllk_8_5e-2": ("aniso2d", 2048, "AINV",
                                dict(alg="SYM-LLK", fill_in=8, thresh=5e-2)),
    "random1e6_llk_8_5e-2": ("random", 1000000, "AINV",
                             dict(alg="LLK", fill_in=8, thresh=5e-2)),
}


def matrix(kind, size):
    if kind == "pde3d":
        return mld.gen_pde3d(size)
    if kind == "aniso2d":
        return mld.gen_aniso2d(size, 1e-3)
    return mld.gen_random(size, seed=40, median=12.0, sigma=1.0, cap=400,
                          band=5000, nlong=2, dom=1.1)


def factors(kind, a, p):
    if kind == "INVK":
        return O.invk_bld(a, p["fill1"], p["fill2"])
    if kind == "INVT":
        return O.invt_bld(a, p["fill"], p["thresh"], p["invfill"],
                          p["invthresh"])
    return O.ainv_bld(a, p["alg"], p["fill_in"], p["thresh"])


def main():
    out = {}
    cache = {}
    for name, (mk, size, kind, p) in CASES.items():
        t0 = time.time()
        key = (mk, size)
        if key not in cache:
            cache.clear()
            cache[key] = matrix(mk, size)
        a = cache[key]
        w, z, d = factors(kind, a, p)
        rec = {"matrix": mk, "size": size, "kind": kind, "params": p,
               "a_nnz": int(a.nnz), "a": R.csr_digest(a)}
        rec.update(R.factor_digests(w, z, d))
        out[name] = rec
        print(f"{name}: {time.time() - t0:.1f} s, nnz W {w.nnz} Z {z.nnz}",
              flush=True)
    with open(os.path.join(HERE, "fullsize_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
