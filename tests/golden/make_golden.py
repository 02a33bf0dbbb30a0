"""Writes tests/golden/golden_small.npz.

Contents: a 5^3 convection-diffusion matrix built by the independent numpy
generator (tests/refcheck.py), its exact inverse from numpy, and the
iteration counts scipy's BiCGSTAB takes with the config-C1 AINV settings
(LLK, fill 8, drop 0.1, rtol 1e-9, b = 1) and the config-C2 INVK settings
(ILU(1), inverse level 1, rtol 1e-8, b = A 1), the preconditioner being the
oracle's factors.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402
import refcheck as R  # noqa: E402


def scipy_iters(a, w, z, d, b, rtol):
    n = a.shape[0]
    M = spla.LinearOperator((n, n), matvec=lambda v: O.apply(w, z, d, v))
    count = [0]

    def cb(_):
        count[0] += 1
    x, info = spla.bicgstab(a, b, rtol=rtol, atol=0.0, M=M, maxiter=1000,
                            callback=cb)
    assert info == 0
    return count[0]


def main():
    a = R.pde3d_numpy(5)
    n = a.shape[0]
    inv = np.linalg.inv(a.toarray())
    w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    it_ainv = scipy_iters(a, w, z, d, np.ones(n), 1e-9)
    w, z, d = O.invk_bld(a, 1, 1)
    it_invk = scipy_iters(a, w, z, d, a @ np.ones(n), 1e-8)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"),
                        n=n, a_irp=a.indptr.astype(np.int64),
                        a_ja=a.indices.astype(np.int32), a_val=a.data,
                        ainv=inv, scipy_bicgstab_iters_ainv=it_ainv,
                        scipy_bicgstab_iters_invk=it_invk)
    print("wrote golden_small.npz", n, it_ainv, it_invk)


if __name__ == "__main__":
    main()
