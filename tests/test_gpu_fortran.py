"""The Fortran iso_c_binding layer end to end on the GPU: the sample driver
(mld-ainv_amd/fortran/ppde3d_hip.f90, built by __graft_entry__.build())
builds the preconditioner through the Fortran solver objects and solves with
a Krylov method; its iteration count must match the oracle's within +-1."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mld-ainv_amd", "fortran", "build", "ppde3d_hip")

pytestmark = pytest.mark.gpu


def run_driver(deck):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (run __graft_entry__.build())")
    out = subprocess.run([DRIVER], input=deck, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = {}
    for line in out.stdout.splitlines():
        m = re.match(r"\s*(.+?)\s*:\s*(\S+)\s*$", line)
        if m:
            fields[m.group(1)] = m.group(2)
    return fields, out.stdout


@pytest.mark.parametrize("ptype,params", [
    ("INVK", dict(f1=1, f2=1, t1=0.0, t2=0.0)),
    ("AINV", dict(f1=8, f2=0, t1=0.1, t2=0.0)),
    ("INVT", dict(f1=10, f2=10, t1=1e-2, t2=1e-2))])
def test_fortran_driver_matches_oracle(mld, O, ptype, params):
    N = 12
    deck = "\n".join(["BICGSTAB", ptype, str(N), "1.d-8", "500",
                      str(params["f1"]), str(params["f2"]),
                      repr(params["t1"]), repr(params["t2"]), "LLK", "F64"])
    f, text = run_driver(deck + "\n")
    its = int(f["Number of iterations"])
    a = mld.gen_pde3d(N)
    b = a @ np.ones(a.shape[0])
    if ptype == "INVK":
        w, z, d = O.invk_bld(a, 1, 1)
    elif ptype == "INVT":
        w, z, d = O.invt_bld(a, 10, 1e-2, 10, 1e-2)
    else:
        w, z, d = O.ainv_bld(a, "LLK", 8, 0.1)
    _, it_ref, _ = O.krylov("BICGSTAB", a, w, z, d, b, 1e-8)
    assert abs(its - it_ref) <= 1, (its, it_ref, text)
    assert float(f["Convergence indicator on exit"]) <= 1e-8
    assert float(f["Max error |x - 1|"]) < 1e-5
    assert int(f["Total nonzeros for PREC"]) == a.shape[0] + w.nnz + z.nnz


@pytest.mark.parametrize("method,irst", [("RGMRES", 30), ("RGMRES", 2),
                                         ("BICGSTABL", 2), ("CGS", 0)])
def test_fortran_driver_methods_and_irst(mld, O, method, irst):
    """The deck's method and IRST lines reach psb_krylov's counterpart."""
    N = 12
    lines = [method, "INVK", str(N), "1.d-8", "800", "1", "1", "0.0", "0.0",
             "LLK", "F64"]
    if irst:
        lines.append(str(irst))
    f, text = run_driver("\n".join(lines) + "\n")
    its = int(f["Number of iterations"])
    a = mld.gen_pde3d(N)
    b = a @ np.ones(a.shape[0])
    w, z, d = O.invk_bld(a, 1, 1)
    _, it_ref, _ = O.krylov(method, a, w, z, d, b, 1e-8, 800,
                            irst=irst if irst else
                            (10 if method == "RGMRES" else 2))
    tol = irst if method == "BICGSTABL" else 1
    assert abs(its - it_ref) <= tol, (its, it_ref, text)
    assert float(f["Max error |x - 1|"]) < 1e-5


@pytest.mark.parametrize("sub,pos,mltype", [("INVK", "POST", "MULT"),
                                            ("AINV", "TWOSIDE", "MULT"),
                                            ("INVK", "POST", "ADD")])
def test_fortran_driver_multilevel(mld, O, sub, pos, mltype):
    """The ML deck fields of tests/pdegen/runs/ppde.inp (smoother factors,
    sweeps, levels, aggregation, correction, side, coarsest-level factors and
    sweeps, threshold) reach the multilevel preconditioner through the
    Fortran type mld_d_ml_prec_type; iterations match the oracle's cycle."""
    N = 14
    fill1, fill2, thr = (1, 1, 0.0) if sub == "INVK" else (6, 0, 5e-2)
    lines = ["BICGSTAB", "ML", str(N), "1.d-8", "300", str(fill1),
             str(fill2), repr(thr), "0.0", "LLK", "F64", "0",
             sub, "1", "3", "SMOOTHED", mltype, pos, sub, "4", "0.05"]
    f, text = run_driver("\n".join(lines) + "\n")
    its = int(f["Number of iterations"])
    a = mld.gen_pde3d(N)
    b = a @ np.ones(a.shape[0])
    spec = ({"kind": "INVK", "fill_in": 1, "inv_fill": 1} if sub == "INVK"
            else {"kind": "AINV", "alg": "LLK", "fill_in": 6, "thresh": 5e-2})
    H = O.MlHierarchy(a, nlev=3, theta=0.05, sub=spec, coarse=spec,
                      ml_type=mltype, pos=pos, sweeps=1, csweeps=4)
    _, it_ref, _ = O.ml_krylov("BICGSTAB", a, H, b, 1e-8, 300)
    assert abs(its - it_ref) <= 1, (its, it_ref, text)
    assert float(f["Max error |x - 1|"]) < 1e-5
    assert "Multilevel" in text
