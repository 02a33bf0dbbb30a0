"""C-ABI surface on the CPU: the library loads, exports every entry point
declared in include/*.h, and the host-side logic of the solver objects
(setters, check, descr, defaults, error codes) behaves as the reference's
(no GPU needed: none of these calls touches the device)."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ainv_\w+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(mld):
    names = declared_functions()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", mld.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    lib = mld.lib()
    for n in names:
        getattr(lib, n)


def test_defaults_setters_and_check(mld):
    a = mld.MldDAinvSolver()
    # d_ainv_solver_default: alg LLK, fill_in 0, thresh 0
    assert a.geti("AINV_ALG") == mld.ALG_LLK
    assert a.geti("SUB_FILLIN") == 0 and a.getr("SUB_ILUTHRS") == 0.0
    a.set("ainv_alg", "sym-llk")  # keys and values are case-insensitive
    assert a.geti("AINV_ALG") == mld.ALG_S_LLK
    a.set("AINV_ALG", "STAB-LLK")
    assert a.geti("AINV_ALG") == mld.ALG_S_FT_LLK
    a.set("SUB_ILUTHRS", -1.0)
    a.check()  # fill_in 0 -> 1 (is_positive_nz_min), thresh -1 -> 0
    assert a.geti("SUB_FILLIN") == 1 and a.getr("SUB_ILUTHRS") == 0.0
    a.set("SOME_UNKNOWN_KEY", 3)  # passed to the base solver, ignored
    assert "Stabilized Left-looking" in a.descr()
    k = mld.MldDInvkSolver()
    k.set("SUB_FILLIN", -2)
    k.set("INV_FILLIN", 3)
    k.check()
    assert k.geti("SUB_FILLIN") == 0 and k.geti("INV_FILLIN") == 3
    assert "INVK Approximate Inverse with ILU(N)" in k.descr()
    t = mld.MldDInvtSolver()
    t.set("SUB_ILUTHRS", 1e-2)
    t.set("INV_THRESH", 5e-2)
    t.set("FACTOR_PREC", "bf16")
    assert t.getr("INV_THRESH") == 5e-2
    assert t.geti("FACTOR_PREC") == mld.PREC_BF16
    assert "Inverse fill threshold" in t.descr()
    c = mld.MldDInvtSolver()
    t.clone_settings(c)
    assert c.getr("INV_THRESH") == 5e-2
    assert mld.MldDAinvSolver.stringval("MLK") == mld.ALG_MLK
    assert mld.MldDAinvSolver.stringval("nope") == -1
    assert mld.MldDAinvSolver.algname(mld.ALG_LLK).strip() == "Left-looking"


def test_argument_errors_without_device(mld):
    t = mld.MldDInvtSolver()
    with pytest.raises(mld.AinvError) as e:
        t.set("FACTOR_PREC", "FP8")
    assert e.value.code == mld.ERR_ARG
    with pytest.raises(mld.AinvError) as e:
        t.geti("NOT_A_KEY")
    assert e.value.code == mld.ERR_ARG
    assert t.get_wrksz() == 2
    assert t.get_nzeros() == 0  # not built
    with pytest.raises(mld.AinvError) as e:
        mld.MldDInvkSolver().export("W")
    assert e.value.code == mld.ERR_STATE


def test_slice_sort_key_without_device(mld):
    # SLICE_SORT: -1 (window from the row lengths), 0 (rows in order) or a
    # multiple of 64; set before a build it touches no device
    a = mld.MldDAinvSolver()
    assert a.geti("SLICE_SORT") == -1
    a.set("SLICE_SORT", 256)
    assert a.geti("SLICE_SORT") == 256
    c = mld.MldDAinvSolver()
    a.clone_settings(c)
    assert c.geti("SLICE_SORT") == 256
    for bad in (-2, 1, 100):
        with pytest.raises(mld.AinvError) as e:
            a.set("SLICE_SORT", bad)
        assert e.value.code == mld.ERR_ARG
    assert a.geti("SLICE_SORT") == 256


def test_fortran_layer_is_built_and_binds_the_abi():
    """amdflang builds the iso_c_binding module and the sample driver;
    the driver references the C entry points it binds."""
    fdir = os.path.join(ROOT, "mld-ainv_amd", "fortran", "build")
    drv = os.path.join(fdir, "ppde3d_hip")
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not available")
    assert os.path.exists(drv), "run __graft_entry__.build()"
    for mod in ("mld_ainv_hip_cbind.mod", "mld_d_ainv_hip_mod.mod"):
        assert os.path.exists(os.path.join(fdir, mod))
    out = subprocess.run(["nm", "-u", drv], capture_output=True, text=True,
                         check=True).stdout
    for sym in ("ainv_build_csr", "ainv_krylov", "ainv_cseti", "ainv_descr",
                "ainv_gen_pde3d"):
        assert sym in out


def test_matrix_market_reader(mld, tmp_path):
    """ainv_mm_read / ainv_mm_read_vector against scipy's writer: general,
    symmetric (expanded), pattern and integer fields, duplicates summed;
    malformed files are errors, not crashes."""
    import scipy.io
    import scipy.sparse as sp
    rng = np.random.default_rng(7)
    a = sp.random(40, 40, density=0.1, random_state=3, format="csr")
    a = a + sp.eye(40)
    scipy.io.mmwrite(str(tmp_path / "g.mtx"), a)
    b = mld.mm_read(str(tmp_path / "g.mtx"))
    assert abs(b - a).max() == 0
    s = (a + a.T).tocsr()
    scipy.io.mmwrite(str(tmp_path / "s.mtx"), sp.tril(s), symmetry="symmetric")
    assert abs(mld.mm_read(str(tmp_path / "s.mtx")) - s).max() == 0
    pat = (a != 0).astype(np.int64)
    scipy.io.mmwrite(str(tmp_path / "p.mtx"), pat, field="pattern")
    assert abs(mld.mm_read(str(tmp_path / "p.mtx")) - pat).max() == 0
    (tmp_path / "d.mtx").write_text(
        "%%MatrixMarket matrix coordinate real general\n% dup\n"
        "2 2 3\n1 1 1.5\n1 1 2.0\n2 2 -1\n")
    dd = mld.mm_read(str(tmp_path / "d.mtx")).toarray()
    assert dd.tolist() == [[3.5, 0.0], [0.0, -1.0]]
    v = rng.standard_normal(17)
    scipy.io.mmwrite(str(tmp_path / "v.mtx"), v.reshape(-1, 1))
    assert np.array_equal(mld.mm_read_vector(str(tmp_path / "v.mtx")), v)
    (tmp_path / "bad.mtx").write_text(
        "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1\n")
    with pytest.raises(mld.AinvError):
        mld.mm_read(str(tmp_path / "bad.mtx"))
    with pytest.raises(mld.AinvError):
        mld.mm_read(str(tmp_path / "missing.mtx"))
