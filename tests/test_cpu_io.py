"""CPU tests of the file formats at the boundary: MatLoad/VecLoad/VecView in
PETSc binary form and the text of EPSValuesView, against the reference's data
files and golden outputs (src/eps/tests/output/test23_*.out)."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
RDB200 = os.path.join(REF, "share/slepc/datafiles/matrices/rdb200.petsc")


def write_petsc_aij(path, n, rowptr, col, val):
    """An independent writer of the PETSc binary AIJ layout (numpy big-endian)."""
    with open(path, "wb") as f:
        np.array([1211216, n, n, len(col)], ">i4").tofile(f)
        np.diff(rowptr).astype(">i4").tofile(f)
        np.asarray(col, ">i4").tofile(f)
        np.asarray(val, ">f8").tofile(f)


def test_mat_load_round_trip(tmp_path):
    import slepc_amd as sa
    d = np.load(os.path.join(HERE, "golden", "rdb200.npz"))
    p = str(tmp_path / "rdb200.petsc")
    write_petsc_aij(p, int(d["n"]), d["rowptr"], d["col"], d["val"])
    A = sa.load_petsc_binary(p)
    assert A.n == 200 and A.nnz == 1120
    assert np.array_equal(A.rowptr, d["rowptr"])
    assert np.array_equal(A.col, d["col"])
    assert np.array_equal(A.val, d["val"])        # bit-exact


@pytest.mark.skipif(not os.path.exists(RDB200), reason="reference data file not present")
def test_mat_load_reference_file():
    """The reference's own rdb200.petsc, read by the library, equals the
    fixture made from it by the independent Python reader."""
    import slepc_amd as sa
    d = np.load(os.path.join(HERE, "golden", "rdb200.npz"))
    A = sa.load_petsc_binary(RDB200)
    assert np.array_equal(A.rowptr, d["rowptr"]) and np.array_equal(A.col, d["col"])
    assert np.array_equal(A.val, d["val"])


def test_mat_load_rejects_bad_files(tmp_path):
    import slepc_amd as sa
    p = str(tmp_path / "bad.petsc")
    np.array([1211214, 3, 3, 0], ">i4").tofile(p)           # a Vec class id
    with pytest.raises(sa.SlepcAmdError, match="class id"):
        sa.load_petsc_binary(p)
    write_petsc_aij(p, 3, np.array([0, 1, 2, 3]), [0, 1, 5], [1.0, 2.0, 3.0])
    with pytest.raises(sa.SlepcAmdError, match="column index"):
        sa.load_petsc_binary(p)
    with open(p, "wb") as f:
        np.array([1211216, 3, 3, 3, 1, 1], ">i4").tofile(f)   # truncated
    with pytest.raises(sa.SlepcAmdError, match="truncated"):
        sa.load_petsc_binary(p)


def test_vec_write_and_load(tmp_path):
    import ctypes as C
    import slepc_amd as sa
    p = str(tmp_path / "vecs.bin")
    x0 = np.linspace(-1, 1, 7)
    x1 = np.arange(3, dtype=float) * np.pi
    L = sa.lib()
    sa.check(L.sa_vec_write_petsc_binary(p.encode(), 7, x0.ctypes.data, 0))
    sa.check(L.sa_vec_write_petsc_binary(p.encode(), 3, x1.ctypes.data, 1))
    raw = open(p, "rb").read()
    expect = (np.array([1211214, 7], ">i4").tobytes() + x0.astype(">f8").tobytes()
              + np.array([1211214, 3], ">i4").tobytes() + x1.astype(">f8").tobytes())
    assert raw == expect
    assert np.array_equal(sa.load_petsc_vec(p, 0), x0)
    assert np.array_equal(sa.load_petsc_vec(p, 1), x1)
    with pytest.raises(sa.SlepcAmdError):
        sa.load_petsc_vec(p, 2)
    n = C.c_int64()
    assert L.sa_vec_load_petsc_binary(p.encode(), 0, C.byref(n), None) == 0 and n.value == 7


def test_values_view_ascii_matches_test23_1():
    """test23_1.out, lines 'Eigenvalues = ' ... for the diagonal matrix 1..6
    solved for the largest real values."""
    import slepc_amd as sa
    txt = sa.format_eigenvalues([6.0, 5.0, 4.0, 3.0, 2.0, 1.0], fmt="ascii")
    assert txt == "Eigenvalues = \n   6.00000\n   5.00000\n   4.00000\n   3.00000\n   2.00000\n   1.00000\n\n"
    path = os.path.join(REF, "src/eps/tests/output/test23_1.out")
    if os.path.exists(path):
        assert txt in open(path).read()


def test_values_view_complex_pairs_match_test9():
    """test9_7_real.out lists the real eigenvalues 1, -1, 0.97137, -0.97137;
    complex values print as re%+.5fi, and a part below 1e-10 times the other
    prints as zero (SlepcPrintEigenvalueASCII, slepcimpl.h:92-112)."""
    import slepc_amd as sa
    txt = sa.format_eigenvalues([1.0, -1.0, 0.9713712, -0.9713712], [0, 0, 0, 0])
    assert txt == "Eigenvalues = \n   1.00000\n   -1.00000\n   0.97137\n   -0.97137\n\n"
    txt = sa.format_eigenvalues([0.5, 0.5, 2.0, 1e-12], [0.25, -0.25, 1e-12, 3.0])
    assert txt == "Eigenvalues = \n   0.50000+0.25000i\n   0.50000-0.25000i\n   2.00000\n   0.00000+3.00000i\n\n"


def test_values_view_matlab_matches_test23_2():
    """test23_2.out: -eps_view_values ::ascii_matlab on the 20x20 diagonal
    matrix with nev 1.  The reference's viewer prints these lines at tab level
    one (a leading space), which the text itself does not contain."""
    import slepc_amd as sa
    txt = sa.format_eigenvalues([20.0], fmt="matlab", name="eps")
    assert txt == "Lambda_eps = [\n2.0000000000000000e+01\n];\n"
    path = os.path.join(REF, "src/eps/tests/output/test23_2.out")
    if os.path.exists(path):
        gold = open(path).read()
        assert "".join(" " + ln + "\n" for ln in txt.splitlines()) in gold
