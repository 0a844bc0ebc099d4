"""CPU tests: the oracle pinned against the reference's golden outputs, closed
forms and independent library results; host-side generators of the product
bit-exact against the oracle; the C ABI exports what include/slepc_amd.h declares."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_outputs.json")))
LIBRES = json.load(open(os.path.join(HERE, "golden", "library_results.json")))


def lap_exact(nx, ny):
    return np.sort([4 - 2 * np.cos(i * np.pi / (ny + 1)) - 2 * np.cos(j * np.pi / (nx + 1))
                    for i in range(1, ny + 1) for j in range(1, nx + 1)])[::-1]


# ------------------------------------------------------------------ BV ops (test1.c golden)
def test_oracle_bv_test1(orc):
    g = GOLD["bv_test1"]
    X0 = np.array(g["X0"])
    Y = np.array(g["Y_after_mult"])
    z = np.array([2.0, -1.0, 0.5, -0.25, 0.125])
    y0 = orc.bv_multvec(X0, -1.0, z, 1.0, Y[:, 0])
    assert np.array_equal(y0, np.array(g["Y_after_multvec"])[:, 0])
    assert np.array_equal(orc.bv_dotvec(X0, y0), np.array(g["z_after_dotvec"]))
    X = orc.bv_multinplace(X0, np.array(g["Q"]), 1, 3)
    X *= 2.0
    assert np.array_equal(X, np.array(g["X_after_multinplace_scale"]))
    assert abs(np.linalg.norm(X[:, 0]) - g["norm2_X0"]) < 5e-5
    assert abs(np.linalg.norm(X) - g["normF_X"]) < 5e-5


def test_oracle_cgs_dgks(orc):
    """CGS+DGKS: orthonormal to < 100 eps (test2.c criterion); h holds the
    projection coefficients and the norm; an exactly dependent column is
    reduced to rounding noise."""
    rng = np.random.default_rng(3)
    n, k = 300, 10
    X = rng.standard_normal((n, k))
    X[:, 6] = X[:, 2] + 1e-10 * X[:, 6]
    V = X.copy(order="F")
    for j in range(k):
        v_before = V[:, j].copy()
        V, nrm, lind, h = orc.bv_orthonormalize_column(V, j)
        assert not lind
        # v_before = sum_i h_i v_i + nrm * v_j
        recon = V[:, :j] @ h[:j] + h[j] * V[:, j]
        assert np.linalg.norm(recon - v_before) < 1e-12 * np.linalg.norm(v_before)
    assert np.linalg.norm(V.T @ V - np.eye(k), 1) < 100 * 2.2e-16


def test_oracle_arrow_tridiag(orc):
    """DSArrowTridiag: the orthogonal Q reduces the arrowhead to tridiagonal
    form with the same spectrum."""
    rng = np.random.default_rng(1)
    n = 9
    d = rng.standard_normal(n)
    e = rng.standard_normal(n - 1)
    A = np.diag(d)
    A[n - 1, :n - 1] = e
    A[:n - 1, n - 1] = e
    dd, ee, Q = orc.arrow_tridiag(d, e)
    T = np.diag(dd) + np.diag(ee[:n - 1], 1) + np.diag(ee[:n - 1], -1)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-14
    assert np.linalg.norm(Q.T @ A @ Q - T) < 1e-13 * np.linalg.norm(A)


def test_oracle_ds_test3_golden(orc):
    """src/sys/classes/ds/tests/test3.c: d_i = i+1, e_i = 1 for i >= l, n=9 l=2 k=5."""
    g = GOLD["ds_test3"]
    n, l, k = g["n"], g["l"], g["k"]
    d = np.arange(1, n + 1, dtype=float)
    e = np.array([0.0 if i < l else 1.0 for i in range(n - 1)])
    wr = np.zeros(n + 1)
    Q = np.zeros((n + 1) * (n + 1))
    rc = orc.lib().orc_ds_hep_compact(n, l, k, d.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p),
                                      1, wr.ctypes.data_as(C.c_void_p), Q.ctypes.data_as(C.c_void_p))
    assert rc == 0
    assert np.allclose(np.round(wr[:n], 5), g["eigs"], atol=1e-12)
    # the same numbers from an independent dense eigensolver (arrowhead + tridiagonal)
    A = np.diag(d)
    for i in range(l, k):
        A[k, i] = A[i, k] = e[i]
    for i in range(k, n - 1):
        A[i, i + 1] = A[i + 1, i] = e[i]
    w = np.linalg.eigvalsh(A[l:, l:])
    assert np.allclose(np.sort(wr[l:n]), np.sort(w), atol=1e-12)


# ------------------------------------------------------------------ whole solves vs goldens
def test_oracle_ex2_golden(orc):
    g = GOLD["ex2_1"]
    A = orc.lap2d(g["n"])
    r = orc.eps_solve(A, g["nev"], g["ncv"], 1e-8, "LM", True, v0=orc.rand_uniform(0, A.n))
    assert np.allclose(np.round(r["eigr"][:4], 5), g["values"], atol=1e-12)
    ex = lap_exact(g["n"], g["n"])
    for v in r["eigr"]:
        assert np.min(np.abs(ex - v)) < 1e-8 * v
    for i in range(r["nconv"]):
        assert orc.residual_rel(A, r["eigr"][i], 0.0, r["vecs"][:, i]) < 1e-8


def test_oracle_ex5_golden_nhep(orc):
    g = GOLD["ex5_1"]
    A = orc.markov(g["m"])
    v0 = np.zeros(A.n)
    v0[:3] = 1
    r = orc.eps_solve(A, g["nev"], 0, 1e-8, "LR", False, v0=v0)
    assert np.allclose(np.round(r["eigr"][:4], 5), g["values"], atol=1e-12)
    assert abs(r["eigr"][0] - 1.0) < 1e-10          # stochastic matrix: rightmost = 1
    for i in range(r["nconv"]):
        assert orc.residual_rel(A, r["eigr"][i], r["eigi"][i], r["vecs"][:, i]) < 1e-8


def test_oracle_ex4_rdb200_golden(orc):
    g = GOLD["ex4_1"]
    d = np.load(os.path.join(HERE, "golden", "rdb200.npz"))
    A = orc.CSR(int(d["n"]), d["rowptr"], d["col"], d["val"])
    r = orc.eps_solve(A, g["nev"], 0, 1e-8, "LM", False, v0=orc.rand_uniform(0, A.n))
    assert np.allclose(np.round(r["eigr"][:4], 5), g["values"], atol=1e-12)
    assert r["its"] == g["its"]
    w = np.linalg.eigvals(A.to_dense())
    for v in r["eigr"]:
        assert np.min(np.abs(w - v)) < 1e-8 * abs(v)


def test_oracle_test27_closed_form(orc):
    g = GOLD["test27_1"]
    A = orc.lap2d(g["n"])
    r = orc.eps_solve(A, g["nev"], 0, 1e-10, "SR", True, v0=orc.rand_uniform(0, A.n))
    ex = lap_exact(10, 10)
    for v in r["eigr"]:
        assert np.min(np.abs(ex - v)) < 1e-9
    # distinct values printed by the reference test appear among the computed ones
    for v in sorted(set(g["values"]))[:3]:
        assert np.min(np.abs(r["eigr"] - v)) < 1e-5


def test_oracle_convdiff_closed_form(orc):
    """Largest real eigenvalue of the 7-point convection-diffusion operator:
    6 + 2 sum_d sqrt((1+0.1 b_d)(1-0.1 b_d)) cos(pi/(N+1))."""
    N = 14
    A = orc.convdiff3d(N)
    r = orc.eps_solve(A, 20, 64, 1e-10, "LR", False, v0=orc.rand_uniform(0, A.n))
    b = np.array([1.0, 0.5, 0.25])
    top = 6 + 2 * np.sum(np.sqrt((1 + 0.1 * b) * (1 - 0.1 * b))) * np.cos(np.pi / (N + 1))
    assert abs(r["eigr"][0] - top) < 1e-10 * top
    assert np.max(np.abs(r["eigi"])) < 1e-8


def test_oracle_library_results(orc):
    """Independent dense LAPACK results (written to fixtures by make_fixtures.py)."""
    for f in LIBRES["fv2d"]:
        A = orc.fv2d(f["nx"], f["ny"], np.array(f["theta"]))
        assert abs(A.sigma - f["sigma"]) <= 1e-13 * f["sigma"]
        r = orc.eps_solve(A, 16, 64, 1e-10, "LR", True, v0=orc.rand_uniform(0, A.n))
        ref = np.array(f["largest16"])
        for v in r["eigr"][:16]:
            assert np.min(np.abs(ref - v)) < 1e-9 * v
    C12 = orc.convdiff3d(12)
    r = orc.eps_solve(C12, 20, 64, 1e-10, "LR", False, v0=orc.rand_uniform(0, C12.n))
    ref = np.array(LIBRES["convdiff12"]["largest_real20"])
    for v in r["eigr"][:20]:
        assert np.min(np.abs(ref - v)) < 1e-9 * v
    T = orc.stencil27(10, seed=1)
    r = orc.eps_solve(T, 32, 96, 1e-10, "LM", True, v0=orc.rand_uniform(0, T.n))
    ref = np.array(LIBRES["stencil27_10"]["largest_mag32"])
    for v in r["eigr"][:32]:
        assert np.min(np.abs(ref - v)) < 1e-9 * abs(v)


def test_oracle_planted_mistake_is_caught(orc):
    """A wrong stencil constant and a dropped neighbour term (planted in the
    oracle's input, on the CPU) move the computed spectrum off the closed
    form: the closed-form checks above are able to fail.  (Flipping every
    off-diagonal sign would not: that matrix is similar to the original.)"""
    A = orc.lap2d(12)
    A.val = np.where(A.val == 4.0, 3.999, A.val)     # planted mistake 1
    A.val[A.rowptr[5]] = 0.0                          # planted mistake 2
    r = orc.eps_solve(A, 4, 0, 1e-10, "SR", True, v0=orc.rand_uniform(0, A.n))
    ex = lap_exact(12, 12)
    assert min(np.min(np.abs(ex - v)) for v in r["eigr"]) > 1e-6 or \
        max(np.min(np.abs(ex - v)) for v in r["eigr"]) > 1e-6


# ------------------------------------------------------------------ product host code (no GPU)
def _same(a, b):
    return (a.n == b.n and np.array_equal(a.rowptr, b.rowptr) and np.array_equal(a.col, b.col)
            and a.val.tobytes() == b.val.tobytes())


def test_generators_bit_exact(orc):
    import slepc_amd as sa
    assert _same(sa.lap2d(17, 13), orc.lap2d(17, 13))
    assert _same(sa.lap2d(31, 40, 10.0, 7), orc.lap2d_pot(31, 40, 7, 10.0))
    assert _same(sa.convdiff3d(9, 7, 5), orc.convdiff3d(9, 7, 5))
    assert _same(sa.stencil27(6, 5, 4, 1), orc.stencil27(6, 5, 4, 1))
    assert _same(sa.markov(15), orc.markov(15))
    for s in range(3):
        th = orc.fv_theta(5, s)
        assert np.array_equal(th, sa.fv_theta(5, s))
        a, b = sa.fv2d(40, 33, th), orc.fv2d(40, 33, th)
        assert _same(a, b) and a.sigma == b.sigma
    assert np.array_equal(sa.rand_uniform(0, 1000), orc.rand_uniform(0, 1000))


def test_generators_full_size_bit_exact(orc):
    """Workload matrices at their full sizes, assembled by the product and the oracle."""
    import slepc_amd as sa
    a, b = sa.convdiff3d(128), orc.convdiff3d(128)
    assert a.nnz == 14581760 and _same(a, b)
    del a, b
    a, b = sa.stencil27(128, seed=1), orc.stencil27(128, seed=1)
    assert a.nnz == 55742968 and _same(a, b)
    del a, b
    th = orc.fv_theta(5, 0)
    a, b = sa.fv2d(256, 256, th), orc.fv2d(256, 256, th)
    assert a.nnz == 326656 and _same(a, b)
    assert sa.lap2d(100).nnz == 49600


def test_csr_assemble_insert_values():
    """MatSetValues(INSERT_VALUES)+MatAssemblyEnd: sorted columns, last value wins."""
    import slepc_amd as sa
    rows = [2, 0, 0, 1, 0, 2, 0]
    cols = [1, 2, 0, 1, 2, 0, 1]
    vals = [5., 1., 2., 3., 4., 6., 7.]
    A = sa.csr_assemble(3, rows, cols, vals)
    assert A.rowptr.tolist() == [0, 3, 4, 6]
    assert A.col.tolist() == [0, 1, 2, 1, 0, 1]
    assert A.val.tolist() == [2., 7., 4., 3., 6., 5.]
    with pytest.raises(sa.SlepcAmdError):
        sa.csr_assemble(3, [3], [0], [1.0])


def test_product_ds_hep_golden():
    """The product's host DS (sa_ds_hep_compact) against ds/tests test3 golden."""
    import slepc_amd as sa
    g = GOLD["ds_test3"]
    n, l, k = g["n"], g["l"], g["k"]
    d = np.arange(1, n + 1, dtype=float)
    e = np.array([0.0 if i < l else 1.0 for i in range(n - 1)])
    wr = np.zeros(n + 1)
    L = sa.lib()
    L.sa_ds_hep_compact.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p]
    assert L.sa_ds_hep_compact(n, l, k, d.ctypes.data, e.ctypes.data, 1, wr.ctypes.data, None) == 0
    assert np.allclose(np.round(wr[:n], 5), g["eigs"], atol=1e-12)


def test_c_abi_exports_every_declared_symbol():
    import slepc_amd as sa
    hdr = open(os.path.join(ROOT, "include", "slepc_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = sorted(set(re.findall(r"\b(sa_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) > 40
    L = sa.lib()
    missing = [nm for nm in names if not hasattr(L, nm)]
    assert not missing, missing


# ------------------------------------------------------------------ block BV operations (oracle)
def _test11_matrix(n=20, k=8):
    """The input of the reference's bv/tests/test11.c (lines 76-89):
    X(i+j, j) = (3i + j - 2) / (2(i + j + 1)) for i <= n/2, i + j < n."""
    X = np.zeros((n, k))
    for j in range(k):
        for i in range(n // 2 + 1):
            if i + j < n:
                X[i + j, j] = (3.0 * i + j - 2) / (2 * (i + j + 1))
    return X


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_oracle_bv_orthogonalize_test11(orc, method):
    """bv/tests/test11.c (n=20, l=2, k=8; output test11_1.out / test11_6.out):
    orthogonalizing the leading columns, then the active ones, leaves the level
    of orthogonality of Q1, Q2 and Q and the residual ||X - Q R|| below
    100 eps, for gs, chol, svqb, tsqr and tsqrchol.  For all but svqb R is
    also the textbook QR factor (numpy's, up to the signs of its rows)."""
    X = _test11_matrix()
    l, k = 2, 8
    V1, R1 = orc.bv_orthogonalize(X.copy(), 0, l, method)
    V, R2 = orc.bv_orthogonalize(V1, l, k, method)
    R = np.zeros((k, k))
    R[:l, :l] = R1[:l, :l]
    R[:, l:] = R2[:, l:]
    eps = 2.220446049250313e-16
    assert np.linalg.norm(V[:, :l].T @ V[:, :l] - np.eye(l)) < 100 * eps
    assert np.linalg.norm(V[:, l:].T @ V[:, l:] - np.eye(k - l)) < 100 * eps
    assert np.linalg.norm(V.T @ V - np.eye(k)) < 100 * eps
    assert np.linalg.norm(X - V @ R) < 100 * eps
    if method != 2:
        Rn = np.linalg.qr(X)[1]
        assert np.allclose(np.abs(R), np.abs(Rn), atol=1e-13)
        assert np.allclose(np.tril(R, -1), 0.0)


@pytest.mark.parametrize("method", [3, 4])
def test_oracle_tsqr_test11_12(orc, method):
    """bv/tests/test11.c -n 180 -l 0 -k 7 -bv_orthog_block tsqr
    (test11_12.out): orthogonality and ||X - Q R|| below 100 eps; on one
    process Q and R are LAPACK's Householder QR, the same factors, signs
    included, as numpy's QR (a separate LAPACK build: equal to rounding)."""
    rng = np.random.default_rng(12)
    X = rng.standard_normal((180, 7))
    V, R = orc.bv_orthogonalize(X.copy(), 0, 7, method)
    eps = 2.220446049250313e-16
    assert np.linalg.norm(V.T @ V - np.eye(7)) < 100 * eps
    assert np.linalg.norm(X - V @ R) < 100 * eps * np.linalg.norm(X)
    Qn, Rn = np.linalg.qr(X)
    assert np.max(np.abs(R - Rn)) < 1e-13 * np.abs(Rn).max()
    assert np.max(np.abs(V - Qn)) < 1e-13


def test_oracle_bv_orthogonalize_breakdown(orc):
    """A zero column stops gs with an error (BVOrthogonalize_GS, bvorthog.c:537)."""
    X = _test11_matrix()
    X[:, 3] = 0.0
    with pytest.raises(RuntimeError):
        orc.bv_orthogonalize(X, 0, 8, 0)


def test_oracle_bv_matmult_test7(orc):
    """bv/tests/test7.c (m=n=10, k=5; output test7_1.out "Norm of error: 0."):
    X(:, j) = ones in rows 0..j, B the 1-D Laplacian (2, -1): BVMatMult equals
    the dense product exactly."""
    n, k = 10, 5
    X = np.zeros((n, k))
    for j in range(k):
        X[:j + 1, j] = 1.0
    rp = [0]
    col, val = [], []
    for i in range(n):
        for c, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= c < n:
                col.append(c)
                val.append(v)
        rp.append(len(col))
    A = orc.CSR(n, np.array(rp, np.int32), np.array(col, np.int32), np.array(val))
    Y = orc.bv_matmult(A, X)
    assert np.array_equal(Y, A.to_dense() @ X)


def test_oracle_bv_dot_matches_numpy(orc):
    rng = np.random.default_rng(3)
    X, Y = rng.standard_normal((300, 7)), rng.standard_normal((300, 5))
    M = orc.bv_dot(X, Y)
    assert np.max(np.abs(M - Y.T @ X)) < 1e-13 * np.abs(Y.T @ X).max()


# ------------------------------------------------------------------ EPSARNOLDI (oracle)
def _lap1d(orc, n):
    rp, col, val = [0], [], []
    for i in range(n):
        for c, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= c < n:
                col.append(c)
                val.append(v)
        rp.append(len(col))
    return orc.CSR(n, np.array(rp, np.int32), np.array(col, np.int32), np.array(val))


# eps/tests/output/test2_1.out (run with -eps_type arnoldi, among others)
TEST2_LARGEST = [3.98974, 3.95906, 3.90828, 3.83792]
TEST2_SMALLEST = [0.01026, 0.04094, 0.09172, 0.16208]


def _solver(name):
    """"lanczos-<type>" -> EPSLANCZOS with -eps_lanczos_reorthog <type>."""
    if name.startswith("lanczos-"):
        return dict(solver="lanczos", reorthog=name[8:])
    return dict(solver=name)


@pytest.mark.parametrize("solver", ["arnoldi", "lanczos", "lanczos-local", "lanczos-periodic", "lanczos-partial",
                                    "lanczos-selective"])
@pytest.mark.parametrize("which", ["LR", "SR"])
def test_oracle_explicit_restart_test2(orc, which, solver):
    """eps/tests/test2.c with -eps_type arnoldi and with -eps_type lanczos
    -eps_lanczos_reorthog full / local / periodic / partial (test2_2) /
    selective (test2_2_selective): 1-D Laplacian n = 30, nev 4,
    tol max(1000 eps, 1e-9): the values printed by the reference (test2_1.out)
    and the closed form 2 - 2 cos(k pi / 31)."""
    A = _lap1d(orc, 30)
    tol = max(1000 * 2.220446049250313e-16, 1e-9)
    r = orc.eps_solve(A, 4, 0, tol, which, True, v0=orc.rand_uniform(0, A.n), **_solver(solver))
    assert r["nconv"] >= 4
    gold = TEST2_LARGEST if which == "LR" else TEST2_SMALLEST
    assert list(np.round(r["eigr"][:4], 5)) == gold
    ex = np.sort(2 - 2 * np.cos(np.arange(1, 31) * np.pi / 31))
    ex = ex[::-1][:4] if which == "LR" else ex[:4]
    assert np.max(np.abs(r["eigr"][:4] - ex)) < 1e-9


def test_oracle_arnoldi_nhep(orc):
    """Arnoldi on non-symmetric operators.  Convection-diffusion at 12^3, 20
    largest real, ncv 40: the values are dense LAPACK's (explicit restart
    from one Schur vector needs ~600 restarts here; with nev 6, ncv 30 it
    stalls at 4, so max_it is raised).  The Markov model of ex5/test9 (m = 15,
    ncv 8): 1 and 0.97137 as printed in test9_1.out, and the dense values."""
    A = orc.convdiff3d(12)
    r = orc.eps_solve(A, 20, 40, 1e-10, "LR", False, v0=orc.rand_uniform(0, A.n), solver="arnoldi", max_it=3000)
    assert r["nconv"] >= 20
    ref = np.array(LIBRES["convdiff12"]["largest_real20"])
    for v in r["eigr"][:20]:
        assert np.min(np.abs(ref - v)) < 1e-9 * v
    M = orc.markov(15)
    r = orc.eps_solve(M, 4, 8, 1e-8, "LR", False, v0=orc.rand_uniform(0, M.n), solver="arnoldi", max_it=300)
    assert r["nconv"] >= 4
    assert list(np.round(r["eigr"][:2], 5)) == [1.0, 0.97137]
    w = np.sort(np.linalg.eigvals(M.to_dense()).real)[::-1][:4]
    assert np.max(np.abs(r["eigr"][:4] - w)) < 1e-7


def test_oracle_lanczos_stencil27(orc):
    """Lanczos with full reorthogonalization on the 27-point 10^3 operator,
    8 largest magnitude, ncv 24: the values are dense LAPACK's."""
    T = orc.stencil27(10, seed=1)
    r = orc.eps_solve(T, 8, 24, 1e-10, "LM", True, v0=orc.rand_uniform(0, T.n), solver="lanczos", max_it=2000)
    assert r["nconv"] >= 8
    ref = np.array(LIBRES["stencil27_10"]["largest_mag32"])
    for v in r["eigr"][:8]:
        assert np.min(np.abs(ref - v)) < 1e-12 * abs(v)


# eps/tests/test8.c (2-D Laplacian of ex2, matrix-free; the same values for
# -eps_type lanczos -eps_lanczos_reorthog local): test8_1.out, test8_2.out,
# test8_3.out -- grid, nev, ncv, mpd, which, max_it, printed values
TEST8 = {
    "1": (20, 4, 11, 0, "LM", 40000, [7.95532, 7.88881, 7.88881, 7.82229]),
    "2": (10, 4, 0, 0, "SR", 600, [0.16203, 0.39851, 0.39851, 0.63499]),
    "3": (10, 12, 21, 9, "SR", 1000, [0.16203, 0.39851, 0.39851, 0.63499, 0.77129, 0.77129, 1.00777, 1.00777,
                                      1.25018, 1.25018, 1.38056, 1.48666]),
}


@pytest.mark.parametrize("reorthog", ["local", "periodic", "partial", "selective"])
@pytest.mark.parametrize("case", sorted(TEST8))
def test_oracle_lanczos_local_test8(orc, case, reorthog):
    """EPSLanczos with local reorthogonalization on the cases of test8.c
    (tol 1e-8; case 3 sets mpd 9, so ncv = nev + mpd = 21 as
    EPSSetDimensions_Default gives): the printed values, double eigenvalues
    included, and the closed form 4 - 2 cos(i pi/(m+1)) - 2 cos(j pi/(m+1)).
    The periodic, partial and selective variants must find the same values."""
    m, nev, ncv, mpd, which, max_it, gold = TEST8[case]
    A = orc.lap2d(m)
    r = orc.eps_solve(A, nev, ncv, 1e-8, which, True, max_it=max_it, mpd=mpd, solver="lanczos", reorthog=reorthog)
    assert r["nconv"] >= nev
    assert list(np.round(r["eigr"][:nev], 5)) == gold
    c = 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    ex = np.sort((4 - c[:, None] - c[None, :]).ravel())
    ex = ex[::-1][:nev] if which == "LM" else ex[:nev]
    assert np.max(np.abs(r["eigr"][:nev] - ex)) < 1e-7 * np.max(np.abs(ex))


def test_oracle_local_sweep_is_local(orc):
    """The oracle's local Lanczos sweep (lanczos.c:89-118) on the 27-point
    10^3 operator, 80 steps from a random vector: each new vector is
    orthogonal to the two before it to rounding (the three-term recurrence),
    the Lanczos relation A V_m = V_m T_m + beta v_{m+1} e_m^T holds, and,
    as Ritz values converge, the basis loses global orthogonality (Paige):
    a sweep that orthogonalised against every column would keep ||V^T V - I||
    near 1e-15, so the last check fails if the mask is dropped."""
    T = orc.stencil27(10, seed=1)
    V, al, be = orc.lanczos_local_sweep(T, 80, orc.rand_uniform(3, T.n))
    m = len(al)
    assert m == 80
    for j in range(1, m + 1):
        for i in (j - 1, j - 2):
            if i >= 0:
                assert abs(V[:, j] @ V[:, i]) < 1e-12
    A = T.to_dense()
    Tm = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    R = A @ V[:, :m] - V[:, :m] @ Tm
    R[:, m - 1] -= be[m - 1] * V[:, m]
    assert np.abs(R).max() < 1e-10 * np.abs(A).max()
    assert np.abs(V.T @ V - np.eye(m + 1)).max() > 1e-6


def _tridiag_random(orc, n=30, seed=5):
    """eps/tests/test3.c's first matrix: tridiag(-1, d, -1) with a random
    diagonal d (here numpy's, seed 5: PetscRandom's stream is not available)."""
    d = np.random.default_rng(seed).random(n)
    rp, ci, va = [0], [], []
    for i in range(n):
        for c, v in ((i - 1, -1.0), (i, d[i]), (i + 1, -1.0)):
            if 0 <= c < n:
                ci.append(c)
                va.append(v)
        rp.append(len(ci))
    return orc.CSR(n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))


@pytest.mark.parametrize("case", ["test2_LR", "test2_interior", "test8_2", "test3_tridiag"])
def test_oracle_lanczos_delayed(orc, case):
    """EPSLanczos with delayed reorthogonalization (EPSDelayedArnoldi) where
    it converges: test2 largest and interior (sinvert at 2.1) values of
    test2_1.out, test8_2.out, and the smallest four of a test3-like
    tridiagonal at test3_2_lanczos_delayed's tol 1e-8 against dense LAPACK,
    restarting exactly as full reorthogonalization does there."""
    kw = {}
    if case.startswith("test2"):
        A, nev, ncv, mpd, tol, mi = _lap1d(orc, 30), 4, 0, 0, 1e-9, 0
        which = "LR" if case == "test2_LR" else "TM"
        if which == "TM":
            kw = dict(target=2.1, st="sinvert")
        gold = TEST2_LARGEST if which == "LR" else TEST2_INTERIOR[30]
    elif case == "test8_2":
        m, nev, ncv, mpd, which, mi, gold = TEST8["2"]
        A, tol = orc.lap2d(m), 1e-8
    else:
        A, nev, ncv, mpd, tol, mi, which = _tridiag_random(orc), 4, 0, 0, 1e-8, 500, "SR"
        gold = list(np.round(np.sort(np.linalg.eigvalsh(A.to_dense()))[:4], 5))
    v0 = orc.rand_uniform(0, A.n)
    r = orc.eps_solve(A, nev, ncv, tol, which, True, v0=v0, max_it=mi, mpd=mpd, solver="lanczos",
                      reorthog="delayed", **kw)
    assert r["nconv"] >= nev
    assert np.allclose(r["eigr"][:nev], gold, atol=5e-6)
    if case == "test3_tridiag":
        f = orc.eps_solve(A, nev, ncv, tol, which, True, v0=v0, max_it=mi, solver="lanczos")
        assert f["its"] == r["its"]


@pytest.mark.parametrize("keep", [0.5, 0.2])
@pytest.mark.parametrize("lock", [1, 0])
@pytest.mark.parametrize("case", ["test2_LR", "test2_SR", "test2_interior", "test9_1"])
def test_oracle_krylovschur_locking_restart(orc, case, lock, keep):
    """Krylov-Schur with -eps_krylovschur_locking {0,1} (test2.c, test9.c
    suffix 2 with -eps_balance none) and -eps_krylovschur_restart: the
    values of test2_1.out (interior with harmonic extraction) and of
    test9_1.out (Markov m = 15, nev 4, max_it 1500: 1, -1, 0.97137,
    -0.97137)."""
    kw = {}
    if case.startswith("test2"):
        A, nev, tol, mi, herm = _lap1d(orc, 30), 4, 1e-9, 0, True
        which = {"test2_LR": "LR", "test2_SR": "SR", "test2_interior": "TM"}[case]
        if which == "TM":
            kw = dict(extraction="harmonic", target=2.1)
        gold = {"LR": TEST2_LARGEST, "SR": TEST2_SMALLEST, "TM": TEST2_INTERIOR[30]}[which]
    else:
        A, nev, tol, mi, herm, which = orc.markov(15), 4, 1e-8, 1500, False, "LM"
        gold = [1.0, -1.0, 0.97137, -0.97137]
    r = orc.eps_solve(A, nev, 0, tol, which, herm, v0=orc.rand_uniform(0, A.n), max_it=mi, lock=lock, keep=keep,
                      **kw)
    assert r["nconv"] >= nev
    got = list(np.round(r["eigr"][:nev], 5))
    if case == "test9_1":   # 1 and -1 tie in magnitude: their order depends on rounding
        got, gold = sorted(got), sorted(gold)
    assert got == gold


@pytest.mark.parametrize("lock", [1, 0])
@pytest.mark.parametrize("balance", ["oneside", "twoside"])
def test_oracle_balance_test9(orc, balance, lock):
    """eps/tests/test9.c suffix 2: -eps_balance {oneside,twoside}
    -eps_krylovschur_locking {0,1} -eps_nev 4 -eps_max_it 1500 on the Markov
    model m = 15: the values of test9_1.out (the +-1 tie compared as a set),
    and the returned vectors, divided by D and normalised, are eigenvectors
    of the original matrix (true relative residual < 1e-7)."""
    A = orc.markov(15)
    r = orc.eps_solve(A, 4, 0, 1e-8, "LM", False, v0=orc.rand_uniform(0, A.n), max_it=1500, balance=balance,
                      lock=lock, nvec=4)
    assert r["nconv"] >= 4
    assert sorted(np.round(r["eigr"][:4], 5)) == sorted([1.0, -1.0, 0.97137, -0.97137])
    for i in range(4):
        x = r["vecs"][:, i]
        assert abs(np.linalg.norm(x) - 1.0) < 1e-12
        assert orc.residual_rel(A, r["eigr"][i], 0.0, x) < 1e-7


def test_oracle_arnoldi_delayed_test9_3(orc):
    """eps/tests/test9.c suffix 3: -eps_type arnoldi -eps_arnoldi_delayed
    -eps_largest_real -eps_nev 3 -eps_tol 1e-7 on the Markov model m = 15:
    the values of test9_3.out, and the same restarts as without delay."""
    A = orc.markov(15)
    v0 = orc.rand_uniform(0, A.n)
    r = orc.eps_solve(A, 3, 0, 1e-7, "LR", False, v0=v0, solver="arnoldi", reorthog="delayed")
    f = orc.eps_solve(A, 3, 0, 1e-7, "LR", False, v0=v0, solver="arnoldi")
    assert r["nconv"] >= 3
    assert list(np.round(r["eigr"][:3], 5)) == [1.0, 0.97137, 0.90423]
    assert r["its"] == f["its"]


def test_oracle_lanczos_partial_reorthogonalizes(orc):
    """Periodic and partial reorthogonalization run extra Gram-Schmidt passes
    when the omega estimate of the loss of orthogonality exceeds
    sqrt(eps/ncv), selective when Ritz vectors of the sweep converge: on the 27-point 10^3 operator (8 largest magnitude, ncv 24)
    they take more passes per operator application than the local variant
    (at most three per application with refinement), and fewer than twice
    as many (the reorthogonalization is occasional, not every step)."""
    T = orc.stencil27(10, seed=1)
    per = {}
    for ro in ["local", "periodic", "partial", "selective"]:
        r = orc.eps_solve(T, 8, 24, 1e-10, "LM", True, v0=orc.rand_uniform(0, T.n), solver="lanczos",
                          max_it=2000, reorthog=ro)
        assert r["nconv"] >= 8
        per[ro] = r["passes"] / r["matvecs"]
    for ro in ["periodic", "partial", "selective"]:
        assert per["local"] < per[ro] < 2 * per["local"]


TEST2_INTERIOR = {30: [2.10130, 1.89870, 2.30286, 2.50131], 32: [2.09516, 2.28463, 1.90484, 2.47152]}


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi"])
@pytest.mark.parametrize("n", [30, 32])
def test_oracle_harmonic_test2(orc, n, solver):
    """eps/tests/test2.c, interior eigenvalues: target 2.1, which = target
    magnitude, harmonic extraction for krylovschur and arnoldi (test2_1.out
    for n = 30, test2_3.out for n = 32): the printed values in the printed
    order, and the closed form 2 - 2 cos(k pi / (n+1)) of the four
    eigenvalues closest to 2.1."""
    A = _lap1d(orc, n)
    tol = max(1000 * 2.220446049250313e-16, 1e-9)
    r = orc.eps_solve(A, 4, 0, tol, "TM", True, v0=orc.rand_uniform(0, n), solver=solver,
                      extraction="harmonic", target=2.1)
    assert r["nconv"] >= 4
    assert np.allclose(r["eigr"][:4], TEST2_INTERIOR[n], atol=5e-6)
    exact = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    near = exact[np.argsort(np.abs(exact - 2.1))[:4]]
    assert np.allclose(r["eigr"][:4], near, rtol=1e-9)


def test_oracle_harmonic_nonsymmetric(orc):
    """Harmonic Krylov-Schur on the non-symmetric 7-point convection-diffusion
    operator (8^3), 6 eigenvalues closest to 3.0: dense LAPACK's eigenvalues
    closest to the target, in that order."""
    A = orc.convdiff3d(8)
    r = orc.eps_solve(A, 6, 24, 1e-10, "TM", False, v0=orc.rand_uniform(0, A.n), extraction="harmonic",
                      target=3.0, max_it=500)
    assert r["nconv"] >= 6
    ev = np.linalg.eigvals(A.to_dense())
    ev = ev[np.argsort(np.abs(ev - 3.0), kind="stable")]
    got = r["eigr"][:6] + 1j * r["eigi"][:6]
    for z in got:
        assert np.min(np.abs(ev[:8] - z)) < 1e-8
    assert np.all(np.diff(np.abs(got - 3.0)) >= -1e-9)


@pytest.mark.parametrize("case", ["markov15", "convdiff8"])
def test_oracle_twosided(orc, case):
    """Two-sided Krylov-Schur (ks-twosided.c): eigenvalues equal the one-sided
    solve (ex5_1.out holds the same values for -eps_two_sided 0 and 1 on the
    Markov model); left eigenvectors satisfy y^T A = lambda y^T to the
    tolerance, have unit norm and are biorthogonal to the right ones (Y^T X
    diagonal)."""
    if case == "markov15":
        A, nev, ncv, tol, which = orc.markov(15), 4, 0, 1e-8, "LR"
    else:
        A, nev, ncv, tol, which = orc.convdiff3d(8), 6, 24, 1e-10, "LR"
    v0 = orc.rand_uniform(0, A.n)
    r = orc.eps_solve(A, nev, ncv, tol, which, False, v0=v0, twosided=True, nvec=nev)
    r1 = orc.eps_solve(A, nev, ncv, tol, which, False, v0=v0)
    assert r["nconv"] >= nev
    assert np.allclose(r["eigr"][:nev], r1["eigr"][:nev], rtol=1e-9)
    if case == "markov15":
        assert list(np.round(r["eigr"][:nev], 5)) == [1.0, 0.97137, 0.90423, 0.85714]
    D = A.to_dense()
    X, Y = r["vecs"][:, :nev], r["lvecs"][:, :nev]
    for i in range(nev):
        lam = r["eigr"][i]
        assert abs(np.linalg.norm(Y[:, i]) - 1.0) < 1e-12
        assert np.linalg.norm(D.T @ Y[:, i] - lam * Y[:, i]) < 10 * tol * abs(lam)
    G = Y.T @ X
    assert np.max(np.abs(G - np.diag(np.diag(G)))) < 1e-6
    assert np.min(np.abs(np.diag(G))) > 1e-3


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi"])
def test_oracle_sinvert_test9_5(orc, solver):
    """eps/tests/test9.c -eps_nev 3 -st_type sinvert -eps_target .5 on the
    Markov model m = 15 (test9_5.out): shift-and-invert with an exact solve
    (the reference's default PREONLY + LU) gives the printed values in the
    printed order, and dense LAPACK's eigenvalues closest to 0.5."""
    A = orc.markov(15)
    r = orc.eps_solve(A, 3, 0, 1e-8, "TM", False, v0=orc.rand_uniform(0, A.n), target=0.5, st="sinvert",
                      solver=solver)
    assert r["nconv"] >= 3
    assert list(np.round(r["eigr"][:3], 5)) == [0.49842, 0.51928, 0.47897]
    ev = np.linalg.eigvals(A.to_dense())
    near = ev[np.argsort(np.abs(ev - 0.5))[:3]]
    assert np.allclose(r["eigr"][:3], near.real, rtol=1e-9)


@pytest.mark.parametrize("solver", ["lanczos", "lanczos-local", "lanczos-periodic", "lanczos-partial",
                                    "lanczos-selective", "krylovschur"])
def test_oracle_sinvert_test2_interior(orc, solver):
    """eps/tests/test2.c interior eigenvalues with Lanczos and shift-and-invert
    at target 2.1 (A - 2.1 I indefinite; test2_1.out) and the closed form.
    With local reorthogonalization the operator's dominant eigenvalue
    (1/(2.1013 - 2.1)) comes back as ghost copies within one sweep; they are
    purged as repeated or spurious, so each eigenvalue appears once."""
    A = _lap1d(orc, 30)
    r = orc.eps_solve(A, 4, 0, 1e-9, "TM", True, v0=orc.rand_uniform(0, 30), target=2.1, st="sinvert",
                      **_solver(solver))
    assert r["nconv"] >= 4
    assert np.allclose(r["eigr"][:4], TEST2_INTERIOR[30], atol=5e-6)
    exact = 2 - 2 * np.cos(np.arange(1, 31) * np.pi / 31)
    assert np.allclose(r["eigr"][:4], exact[np.argsort(np.abs(exact - 2.1))[:4]], rtol=1e-9)


def _complex_csr(orc, n=300, seed=3):
    """Sparse non-symmetric matrix with complex conjugate eigenvalue pairs at
    the wanted end of the spectrum (7 nonzeros per row, diagonal shifted by 4)."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for r in range(n):
        cols = sorted(set([r, (r + 1) % n, (r - 3) % n] + list(rng.choice(n, 3, replace=False))))
        for c in cols:
            ci.append(c)
            va.append(rng.standard_normal() + (4.0 if c == r else 0.0))
        rp.append(len(ci))
    return orc.CSR(n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))


@pytest.mark.parametrize("variant", ["krylovschur", "twosided", "arnoldi", "sinvert"])
def test_oracle_complex_pairs(orc, variant):
    """Conjugate pairs through every non-symmetric path: the values equal dense
    LAPACK's in the criterion's order (largest magnitude; closest to 4 with
    shift-and-invert), each pair's true residual ||A x - lambda x|| / |lambda|
    is below 1e-8 (x = re + i im from the two columns), and with two sides
    the left pair satisfies A^T y = conj(lambda) y."""
    A = _complex_csr(orc)
    D = A.to_dense()
    ev = np.linalg.eigvals(D)
    kw = dict(twosided=True) if variant == "twosided" else dict(solver="arnoldi") if variant == "arnoldi" else {}
    if variant == "sinvert":
        r = orc.eps_solve(A, 4, 0, 1e-10, "TM", False, v0=orc.rand_uniform(0, A.n), target=4.0, st="sinvert", nvec=5)
        ref = ev[np.argsort(np.abs(ev - 4.0), kind="stable")]
        nev = 4
    else:
        r = orc.eps_solve(A, 6, 24, 1e-10, "LM", False, v0=orc.rand_uniform(0, A.n), max_it=500, nvec=7, **kw)
        ref = ev[np.argsort(-np.abs(ev), kind="stable")]
        nev = 6
    got = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.any(r["eigi"][:nev] != 0.0)
    for z in got:
        assert np.min(np.abs(ref[:nev + 2] - z)) < 1e-9 * abs(z)
    i = 0
    while i < nev:
        lr, li = r["eigr"][i], r["eigi"][i]
        if li != 0.0:   # the pair's second column (one past nev when the pair straddles it)
            xr, xi = r["vecs"][:, i], r["vecs"][:, i + 1]
            assert orc.residual_rel(A, lr, li, xr, xi) < 1e-8
            if variant == "twosided":
                yr, yi = r["lvecs"][:, i], r["lvecs"][:, i + 1]
                u1 = D.T @ yr - lr * yr - li * yi
                u2 = D.T @ yi - lr * yi + li * yr
                assert np.hypot(np.linalg.norm(u1), np.linalg.norm(u2)) < 1e-8 * abs(lr + 1j * li)
            i += 2
        else:
            assert orc.residual_rel(A, lr, 0.0, r["vecs"][:, i]) < 1e-8
            i += 1


# ------------------------------------------------------------------ EPSSetTrueResidual
def _brusselator(orc, N=30):
    """test22.c's Brusselator wave model (a shell matrix there, assembled
    here): A = [tau1 T + (beta-1) I, alpha^2 I; -beta I, tau2 T - alpha^2 I],
    T = tridiag(1, -2, 1), tau_i = delta_i / (h L)^2, h = 1/(N+1)."""
    alpha, beta, d1, d2, L = 2.0, 5.45, 0.008, 0.004, 0.51302
    h = 1.0 / (N + 1)
    t1, t2 = d1 / (h * L) ** 2, d2 / (h * L) ** 2
    D = np.zeros((2 * N, 2 * N))
    for i in range(N):
        D[i, i] = -2 * t1 + beta - 1.0
        D[N + i, N + i] = -2 * t2 - alpha * alpha
        D[i, N + i] = alpha * alpha
        D[N + i, i] = -beta
        if i > 0:
            D[i, i - 1], D[N + i, N + i - 1] = t1, t2
        if i < N - 1:
            D[i, i + 1], D[N + i, N + i + 1] = t1, t2
    rows, cols = np.nonzero(D)
    rp = np.zeros(2 * N + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    return orc.CSR(2 * N, np.cumsum(rp).astype(np.int32), cols.astype(np.int32), D[rows, cols].copy())


TEST22_VALUES = sorted([(0.00019, 2.13938), (0.00019, -2.13938), (-0.67192, 2.52712), (-0.67192, -2.52712)])


@pytest.mark.parametrize("trueres", [False, True])
def test_oracle_test22_true_residual(orc, trueres):
    """test22.c -eps_nev 4 -eps_true_residual {0,1} (test22_1.out and its
    _alt orders): the two rightmost conjugate pairs of the Brusselator, and
    every error estimate below 5 |lambda| tol as the test checks."""
    A = _brusselator(orc)
    r = orc.eps_solve(A, 4, 0, 1e-8, "LR", False, v0=orc.rand_uniform(0, A.n), trueres=trueres)
    assert r["nconv"] >= 4
    got = sorted((round(float(a), 5), round(float(b), 5)) for a, b in zip(r["eigr"][:4], r["eigi"][:4]))
    assert got == TEST22_VALUES
    assert np.all(r["errest"][:4] < 5 * np.hypot(r["eigr"][:4], r["eigi"][:4]) * 1e-8)


def test_oracle_true_residual_test9_4(orc):
    """test9.c suffix 4 (-eps_nev 4 -eps_true_residual) prints test9_1.out."""
    A = orc.markov(15)
    r = orc.eps_solve(A, 4, 0, 1e-8, "LM", False, v0=orc.rand_uniform(0, A.n), trueres=True)
    assert list(np.round(r["eigr"][:4], 5)) == [1.0, -1.0, 0.97137, -0.97137]


@pytest.mark.parametrize("case", ["markov15", "convdiff10", "lap2d_pot", "sinvert_markov"])
def test_oracle_true_residual_is_the_residual_of_the_returned_vector(orc, case):
    """With the true residual the error estimate of a returned pair is
    ||A x - lambda x|| / |lambda| of the vector itself: recomputed here with
    dense numpy products (no oracle code), it agrees to rounding, and every
    pair meets the tolerance.  Without it the estimate is beta |y_last|,
    which differs from the residual in the trailing digits (the check can
    tell the two apart)."""
    kw = {}
    if case == "markov15":
        A, nev, ncv, tol, which, herm = orc.markov(15), 4, 0, 1e-8, "LM", False
    elif case == "convdiff10":
        A, nev, ncv, tol, which, herm = orc.convdiff3d(10), 8, 32, 1e-10, "LR", False
    elif case == "lap2d_pot":
        A, nev, ncv, tol, which, herm = orc.lap2d_pot(30, 27, 2, 10.0), 8, 24, 1e-10, "LM", True
    else:
        A, nev, ncv, tol, which, herm = orc.markov(15), 3, 0, 1e-8, "TM", False
        kw = dict(target=0.5, st="sinvert")
    D = A.to_dense()
    v0 = orc.rand_uniform(0, A.n)
    diffs = {}
    for tr in (True, False):
        r = orc.eps_solve(A, nev, ncv, tol, which, herm, v0=v0, nvec=nev, trueres=tr, **kw)
        assert r["nconv"] >= nev
        rel = []
        for i in range(nev):
            assert r["eigi"][i] == 0.0
            x, lam = r["vecs"][:, i], r["eigr"][i]
            rel.append(np.linalg.norm(D @ x - lam * x) / abs(lam) / np.linalg.norm(x))
        rel = np.array(rel)
        # both sides are relative to |lambda|; rounding in a residual norm is
        # about eps ||A|| ||x|| / |lambda|
        diffs[tr] = np.abs(rel - r["errest"][:nev])
        if tr:
            assert np.all(rel < tol)
            assert np.all(diffs[tr] < 1e-14)
    if case in ("convdiff10", "lap2d_pot"):
        assert diffs[False].max() > 1e-14


def test_oracle_multvec_beta_zero_does_not_read_y(orc):
    """BVMultVec with beta = 0 (gemv semantics, bvblas.c:56-67): y is not read,
    so NaN garbage in it cannot reach the result (the true-residual branch
    forms its Ritz vector into uninitialised memory this way)."""
    n, k = 1500, 5
    V = np.random.default_rng(1).standard_normal((n, k))
    q = np.arange(1.0, k + 1)
    y = orc.bv_multvec(V, 2.0, q, 0.0, np.full(n, np.nan))
    assert np.allclose(y, 2.0 * V @ q, rtol=1e-14, atol=1e-13)


# ------------------------------------------------------------------ refinement types (BVSetOrthogonalization)
def _bv_test8_matrix(n=20, k=8):
    """The columns bv/tests/test8.c fills (lines 47-58), same as test11.c."""
    return _test11_matrix(n, k)


@pytest.mark.parametrize("refine", ["ifneeded", "never", "always"])
def test_oracle_bv_test8_some_column(orc, refine):
    """bv/tests/test8.c (-bv_orthog_refine {ifneeded,never,always}, output
    test8_1.out for all three): MGS orthonormalisation of the first k-1
    columns, then BVOrthogonalizeSomeColumn of the last against the odd
    columns only; the printed V^T v_{k-1} (entries below 5 eps set to 0) is
    0.700933 0 0.846442 0 0.902098 0 0.967361.  The pass counts follow the
    refinement (never: one per column, always: two)."""
    n, k = 20, 8
    V = _bv_test8_matrix(n, k)
    passes = 0
    for j in range(k - 1):
        V, h, nrm, lind, p = orc.bv_orthogonalize_some_column(V, j, [1] * j, refine)
        V[:, j] /= nrm
        passes += p
    which = [1 if i % 2 else 0 for i in range(k - 1)]
    V, h, nrm, lind, p = orc.bv_orthogonalize_some_column(V, k - 1, which, refine)
    z = V[:, :k - 1].T @ V[:, k - 1]
    z[np.abs(z) < 5 * np.finfo(float).eps] = 0.0
    assert ["%g" % v for v in z] == ["0.700933", "0", "0.846442", "0", "0.902098", "0", "0.967361"]
    if refine == "never":
        assert passes == k - 1 and p == 1
    if refine == "always":
        assert passes == 2 * (k - 1) and p == 2


def test_oracle_refine_always_test8(orc):
    """eps/tests/test8.c suffix 1_krylovschur_vecs (-bv_orthog_refine always
    -eps_ncv 10, output test8_1.out: 7.95532, 7.88881, 7.88881, 7.82229) with
    two Gram-Schmidt passes per operator application; refinement never gives
    the same values with one pass per application.  Which copies of the
    double eigenvalue 7.88881 a Krylov method returns depends on rounding and
    on the start vector (PetscRandom in the reference), so the distinct
    printed values and the closed-form spectrum are checked."""
    A = orc.lap2d(20)
    c = 2 * np.cos(np.arange(1, 21) * np.pi / 21)
    ex = np.sort((4 - c[:, None] - c[None, :]).ravel())
    for refine, ppm in (("always", 2.0), ("never", 1.0)):
        r = orc.eps_solve(A, 4, 10, 1e-8, "LM", True, refine=refine, max_it=40000)
        assert r["nconv"] >= 4
        got = np.round(r["eigr"][:4], 5)
        assert {7.95532, 7.88881, 7.82229} <= set(got.tolist())
        assert all(np.min(np.abs(ex - v)) < 1e-7 for v in r["eigr"][:4])
        # start-vector orthonormalisations take passes of their own
        assert abs(r["passes"] / r["matvecs"] - ppm) < 0.05


def test_oracle_arnoldi_delayed1_test9_3(orc):
    """eps/tests/test9.c suffix 3 with -bv_orthog_refine never: EPSARNOLDI
    delayed runs EPSDelayedArnoldi1 (arnoldi.c:84-87), whose basis carries no
    reorthogonalisation; the printed values of test9_3.out (1.00000,
    0.97137, 0.90423) are shared with refine ifneeded."""
    A = orc.markov(15)
    v0 = orc.rand_uniform(0, A.n)
    r = orc.eps_solve(A, 3, 0, 1e-7, "LR", False, v0=v0, solver="arnoldi", reorthog="delayed", refine="never")
    assert r["nconv"] >= 3
    assert list(np.round(r["eigr"][:3], 5)) == [1.0, 0.97137, 0.90423]
    r2 = orc.eps_solve(A, 3, 0, 1e-7, "LR", False, v0=v0, solver="arnoldi", reorthog="delayed")
    assert r["its"] != r2["its"] or not np.array_equal(r["eigr"], r2["eigr"])   # a different sweep ran


def test_oracle_lanczos_delayed1_test2(orc):
    """EPSLanczos with delayed reorthogonalization and refinement never
    (lanczos.c:500, EPSDelayedArnoldi1) on test2's 1-D Laplacian (n = 30):
    the converged values are the largest of test2_1.out and of the closed
    form.  The reference has no test of this combination.  Its sweep leaves
    the last column unnormalised and reports beta = ||v_{m-1}|| for the
    bound of the last Ritz vector; without reorthogonalisation the fourth
    value of this case stagnates (100 restarts, 3 of 4 converged), which is
    recorded here as the restatement's behaviour, parity with the reference
    unpinned beyond the converged values."""
    from test_cpu_oracle import TEST2_LARGEST
    A = _lap1d(orc, 30)
    r = orc.eps_solve(A, 4, 0, 1e-9, "LR", True, v0=orc.rand_uniform(0, 30), solver="lanczos",
                      reorthog="delayed", refine="never")
    nc = r["nconv"]
    assert nc >= 3
    assert list(np.round(r["eigr"][:nc], 5)) == TEST2_LARGEST[:nc]
    ex = np.sort(2 - 2 * np.cos(np.arange(1, 31) * np.pi / 31))[::-1][:nc]
    assert np.max(np.abs(r["eigr"][:nc] - ex)) < 1e-8


@pytest.mark.parametrize("twoside", [False, True])
def test_oracle_balance_one_iteration_closed_form(orc, twoside):
    """EPSBuildBalance_Krylov (epsdefault.c:443-504) after one iteration from
    D = I and the +-1 vector z: oneside D = 1 / |A z|; twoside D = sqrt(|A^T z|
    / |A z|) where |A z| exceeds cutoff ||A z||_inf.  z is drawn as the
    oracle does (U[-1,1) seed 0xBA1A mapped to [0,1), -1 below 0.5), so a
    dropped transpose, a missing square root or a wrong sign split fails."""
    A = orc.markov(15)
    Ad = A.to_dense()
    u = orc.rand_uniform(0xBA1A, A.n)
    z = np.where(0.5 * (u + 1.0) < 0.5, -1.0, 1.0)
    assert 0.35 < np.mean(z < 0) < 0.65                  # an even split of signs
    p = np.abs(Ad @ z)
    if twoside:
        r = np.abs(Ad.T @ z)
        ok = (p > 1e-8 * p.max()) & (r != 0)
        ref = np.where(ok, np.sqrt(r / np.where(ok, p, 1.0)), 1.0)
    else:
        # rows of the Markov matrix where A z cancels to rounding level (their
        # D is 1/rounding, order-dependent) are left out
        ok = p > 1e-8 * p.max()
        ref = np.where(ok, 1.0 / np.where(ok, p, 1.0), 1.0)
    D = orc.build_balance(A, twoside, its=1)
    assert np.allclose(D[ok], ref[ok], rtol=1e-13, atol=0) and ok.sum() > 0.8 * A.n


def _dense_csr(orc, A):
    n = A.shape[0]
    rp = np.arange(0, n * n + 1, n, dtype=np.int32)
    ci = np.tile(np.arange(n, dtype=np.int32), n)
    return orc.CSR(n, rp, ci, np.ascontiguousarray(A, dtype=np.float64).ravel())


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_oracle_small_problems_every_eigenvalue(orc, n):
    """The small-problem safeguard of EPSSetUp (epssetup.c:342-345): nev and a
    set ncv above n are clamped to n, EPSSetDimensions_Default then accepts
    ncv = nev = n (epssetup.c:679), and the Arnoldi sweep breaks down once the
    Krylov space is the whole space (beta = 0, every estimate zero), so one
    restart returns all n eigenvalues of the dense matrix (numpy's eigvalsh /
    eigvals), conjugate pairs included."""
    rng = np.random.default_rng(20 + n)
    B = rng.standard_normal((n, n))
    v0 = orc.rand_uniform(0, n)
    S = B + B.T
    w = np.linalg.eigvalsh(S)
    for nev, ncv in ((n, 0), (n + 3, 0), (max(1, n - 1), 50)):
        r = orc.eps_solve(_dense_csr(orc, S), nev, ncv, 1e-10, "LM", True, v0=v0)
        assert r["nconv"] == n and r["its"] == 1 and r["reason"] > 0
        assert np.allclose(np.sort(r["eigr"]), w, rtol=0, atol=1e-13 * np.abs(w).max())
    r = orc.eps_solve(_dense_csr(orc, B), n, 0, 1e-10, "LM", False, v0=v0)
    ev = np.linalg.eigvals(B)
    got = np.sort_complex(r["eigr"][:n] + 1j * r["eigi"][:n])
    assert r["nconv"] == n and np.allclose(got, np.sort_complex(ev), rtol=0, atol=1e-12 * np.abs(ev).max())


def test_oracle_rejects_dimensions_as_the_reference():
    """EPSSetDimensions_Default (epssetup.c:679): ncv = nev on a larger
    operator is rejected; krylovschur.c:88: ncv above nev + mpd is rejected."""
    from oracle import oracle as orc
    A = orc.lap2d(10)
    v0 = orc.rand_uniform(0, A.n)
    with pytest.raises(RuntimeError, match="rc=-16"):
        orc.eps_solve(A, 10, 10, 1e-8, "LM", True, v0=v0)
    with pytest.raises(RuntimeError, match="rc=-16"):
        orc.eps_solve(A, 10, 30, 1e-8, "LM", True, v0=v0, mpd=15)
    r = orc.eps_solve(A, 10, 11, 1e-8, "LM", True, v0=v0, max_it=5000)   # ncv = nev + 1 is accepted
    assert r["nconv"] >= 10


def _scaled_lap(orc, m, scale):
    A = orc.lap2d(m)
    return orc.CSR(A.n, A.rowptr, A.col, A.val * scale)


@pytest.mark.parametrize("conv,which", [("abs", "LM"), ("rel", "LM"), ("norm", "SM")])
def test_oracle_convergence_criteria(orc, conv, which):
    """EPSConvergedAbsolute / Relative / Norm (epsdefault.c:224-256) with true
    residuals on the 2-D Laplacian scaled by 1e-4 (|lambda| and ||A||_inf =
    8e-4 far from 1), tol 1e-8: every estimate equals the explicit residual
    ||A x - lambda x|| (dense numpy) divided by 1, |lambda| or ||A||_inf, and
    meets tol.  The absolute test accepts pairs whose relative residual is
    about 1e-5, and the norm test on the smallest eigenvalues accepts
    relative residuals far above 1e-8: a criterion that divided by the wrong
    quantity fails one of these."""
    A = _scaled_lap(orc, 20, 1e-4)
    Ad = A.to_dense()
    nrma = np.max(np.abs(Ad).sum(axis=1))
    v0 = orc.rand_uniform(0, A.n)
    r = orc.eps_solve(A, 4, 20, 1e-8, which, True, v0=v0, trueres=True, conv=conv, max_it=5000, nvec=4)
    assert r["nconv"] >= 4
    rel = []
    for i in range(4):
        lam, x = r["eigr"][i], r["vecs"][:, i] / np.linalg.norm(r["vecs"][:, i])
        res = np.linalg.norm(Ad @ x - lam * x)
        denom = {"abs": 1.0, "rel": abs(lam), "norm": nrma}[conv]
        assert r["errest"][i] <= 1e-8
        assert abs(r["errest"][i] - res / denom) <= 1e-4 * res / denom
        rel.append(res / abs(lam))
    if conv == "abs":
        assert max(rel) > 1e-6
    elif conv == "norm":
        assert max(rel) > 1e-7
    else:
        assert max(rel) <= 1e-8


def _right_of(t):
    """test11.c's MyEigenSort: eigenvalues to the right of t first, then the
    closest to t."""
    def cmp(ar, ai, br, bi):
        ra, rb = t < ar, t < br
        if ra == rb:
            da, db = np.hypot(ar - t, ai), np.hypot(br - t, bi)
            return -1 if da < db else (1 if da > db else 0)
        return -1 if ra else 1
    return cmp


def test_oracle_user_comparison_test11(orc):
    """eps/tests/test11.c suffix 1 (Krylov-Schur, -st_type sinvert, shift 0.5,
    start vector of ones, the user's order "closest to the right of 0.5"):
    test11_1.out prints 0.51928, 0.55740, 0.57028, 0.57143."""
    A = orc.markov(15)
    r = orc.eps_solve_user_order(A, 4, _right_of(0.5), tol=1e-10, hermitian=False, v0=np.ones(A.n),
                                 st="sinvert", target=0.5)
    assert r["nconv"] >= 4
    assert ["%.5f" % v for v in r["eigr"][:4]] == ["0.51928", "0.55740", "0.57028", "0.57143"]
