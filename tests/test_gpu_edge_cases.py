"""Setup checks and small problems of the EPS interface, as the reference
sets them up.  EPSSetDimensions rejects nev <= 0 (epsopts.c:409).  At setup
the small-problem safeguard clamps nev and a set ncv to n (epssetup.c:342-345);
EPSSetDimensions_Default then requires ncv >= nev+1 unless ncv = nev = n
(epssetup.c:679), Krylov-Schur requires ncv <= nev+mpd (krylovschur.c:88), and
no more initial vectors than ncv are accepted (epssetup.c:423).  Operators of
1 to 8 rows run the whole path: the Arnoldi sweep breaks down once the Krylov
space is the whole space, and one restart returns every eigenvalue of the
dense matrix, for each solver, in FP64 and mixed precision, one- and
two-sided."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 8]


def _dense(sa, orc, A):
    n = A.shape[0]
    rp = np.arange(0, n * n + 1, n, dtype=np.int32)
    ci = np.tile(np.arange(n, dtype=np.int32), n)
    val = np.ascontiguousarray(A, dtype=np.float64).ravel()
    return sa.CSR(n, rp, ci, val), orc.CSR(n, rp, ci, val.copy())


def _eps(sa, Ah, herm, nev, ncv, solver="krylovschur", mixed=False, twosided=False, v0=None):
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(Ah, sa.FP64), sa.Mat(Ah, sa.FP32) if mixed else None)
    E.setProblemType(herm)
    E.setWhichEigenpairs("LM")
    E.setDimensions(nev, ncv)
    E.setTolerances(1e-10, 0)
    if mixed:
        E.setMixedPrecision(True, 1e-5)
    if twosided:
        E.setTwoSided(True)
    if v0 is not None:
        E.setInitialSpace(v0)
    return E


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi", "lanczos"])
@pytest.mark.parametrize("n", SIZES)
def test_small_symmetric_every_eigenvalue(sa, orc, n, solver):
    """nev = n, nev = n + 3 (clamped to n) and nev = n - 1 with ncv 50
    (clamped to n): every eigenvalue of the dense symmetric matrix, equal to
    the oracle's and to numpy's eigvalsh, in one restart, each pair within
    the tolerance."""
    rng = np.random.default_rng(20 + n)
    B = rng.standard_normal((n, n))
    S = B + B.T
    Ah, Ao = _dense(sa, orc, S)
    v0 = orc.rand_uniform(0, n)
    w = np.linalg.eigvalsh(S)
    for nev, ncv in ((n, 0), (n + 3, 0), (max(1, n - 1), 50)):
        E = _eps(sa, Ah, True, nev, ncv, solver, v0=v0)
        E.solve()
        r = orc.eps_solve(Ao, nev, ncv, 1e-10, "LM", True, v0=v0, solver=solver)
        assert E.getConverged() == n == r["nconv"] and E.getIterationNumber() == r["its"] == 1
        vals = np.sort([E.getEigenpair(i, False)[0] for i in range(n)])
        assert np.allclose(vals, w, rtol=0, atol=1e-13 * np.abs(w).max())
        assert np.allclose(vals, np.sort(r["eigr"]), rtol=0, atol=1e-13 * np.abs(w).max())
        for i in range(n):
            assert E.computeError(i) <= 1e-10


@pytest.mark.parametrize("twosided", [False, True])
@pytest.mark.parametrize("n", SIZES)
def test_small_nonsymmetric_every_eigenvalue(sa, orc, n, twosided):
    """Non-symmetric dense matrices (complex conjugate pairs from n = 3 on):
    every eigenvalue equals numpy's eigvals and the oracle's, one- and
    two-sided; the two-sided solve also returns left vectors within the
    tolerance."""
    rng = np.random.default_rng(40 + n)
    B = rng.standard_normal((n, n))
    Ah, Ao = _dense(sa, orc, B)
    v0 = orc.rand_uniform(0, n)
    E = _eps(sa, Ah, False, n, 0, twosided=twosided, v0=v0)
    E.solve()
    r = orc.eps_solve(Ao, n, 0, 1e-10, "LM", False, v0=v0, twosided=twosided)
    assert E.getConverged() == n == r["nconv"]
    got = np.sort_complex(np.array([complex(*E.getEigenpair(i, False)[:2]) for i in range(n)]))
    ev = np.sort_complex(np.linalg.eigvals(B))
    scale = np.abs(ev).max()
    assert np.allclose(got, ev, rtol=0, atol=1e-12 * scale)
    assert np.allclose(got, np.sort_complex(r["eigr"] + 1j * r["eigi"]), rtol=0, atol=1e-12 * scale)
    for i in range(n):
        assert E.computeError(i) <= 1e-10


def test_small_mixed_precision(sa, orc):
    """Mixed FP32 -> FP64 on an 8-row symmetric matrix: the FP32 phase and the
    FP64 polish both stop at the breakdown, and every eigenvalue is exact to
    FP64 rounding."""
    rng = np.random.default_rng(7)
    B = rng.standard_normal((8, 8))
    S = B + B.T
    Ah, _ = _dense(sa, orc, S)
    E = _eps(sa, Ah, True, 8, 0, mixed=True, v0=orc.rand_uniform(0, 8))
    E.solve()
    assert E.getConverged() == 8
    vals = np.sort([E.getEigenpair(i, False)[0] for i in range(8)])
    w = np.linalg.eigvalsh(S)
    assert np.allclose(vals, w, rtol=0, atol=1e-13 * np.abs(w).max())


def test_setup_rejects_what_the_reference_rejects(sa, orc):
    """nev <= 0, ncv = nev on a larger operator, ncv > nev + mpd and more
    initial vectors than ncv are errors with the reference's messages (and
    ncv above the fused kernels' 126 an error of its own); the same
    EPS then solves once the dimensions are valid."""
    Ah = sa.lap2d(10)
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(True)
    with pytest.raises(RuntimeError, match="nev must be positive"):
        E.setDimensions(0, 10)
    E.setDimensions(10, 10)
    with pytest.raises(RuntimeError, match="ncv must be at least nev\\+1"):
        E.solve()
    E.setDimensions(10, 30, 15)
    with pytest.raises(RuntimeError, match="must not be larger than nev\\+mpd"):
        E.solve()
    E.setDimensions(4, 10)
    E.setInitialSpace(orc.rand_uniform(3, 12 * Ah.n).reshape(12, Ah.n))
    with pytest.raises(RuntimeError, match="initial vectors is larger than ncv"):
        E.solve()
    E.setInitialSpace(orc.rand_uniform(0, Ah.n))
    E.solve()
    assert E.getConverged() >= 4
    for i in range(4):
        assert E.computeError(i) <= 1e-8
    E2 = sa.EPS()                                 # 144 rows: ncv 127 is not clamped
    E2.setOperators(sa.Mat(sa.lap2d(12)))
    E2.setProblemType(True)
    E2.setDimensions(10, 127)
    with pytest.raises(RuntimeError, match="ncv above 126"):
        E2.solve()


def _csr4():
    rp = np.array([0, 2, 4, 6, 8], np.int32)
    ci = np.array([0, 1, 1, 2, 2, 3, 0, 3], np.int32)
    return rp, ci, np.arange(1.0, 9.0)


@pytest.mark.parametrize("bad,msg", [("col_range", "column index out of range"),
                                     ("col_negative", "column index out of range"),
                                     ("col_order", "must ascend within a row"),
                                     ("rowptr_order", "row pointer not monotone"),
                                     ("rowptr_nnz", "row pointer does not match nnz"),
                                     ("rowptr_start", "row pointer does not match nnz")])
def test_csr_rejected_before_upload(sa, bad, msg):
    """MatSetValues / MatAssemblyEnd reject what sa_mat_create_csr rejects on
    the host, before anything reaches the device: column indices outside
    [0, n), columns not ascending within a row (the SpMV kernels and the
    transpose assume sorted rows), row pointers that decrease or do not
    span [0, nnz].  A valid operator created afterwards multiplies exactly."""
    from types import SimpleNamespace
    rp, ci, val = _csr4()
    nnz = 8
    if bad == "col_range":
        ci[5] = 4
    elif bad == "col_negative":
        ci[0] = -1
    elif bad == "col_order":
        ci[2], ci[3] = 2, 1
    elif bad == "rowptr_order":
        rp[2] = 1
    elif bad == "rowptr_nnz":
        nnz = 7
    else:
        rp[0] = 1
    with pytest.raises(RuntimeError, match=msg):
        sa.Mat(SimpleNamespace(n=4, rowptr=rp, col=ci, val=val, nnz=nnz))
    rp, ci, val = _csr4()
    M = sa.Mat(sa.CSR(4, rp, ci, val))
    x = np.array([1.0, -2.0, 3.0, 5.0])
    A = np.zeros((4, 4))
    for r in range(4):
        A[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert np.array_equal(M.mult_host(x), A @ x)


def test_bv_arguments_rejected(sa):
    """BV calls with columns outside the basis or an inverted active range
    are errors (BVSetActiveColumns, BVGetColumn argument checks), and the
    basis is usable afterwards."""
    V = sa.BV(1000, 8)
    with pytest.raises(RuntimeError, match="invalid active columns"):
        V.set_active_columns(3, 2)
    with pytest.raises(RuntimeError, match="invalid active columns"):
        V.set_active_columns(0, 9)
    with pytest.raises(RuntimeError, match="invalid column"):
        V.norm_column(8)
    with pytest.raises(RuntimeError, match="invalid column"):
        V.copy_column(0, -1)
    with pytest.raises(RuntimeError, match="too many columns"):
        sa.BV(1000, 128)
    x = np.linspace(-1.0, 1.0, 1000)
    V.set_column(2, x)
    assert abs(V.norm_column(2) - np.linalg.norm(x)) < 1e-13 * np.linalg.norm(x)


@pytest.mark.parametrize("bad,msg", [("col_range", "column index out of range in batch"),
                                     ("col_negative", "column index out of range in batch"),
                                     ("col_order", "must ascend within a row"),
                                     ("rowptr_order", "row pointer not monotone in batch"),
                                     ("rowptr_start", "must start at 0")])
def test_batch_csr_rejected(sa, bad, msg):
    """The batched solver checks every problem's CSR as a single operator is
    checked, before anything is allocated or uploaded (here the second of two
    problems is malformed)."""
    from types import SimpleNamespace
    from slepc_amd.batch import Batch
    rp, ci, val = _csr4()
    if bad == "col_range":
        ci[5] = 4
    elif bad == "col_negative":
        ci[0] = -1
    elif bad == "col_order":
        ci[2], ci[3] = 2, 1
    elif bad == "rowptr_order":
        rp[2] = 1
    else:
        rp[0] = 1
    with pytest.raises(RuntimeError, match=msg):
        Batch([sa.lap2d(6), SimpleNamespace(n=4, rowptr=rp, col=ci, val=val)])


def test_batch_dimensions_and_small_problems(sa, orc):
    """sa_batch_solve needs nev < ncv <= min(64, n) for every problem (the
    lock-step kernels hold at most 64 columns; ncv = nev is rejected as
    EPSSetDimensions_Default rejects it); a batch of a 16-row and a 64-row
    problem then solves, each problem equal to the oracle run alone."""
    from slepc_amd.batch import Batch
    csrs = [sa.lap2d(4, pot_scale=10.0, seed=1), sa.lap2d(8, pot_scale=10.0, seed=2)]
    B = Batch(csrs)
    with pytest.raises(RuntimeError, match="invalid ncv"):
        B.solve(8, 24, 1e-10, "LM")          # ncv > n of the first problem
    with pytest.raises(RuntimeError, match="invalid ncv"):
        B.solve(8, 8, 1e-10, "LM")
    v0 = [orc.rand_uniform(orc.sub_seed(5, p), A.n) for p, A in enumerate(csrs)]
    B.solve(4, 12, 1e-10, "LM", v0=v0)
    err = B.errors()
    for p, A in enumerate(csrs):
        nconv, its, reason = B.converged(p)
        assert reason == 1 and nconv >= 4
        Ao = orc.CSR(A.n, A.rowptr, A.col, A.val)
        r = orc.eps_solve(Ao, 4, 12, 1e-10, "LM", True, v0=v0[p])
        vals = np.sort([B.eigenpair(p, i, False)[0] for i in range(4)])
        assert np.allclose(vals, np.sort(r["eigr"][:4]), rtol=1e-10, atol=0)
        w = np.linalg.eigvalsh(Ao.to_dense())
        assert np.allclose(np.sort(vals), np.sort(w[np.argsort(-np.abs(w))[:4]]), rtol=1e-10, atol=0)
        assert np.all(err[p] <= 1e-10)


def test_failed_allocation_reported_and_library_usable(sa):
    """A basis too large for the device (32 TiB) fails in hipMalloc: the call
    returns the error, leaves no half-built basis behind, and later bases
    and solves work."""
    with pytest.raises(RuntimeError, match="hipMalloc"):
        sa.BV(1 << 40, 4)
    V = sa.BV(5000, 4)
    x = np.cos(np.arange(5000.0))
    V.set_column(1, x)
    assert abs(V.norm_column(1) - np.linalg.norm(x)) < 1e-13 * np.linalg.norm(x)
    E = sa.EPS()
    E.setOperators(sa.Mat(sa.lap2d(20)))
    E.setProblemType(True)
    E.setDimensions(4, 12)
    E.setTolerances(1e-10, 5000)
    E.solve()
    assert E.getConverged() >= 4


def test_wide_fp64_basis_beyond_64kb_lds(sa, orc):
    """FP64 bases wider than 108 columns need more than 64 KB of LDS in the
    Gram-Schmidt slot kernels (the tile of column c takes (66 c + 320) x 8
    bytes), which gfx950 allows a workgroup (160 KB).  A 127-column basis
    orthonormalised column by column equals the oracle to rounding, and a
    Krylov-Schur solve with ncv 120 returns the oracle's values and the
    closed-form spectrum of the 2-D Laplacian."""
    n, k = 3000, 127
    X0 = np.random.default_rng(12).standard_normal((n, k))
    V = sa.BV(n, k, sa.FP64)
    V.set_matrix(X0)
    Vo = X0.copy(order="F")
    for j in range(k):
        nrm, lind = V.orthonormalize_column(j)
        Vo, nrmo, lindo, ho = orc.bv_orthonormalize_column(Vo, j)
        assert not lind and not lindo and abs(nrm - nrmo) <= 1e-12 * abs(nrmo)
        if j >= 100:
            h = V.coefficients(j, j + 1)[:j + 1, 0]
            assert np.max(np.abs(h - ho)) <= 1e-12 * np.abs(ho).max()
    Q = V.get_matrix(k)
    assert np.max(np.abs(Q - Vo)) < 1e-11
    assert np.linalg.norm(Q.T @ Q - np.eye(k), 1) < 100 * 2.2e-16 * k
    nx, ny = 30, 31
    Ah = sa.lap2d(nx, ny)
    Ao = orc.CSR(Ah.n, Ah.rowptr, Ah.col, Ah.val)
    v0 = orc.rand_uniform(0, Ah.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(True)
    E.setDimensions(40, 120)
    E.setTolerances(1e-10, 1000)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, 40, 120, 1e-10, "LM", True, v0=v0, max_it=1000)
    assert E.getConverged() >= 40 and r["nconv"] >= 40
    vals = np.sort([E.getEigenpair(i, False)[0] for i in range(40)])
    assert np.allclose(vals, np.sort(r["eigr"][:40]), rtol=1e-10, atol=0)
    i, j = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1))
    w = np.sort((4 - 2 * np.cos(i * np.pi / (nx + 1)) - 2 * np.cos(j * np.pi / (ny + 1))).ravel())
    assert np.allclose(vals, w[-40:], rtol=1e-10, atol=0)
    for q in range(40):
        assert E.computeError(q) <= 1e-9
