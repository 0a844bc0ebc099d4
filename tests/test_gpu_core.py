"""GPU parity tests: HIP SpMV / BV kernels / Krylov-Schur against the CPU oracle
and the reference's golden outputs.  All calls go through the C ABI."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))
LIBRES = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "library_results.json")))


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


# ------------------------------------------------------------------ SpMV
@pytest.mark.parametrize("gen", ["lap2d", "convdiff3d", "stencil27", "markov"])
def test_spmv_fp64_matches_oracle(sa, orc, gen):
    A = {"lap2d": lambda m: m.lap2d(37, 23), "convdiff3d": lambda m: m.convdiff3d(11, 9, 7),
         "stencil27": lambda m: m.stencil27(9, 8, 7), "markov": lambda m: m.markov(20)}[gen]
    Ah, Ao = A(sa), A(orc)
    M = sa.Mat(Ah, sa.FP64)
    # integer-valued x on an integer stencil: products and sums are exact
    xi = np.arange(Ah.n, dtype=np.float64) % 7 - 3
    y = M.mult_host(xi)
    yo = Ao.matvec(xi)
    if gen == "lap2d":
        assert np.array_equal(y, yo)
    x = orc.rand_uniform(3, Ah.n)
    y, yo = M.mult_host(x), Ao.matvec(x)
    # one rounding per product and per addition: |err| <= nnz_row * eps * sum|a||x|
    absA = np.abs(Ao.to_dense()) if Ah.n <= 2000 else None
    bound = 32 * 2.2e-16 * (absA @ np.abs(x) if absA is not None else np.abs(yo).max() + 1)
    assert np.all(np.abs(y - yo) <= bound + 1e-300)


def _ragged(sa, n, lens, seed):
    """CSR with the given row lengths, random ascending columns and small
    integer values (so that every product and partial sum is exact)."""
    rng = np.random.default_rng(seed)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=int(k), replace=False)) for k in lens]).astype(np.int32)
    val = rng.integers(-3, 4, size=int(rp[-1])).astype(np.float64)
    return sa.CSR(n, rp, col, val)


def _ragged_fast(sa, n, lens, seed):
    """As _ragged for large n: ascending columns from random gaps after a
    random start (vectorised)."""
    rng = np.random.default_rng(seed)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    gap = max(2, n // (4 * int(max(lens))))
    steps = np.cumsum(rng.integers(1, gap, nnz))
    within = steps - np.repeat(steps[rp[:-1]], lens)
    start = rng.integers(0, max(1, n - int(max(lens)) * gap), n)
    col = (np.repeat(start, lens) + within).astype(np.int32)
    val = rng.integers(-3, 4, size=nnz).astype(np.float64)
    return sa.CSR(n, rp.astype(np.int32), col, val)


def _stream_rows(rp, n):
    """The row block the library picks for the stream SpMV (256, 128, 64 or
    32 rows of at most 2,048 nonzeros, the largest giving 1,024 blocks; 0 when
    none fits)."""
    best = 0
    for rb in (256, 128, 64, 32):
        idx = np.minimum(np.arange(0, n + rb, rb), n)
        if np.max(rp[idx[1:]] - rp[idx[:-1]]) > 2048:
            continue
        best = rb
        if (n + rb - 1) // rb >= 1024:
            break
    return best


@pytest.fixture
def stream_layout(sa):
    """Every operator created in the test takes the stream layout when a row
    block fits (value/index threshold 0)."""
    old = sa.set_spmv_stream_bytes(0)
    yield
    sa.set_spmv_stream_bytes(old)


# Row-length patterns that select each block size of the stream SpMV (at most
# 2,048 nonzeros per block of 256/128/64/32 rows) and the vector-CSR kernel
# (no block size fits), with empty rows and sizes not divisible by the block.
# "small" takes 32-row blocks (too few rows for larger ones to fill the GPU);
# "large" takes the block size named, over thousands of blocks.  Integer
# data: every sum is exact.
STREAM_CASES = {"rows256": ((0, 8), 1001, 2_000_003, 256), "rows128": ((9, 15), 777, 600_001, 128),
                "rows64": ((20, 30), 555, 300_007, 64), "rows32": ((40, 60), 333, 150_001, 32),
                "vector": ((100, 120), 301, 301, 0), "one_long_row": ((0, 5), 2500, 2500, 0)}


# the vector-CSR cases have one size
CASE_SIZES = [(c, s) for c, v in STREAM_CASES.items() for s in (("small", "large") if v[1] != v[2] else ("small",))]


@pytest.mark.parametrize("case,size", CASE_SIZES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spmv_row_blocks_exact(sa, stream_layout, case, prec, size):
    (lo, hi), n_small, n_large, rb_large = STREAM_CASES[case]
    n = n_small if size == "small" else n_large
    rng = np.random.default_rng(11)
    lens = rng.integers(lo, hi + 1, size=n)
    if case == "one_long_row":
        lens[700] = 2100                  # no block of any size holds this row
    Ah = _ragged(sa, n, lens, 5) if n < 5000 else _ragged_fast(sa, n, lens, 5)
    rb = _stream_rows(Ah.rowptr, n)
    assert rb == (rb_large if size == "large" else (32 if rb_large else 0))
    x = (np.arange(n) % 7 - 3).astype(np.float64)
    row = np.repeat(np.arange(n), lens)
    ref = np.bincount(row, weights=Ah.val * x[Ah.col], minlength=n)
    M = sa.Mat(Ah, sa.FP64 if prec == "fp64" else sa.FP32)
    y = M.mult_host(x if prec == "fp64" else x.astype(np.float32)).astype(np.float64)
    assert np.array_equal(y, ref)


def test_spmv_fp32(sa, orc):
    Ah = sa.stencil27(16, 15, 14, seed=1)
    Ao = orc.stencil27(16, 15, 14, seed=1)
    M = sa.Mat(Ah, sa.FP32)
    x = orc.rand_uniform(4, Ah.n).astype(np.float32)
    y = M.mult_host(x).astype(np.float64)
    yo = Ao.matvec(x.astype(np.float64))
    assert np.max(np.abs(y - yo)) <= 27 * 6e-8 * 60


# ------------------------------------------------------------------ BV golden (test1.c)
def test_bv_test1_golden(sa):
    g = GOLD["bv_test1"]
    X0 = np.array(g["X0"])
    Ymult = np.array(g["Y_after_mult"])
    for prec in (sa.FP64, sa.FP32):
        X = sa.BV(10, 5, prec)
        X.set_matrix(X0)
        # BVMultVec(X,-1,1,Y[0],z), z = 2, -1, .5, -.25, .125 (test1.c:104-110)
        z = np.array([2.0, -1.0, 0.5, -0.25, 0.125])
        Y = sa.BV(10, 1, prec)
        Y.set_column(0, Ymult[:, 0])
        X.multvec(-1.0, 1.0, Y.column_ptr(0), z)
        y0 = Y.get_column(0)
        assert np.array_equal(y0, np.array(g["Y_after_multvec"])[:, 0])
        # BVDotVec(X, Y[0]) (test1.c:130-134)
        q = X.dotvec(Y.column_ptr(0))
        assert np.array_equal(q, np.array(g["z_after_dotvec"]))
        # BVMultInPlace(X,Q,1,l=3) then BVScale(X,2) (test1.c:146-147)
        Q = np.array(g["Q"])
        X.multinplace(Q, 1, 3)
        X.scale_column(-1, 2.0)
        assert np.array_equal(X.get_matrix(5), np.array(g["X_after_multinplace_scale"]))
        assert abs(X.norm_column(0) - g["norm2_X0"]) < 5e-5
        assert abs(X.norm() - g["normF_X"]) < 5e-5


def test_bv_multvec_beta_zero_does_not_read_y(sa):
    """BVMultVec with beta = 0 follows gemv (bvblas.c:56-67): y is not read, so
    NaN in it cannot reach the result."""
    n, k = 5000, 6
    V = sa.BV(n, k)
    X = np.random.default_rng(2).standard_normal((n, k))
    V.set_matrix(X)
    V.set_active_columns(0, k)
    q = np.arange(1.0, k + 1)
    y = sa.DeviceArray.from_host(np.full(n, np.nan))
    V.multvec(2.0, 0.0, y.ptr, q)
    got = y.to_host(np.float64, n)
    assert np.allclose(got, 2.0 * X @ q, rtol=1e-13, atol=1e-12)


def test_bv_orthogonality_level(sa):
    """test2.c criterion: CGS with refinement gives ||X^T X - I||_1 < 100 eps."""
    n, k = 20, 8
    rng = np.random.default_rng(7)
    X0 = rng.standard_normal((n, k))
    X = sa.BV(n, k, sa.FP64)
    X.set_matrix(X0)
    for j in range(k):
        X.orthonormalize_column(j)
    Q = X.get_matrix(k)
    assert np.linalg.norm(Q.T @ Q - np.eye(k), 1) < 100 * 2.2e-16
    Qo = np.linalg.qr(X0)[0]
    assert np.linalg.norm(np.abs(Q.T @ Qo) - np.eye(k)) < 1e-12


# ------------------------------------------------------------------ Gram-Schmidt vs oracle
@pytest.mark.parametrize("n", [1000, 70001])
def test_orthonormalize_matches_oracle(sa, orc, n):
    """Well-conditioned columns: GPU CGS/DGKS equals the oracle to rounding."""
    k = 12
    rng = np.random.default_rng(11)
    X0 = rng.standard_normal((n, k))
    V = sa.BV(n, k, sa.FP64)
    V.set_matrix(X0)
    Vo = X0.copy(order="F")
    for j in range(k):
        nrm, lind = V.orthonormalize_column(j)
        Vo, nrmo, lindo, ho = orc.bv_orthonormalize_column(Vo, j)
        assert lind == bool(lindo)
        assert abs(nrm - nrmo) <= 1e-12 * abs(nrmo)
        h = V.coefficients(j, j + 1)[:j + 1, 0]
        assert np.max(np.abs(h - ho)) <= 1e-12 * np.abs(ho).max()
    Q = V.get_matrix(k)
    assert np.max(np.abs(Q - Vo)) < 1e-12
    assert np.linalg.norm(Q.T @ Q - np.eye(k), 1) < 100 * 2.2e-16


def test_orthonormalize_dgks_refinement(sa, orc):
    """A nearly dependent column triggers the DGKS second pass; the result is
    orthogonal to working precision and lindep agrees with the oracle."""
    n, k = 5000, 8
    rng = np.random.default_rng(5)
    X0 = rng.standard_normal((n, k))
    X0[:, 5] = X0[:, 4] + 1e-9 * X0[:, 5]
    X0[:, 7] = X0[:, 0] * 2.0                    # exactly dependent: lindep
    V = sa.BV(n, k, sa.FP64)
    V.set_matrix(X0)
    Vo = X0.copy(order="F")
    scale = np.linalg.norm(X0, axis=0)
    for j in range(k):
        nrm, lind = V.orthonormalize_column(j)
        Vo, nrmo, lindo, _ = orc.bv_orthonormalize_column(Vo, j)
        assert lind == bool(lindo), j
        if j == 7:      # exact dependence: only rounding noise is left
            assert nrm < 1e-13 * scale[j] and nrmo < 1e-13 * scale[j]
        elif not lind:
            assert abs(nrm - nrmo) <= 1e-6 * abs(nrmo)
    Q = V.get_matrix(7)
    assert np.linalg.norm(Q.T @ Q - np.eye(7), 1) < 100 * 2.2e-16


def test_arnoldi_relation(sa, orc):
    """BVMatArnoldi: A V_m - V_m H_m = beta v_{m+1} e_m^T (bvkrylov.c:36-37)."""
    Ah = sa.convdiff3d(10, 9, 8)
    A = sa.Mat(Ah)
    m = 30
    V = sa.BV(Ah.n, m + 1, sa.FP64)
    V.set_column(0, orc.rand_uniform(0, Ah.n))
    V.orthonormalize_column(0)
    mm, beta, bd = V.mat_arnoldi(A, 0, m)
    assert mm == m and not bd
    Vm = V.get_matrix(m + 1)
    H = V.coefficients(1, m + 1)            # column c holds H(:, c-1)
    Ad = orc.convdiff3d(10, 9, 8).to_dense()
    R = Ad @ Vm[:, :m] - Vm[:, :m + 1] @ np.vstack([H[:m + 1, :m]])
    assert np.linalg.norm(R) < 1e-11 * np.linalg.norm(Ad)
    assert abs(H[m, m - 1] - beta) < 1e-14 * beta
    assert np.linalg.norm(Vm.T @ Vm - np.eye(m + 1)) < 1e-13


# Shapes: ragged n/K/N in place (Q staged in LDS); more 64-row blocks than
# workgroups (grid-stride loop); 64x64 (4 output tiles); a Q slice too large
# for LDS (the tile-in-LDS kernel); K = 3, N = 1; several blocks per
# workgroup of two k batches each (the load pipeline crosses blocks); a
# basis above 256 MB in FP64 (non-temporal loads).
@pytest.mark.parametrize("shape", [(5000, 41, 3, 37, 5, 29), (70001, 20, 0, 20, 0, 10), (1000, 70, 0, 64, 0, 64),
                                   (300, 127, 2, 122, 10, 110), (777, 10, 0, 3, 0, 1),
                                   (300001, 45, 2, 42, 3, 23), (2_100_001, 21, 0, 21, 0, 9)])
@pytest.mark.parametrize("prec", [0, 1])
def test_multinplace_matches_oracle(sa, orc, prec, shape):
    n, m, l, k, s0, e0 = shape
    rng = np.random.default_rng(2)
    X0 = rng.standard_normal((n, m))
    Q = rng.standard_normal((m, m))
    V = sa.BV(n, m, prec)
    V.set_matrix(X0)
    V.set_active_columns(l, k)
    V.multinplace(Q, s0, e0)
    out = V.get_matrix(m)
    ref = orc.bv_multinplace(X0 if prec == 0 else X0.astype(np.float32).astype(np.float64), Q, s0, e0, l=l, k=k)
    tol = 1e-12 if prec == 0 else 2e-5
    assert np.max(np.abs(out - ref)) <= tol * np.abs(ref).max()


# ------------------------------------------------------------------ full solves
def _solve(sa, Ah, nev, ncv, tol, which, herm, v0=None, mixed=False, A32=None):
    A = sa.Mat(Ah, sa.FP64)
    E = sa.EPS()
    E.setOperators(A, A32)
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    E.setDimensions(nev, ncv)
    E.setTolerances(tol)
    if v0 is not None:
        E.setInitialSpace(v0)
    if mixed:
        E.setMixedPrecision(True, 1e-5)
    E.solve()
    return E, A


def test_ex2_golden(sa, orc):
    g = GOLD["ex2_1"]
    Ah = sa.lap2d(72)
    v0 = orc.rand_uniform(0, Ah.n)
    E, _ = _solve(sa, Ah, 4, 20, 1e-8, "LM", True, v0)
    vals = [E.getEigenpair(i, False)[0] for i in range(E.getConverged())]
    r = orc.eps_solve(orc.lap2d(72), 4, 20, 1e-8, "LM", True, v0=v0)
    assert E.getConverged() == r["nconv"]
    assert E.getIterationNumber() == r["its"]
    assert _rel(vals, r["eigr"]) < 1e-10
    assert np.allclose(np.round(vals[:4], 5), g["values"], atol=1e-12)
    for i in range(E.getConverged()):
        assert E.computeError(i) < 1e-8


def test_ex5_golden_nhep(sa, orc):
    g = GOLD["ex5_1"]
    Ah = sa.markov(15)
    v0 = np.zeros(Ah.n)
    v0[:3] = 1.0
    E, _ = _solve(sa, Ah, 4, 0, 1e-8, "LR", False, v0)
    r = orc.eps_solve(orc.markov(15), 4, 0, 1e-8, "LR", False, v0=v0)
    vals = [E.getEigenpair(i, False)[0] for i in range(E.getConverged())]
    assert E.getConverged() == r["nconv"]
    assert _rel(vals, r["eigr"]) < 1e-10
    assert np.allclose(np.round(vals[:4], 5), g["values"], atol=1e-12)
    for i in range(E.getConverged()):
        assert E.computeError(i) < 1e-8


def test_ex4_rdb200_golden(sa, orc):
    g = GOLD["ex4_1"]
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "rdb200.npz"))
    Ah = sa.CSR(int(d["n"]), d["rowptr"], d["col"], d["val"])
    v0 = orc.rand_uniform(0, Ah.n)
    E, _ = _solve(sa, Ah, 4, 0, 1e-8, "LM", False, v0)
    assert E.getIterationNumber() == g["its"]
    vals = [E.getEigenpair(i, False) for i in range(4)]
    re = [v[0] for v in vals]
    assert np.allclose(np.round(re, 5), g["values"], atol=1e-12)
    for i in range(E.getConverged()):
        assert E.computeError(i) < 1e-8


def test_c1_matches_oracle(sa, orc):
    """Workload 1: 100x100 Laplacian, nev 10, ncv 32, tol 1e-8, U[-1,1] seed 0.
    The spectrum is degenerate, so the restart count depends on rounding (the
    oracle itself changes with its thread count); parity is on the converged
    values (exact closed forms), the residuals and the invariant subspace of
    the simple eigenvalues."""
    Ah = sa.lap2d(100)
    v0 = orc.rand_uniform(0, Ah.n)
    E, _ = _solve(sa, Ah, 10, 32, 1e-8, "LM", True, v0)
    r = orc.eps_solve(orc.lap2d(100), 10, 32, 1e-8, "LM", True, v0=v0, nthreads=1)
    nc = E.getConverged()
    assert nc == r["nconv"]
    assert abs(E.getIterationNumber() - r["its"]) <= 0.3 * r["its"]
    vals, X = [], []
    for i in range(nc):
        re, im, xr, _ = E.getEigenpair(i)
        vals.append(re)
        X.append(xr)
        assert E.computeError(i) <= 1e-8
    n = 100
    ev = np.array([4 - 2 * np.cos(i * np.pi / (n + 1)) - 2 * np.cos(j * np.pi / (n + 1))
                   for i in range(1, n + 1) for j in range(1, n + 1)])
    for v in vals:
        assert np.min(np.abs(ev - v)) < 1e-10 * v
    assert np.allclose(np.unique(np.round(vals, 8)), np.unique(np.round(r["eigr"], 8)), rtol=0, atol=1e-12)
    # simple eigenvalues: eigenvectors agree up to sign (angle < 1e-6)
    X = np.array(X).T
    for i in range(nc):
        if np.min(np.abs(np.delete(ev, np.argmin(np.abs(ev - vals[i]))) - vals[i])) > 1e-6:
            j = int(np.argmin(np.abs(r["eigr"] - vals[i])))
            c = abs(X[:, i] @ r["vecs"][:, j])
            assert np.arccos(min(1.0, c)) < 1e-6


def test_convdiff_nhep_library(sa, orc):
    """Nonsymmetric 7-point convection-diffusion at 12^3, 20 largest real,
    ncv 64, tol 1e-10 against dense LAPACK eigenvalues: the spectrum has no
    multiple eigenvalues (the closed form 6 + 2 sum_d s_d cos(k_d pi/13) has
    its 21 largest values at least 1.6e-4 apart), so the 20 computed values,
    sorted, are the 20 largest in order, and every returned pair meets the
    tolerance with the explicit-residual convergence test."""
    Ah = sa.convdiff3d(12)
    v0 = orc.rand_uniform(0, Ah.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(False)
    E.setWhichEigenpairs("LR")
    E.setDimensions(20, 64)
    E.setTolerances(1e-10, 0)
    E.setTrueResidual(True)
    E.setInitialSpace(v0)
    E.solve()
    nc = E.getConverged()
    assert nc >= 20
    vals = np.sort([E.getEigenpair(i, False)[0] for i in range(20)])[::-1]
    ref = np.array(LIBRES["convdiff12"]["largest_real20"])
    assert np.max(np.abs(vals - ref) / ref) < 1e-12
    for i in range(nc):
        assert E.computeError(i) <= 1e-10


def test_mixed_precision_c1(sa, orc):
    Ah = sa.lap2d(100)
    v0 = orc.rand_uniform(0, Ah.n)
    A32 = sa.Mat(Ah, sa.FP32)
    E, _ = _solve(sa, Ah, 10, 32, 1e-10, "LM", True, v0, mixed=True, A32=A32)
    r = orc.eps_solve(orc.lap2d(100), 10, 32, 1e-10, "LM", True, v0=v0)
    nc = E.getConverged()
    assert nc >= 10
    st = E.getStats()
    assert st["matvecs32"] > 0
    n = 100
    ev = np.array([4 - 2 * np.cos(i * np.pi / (n + 1)) - 2 * np.cos(j * np.pi / (n + 1))
                   for i in range(1, n + 1) for j in range(1, n + 1)])
    for i in range(nc):
        v = E.getEigenpair(i, False)[0]
        assert np.min(np.abs(ev - v)) < 1e-10 * v
        assert E.computeError(i) <= 1e-10


def test_concurrent_solves_match_sequential(sa):
    """Solves running at once on several host threads (one HIP stream each,
    sharing the operators) give the same eigenvalues and iteration counts as
    the same solves one after the other."""
    import concurrent.futures as cf
    probs = [sa.Mat(sa.fv2d(48, 48, sa.fv_theta(5, s))) for s in range(6)]
    n = 48 * 48
    v0 = sa.rand_uniform(0, n)

    def solve(i):
        E = sa.EPS()
        E.setOperators(probs[i])
        E.setProblemType(True)
        E.setWhichEigenpairs("LR")
        E.setDimensions(8, 32)
        E.setTolerances(1e-10)
        E.setInitialSpace(v0)
        E.solve()
        vals = [E.getEigenpair(j, vectors=False)[0] for j in range(E.getConverged())]
        errs = [E.computeError(j) for j in range(min(8, E.getConverged()))]
        return vals, E.getIterationNumber(), errs

    seq = [solve(i) for i in range(len(probs))]
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        conc = list(ex.map(solve, range(len(probs))))
    for (v1, it1, e1), (v2, it2, e2) in zip(seq, conc):
        assert it1 == it2
        assert np.array_equal(v1, v2)          # same kernels, same order: identical bits
        assert max(e2) <= 1e-10


def test_concurrency_hint_partition(sa, orc):
    """sa_set_concurrent_solves(4): bases created afterwards use two 64-row
    chunks per Gram-Schmidt workgroup.  Concurrent solves under the hint give
    the oracle's eigenvalues (to 1e-10 relative) with every pair at tol, and
    the same bits as the hinted solves run one after the other."""
    import concurrent.futures as cf
    g = 96
    hs = [sa.fv2d(g, g, sa.fv_theta(5, s)) for s in range(4)]
    probs = [sa.Mat(h) for h in hs]
    n = g * g
    v0 = sa.rand_uniform(0, n)

    def solve(i):
        E = sa.EPS()
        E.setOperators(probs[i])
        E.setProblemType(True)
        E.setWhichEigenpairs("LR")
        E.setDimensions(8, 32)
        E.setTolerances(1e-10, 20000)
        E.setTrueResidual(True)
        E.setInitialSpace(v0)
        E.solve()
        vals = [E.getEigenpair(j, vectors=False)[0] for j in range(E.getConverged())]
        errs = [E.computeError(j) for j in range(min(8, E.getConverged()))]
        return vals, E.getIterationNumber(), errs

    prev = sa.set_concurrent_solves(4)
    try:
        seq = [solve(i) for i in range(len(probs))]
        with cf.ThreadPoolExecutor(max_workers=4) as ex:
            conc = list(ex.map(solve, range(len(probs))))
    finally:
        sa.set_concurrent_solves(prev)
    for i, ((v1, it1, e1), (v2, it2, e2)) in enumerate(zip(seq, conc)):
        assert it1 == it2 and np.array_equal(v1, v2)
        assert len(e2) == 8 and max(e2) <= 1e-10
        h = hs[i]
        ref = orc.eps_solve(orc.CSR(h.n, h.rowptr, h.col, h.val), 8, 32, 1e-10, "LR", True, v0=v0, nvec=0,
                            trueres=True, max_it=20000)
        r = np.sort(ref["eigr"][:8])[::-1]
        assert np.allclose(np.sort(v2[:8])[::-1], r, rtol=1e-10, atol=0)


# ------------------------------------------------------------------ EPSARNOLDI
@pytest.mark.parametrize("solver,case", [("arnoldi", "test2_LR"), ("arnoldi", "test2_SR"), ("arnoldi", "markov15"),
                                         ("arnoldi", "convdiff12"), ("lanczos", "test2_LR"), ("lanczos", "test2_SR"),
                                         ("lanczos", "stencil27_10")])
def test_explicit_restart_matches_oracle(sa, orc, solver, case):
    """Explicitly restarted Arnoldi (EPSARNOLDI) and Lanczos with full
    reorthogonalization (EPSLANCZOS) through the C ABI against the oracle's
    restatements from the same start vector: eigenvalues, restart count and
    residuals.  test2: 1-D Laplacian n = 30 (values of test2_1.out);
    markov15: NHEP, ncv 8; convdiff12 (NHEP, 20 largest real, ncv 40) and
    stencil27_10 (8 largest magnitude, ncv 24) are long explicit-restart
    trajectories: values and residuals only."""
    from test_cpu_oracle import _lap1d, TEST2_LARGEST, TEST2_SMALLEST
    if case.startswith("test2"):
        Ao, which, herm, nev, ncv, tol, mi = _lap1d(orc, 30), case[-2:], True, 4, 0, 1e-9, 0
    elif case == "markov15":
        Ao, which, herm, nev, ncv, tol, mi = orc.markov(15), "LR", False, 4, 8, 1e-8, 300
    elif case == "stencil27_10":
        Ao, which, herm, nev, ncv, tol, mi = orc.stencil27(10, seed=1), "LM", True, 8, 24, 1e-10, 2000
    else:
        Ao, which, herm, nev, ncv, tol, mi = orc.convdiff3d(12), "LR", False, 20, 40, 1e-10, 3000
    Ah = sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)
    v0 = orc.rand_uniform(0, Ao.n)
    A = sa.Mat(Ah)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(A)
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    E.setDimensions(nev, ncv if ncv else 0)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, ncv, tol, which, herm, v0=v0, solver=solver, max_it=mi)
    nc = E.getConverged()
    assert nc >= nev and r["nconv"] >= nev
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(nev)])
    assert np.max(np.abs(vals - r["eigr"][:nev]) / np.abs(r["eigr"][:nev])) < 1e-9
    if case == "test2_LR":
        assert list(np.round(vals, 5)) == TEST2_LARGEST
    if case == "test2_SR":
        assert list(np.round(vals, 5)) == TEST2_SMALLEST
    # the short trajectories restart exactly as the oracle's; the ~600-restart
    # convdiff12 run crawls through near-stagnation, where rounding decides
    # when each pair converges, so only its values and residuals are compared
    if case not in ("convdiff12", "stencil27_10"):
        assert abs(E.getIterationNumber() - r["its"]) <= 1
    for i in range(nev):
        assert E.computeError(i) < 10 * tol


@pytest.mark.parametrize("reorthog", ["local", "periodic", "partial", "selective"])
@pytest.mark.parametrize("case", ["test2_LR", "test2_SR", "test2_interior", "test8_1", "test8_2", "test8_3",
                                  "stencil27_10"])
def test_lanczos_local_matches_oracle(sa, orc, case, reorthog):
    """EPSLanczos with local reorthogonalization (MGS against the locked
    columns and the two previous Lanczos vectors, purging of repeated and
    spurious Ritz values), with periodic / partial reorthogonalization
    (local steps plus MGS against the sweep's columns when the omega
    estimate calls for it) and with selective reorthogonalization (against
    the sweep's nearly converged Ritz vectors) through the C ABI against the oracle from the same
    start vector: eigenvalues to 1e-9 and the values printed by the
    reference.  test2: 1-D Laplacian n = 30 (test2_1.out, the interior case
    with shift-and-invert at 2.1, where ghost copies of the dominant value
    are purged); test8_*: 2-D Laplacian (test8_1.out .. test8_3.out, double
    eigenvalues included); stencil27_10: 8 largest magnitude, ncv 24.
    Without orthogonality the restart counts depend on rounding, so they are
    not compared."""
    from test_cpu_oracle import _lap1d, TEST2_LARGEST, TEST2_SMALLEST, TEST2_INTERIOR, TEST8
    sinv, gold, mpd, v0 = False, None, 0, None
    if case.startswith("test2"):
        Ao, nev, ncv, tol, mi = _lap1d(orc, 30), 4, 0, 1e-9, 0
        which = {"test2_LR": "LR", "test2_SR": "SR", "test2_interior": "TM"}[case]
        gold = {"LR": TEST2_LARGEST, "SR": TEST2_SMALLEST, "TM": TEST2_INTERIOR[30]}[which]
        sinv = which == "TM"
        v0 = orc.rand_uniform(0, Ao.n)
    elif case.startswith("test8"):
        m, nev, ncv, mpd, which, mi, gold = TEST8[case[-1]]
        Ao, tol = orc.lap2d(m), 1e-8
    else:
        Ao, which, nev, ncv, tol, mi = orc.stencil27(10, seed=1), "LM", 8, 24, 1e-10, 2000
        v0 = orc.rand_uniform(0, Ao.n)
    if v0 is None:
        v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType("lanczos")
    with pytest.raises(Exception):   # not an EPSLanczosReorthogType
        sa._lib.check(sa._lib.lib().sa_eps_lanczos_set_reorthog(E.h, 6))
    E.setLanczosReorthog(reorthog)
    assert E.getLanczosReorthog() == reorthog
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(True)
    E.setWhichEigenpairs(which)
    if sinv:
        E.setTarget(2.1)
        E.setST("sinvert")
        E.setSTKSP("jacobi", 1e-13, 20000)
    E.setDimensions(nev, ncv, mpd)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    E.solve()
    kw = dict(target=2.1, st="sinvert") if sinv else {}
    r = orc.eps_solve(Ao, nev, ncv, tol, which, True, v0=v0, solver="lanczos", reorthog=reorthog, max_it=mi,
                      mpd=mpd, **kw)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(nev)])
    assert np.max(np.abs(vals - r["eigr"][:nev]) / np.abs(r["eigr"][:nev])) < 1e-9
    if gold is not None:
        assert np.allclose(vals, gold, atol=5e-6)
    bound = 10 * tol * (max(1.0, 4.2 / np.min(np.abs(vals))) if sinv else 1.0)
    for i in range(nev):
        assert E.computeError(i) < bound


@pytest.mark.parametrize("case", ["test2_LR", "test2_interior", "test8_2", "test3_tridiag"])
def test_lanczos_delayed_matches_oracle(sa, orc, case):
    """EPSLanczos with delayed reorthogonalization through the C ABI against
    the oracle from the same start vector, on the cases where the delayed
    scheme converges (see test_oracle_lanczos_delayed): eigenvalues to 1e-9
    and the printed / dense values."""
    from test_cpu_oracle import _lap1d, _tridiag_random, TEST2_LARGEST, TEST2_INTERIOR, TEST8
    sinv, mpd = False, 0
    if case.startswith("test2"):
        Ao, nev, ncv, tol, mi = _lap1d(orc, 30), 4, 0, 1e-9, 0
        which = "LR" if case == "test2_LR" else "TM"
        sinv = which == "TM"
        gold = TEST2_LARGEST if which == "LR" else TEST2_INTERIOR[30]
    elif case == "test8_2":
        m, nev, ncv, mpd, which, mi, gold = TEST8["2"]
        Ao, tol = orc.lap2d(m), 1e-8
    else:
        Ao, nev, ncv, tol, mi, which = _tridiag_random(orc), 4, 0, 1e-8, 500, "SR"
        gold = np.sort(np.linalg.eigvalsh(Ao.to_dense()))[:4]
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType("lanczos")
    E.setLanczosReorthog("delayed")
    assert E.getLanczosReorthog() == "delayed"
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(True)
    E.setWhichEigenpairs(which)
    if sinv:
        E.setTarget(2.1)
        E.setST("sinvert")
        E.setSTKSP("jacobi", 1e-13, 20000)
    E.setDimensions(nev, ncv, mpd)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    E.solve()
    kw = dict(target=2.1, st="sinvert") if sinv else {}
    r = orc.eps_solve(Ao, nev, ncv, tol, which, True, v0=v0, solver="lanczos", reorthog="delayed", max_it=mi,
                      mpd=mpd, **kw)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(nev)])
    assert np.max(np.abs(vals - r["eigr"][:nev]) / np.abs(r["eigr"][:nev])) < 1e-9
    assert np.allclose(vals, gold, atol=5e-6)


@pytest.mark.parametrize("keep", [0.5, 0.2])
@pytest.mark.parametrize("lock", [True, False])
@pytest.mark.parametrize("case", ["test2_LR", "test2_interior", "test9_1"])
def test_krylovschur_locking_restart_matches_oracle(sa, orc, case, lock, keep):
    """EPSKrylovSchurSetLocking / SetRestart through the C ABI against the
    oracle from the same start vector: eigenvalues to 1e-9, the printed
    values (test2_1.out, test9_1.out) and the restart count."""
    from test_cpu_oracle import _lap1d, TEST2_LARGEST, TEST2_INTERIOR
    harm = False
    if case.startswith("test2"):
        Ao, nev, tol, mi, herm = _lap1d(orc, 30), 4, 1e-9, 0, True
        which = "LR" if case == "test2_LR" else "TM"
        harm = which == "TM"
        gold = TEST2_LARGEST if which == "LR" else TEST2_INTERIOR[30]
    else:
        Ao, nev, tol, mi, herm, which = orc.markov(15), 4, 1e-8, 1500, False, "LM"
        gold = [1.0, -1.0, 0.97137, -0.97137]
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setKrylovSchurLocking(lock)
    E.setKrylovSchurRestart(keep)
    assert E.getKrylovSchurLocking() == lock and abs(E.getKrylovSchurRestart() - keep) < 1e-15
    with pytest.raises(Exception):
        E.setKrylovSchurRestart(0.95)
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    if harm:
        E.setTarget(2.1)
        E.setExtraction("harmonic")
    E.setDimensions(nev, 0)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    E.solve()
    kw = dict(extraction="harmonic", target=2.1) if harm else {}
    r = orc.eps_solve(Ao, nev, 0, tol, which, herm, v0=v0, max_it=mi, lock=1 if lock else 0, keep=keep, **kw)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(nev)])
    ref = r["eigr"][:nev]
    if case == "test9_1":   # 1 and -1 tie in magnitude: their order depends on rounding
        vals, ref, gold = np.sort(vals), np.sort(ref), sorted(gold)
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    assert list(np.round(vals, 5)) == list(gold)
    assert abs(E.getIterationNumber() - r["its"]) <= 1


@pytest.mark.parametrize("lock", [True, False])
@pytest.mark.parametrize("balance", ["oneside", "twoside"])
def test_balance_matches_oracle(sa, orc, balance, lock):
    """EPSSetBalance through the C ABI (test9.c suffix 2: Markov m = 15,
    nev 4, max_it 1500) against the oracle's balancing: eigenvalues of
    test9_1.out (the +-1 tie as a set), the oracle's values to 10 tol, and
    eigenvectors of the original matrix (EPSComputeError with A, unit norm)."""
    Ao = orc.markov(15)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(False)
    E.setWhichEigenpairs("LM")
    E.setBalance(balance)
    E.setKrylovSchurLocking(lock)
    E.setDimensions(4, 0)
    E.setTolerances(1e-8, 1500)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, 4, 0, 1e-8, "LM", False, v0=v0, max_it=1500, balance=balance, lock=1 if lock else 0)
    assert E.getConverged() >= 4 and r["nconv"] >= 4
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(4)])
    # D is built from sums taken in another order than the oracle's, so the
    # balanced operators differ in rounding and the values agree to the
    # convergence tolerance (1e-8), not beyond
    assert np.max(np.abs(np.sort(vals) - np.sort(r["eigr"][:4]))) < 1e-7
    assert sorted(np.round(vals, 5)) == sorted([1.0, -1.0, 0.97137, -0.97137])
    assert abs(E.getIterationNumber() - r["its"]) <= 2
    for i in range(4):
        x = E.getEigenpair(i)[2]
        assert abs(np.linalg.norm(x) - 1.0) < 1e-12
        assert E.computeError(i) < 1e-7


@pytest.mark.parametrize("twoside", [False, True])
@pytest.mark.parametrize("case", ["markov15", "convdiff8"])
def test_balance_vector_matches_oracle(sa, orc, case, twoside):
    """The balancing diagonal D itself (sa_mat_build_balance,
    EPSBuildBalance_Krylov epsdefault.c:443-504): the device D equals the
    oracle's from the same +-1 vectors after 5 iterations (to rounding of the
    SpMV sums), and after one iteration equals the closed form 1/|A z|
    (oneside) or sqrt(|A^T z| / |A z|) (twoside).  B = D A D^-1 is similar to
    A for any D, so the solve's eigenvalues alone cannot catch a wrong D."""
    Ao = orc.markov(15) if case == "markov15" else orc.convdiff3d(8)
    M = sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val))
    D = M.build_balance(twoside, its=5)
    Do = orc.build_balance(Ao, twoside, its=5)
    good = np.abs(Do) < 1e8                     # rows whose |A z| cancelled to rounding are left out
    assert good.sum() > 0.8 * Ao.n
    assert np.allclose(D[good], Do[good], rtol=1e-11, atol=0)
    D1 = M.build_balance(twoside, its=1)
    Ad = Ao.to_dense()
    u = orc.rand_uniform(0xBA1A, Ao.n)
    z = np.where(0.5 * (u + 1.0) < 0.5, -1.0, 1.0)
    p = np.abs(Ad @ z)
    rt = np.abs(Ad.T @ z)
    ok = (p > 1e-8 * p.max()) & ((rt > 1e-8 * rt.max()) if twoside else True)
    ref = np.sqrt(rt / np.where(ok, p, 1.0)) if twoside else 1.0 / np.where(ok, p, 1.0)
    assert ok.sum() > 0.8 * Ao.n
    assert np.allclose(D1[ok], ref[ok], rtol=1e-13, atol=0)


def test_arnoldi_delayed_matches_oracle(sa, orc):
    """EPSArnoldi with delayed reorthogonalization (test9.c suffix 3: Markov
    m = 15, 3 largest real, tol 1e-7) through the C ABI against the oracle:
    the values of test9_3.out in order and the restart count."""
    Ao = orc.markov(15)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType("arnoldi")
    E.setArnoldiDelayed(True)
    assert E.getArnoldiDelayed()
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(False)
    E.setWhichEigenpairs("LR")
    E.setDimensions(3, 0)
    E.setTolerances(1e-7, 0)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, 3, 0, 1e-7, "LR", False, v0=v0, solver="arnoldi", reorthog="delayed")
    assert E.getConverged() >= 3
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(3)])
    assert list(np.round(vals, 5)) == [1.0, 0.97137, 0.90423]
    assert np.max(np.abs(vals - r["eigr"][:3])) < 1e-9
    assert abs(E.getIterationNumber() - r["its"]) <= 1


@pytest.mark.parametrize("solver,case", [("krylovschur", "test2_30"), ("krylovschur", "test2_32"),
                                         ("arnoldi", "test2_30"), ("arnoldi", "test2_32"),
                                         ("krylovschur", "convdiff8")])
def test_harmonic_matches_oracle(sa, orc, solver, case):
    """Harmonic extraction (EPS_HARMONIC) with which = target magnitude
    through the C ABI against the oracle from the same start vector.  test2:
    interior eigenvalues of the 1-D Laplacian at target 2.1 (test2_1.out for
    n = 30, test2_3.out for n = 32; printed values and order, restart counts
    equal); convdiff8: the non-symmetric 7-point operator on 8^3, 6 values
    closest to 3.0 (values and residuals)."""
    from test_cpu_oracle import _lap1d, TEST2_INTERIOR
    if case.startswith("test2"):
        nn = int(case[-2:])
        Ao, herm, nev, ncv, tol, mi, tgt = _lap1d(orc, nn), True, 4, 0, 1e-9, 0, 2.1
    else:
        Ao, herm, nev, ncv, tol, mi, tgt = orc.convdiff3d(8), False, 6, 24, 1e-10, 500, 3.0
    Ah = sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)
    v0 = orc.rand_uniform(0, Ao.n)
    A = sa.Mat(Ah)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(A)
    E.setProblemType(herm)
    E.setWhichEigenpairs("TM")
    E.setTarget(tgt)
    E.setExtraction("harmonic")
    E.setDimensions(nev, ncv if ncv else 0)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, ncv, tol, "TM", herm, v0=v0, solver=solver, max_it=mi, extraction="harmonic",
                      target=tgt)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    if case.startswith("test2"):
        assert np.allclose(vals.real, TEST2_INTERIOR[nn], atol=5e-6)
        assert abs(E.getIterationNumber() - r["its"]) <= 1
    # convdiff8's targets sit in a cluster of close eigenvalues: the restart
    # count follows rounding (the oracle alone gives 165 or 191 restarts with
    # different host thread counts), so only values and residuals are compared
    for i in range(nev):
        assert E.computeError(i) < 10 * tol


@pytest.mark.parametrize("case", ["markov15", "convdiff8", "convdiff12"])
def test_twosided_matches_oracle(sa, orc, case):
    """Two-sided Krylov-Schur through the C ABI against the oracle from the
    same start vectors: eigenvalues (ex5_1.out on the Markov model), restart
    count, right and left eigenvectors (up to sign), and EPSComputeError,
    which takes the larger of the right and left residuals."""
    if case == "markov15":
        Ao, nev, ncv, tol = orc.markov(15), 4, 0, 1e-8
    elif case == "convdiff8":
        Ao, nev, ncv, tol = orc.convdiff3d(8), 6, 24, 1e-10
    else:
        Ao, nev, ncv, tol = orc.convdiff3d(12), 8, 32, 1e-10
    Ah = sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(False)
    E.setWhichEigenpairs("LR")
    E.setDimensions(nev, ncv)
    E.setTolerances(tol, 0)
    E.setInitialSpace(v0)
    E.setTwoSided(True)
    E.solve()
    r = orc.eps_solve(Ao, nev, ncv, tol, "LR", False, v0=v0, twosided=True, nvec=nev)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(nev)])
    assert np.max(np.abs(vals - r["eigr"][:nev]) / np.abs(r["eigr"][:nev])) < 1e-9
    if case == "markov15":
        assert list(np.round(vals, 5)) == [1.0, 0.97137, 0.90423, 0.85714]
    if case != "convdiff12":
        assert abs(E.getIterationNumber() - r["its"]) <= 1
    for i in range(nev):
        x = E.getEigenpair(i)[2]
        y = E.getLeftEigenvector(i)[0]
        xo, yo = r["vecs"][:, i], r["lvecs"][:, i]
        assert min(np.abs(x - xo).max(), np.abs(x + xo).max()) < 1e-6
        assert min(np.abs(y - yo).max(), np.abs(y + yo).max()) < 1e-6
        assert E.computeError(i) < 10 * tol


def test_twosided_reused_over_operators_of_equal_size(sa, orc):
    """One two-sided EPS solved with one operator, then given another of the
    same size (created right after the first is destroyed, so it may reuse
    its address): the left eigenvectors and EPSComputeError (the larger of
    the right and left residuals) belong to the new operator, whose A^T is
    rebuilt, and match the oracle."""
    A1 = orc.convdiff3d(8)
    # the same size and pattern, other values: A2 = 1.5 A1^T + 0.3 I (a real
    # spectrum; its left eigenvectors are A1's right ones, so a left basis
    # swept with the stale A1^T would converge to other vectors)
    D2 = A1.to_dense().T * 1.5 + 0.3 * np.eye(A1.n)
    rows, cols = np.nonzero(D2)
    rp = np.zeros(A1.n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    A2 = orc.CSR(A1.n, np.cumsum(rp).astype(np.int32), cols.astype(np.int32), D2[rows, cols].copy())
    v0 = orc.rand_uniform(0, A1.n)
    E = sa.EPS()
    E.setProblemType(False)
    E.setWhichEigenpairs("LR")
    E.setDimensions(6, 24)
    E.setTolerances(1e-10, 0)
    E.setTwoSided(True)
    M1 = sa.Mat(sa.CSR(A1.n, A1.rowptr, A1.col, A1.val))
    E.setOperators(M1)
    E.setInitialSpace(v0)
    E.solve()
    assert E.getConverged() >= 6
    E.computeError(0)
    M1.destroy()
    M2 = sa.Mat(sa.CSR(A2.n, A2.rowptr, A2.col, A2.val))
    E.setOperators(M2)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(A2, 6, 24, 1e-10, "LR", False, v0=v0, twosided=True, nvec=6)
    assert E.getConverged() >= 6
    Dt = D2.T
    for i in range(6):
        lam, im, x, _ = E.getEigenpair(i)
        assert im == 0.0
        assert abs(lam - r["eigr"][i]) < 1e-9 * abs(lam)
        y = E.getLeftEigenvector(i)[0]
        assert np.linalg.norm(Dt @ y - lam * y) / abs(lam) < 1e-9
        assert E.computeError(i) < 1e-9


@pytest.mark.parametrize("solver,case", [("krylovschur", "markov15"), ("arnoldi", "markov15"),
                                         ("lanczos", "test2"), ("krylovschur", "test2"),
                                         ("krylovschur", "lap2d40")])
def test_sinvert_matches_oracle(sa, orc, solver, case):
    """Shift-and-invert through the C ABI (device BiCGStab with Jacobi, rtol
    1e-13) against the oracle's exact solves (dense LU, the reference's
    PREONLY + LU): eigenvalues in the same order to 1e-9, residuals.
    markov15: test9_5.out (target 0.5); test2: test2_1.out interior values
    (target 2.1, indefinite); lap2d40: smallest 6 of the 2-D Laplacian on a
    40 x 37 grid (target 0, SPD; a non-square grid, so that no eigenvalue
    is double: a single-vector Krylov method finds the second copy of a
    double eigenvalue only through rounding, which differs between exact
    and iterative solves)."""
    from test_cpu_oracle import _lap1d, TEST2_INTERIOR
    if case == "markov15":
        Ao, herm, nev, tol, tgt = orc.markov(15), False, 3, 1e-8, 0.5
    elif case == "test2":
        Ao, herm, nev, tol, tgt = _lap1d(orc, 30), True, 4, 1e-9, 2.1
    else:
        Ao, herm, nev, tol, tgt = orc.lap2d(40, 37), True, 6, 1e-10, 0.0
    Ah = sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(herm)
    E.setWhichEigenpairs("TM")
    E.setTarget(tgt)
    E.setST("sinvert")
    E.setSTKSP("jacobi", 1e-13, 20000)
    E.setDimensions(nev, 0)
    E.setTolerances(tol, 0)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, 0, tol, "TM", herm, v0=v0, target=tgt, st="sinvert", solver=solver)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    if case == "markov15":
        assert list(np.round(vals.real, 5)) == [0.49842, 0.51928, 0.47897]
    if case == "test2":
        assert np.allclose(vals.real, TEST2_INTERIOR[30], atol=5e-6)
    # the convergence test bounds the residual of the transformed problem:
    # ||A x - lambda x|| / |lambda| <= ||A - sigma I|| tol / |lambda|
    nrm = max(np.abs(Ao.val[Ao.rowptr[i]:Ao.rowptr[i + 1]]).sum() for i in range(Ao.n)) + abs(tgt)
    for i in range(nev):
        assert E.computeError(i) < 10 * tol * max(1.0, nrm / abs(vals[i]))


@pytest.mark.parametrize("solver,case", [("krylovschur", "markov15"), ("arnoldi", "markov15"),
                                         ("lanczos", "test2"), ("krylovschur", "lap2d40")])
def test_sinvert_gmres_matches_oracle(sa, orc, solver, case):
    """Shift-and-invert with GMRES(m) and Jacobi (KSPGMRES, PETSc's default
    KSP type: right preconditioning, classical Gram-Schmidt without
    refinement, rtol 1e-13) against the oracle's exact solves: eigenvalues in
    order to 1e-9 and residuals, as for BiCGStab.  Restarted GMRES can stall
    on an indefinite A - sigma I (a numpy GMRES(30) on the Markov matrix at
    0.5 stays at a relative residual of 0.56 from the second cycle on, while
    GMRES(100) converges in 184 iterations), so the Markov case runs m = 100
    and the 1-D Laplacian m = n = 30;
    the positive definite 2-D Laplacian takes the default m = 30."""
    from test_cpu_oracle import _lap1d, TEST2_INTERIOR
    if case == "markov15":
        Ao, herm, nev, tol, tgt, m = orc.markov(15), False, 3, 1e-8, 0.5, 100
    elif case == "test2":
        Ao, herm, nev, tol, tgt, m = _lap1d(orc, 30), True, 4, 1e-9, 2.1, 30
    else:
        Ao, herm, nev, tol, tgt, m = orc.lap2d(40, 37), True, 6, 1e-10, 0.0, 30
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(herm)
    E.setWhichEigenpairs("TM")
    E.setTarget(tgt)
    E.setST("sinvert")
    E.setSTKSP("jacobi", 1e-13, 20000)
    E.setSTKSPType("gmres", m)
    E.setDimensions(nev, 0)
    E.setTolerances(tol, 500)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, 0, tol, "TM", herm, v0=v0, target=tgt, st="sinvert", solver=solver, max_it=500)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    assert E.getStats()["ksp_its"] > 0
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    if case == "markov15":
        assert list(np.round(vals.real, 5)) == [0.49842, 0.51928, 0.47897]
    if case == "test2":
        assert np.allclose(vals.real, TEST2_INTERIOR[30], atol=5e-6)
    nrm = max(np.abs(Ao.val[Ao.rowptr[i]:Ao.rowptr[i + 1]]).sum() for i in range(Ao.n)) + abs(tgt)
    for i in range(nev):
        assert E.computeError(i) < 10 * tol * max(1.0, nrm / abs(vals[i]))


def test_sinvert_gmres_restart_and_arguments(sa, orc):
    """A shorter GMRES restart takes more iterations for the same values (the
    2-D Laplacian at target 0, 6 smallest); restart outside [1, 126] and an
    unknown KSP type are rejected."""
    Ao = orc.lap2d(40, 37)
    v0 = orc.rand_uniform(0, Ao.n)
    out = {}
    for restart in (5, 30):
        E = sa.EPS()
        E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
        E.setProblemType(True)
        E.setWhichEigenpairs("TM")
        E.setTarget(0.0)
        E.setST("sinvert")
        E.setSTKSP("jacobi", 1e-13, 50000)
        E.setSTKSPType("gmres", restart)
        E.setDimensions(6, 0)
        E.setTolerances(1e-10, 500)
        E.setInitialSpace(v0)
        E.solve()
        assert E.getConverged() >= 6
        out[restart] = (np.array([E.getEigenpair(i, False)[0] for i in range(6)]), E.getStats()["ksp_its"])
    assert np.allclose(out[5][0], out[30][0], rtol=1e-9, atol=0)
    assert out[5][1] > out[30][1] > 0
    E = sa.EPS()
    with pytest.raises(sa.SlepcAmdError, match="restart"):
        E.setSTKSPType("gmres", 127)
    with pytest.raises(sa.SlepcAmdError, match="restart"):
        E.setSTKSPType("gmres", 0)
    with pytest.raises(sa.SlepcAmdError, match="unsupported KSP type"):
        sa.check(sa.lib().sa_eps_set_st_ksp_type(E.h, 7, 30))


def test_sinvert_laplacian_smallest(sa, orc):
    """The 2-D Laplacian on a 100 x 97 grid (C1's stencil and size, with
    distinct eigenvalues) with shift-and-invert at target 0: the 10 smallest
    eigenvalues from the closed form 4 - 2 cos(i pi/101) - 2 cos(j pi/98),
    every residual within ||A|| tol / |lambda| (the convergence test bounds
    the residual of the transformed problem)."""
    Ah = sa.lap2d(100, 97)
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(True)
    E.setWhichEigenpairs("TM")
    E.setTarget(0.0)
    E.setST("sinvert")
    E.setSTKSP("jacobi", 1e-12, 20000)
    E.setDimensions(10, 32)
    E.setTolerances(1e-8, 0)
    E.setInitialSpace(orc.rand_uniform(0, Ah.n))
    E.solve()
    assert E.getConverged() >= 10
    kx, ky = np.arange(1, 101) * np.pi / 101, np.arange(1, 98) * np.pi / 98
    exact = np.sort((4 - 2 * np.cos(kx)[:, None] - 2 * np.cos(ky)[None, :]).ravel())[:10]
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(10)])
    assert np.allclose(np.sort(vals), exact, rtol=1e-8)
    for i in range(10):
        assert E.computeError(i) < 10 * 1e-8 * 8.0 / vals[i]


@pytest.mark.parametrize("variant", ["krylovschur", "twosided", "arnoldi", "harmonic", "sinvert"])
def test_complex_pairs_match_oracle(sa, orc, variant):
    """A sparse non-symmetric matrix whose wanted eigenvalues include conjugate
    pairs, through every non-symmetric path of the C ABI against the oracle
    from the same start vector: eigenvalues (real and imaginary parts, in
    order) and EPSComputeError of every pair (right and, two-sided, left
    residuals of the complex vectors)."""
    from test_cpu_oracle import _complex_csr
    Ao = _complex_csr(orc)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(False)
    kw = {}
    if variant in ("sinvert", "harmonic"):
        # shift-and-invert at 7.5, beyond the spectrum: A - 7.5 I is well
        # conditioned for BiCGStab (at 4.0, inside the spectrum, BiCGStab
        # stagnates on this matrix; the oracle's exact LU does not)
        nev, ncv, which, tgt = 4, 0, "TM", 7.5 if variant == "sinvert" else 6.5
        E.setWhichEigenpairs("TM")
        E.setTarget(tgt)
        kw = dict(target=tgt)
        if variant == "sinvert":
            E.setST("sinvert")
            E.setSTKSP("jacobi", 1e-13, 20000)
            kw["st"] = "sinvert"
        else:
            E.setExtraction("harmonic")
            kw["extraction"] = "harmonic"
    else:
        nev, ncv, which = 6, 24, "LM"
        E.setWhichEigenpairs("LM")
        if variant == "twosided":
            E.setTwoSided(True)
            kw = dict(twosided=True)
        if variant == "arnoldi":
            E.setType("arnoldi")
            kw = dict(solver="arnoldi")
    E.setDimensions(nev, ncv)
    E.setTolerances(1e-10, 500)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, ncv, 1e-10, which, False, v0=v0, max_it=500, **kw)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.any(ref.imag != 0.0)
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    for i in range(nev):
        assert E.computeError(i) < 1e-8


# ------------------------------------------------------------------ EPSSetTrueResidual
@pytest.mark.parametrize("case", ["test22", "markov15", "convdiff10", "lap2d_pot", "sinvert_markov", "arnoldi_markov",
                                  "twosided_convdiff8", "complex_arnoldi"])
def test_true_residual_matches_oracle(sa, orc, case):
    """EPSSetTrueResidual through the C ABI against the oracle's branch from the
    same start vector: eigenvalues in order, restart count, and error
    estimates (explicit residuals of the candidate Ritz vectors, formed on
    the device with the restart GEMM and the fused SpMV residual kernel);
    every returned pair meets the tolerance by EPSComputeError.  test22:
    the Brusselator of test22.c -eps_true_residual 1 (test22_1.out values,
    complex pairs)."""
    from test_cpu_oracle import _brusselator, _complex_csr, TEST22_VALUES
    kw, solver, herm, tgt = {}, "krylovschur", False, None
    if case == "test22":
        Ao, nev, ncv, tol, which = _brusselator(orc), 4, 0, 1e-8, "LR"
    elif case in ("markov15", "arnoldi_markov"):
        Ao, nev, ncv, tol, which = orc.markov(15), 4, 0, 1e-8, "LM"
        if case == "arnoldi_markov":
            solver = "arnoldi"
    elif case == "convdiff10":
        Ao, nev, ncv, tol, which = orc.convdiff3d(10), 8, 32, 1e-10, "LR"
    elif case == "lap2d_pot":
        Ao, nev, ncv, tol, which, herm = orc.lap2d_pot(30, 27, 2, 10.0), 8, 24, 1e-10, "LM", True
    elif case == "sinvert_markov":
        Ao, nev, ncv, tol, which, tgt = orc.markov(15), 3, 0, 1e-8, "TM", 0.5
        kw = dict(target=0.5, st="sinvert")
    elif case == "twosided_convdiff8":
        Ao, nev, ncv, tol, which = orc.convdiff3d(8), 6, 24, 1e-10, "LR"
        kw = dict(twosided=True)
    else:
        Ao, nev, ncv, tol, which = _complex_csr(orc), 6, 24, 1e-10, "LM"
        solver = "arnoldi"
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    if tgt is not None:
        E.setTarget(tgt)
        E.setST("sinvert")
        E.setSTKSP("jacobi", 1e-13, 20000)
    if case == "twosided_convdiff8":
        E.setTwoSided(True)
    E.setDimensions(nev, ncv)
    E.setTolerances(tol, 500)
    E.setInitialSpace(v0)
    E.setTrueResidual(True)
    assert E.getTrueResidual()
    E.solve()
    r = orc.eps_solve(Ao, nev, ncv, tol, which, herm, v0=v0, max_it=500, solver=solver, trueres=True, **kw)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-9
    if case == "test22":
        assert sorted((round(v.real, 5), round(v.imag, 5)) for v in vals) == TEST22_VALUES
    # sinvert: iterative solves against the oracle's exact LU; complex_arnoldi:
    # 110 explicit restarts, over which the trajectories part in rounding
    same_path = case not in ("sinvert_markov", "complex_arnoldi")
    if same_path:
        assert abs(E.getIterationNumber() - r["its"]) <= 1
    st = E.getStats()
    assert st["true_residuals"] >= nev
    for i in range(nev):
        est = E.getErrorEstimate(i)
        assert est < tol
        if same_path:
            assert abs(est - r["errest"][i]) < 1e-3 * tol + 1e-14
        if case != "twosided_convdiff8":     # EPSComputeError takes the left residual too
            assert E.computeError(i) <= tol


def test_true_residual_off_keeps_the_estimates(sa, orc):
    """Without EPSSetTrueResidual the estimates stay beta |y_last| (equal to
    the oracle's) and no explicit residual is taken."""
    Ao = orc.convdiff3d(10)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(False)
    E.setWhichEigenpairs("LR")
    E.setDimensions(8, 32)
    E.setTolerances(1e-10, 0)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, 8, 32, 1e-10, "LR", False, v0=v0)
    assert E.getStats()["true_residuals"] == 0
    for i in range(8):
        assert abs(E.getErrorEstimate(i) - r["errest"][i]) < 1e-3 * 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mat_residual_norms_matches_oracle(sa, orc, prec):
    """sa_mat_residual_norms (fused SpMV + axpy + norm) against the oracle's
    EPSComputeResidualNorm_Private on real vectors, a complex pair, a tiny
    eigenvalue (lambda terms dropped) and an imaginary part below |re| eps
    (treated as real); more than 16 vectors span several launches."""
    Ao = orc.convdiff3d(13, 11, 9)
    n = Ao.n
    k = 21
    X = np.stack([orc.rand_uniform(40 + j, n) for j in range(k)], axis=1)
    lr = np.linspace(0.5, 11.0, k)
    li = np.zeros(k)
    li[3], li[4] = 0.75, -0.75          # pair in columns 3, 4
    li[17], li[18] = 2.0, -2.0          # pair in columns 17, 18
    lr[7] = 1e-17                       # |lambda| <= eps: ||A x||
    li[9] = 1e-20                       # below |re| eps: real
    dt = np.float64 if prec == "fp64" else np.float32
    Xd = X.astype(dt).astype(np.float64)
    M = sa.Mat(sa.CSR(n, Ao.rowptr, Ao.col, Ao.val), sa.FP64 if prec == "fp64" else sa.FP32)
    res = M.residual_norms(Xd, lr, li)
    Ao32 = orc.CSR(n, Ao.rowptr, Ao.col, Ao.val.astype(dt).astype(np.float64))
    ref = np.zeros(k)
    i = 0
    while i < k:
        if li[i] != 0.0 and abs(li[i]) >= abs(lr[i]) * 2.220446049250313e-16:
            ref[i] = ref[i + 1] = orc.residual_abs(Ao32, lr[i], li[i], Xd[:, i], Xd[:, i + 1])
            i += 2
        else:
            ref[i] = orc.residual_abs(Ao32, lr[i], 0.0, Xd[:, i])
            i += 1
    tol = 1e-12 if prec == "fp64" else 2e-6
    assert np.max(np.abs(res - ref) / ref) < tol
    assert res[3] == res[4] and res[17] == res[18]


# ------------------------------------------------------------------ shift-and-invert, PREONLY + LU
@pytest.mark.parametrize("solver,case", [("krylovschur", "markov15"), ("arnoldi", "markov15"),
                                         ("lanczos", "test2"), ("krylovschur", "test2"),
                                         ("krylovschur", "lap2d40"), ("krylovschur", "complex_inside"),
                                         ("arnoldi", "complex_inside")])
def test_sinvert_lu_matches_oracle(sa, orc, solver, case):
    """Shift-and-invert with the default KSP, the reference's PREONLY + LU
    (stsles.c:53-54): a banded LU of A - sigma I with partial pivoting
    factored on the device, two substitutions per operator application.
    Both sides solve exactly, so the GPU follows the oracle's trajectory:
    eigenvalues in order to 1e-10, restart counts equal.  complex_inside:
    the non-symmetric matrix with conjugate pairs at a target inside its
    spectrum (A - 4 I indefinite), where BiCGStab stagnates."""
    from test_cpu_oracle import _lap1d, _complex_csr, TEST2_INTERIOR
    if case == "markov15":
        Ao, herm, nev, tol, tgt = orc.markov(15), False, 3, 1e-8, 0.5
    elif case == "test2":
        Ao, herm, nev, tol, tgt = _lap1d(orc, 30), True, 4, 1e-9, 2.1
    elif case == "lap2d40":
        Ao, herm, nev, tol, tgt = orc.lap2d(40, 37), True, 6, 1e-10, 0.0
    else:
        Ao, herm, nev, tol, tgt = _complex_csr(orc), False, 4, 1e-10, 4.0
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(herm)
    E.setWhichEigenpairs("TM")
    E.setTarget(tgt)
    E.setST("sinvert")
    E.setDimensions(nev, 0)
    E.setTolerances(tol, 500)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, nev, 0, tol, "TM", herm, v0=v0, target=tgt, st="sinvert", solver=solver, max_it=500)
    assert E.getConverged() >= nev and r["nconv"] >= nev
    pairs = [E.getEigenpair(i, False) for i in range(nev)]
    vals = np.array([p[0] + 1j * p[1] for p in pairs])
    ref = r["eigr"][:nev] + 1j * r["eigi"][:nev]
    assert np.max(np.abs(vals - ref) / np.abs(ref)) < 1e-10
    assert abs(E.getIterationNumber() - r["its"]) <= 1
    if case == "markov15":
        assert list(np.round(vals.real, 5)) == [0.49842, 0.51928, 0.47897]
    if case == "test2":
        assert np.allclose(vals.real, TEST2_INTERIOR[30], atol=5e-6)
    if case == "complex_inside":
        assert np.any(ref.imag != 0.0)
    nrm = max(np.abs(Ao.val[Ao.rowptr[i]:Ao.rowptr[i + 1]]).sum() for i in range(Ao.n)) + abs(tgt)
    for i in range(nev):
        assert E.computeError(i) < 10 * tol * max(1.0, nrm / abs(vals[i]))


def test_sinvert_lu_rejects_a_band_too_wide(sa):
    """The 2-D Laplacian on 256 x 256 (band 256 on each side) exceeds the
    device LU's work limit: the solve ends with an error naming BiCGStab as
    the alternative, and no eigenpair is reported."""
    A = sa.lap2d(256)
    E = sa.EPS()
    E.setOperators(sa.Mat(A))
    E.setProblemType(True)
    E.setWhichEigenpairs("TM")
    E.setTarget(0.0)
    E.setST("sinvert")
    E.setDimensions(4, 0)
    with pytest.raises(sa.SlepcAmdError, match="too wide"):
        E.solve()
