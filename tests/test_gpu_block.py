"""GPU parity tests of the block BV operations (BVDot, BVMatMult,
BVOrthogonalize) against the oracle and the reference's bv tests 7 and 11.
All calls go through the C ABI."""
import numpy as np
import pytest

from test_cpu_oracle import _test11_matrix

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def _align_signs(Q, R, Qref):
    """SVQB takes its columns from eigenvectors (syev), whose signs follow the
    rounding of the Gram matrix: flip columns of Q (and rows of R, so that
    Q R is unchanged) to match Qref before comparing."""
    sgn = np.where(np.sum(Q * Qref, axis=0) < 0, -1.0, 1.0)
    return Q * sgn, R * sgn[:R.shape[0], None]


def _bv(sa, X, prec=0):
    V = sa.BV(X.shape[0], X.shape[1], prec)
    V.set_matrix(X)
    return V


METHODS = {"gs": 0, "chol": 1, "svqb": 2, "tsqr": 3, "tsqrchol": 4}


@pytest.mark.parametrize("method", list(METHODS))
def test_orthogonalize_test11(sa, orc, method):
    """bv/tests/test11.c (n=20, l=2, k=8): level of orthogonality of Q1, Q2
    and Q and residual ||X - Q R|| below 100 eps; Q and R equal the oracle's."""
    X = _test11_matrix()
    l, k = 2, 8
    V = _bv(sa, X)
    V.set_active_columns(0, l)
    R1 = V.orthogonalize(method)
    V.set_active_columns(l, k)
    R2 = V.orthogonalize(method)
    Q = V.get_matrix(k)
    R = np.zeros((k, k))
    R[:l, :l] = R1[:l, :l]
    R[:, l:] = R2[:, l:]
    assert np.linalg.norm(Q[:, :l].T @ Q[:, :l] - np.eye(l)) < 100 * EPS
    assert np.linalg.norm(Q[:, l:].T @ Q[:, l:] - np.eye(k - l)) < 100 * EPS
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 100 * EPS
    assert np.linalg.norm(X - Q @ R) < 100 * EPS
    m = METHODS[method]
    Vo, Ro1 = orc.bv_orthogonalize(X.copy(), 0, l, m)
    Vo, Ro2 = orc.bv_orthogonalize(Vo, l, k, m)
    Q, R2 = _align_signs(Q, R2, Vo) if method == "svqb" else (Q, R2)
    assert np.max(np.abs(Q - Vo)) < 1e-13
    assert np.max(np.abs(R2[:, l:] - Ro2[:, l:])) < 1e-13


@pytest.mark.parametrize("method", list(METHODS))
def test_orthogonalize_large_matches_oracle(sa, orc, method):
    """100,001 rows, 40 columns, 8 leading ones already orthonormal (more
    64-row blocks than workgroups, BVDot split over two 32-row chunks per
    step): Q and R agree with the oracle, Q is orthonormal and X = Q R."""
    n, k, l = 100001, 40, 8
    rng = np.random.default_rng(9)
    X = rng.standard_normal((n, k))
    X[:, :l] = np.linalg.qr(X[:, :l])[0]
    V = _bv(sa, X)
    V.set_active_columns(l, k)
    R = V.orthogonalize(method)
    Q = V.get_matrix(k)
    m = METHODS[method]
    Vo, Ro = orc.bv_orthogonalize(X.copy(), l, k, m)
    Q, R = _align_signs(Q, R, Vo) if method == "svqb" else (Q, R)
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-13
    Rf = R.copy()
    Rf[:l, :l] = np.eye(l)
    assert np.linalg.norm(X - Q @ Rf) < 1e-12 * np.linalg.norm(X)
    assert np.max(np.abs(Q - Vo)) < 1e-11
    assert np.max(np.abs(R[:, l:] - Ro[:, l:])) < 1e-11 * np.abs(Ro).max()


@pytest.mark.parametrize("method", ["tsqr", "tsqrchol"])
@pytest.mark.parametrize("n,k", [(180, 7), (20, 20), (256 * 40 + 3, 16), (300001, 64), (70000, 33), (257, 64)])
def test_tsqr_shapes_match_oracle(sa, orc, method, n, k):
    """TSQR over tile trees of every depth: bv/tests/test11.c with -n 180 -l 0
    -k 7 (test11_12.out: orthogonality and ||X - QR|| below 100 eps), a square
    block, a last tile with fewer rows than columns, 64 columns over 300,001
    rows (five tree levels), 33 columns (the 64-wide kernel with padding) and
    a two-tile tree.  Q and R equal geqrf/orgqr's (the oracle), signs
    included."""
    rng = np.random.default_rng(n + k)
    X = rng.standard_normal((n, k))
    V = _bv(sa, X)
    V.set_active_columns(0, k)
    R = V.orthogonalize(method)
    Q = V.get_matrix(k)
    Vo, Ro = orc.bv_orthogonalize(X.copy(), 0, k, METHODS[method])
    scale = np.sqrt(n) if n > 1000 else 1.0
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 100 * EPS * scale
    assert np.linalg.norm(X - Q @ R) < 100 * EPS * scale * np.linalg.norm(X)
    assert np.array_equal(np.sign(np.diag(R)), np.sign(np.diag(Ro)))
    assert np.max(np.abs(Q - Vo)) < 1e-11
    assert np.max(np.abs(R - Ro)) < 1e-11 * np.abs(Ro).max()


@pytest.mark.parametrize("method", ["tsqr", "tsqrchol"])
def test_tsqr_panels_match_oracle(sa, orc, method):
    """More than 64 active columns (114 columns after 6 leading orthonormal
    ones over 40,003 rows: panels of 64 and 50; a BV holds up to 127
    columns): Q and R equal geqrf/orgqr's (the oracle), signs included; Q
    orthonormal, X = Q R."""
    n, l, k = 40003, 6, 120
    rng = np.random.default_rng(21)
    X = rng.standard_normal((n, k))
    X[:, :l] = np.linalg.qr(X[:, :l])[0]
    V = _bv(sa, X)
    V.set_active_columns(l, k)
    R = V.orthogonalize(method)
    Q = V.get_matrix(k)
    Vo, Ro = orc.bv_orthogonalize(X.copy(), l, k, METHODS[method])
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-12
    Rf = R.copy()
    Rf[:l, :l] = np.eye(l)
    assert np.linalg.norm(X - Q @ Rf) < 1e-12 * np.linalg.norm(X)
    assert np.array_equal(np.sign(np.diag(R)[l:]), np.sign(np.diag(Ro)[l:]))
    assert np.max(np.abs(Q - Vo)) < 1e-10
    assert np.max(np.abs(R[:, l:] - Ro[:, l:])) < 1e-10 * np.abs(Ro).max()


def test_tsqr_ill_conditioned(sa, orc):
    """Columns with condition number 1e13: TSQR keeps Q orthonormal to
    rounding (Householder), where a Cholesky-based QR loses about 26 digits
    of orthogonality; Q agrees with the oracle's Householder QR."""
    n, k = 50000, 24
    rng = np.random.default_rng(5)
    U = np.linalg.qr(rng.standard_normal((n, k)))[0]
    W = np.linalg.qr(rng.standard_normal((k, k)))[0]
    X = (U * np.logspace(0, -13, k)) @ W.T
    V = _bv(sa, X)
    V.set_active_columns(0, k)
    R = V.orthogonalize("tsqr")
    Q = V.get_matrix(k)
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-13
    assert np.linalg.norm(X - Q @ R) < 1e-14 * np.linalg.norm(X)
    Vo, Ro = orc.bv_orthogonalize(X.copy(), 0, k, 3)
    assert np.array_equal(np.sign(np.diag(R)), np.sign(np.diag(Ro)))


def test_tsqr_fp32(sa, orc):
    """FP32 basis: Q orthonormal to FP32 rounding and close to the oracle's
    FP64 Householder QR of the same (FP32-representable) columns."""
    n, k = 100001, 20
    rng = np.random.default_rng(6)
    X = rng.standard_normal((n, k)).astype(np.float32).astype(np.float64)
    V = _bv(sa, X, 1)
    V.set_active_columns(0, k)
    R = V.orthogonalize("tsqr")
    Q = V.get_matrix(k)
    Vo, Ro = orc.bv_orthogonalize(X.copy(), 0, k, 3)
    assert np.linalg.norm(Q.T @ Q - np.eye(k)) < 1e-4
    assert np.max(np.abs(Q - Vo)) < 1e-4
    assert np.max(np.abs(R - Ro)) < 1e-4 * np.abs(Ro).max()


def test_tsqrchol_dependent_column_raises(sa):
    """A zero column leaves R singular: TSQRCHOL's triangular inverse fails
    with an error (the trtri check of BVMatTriInv_LAPACK_Private, bvlapack.c:208-245), while
    TSQR still returns an orthonormal Q."""
    X = _test11_matrix()
    X[:, 3] = 0.0
    V = _bv(sa, X)
    with pytest.raises(RuntimeError):
        V.orthogonalize("tsqrchol")
    V = _bv(sa, X)
    R = V.orthogonalize("tsqr")
    Q = V.get_matrix(X.shape[1])
    assert np.linalg.norm(Q.T @ Q - np.eye(X.shape[1])) < 100 * EPS
    assert np.linalg.norm(X - Q @ R) < 100 * EPS


def test_orthogonalize_breakdown_raises(sa):
    """A zero column stops gs with an error, as in bvorthog.c:537."""
    X = _test11_matrix()
    X[:, 3] = 0.0
    V = _bv(sa, X)
    with pytest.raises(RuntimeError):
        V.orthogonalize("gs")


@pytest.mark.parametrize("prec", [0, 1])
def test_bv_dot_matches_oracle(sa, orc, prec):
    """BVDot with active ranges (X columns 3..73, Y columns 2..52: blocks of
    64 x 64 and remainders) on 70,001 rows: entries in rows Y.l.., columns
    X.l.. equal Y^T X; the others are untouched."""
    n = 70001
    rng = np.random.default_rng(4)
    Xh, Yh = rng.standard_normal((n, 75)), rng.standard_normal((n, 55))
    if prec == 1:
        Xh, Yh = Xh.astype(np.float32).astype(np.float64), Yh.astype(np.float32).astype(np.float64)
    X, Y = _bv(sa, Xh, prec), _bv(sa, Yh, prec)
    X.set_active_columns(3, 73)
    Y.set_active_columns(2, 52)
    M = X.dot(Y)
    ref = orc.bv_dot(Xh[:, 3:73], Yh[:, 2:52])
    tol = 1e-12 if prec == 0 else 2e-5
    assert np.max(np.abs(M[2:52, 3:73] - ref)) <= tol * np.abs(ref).max()
    assert not M[:2, :].any() and not M[:, :3].any()


def test_bv_matmult_test7_exact(sa, orc):
    """bv/tests/test7.c (m=n=10, k=5, Y with active columns 2..7): BVMatMult
    equals column-by-column MatMult exactly (test7_1.out: "Norm of error: 0.")."""
    n, k = 10, 5
    Xh = np.zeros((n, k))
    for j in range(k):
        Xh[:j + 1, j] = 1.0
    rp, col, val = [0], [], []
    for i in range(n):
        for c, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= c < n:
                col.append(c)
                val.append(v)
        rp.append(len(col))
    Ah = sa.CSR(n, np.array(rp, np.int32), np.array(col, np.int32), np.array(val))
    A = sa.Mat(Ah)
    X = _bv(sa, Xh)
    Y = sa.BV(n, k + 4, 0)
    Y.set_active_columns(2, k + 2)
    X.matmult(A, Y)
    out = Y.get_matrix(k + 4)[:, 2:k + 2]
    for j in range(k):
        assert np.array_equal(out[:, j], A.mult_host(Xh[:, j]))


@pytest.mark.parametrize("prec", [0, 1])
def test_bv_matmult_stencil27(sa, orc, prec):
    """BVMatMult of 11 columns (a pass of 8 and a pass of 3) through the
    27-point operator with its diagonal rounded to integers and integer-valued
    columns, so that every sum is exact in either precision: equal to the
    oracle bit for bit."""
    Ah = sa.stencil27(17, 16, 15, seed=1)
    Ao = orc.stencil27(17, 16, 15, seed=1)
    n = Ah.n
    Xh = (np.arange(n * 11).reshape(n, 11, order="F") % 5 - 2).astype(np.float64)
    Ah.val = np.round(Ah.val)
    Ao.val = np.round(Ao.val)
    A = sa.Mat(Ah, prec)
    X = _bv(sa, Xh, prec)
    Y = sa.BV(n, 11, prec)
    X.matmult(A, Y)
    ref = orc.bv_matmult(Ao, Xh)
    assert np.array_equal(Y.get_matrix(11), ref)


def _random_csr(sa, n, lo, hi, seed):
    """n rows with lo..hi nonzeros each (ragged), random ascending columns
    (random gaps from a random start) and random values."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, n)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    gap = 1000
    steps = np.cumsum(rng.integers(1, gap, nnz))
    row = np.repeat(np.arange(n), lens)
    within = steps - np.repeat(steps[rp[:-1]], lens)          # 0 at each row's first entry
    start = rng.integers(0, n - hi * gap, n)
    col = (start[row] + within).astype(np.int32)
    rp = rp.astype(np.int32)
    val = rng.standard_normal(nnz)
    return sa.CSR(n, rp, col, val)


@pytest.mark.parametrize("case", ["stencil27_80_fp64", "stencil27_80_fp32", "convdiff_96_fp64",
                                  "ragged_12_fp64", "ragged_40_fp64"])
def test_bv_matmult_stream_layout(sa, case):
    """Operators whose values and indices exceed 64 MiB take the stream SpMM
    (row blocks of 256, 128, 64 and 32 rows staged in LDS: 7-point, 9-16,
    27-point and 33-48 nonzeros per row).  On real-valued data each of 11
    columns (passes of 8 and 3) equals MatMult of that column bit for bit,
    since both form a row's sum in the same order."""
    prec = 1 if case.endswith("fp32") else 0
    if case.startswith("stencil27"):
        Ah = sa.stencil27(80, seed=3)
    elif case.startswith("convdiff"):
        Ah = sa.convdiff3d(96)
    elif case.startswith("ragged_12"):
        Ah = _random_csr(sa, 700_000, 9, 16, 5)
    else:
        Ah = _random_csr(sa, 200_000, 33, 48, 6)
    es = 8 if prec == 0 else 4
    assert Ah.nnz * (es + 4) > 64 << 20
    n, nc = Ah.n, 11
    A = sa.Mat(Ah, prec)
    Xh = np.random.default_rng(7).standard_normal((n, nc))
    X = _bv(sa, Xh, prec)
    Y = sa.BV(n, nc, prec)
    X.matmult(A, Y)
    out = Y.get_matrix(nc)
    dt = np.float64 if prec == 0 else np.float32
    for j in range(nc):
        assert np.array_equal(out[:, j], A.mult_host(Xh[:, j].astype(dt))), j


def test_bv_matmult_stream_layout_oracle(sa, orc):
    """The stream SpMM on the 27-point operator at 80^3 with integer-valued
    diagonal and columns (every sum exact): equal to the oracle's BVMatMult."""
    Ah = sa.stencil27(80, seed=1)
    Ao = orc.stencil27(80, seed=1)
    Ah.val = np.round(Ah.val)
    Ao.val = np.round(Ao.val)
    n = Ah.n
    Xh = (np.arange(n * 9).reshape(n, 9, order="F") % 7 - 3).astype(np.float64)
    A = sa.Mat(Ah)
    X = _bv(sa, Xh)
    Y = sa.BV(n, 9, 0)
    X.matmult(A, Y)
    assert np.array_equal(Y.get_matrix(9), orc.bv_matmult(Ao, Xh))
