"""The expected values of the deflation-space tests, pinned without a GPU.

eps/tests/test10.c builds the graph Laplacian of an n x m grid (degree on the
diagonal, -1 per grid edge), deflates its null vector [1 ... 1] with
EPSSetDeflationSpace and asks for the 4 smallest eigenvalues; test10_1.out
prints 0.08101, 0.09789, 0.17890, 0.31749 for -m 11 (n = 10).  The spectrum
is known in closed form, (2 - 2cos(pi i/n)) + (2 - 2cos(pi j/m)), and the
eigenvalues of A on the orthogonal complement of a space C are those of
Q^T A Q for an orthonormal basis Q of that complement; both are checked here
against a dense eigensolver, so a GPU test can compare with them."""
import numpy as np
import scipy.sparse as sp

TEST10_1 = ["0.08101", "0.09789", "0.17890", "0.31749"]   # eps/tests/output/test10_1.out


def graph_laplacian(n, m):
    """test10.c:43-54: node II = i*n + j of an m-row, n-column grid."""
    N = n * m
    rows, cols, vals = [], [], []
    for II in range(N):
        i, j = divmod(II, n)
        w = 0.0
        for cond, J in ((i > 0, II - n), (i < m - 1, II + n), (j > 0, II - 1), (j < n - 1, II + 1)):
            if cond:
                rows.append(II), cols.append(J), vals.append(-1.0)
                w += 1.0
        rows.append(II), cols.append(II), vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def closed_form(n, m):
    a = 2.0 - 2.0 * np.cos(np.pi * np.arange(n) / n)
    b = 2.0 - 2.0 * np.cos(np.pi * np.arange(m) / m)
    return np.sort((a[:, None] + b[None, :]).ravel())


def complement_eigvals(A, C):
    """Eigenvalues of A restricted to the orthogonal complement of span(C)
    (columns), as Q^T A Q with Q from a complete QR of C."""
    Q, _ = np.linalg.qr(C, mode="complete")
    r = np.linalg.matrix_rank(C)
    Qc = Q[:, r:]
    M = Qc.T @ (A @ Qc)
    return np.linalg.eigvals(M)


def test_test10_values_from_the_closed_form():
    A = graph_laplacian(10, 11).toarray()
    ev = np.linalg.eigvalsh(A)
    cf = closed_form(10, 11)
    assert np.allclose(ev, cf, atol=1e-12)
    assert abs(cf[0]) < 1e-15 and cf[1] > 1e-2       # one null vector, the constant
    assert ["%.5f" % v for v in cf[1:5]] == TEST10_1
    # deflating the constant vector removes exactly the zero eigenvalue
    d = np.sort(complement_eigvals(A, np.ones((110, 1))).real)
    assert np.allclose(d, cf[1:], atol=1e-12)


def test_complement_of_a_non_invariant_space():
    """A random space is not invariant: the complement's smallest eigenvalues
    are not eigenvalues of A (the grid's repeated eigenvalues survive in the
    complement, the others move), and all of them interlace A's (Cauchy)."""
    A = graph_laplacian(12, 12).toarray()
    C = np.random.default_rng(3).standard_normal((144, 3))
    d = np.sort(complement_eigvals(A, C).real)
    ev = np.linalg.eigvalsh(A)
    assert len(d) == 141
    assert np.all(d >= ev[:141] - 1e-12) and np.all(d <= ev[3:] + 1e-12)
    assert np.min(np.abs(d[:4, None] - ev[None, :])) > 1e-4
    # a dependent extra column changes nothing
    d2 = np.sort(complement_eigvals(A, np.column_stack([C, C[:, 0] + 2 * C[:, 1]])).real)
    assert np.allclose(d, d2, atol=1e-12)
