"""EPSSetDeflationSpace on the device (sa_eps_set_deflation_space;
src/eps/interface/epssetup.c:412-418,544-583, src/eps/interface/epssolve.c:201-205).

eps/tests/test10.c (test10_1.out, -eps_nev 4 -m 11 -eps_max_it 500 for
krylovschur and arnoldi): the Fiedler vector of a 10 x 11 grid, the constant
null vector deflated.  A space that is not invariant is checked against the
dense eigenvalues of A on its orthogonal complement."""
import numpy as np
import pytest
import scipy.sparse as sp

from test_cpu_deflation import TEST10_1, closed_form, complement_eigvals, graph_laplacian

pytestmark = pytest.mark.gpu


def _mat(sa, S):
    S = sp.csr_matrix(S)
    S.sort_indices()
    return sa.Mat(sa.CSR(S.shape[0], S.indptr.astype(np.int32), S.indices.astype(np.int32),
                         S.data.astype(np.float64)))


def _eps(sa, S, which, nev, solver="krylovschur", herm=True, tol=1e-8, max_it=500):
    A = _mat(sa, S)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(A)
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    E.setDimensions(nev, 0)
    E.setTolerances(tol, max_it)
    return E, A


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi"])
def test_test10_fiedler(sa, solver):
    n = 110
    E, A = _eps(sa, graph_laplacian(10, 11), "SR", 4, solver)
    E.setDeflationSpace(np.ones(n))
    E.solve()
    assert E.getConverged() >= 4
    vals = [E.getEigenpair(i, False)[0] for i in range(4)]
    assert ["%.5f" % v for v in vals] == TEST10_1
    assert np.allclose(vals, closed_form(10, 11)[1:5], rtol=0, atol=1e-9)
    for i in range(4):
        x = E.getEigenpair(i)[2]
        assert abs(x.sum()) < 1e-12 * np.sqrt(n) * np.linalg.norm(x)   # orthogonal to the deflation space
        assert E.computeError(i, sa.ERROR_RELATIVE) < 1e-7
    # the solve removed the space: the next solve finds the null vector
    E.solve()
    vals = [E.getEigenpair(i, False)[0] for i in range(4)]
    assert abs(vals[0]) < 1e-9
    assert ["%.5f" % v for v in vals[1:]] == TEST10_1[:3]


def test_test10_shift_and_invert(sa):
    """Closest to 0.05 with shift-and-invert: without the deflation space the
    zero eigenvalue (distance 0.05) would be the third value."""
    E, A = _eps(sa, graph_laplacian(10, 11), "TM", 4)
    E.setTarget(0.05)
    E.setST("sinvert")
    E.setDeflationSpace(np.ones((1, 110)))
    E.solve()
    assert E.getConverged() >= 4
    vals = [E.getEigenpair(i, False)[0] for i in range(4)]
    assert ["%.5f" % v for v in vals] == TEST10_1
    assert np.allclose(vals, closed_form(10, 11)[1:5], rtol=0, atol=1e-9)


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi"])
def test_non_invariant_space_matches_the_dense_complement(sa, solver):
    S = graph_laplacian(12, 12)
    C = np.random.default_rng(3).standard_normal((144, 3))
    # BVInsertVecs (bvfunc.c:359-374) drops a vector whose norm is zero after
    # orthogonalization; a numerically dependent one keeps its rounding noise
    # (the refinement criterion of bvorthog.c:175-185 holds for noise), so
    # the dropped vector here is exactly zero
    Cd = np.column_stack([C[:, :2], np.zeros(144), C[:, 2]])
    want = np.sort(complement_eigvals(S.toarray(), C).real)[:4]
    E, A = _eps(sa, S, "SR", 4, solver, tol=1e-10)
    E.setDeflationSpace(Cd)                                  # an n x k array: columns
    E.solve()
    assert E.getConverged() >= 4
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(4)])
    assert np.allclose(vals, want, rtol=0, atol=1e-9)
    for i in range(4):
        x = E.getEigenpair(i)[2]
        assert np.linalg.norm(C.T @ x) < 1e-12 * np.linalg.norm(C) * np.linalg.norm(x)


def test_non_hermitian(sa, orc):
    """markov(15) (eps/tutorials/ex5.c's walk), one random vector deflated:
    the rightmost eigenvalues of A on the complement."""
    Ao = orc.markov(15)
    S = sp.csr_matrix((Ao.val, Ao.col, Ao.rowptr), shape=(Ao.n, Ao.n))
    c = np.random.default_rng(7).standard_normal((Ao.n, 1))
    want = complement_eigvals(S.toarray(), c)
    want = want[np.argsort(-want.real)]
    E, A = _eps(sa, S, "LR", 4, herm=False, tol=1e-10, max_it=300)
    E.setDeflationSpace(c)
    E.solve()
    assert E.getConverged() >= 4
    for i in range(4):
        re, im, x, _ = E.getEigenpair(i)
        assert abs(re - want[i].real) < 1e-8
        assert np.min(np.abs(want[:6] - complex(re, im))) < 1e-8
        assert abs(c[:, 0] @ x) < 1e-12 * np.linalg.norm(c) * np.linalg.norm(x)


def test_rejected_configurations(sa):
    S = graph_laplacian(10, 11)
    E, A = _eps(sa, S, "SR", 4, "lanczos")
    E.setDeflationSpace(np.ones(110))
    with pytest.raises(RuntimeError, match="deflation space runs with"):
        E.solve()
    E2, A2 = _eps(sa, S, "SR", 4)
    with pytest.raises(RuntimeError, match="invalid deflation space"):
        sa.check(sa.lib().sa_eps_set_deflation_space(E2.h, -1, np.empty(0)))
    E2.setDeflationSpace(np.ones(110))
    B = _mat(sa, graph_laplacian(10, 12))
    E2.setOperators(B)
    with pytest.raises(RuntimeError, match="operator's length"):
        E2.solve()
