"""EPSMonitor and the stopping test through the C ABI (sa_eps_monitor_set,
sa_eps_set_stopping_test_function; src/eps/interface/epsmon.c:39-46,96-122,
src/eps/interface/epsopts.c:863-885, EPSStoppingBasic epsdefault.c:290-302).

The monitor's per-iteration view (iteration number, converged count, first
Ritz value) is compared with the oracle's trace of the same solve on short
trajectories where the GPU restarts exactly as the oracle does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(orc, case):
    from test_cpu_oracle import _lap1d
    if case == "ks_hep":
        return orc.lap2d(20), "LM", True, 4, 16, 1e-10, 0, "krylovschur"
    if case == "ks_nhep":
        return orc.markov(15), "LR", False, 4, 8, 1e-8, 300, "krylovschur"
    if case == "arnoldi":
        return _lap1d(orc, 30), "LR", True, 4, 0, 1e-9, 0, "arnoldi"
    return _lap1d(orc, 30), "LR", True, 4, 0, 1e-9, 0, "lanczos"


def _eps(sa, Ao, which, herm, nev, ncv, tol, mi, solver, v0):
    A = sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val))
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(A)
    E.setProblemType(herm)
    E.setWhichEigenpairs(which)
    E.setDimensions(nev, ncv)
    E.setTolerances(tol, mi)
    E.setInitialSpace(v0)
    return E, A


@pytest.mark.parametrize("case", ["ks_hep", "ks_nhep", "arnoldi", "lanczos"])
def test_monitor_trace_matches_oracle(sa, orc, case):
    Ao, which, herm, nev, ncv, tol, mi, solver = _case(orc, case)
    v0 = orc.rand_uniform(0, Ao.n)
    E, A = _eps(sa, Ao, which, herm, nev, ncv, tol, mi, solver, v0)
    trace = []
    E.setMonitor(lambda its, nconv, er, ei, err, nest: trace.append((its, nconv, er[0], nest, err.copy())))
    calls2 = []
    E.setMonitor(lambda its, nconv, er, ei, err, nest: calls2.append(its))   # a second monitor, run after the first
    E.solve()
    its = E.getIterationNumber()
    assert [t[0] for t in trace] == list(range(1, its + 1))
    assert calls2 == list(range(1, its + 1))
    r = orc.eps_solve(Ao, nev, ncv, tol, which, herm, v0=v0, solver=solver, max_it=mi, trace=10000)
    assert r["its"] == its
    assert [t[1] for t in trace] == list(r["trace_nconv"])
    e0 = np.array([t[2] for t in trace])
    assert np.max(np.abs(e0 - r["trace_eig0"]) / np.abs(r["trace_eig0"])) < 1e-9
    assert trace[-1][1] >= nev and E.getConverged() >= nev
    # the estimates of the last iteration: the converged ones below tol
    last = trace[-1]
    assert np.all(last[4][:nev] < tol)
    E.cancelMonitor()
    trace.clear()
    E.solve()
    assert trace == []


def test_monitor_error_ends_the_solve(sa, orc):
    Ao, which, herm, nev, ncv, tol, mi, solver = _case(orc, "ks_hep")
    E, A = _eps(sa, Ao, which, herm, nev, ncv, tol, mi, solver, orc.rand_uniform(0, Ao.n))
    seen = []

    def mon(its, nconv, er, ei, err, nest):
        seen.append(its)
        return 7 if its == 3 else 0
    E.setMonitor(mon)
    with pytest.raises(Exception) as ei:
        E.solve()
    assert "monitor" in str(ei.value)
    assert seen == [1, 2, 3]


@pytest.mark.parametrize("case", ["ks_hep", "ks_nhep", "lanczos"])
def test_user_stopping_test(sa, orc, case):
    """A stopping test that answers CONVERGED_USER at iteration 5 ends the
    solve there, exactly where the oracle's basic test ends it with
    max_it = 5 (same converged count and values); restoring the basic test
    gives the full solve again."""
    Ao, which, herm, nev, ncv, tol, mi, solver = _case(orc, case)
    v0 = orc.rand_uniform(0, Ao.n)
    E, A = _eps(sa, Ao, which, herm, nev, ncv, tol, mi, solver, v0)
    calls = []

    def stop(its, max_it, nconv, nev_):
        calls.append((its, nconv, nev_))
        if nconv >= nev_:
            return sa.REASON_CONVERGED_TOL
        return sa.REASON_CONVERGED_USER if its >= 5 else sa.REASON_ITERATING
    E.setStoppingTest(stop)
    E.solve()
    assert E.getIterationNumber() == 5 and E.getConvergedReason() == sa.REASON_CONVERGED_USER
    assert [c[0] for c in calls] == [1, 2, 3, 4, 5] and all(c[2] == nev for c in calls)
    r = orc.eps_solve(Ao, nev, ncv, tol, which, herm, v0=v0, solver=solver, max_it=5)
    assert r["its"] == 5 and r["reason"] == -1
    nc = E.getConverged()
    assert nc == r["nconv"] == calls[-1][1]
    for i in range(nc):
        v = E.getEigenpair(i, False)[0]
        assert abs(v - r["eigr"][i]) <= 1e-9 * abs(r["eigr"][i])
        assert E.computeError(i) < 10 * tol
    E.setStoppingTest(None)
    E.solve()
    assert E.getConvergedReason() == sa.REASON_CONVERGED_TOL and E.getConverged() >= nev


@pytest.mark.parametrize("conv,which", [("abs", "LM"), ("rel", "LM"), ("norm", "SM")])
def test_convergence_criteria_match_oracle(sa, orc, conv, which):
    """EPSSetConvergenceTest on the 2-D Laplacian scaled by 1e-4 with true
    residuals, tol 1e-8: the device solve returns the oracle's values and
    restart count, and each error estimate equals its explicit residual
    (EPSComputeError, absolute) divided by 1, |lambda| or ||A||_inf."""
    from test_cpu_oracle import _scaled_lap
    Ao = _scaled_lap(orc, 20, 1e-4)
    v0 = orc.rand_uniform(0, Ao.n)
    E = sa.EPS()
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(True)
    E.setWhichEigenpairs(which)
    E.setDimensions(4, 20)
    E.setTolerances(1e-8, 5000)
    E.setTrueResidual(True)
    E.setConvergenceTest(conv)
    E.setInitialSpace(v0)
    E.solve()
    r = orc.eps_solve(Ao, 4, 20, 1e-8, which, True, v0=v0, trueres=True, conv=conv, max_it=5000)
    assert E.getConverged() >= 4 and r["nconv"] >= 4
    assert abs(E.getIterationNumber() - r["its"]) <= 1
    vals = np.array([E.getEigenpair(i, False)[0] for i in range(4)])
    assert np.allclose(vals, r["eigr"][:4], rtol=1e-9, atol=0)
    nrma = np.max(np.abs(Ao.to_dense()).sum(axis=1))
    nA, nB = E.getMatrixNorms()
    assert nB == 0.0 and abs(nA - (nrma if conv == "norm" else 0.0)) <= 1e-15 * nrma
    for i in range(4):
        denom = {"abs": 1.0, "rel": abs(vals[i]), "norm": nrma}[conv]
        est = E.getErrorEstimate(i)
        assert est <= 1e-8
        assert abs(est - E.computeError(i, sa.ERROR_ABSOLUTE) / denom) <= 1e-3 * est


def test_convergence_norm_test30_and_backward_error(sa):
    """eps/tests/test30.c suffix 1 (-eps_conv_norm, test30_1.out): the diagonal
    matrix 1..20, 4 eigenvalues closest to 4.8 print 5.00000, 4.00000,
    6.00000, 3.00000 and the computed norm(A) is 20 (here with
    shift-and-invert; the reference test runs GD).  ||B|| stays 0 for the
    standard problem until a backward error is computed, which sets it to 1
    (epssolve.c:801): EPS_ERROR_BACKWARD = ||r|| / (||A|| + |lambda|)."""
    n = 20
    Ah = sa.CSR(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.arange(1.0, n + 1))
    E = sa.EPS()
    E.setOperators(sa.Mat(Ah))
    E.setProblemType(True)
    E.setWhichEigenpairs("TM")
    E.setTarget(4.8)
    E.setST("sinvert")
    E.setDimensions(4, 0)
    E.setConvergenceTest("norm")
    E.solve()
    assert E.getConverged() >= 4
    vals = [E.getEigenpair(i, False)[0] for i in range(4)]
    assert ["%.5f" % v for v in vals] == ["5.00000", "4.00000", "6.00000", "3.00000"]
    assert E.getMatrixNorms() == (20.0, 0.0)
    for i in range(4):
        a = E.computeError(i, sa.ERROR_ABSOLUTE)
        b = E.computeError(i, sa.ERROR_BACKWARD)
        assert b == a / (20.0 + abs(vals[i]))
    assert E.getMatrixNorms() == (20.0, 1.0)


def test_user_convergence_function(sa, orc):
    """EPSSetConvergenceTestFunction: a user function computing the relative
    estimate gives the default test's values and restart count bit for bit;
    a function that fails ends the solve with an error; None restores the
    relative test."""
    Ao = orc.lap2d(20)
    v0 = orc.rand_uniform(0, Ao.n)

    def run(fn):
        E = sa.EPS()
        E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
        E.setProblemType(True)
        E.setDimensions(6, 20)
        E.setTolerances(1e-10, 0)
        E.setInitialSpace(v0)
        if fn != "default":
            E.setConvergenceTestFunction(fn)
        E.solve()
        return [E.getEigenpair(i, False)[0] for i in range(6)], E.getIterationNumber()

    v1, i1 = run("default")
    calls = []

    def rel(er, ei, res):
        calls.append(res)
        return res / np.hypot(er, ei)
    v2, i2 = run(rel)
    assert v1 == v2 and i1 == i2 and len(calls) > 0
    with pytest.raises(sa.SlepcAmdError, match="convergence test returned an error"):
        run(lambda er, ei, res: 1 / 0)
    v3, i3 = run(None)
    assert v3 == v1 and i3 == i1
    E = sa.EPS()
    with pytest.raises(sa.SlepcAmdError, match="EPSSetConvergenceTestFunction"):
        sa.check(sa.lib().sa_eps_set_convergence_test(E.h, 3))


@pytest.mark.parametrize("solver", ["krylovschur", "arnoldi"])
def test_user_comparison_test11(sa, orc, solver):
    """EPSSetEigenvalueComparison, eps/tests/test11.c suffix 1 (-eps_type
    {krylovschur, arnoldi} -st_type sinvert, shift 0.5, ones as start
    vector): the eigenvalues closest to the right of 0.5 print 0.51928,
    0.55740, 0.57028, 0.57143 (test11_1.out), equal to the oracle run with the
    same comparison in values and restarts."""
    from test_cpu_oracle import _right_of
    Ao = orc.markov(15)
    E = sa.EPS()
    E.setType(solver)
    E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
    E.setProblemType(False)
    E.setDimensions(4, 0)
    E.setTolerances(1e-10, 0)
    E.setTarget(0.5)
    E.setST("sinvert", 0.5)
    E.setEigenvalueComparison(_right_of(0.5))
    E.setInitialSpace(np.ones(Ao.n))
    E.solve()
    r = orc.eps_solve_user_order(Ao, 4, _right_of(0.5), tol=1e-10, hermitian=False, v0=np.ones(Ao.n),
                                 st="sinvert", target=0.5, solver=solver)
    assert E.getConverged() >= 4
    vals = [E.getEigenpair(i, False)[0] for i in range(4)]
    assert ["%.5f" % v for v in vals] == ["0.51928", "0.55740", "0.57028", "0.57143"]
    assert np.allclose(vals, r["eigr"][:4], rtol=1e-10, atol=0)
    assert abs(E.getIterationNumber() - r["its"]) <= 1


def test_user_comparison_as_largest_magnitude_and_errors(sa, orc):
    """A comparison that orders by decreasing magnitude reproduces which = LM
    bit for bit; a comparison that fails ends the solve with an error; None
    restores largest magnitude."""
    Ao = orc.markov(15)
    v0 = orc.rand_uniform(0, Ao.n)

    def run(cmp):
        E = sa.EPS()
        E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
        E.setProblemType(False)
        E.setDimensions(4, 0)
        E.setTolerances(1e-10, 0)
        if cmp != "LM":
            E.setEigenvalueComparison(cmp)
        E.setInitialSpace(v0)
        E.solve()
        return [E.getEigenpair(i, False)[:2] for i in range(4)], E.getIterationNumber()

    def lm(ar, ai, br, bi):
        a, b = np.hypot(ar, ai), np.hypot(br, bi)
        return 1 if a < b else (-1 if a > b else 0)
    assert run(lm) == run("LM")
    with pytest.raises(sa.SlepcAmdError, match="comparison returned an error"):
        run(lambda ar, ai, br, bi: 1 / 0)
    assert run(None) == run("LM")


def test_user_comparisons_on_concurrent_threads(sa, orc):
    """Each host thread's solve uses its own comparison (the comparison is
    installed per thread for the duration of a solve): four concurrent
    solves, two ordering by largest magnitude and two by smallest real part,
    each give the same values and restart count as the same solve run alone."""
    import concurrent.futures as cf
    Ao = orc.lap2d(20)
    v0 = orc.rand_uniform(0, Ao.n)

    def lm(ar, ai, br, bi):
        a, b = np.hypot(ar, ai), np.hypot(br, bi)
        return 1 if a < b else (-1 if a > b else 0)

    def sr(ar, ai, br, bi):
        return 1 if ar > br else (-1 if ar < br else 0)

    def run(cmp):
        E = sa.EPS()
        E.setOperators(sa.Mat(sa.CSR(Ao.n, Ao.rowptr, Ao.col, Ao.val)))
        E.setProblemType(True)
        E.setDimensions(4, 16)
        E.setTolerances(1e-10, 5000)
        E.setEigenvalueComparison(cmp)
        E.setInitialSpace(v0)
        E.solve()
        return [E.getEigenpair(i, False)[0] for i in range(4)], E.getIterationNumber()

    alone = {"lm": run(lm), "sr": run(sr)}
    assert alone["lm"][0][0] > 7.9 and alone["sr"][0][0] < 0.05
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        futs = [(k, ex.submit(run, f)) for k, f in (("lm", lm), ("sr", sr), ("lm", lm), ("sr", sr))]
        for k, f in futs:
            assert f.result() == alone[k]
