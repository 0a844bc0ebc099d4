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
