"""GPU tests of the batched solver: every problem of an uneven batch agrees
with the CPU oracle run on that problem alone (same start vector, same
algorithm), in FP64 and in mixed precision, with either sweep kernel (a
workgroup per problem, or lock-step over row chunks) and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(sa, orc):
    from slepc_amd.batch import uneven_laplacians
    P = uneven_laplacians(P=7, seed=2, lo=20, hi=70)
    return P


@pytest.mark.parametrize("sweep", ["persistent", "wide", "auto"])
@pytest.mark.parametrize("mixed", [False, True])
def test_batched_matches_oracle(sa, orc, mixed, sweep):
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]
    B = Batch(csrs)
    B.set_sweep(sweep)
    B.solve(8, 24, 1e-10, "LM", mixed=mixed, v0=v0)
    info = B.sweep_info()
    assert info["mode"] == (sweep if sweep != "auto" else info["mode"]) and info["mode"] in ("persistent", "wide")
    if sweep == "auto":
        assert info["ms_persistent"] > 0 and info["ms_wide"] > 0
    err = B.errors()
    for p, A in enumerate(csrs):
        nconv, its, reason = B.converged(p)
        assert reason == 1 and nconv >= 8
        Ao = orc.CSR(A.n, A.rowptr, A.col, A.val)
        r = orc.eps_solve(Ao, 8, 24, 1e-10, "LM", True, v0=v0[p])
        vals = np.array([B.eigenpair(p, i, False)[0] for i in range(8)])
        w = np.linalg.eigvalsh(Ao.to_dense())
        for v in vals:
            assert np.min(np.abs(w - v)) < 1e-10 * abs(v)
        assert np.allclose(np.sort(vals), np.sort(r["eigr"][:8]), rtol=1e-10, atol=0)
        if not mixed:
            assert abs(its - r["its"]) <= max(2, 0.2 * r["its"])
        assert np.all(err[p] <= 1e-10), err[p]
    st = B.stats()
    assert st["matvecs"] > 0 and (st["matvecs32"] > 0) == mixed


def test_batched_c5_shapes(sa):
    """The 512-problem family: sizes span 1k..65k rows and all converge."""
    from slepc_amd.batch import Batch, uneven_laplacians
    csrs = uneven_laplacians(P=512, seed=2)
    ns = np.array([A.n for A in csrs])
    assert ns.min() >= 32 * 32 and ns.max() <= 256 * 256
    B = Batch(csrs[:64])
    B.solve(8, 24, 1e-10, "LM", mixed=True)
    err = B.errors()
    assert np.all(err <= 1e-10)


def test_batched_fv_family_ncv64_matches_single_solves(sa, orc):
    """The parameterised FV family at ncv 64 (lock-step sweep, the one-workgroup
    sweep holds at most 63 columns): every problem matches the single-problem
    solver and, on the smallest grid, the CPU oracle."""
    from slepc_amd.batch import Batch
    thetas = [sa.fv_theta(5, s) for s in range(4)]
    csrs = [sa.fv2d(g, g, th) for g, th in zip((24, 32, 40, 48), thetas)]
    v0 = [sa.rand_uniform(0, A.n) for A in csrs]
    B = Batch(csrs)
    B.solve(16, 64, 1e-10, "LR", v0=v0)
    assert B.sweep_info()["mode"] == "wide"
    err = B.errors()
    assert np.all(err <= 1e-10), err
    for p, A in enumerate(csrs):
        nconv, its, reason = B.converged(p)
        assert reason == 1 and nconv >= 16
        vals = np.sort([B.eigenpair(p, i, False)[0] for i in range(16)])
        E = sa.EPS()
        E.setOperators(sa.Mat(A))
        E.setProblemType(True)
        E.setWhichEigenpairs("LR")
        E.setDimensions(16, 64)
        E.setTolerances(1e-10)
        E.setInitialSpace(v0[p])
        E.solve()
        single = np.sort([E.getEigenpair(i, vectors=False)[0] for i in range(16)])
        assert np.allclose(vals, single, rtol=1e-10, atol=0)
        if p == 0:
            r = orc.eps_solve(orc.CSR(A.n, A.rowptr, A.col, A.val), 16, 64, 1e-10, "LR", True, v0=v0[p])
            assert np.allclose(vals, np.sort(r["eigr"][:16]), rtol=1e-10, atol=0)


@pytest.mark.parametrize("sweep", ["persistent", "wide"])
@pytest.mark.parametrize("mixed", [False, True])
def test_batched_true_residual_matches_oracle(sa, orc, mixed, sweep):
    """EPSSetTrueResidual for the whole batch (candidate Ritz vectors of all
    problems formed and evaluated on the device at once): every problem
    gives the oracle's values with its true-residual branch, the error
    estimates are explicit residuals (equal to EPSComputeError of the
    returned pair to rounding), and every pair meets the tolerance."""
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]
    B = Batch(csrs)
    B.set_sweep(sweep)
    B.set_true_residual(True)
    B.solve(8, 24, 1e-10, "LM", mixed=mixed, v0=v0)
    err = B.errors()
    for p, A in enumerate(csrs):
        nconv, its, reason = B.converged(p)
        assert reason == 1 and nconv >= 8
        Ao = orc.CSR(A.n, A.rowptr, A.col, A.val)
        r = orc.eps_solve(Ao, 8, 24, 1e-10, "LM", True, v0=v0[p], trueres=True)
        vals = np.array([B.eigenpair(p, i, False)[0] for i in range(8)])
        assert np.allclose(np.sort(vals), np.sort(r["eigr"][:8]), rtol=1e-10, atol=0)
        if not mixed:
            assert abs(its - r["its"]) <= max(2, 0.2 * r["its"])
        assert np.all(err[p] <= 1e-10), err[p]
