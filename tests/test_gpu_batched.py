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


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_graph_replay_equals_direct_launches(sa, orc, mixed):
    """Lock-step restart sweeps replayed from captured HIP graphs give the
    same bits as the same sweeps launched one kernel at a time.  A second
    solve of the same batch replays the sweeps captured by the first (the
    signatures repeat), with the solver's own host work and device kernels
    between the replays; every replay feeds the next restart, so any wrong
    replay would change the trajectory."""
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]

    def run(B):
        B.solve(8, 24, 1e-10, "LM", mixed=mixed, v0=v0)
        vals = [[B.eigenpair(p, i, False)[0] for i in range(8)] for p in range(len(csrs))]
        its = [B.converged(p)[1] for p in range(len(csrs))]
        return np.array(vals), its, B.errors()

    B = Batch(csrs)
    B.set_sweep("wide")
    B.set_true_residual(True)
    v1, it1, e1 = run(B)
    g1 = B.graph_stats()
    assert g1["captures"] >= 1 and g1["cached"] >= 1
    v2, it2, e2 = run(B)
    g2 = B.graph_stats()
    assert g2["replays"] > g1["replays"]
    assert np.array_equal(v1, v2) and it1 == it2
    B.set_graphs(False)
    v3, it3, e3 = run(B)
    assert B.graph_stats() == g2                     # nothing captured or replayed
    assert np.array_equal(v1, v3) and it1 == it3
    assert np.all(e1 <= 1e-10) and np.array_equal(e1, e3)
    for p, A in enumerate(csrs[:3]):
        r = orc.eps_solve(orc.CSR(A.n, A.rowptr, A.col, A.val), 8, 24, 1e-10, "LM", True, v0=v0[p], trueres=True)
        assert np.allclose(np.sort(v1[p]), np.sort(r["eigr"][:8]), rtol=1e-10, atol=0)


def test_batched_schedule_autotuner(sa, orc):
    """The schedule autotuner times the one-workgroup sweep and three
    lock-step tile schedules on the first restart round and keeps the
    fastest; the solve then matches the oracle and a later solve reuses the
    choice without timing again."""
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]
    B = Batch(csrs)
    B.set_sweep("auto")
    B.solve(8, 24, 1e-10, "LM", v0=v0)
    sc = B.schedule()
    timed = [t for t in sc["ms"] if t > 0]
    assert len(timed) >= 3
    slot = 0 if sc["kind"] == "persistent" else None
    if slot is not None:
        assert sc["ms"][0] == min(timed)
    else:
        assert min(timed) in sc["ms"][1:] and sc["lanes"] in (4, 8, 16, 32) and sc["nt"] in (0, 1)
    err = B.errors()
    assert np.all(err <= 1e-10)
    for p, A in enumerate(csrs[:3]):
        r = orc.eps_solve(orc.CSR(A.n, A.rowptr, A.col, A.val), 8, 24, 1e-10, "LM", True, v0=v0[p])
        vals = np.sort([B.eigenpair(p, i, False)[0] for i in range(8)])
        assert np.allclose(vals, np.sort(r["eigr"][:8]), rtol=1e-10, atol=0)
    B.solve(8, 24, 1e-10, "LM", v0=v0)       # reuses the tuned schedule
    assert B.schedule()["kind"] == sc["kind"] and B.schedule()["lanes"] == sc["lanes"]


def test_batched_ncv_autotuner_and_per_problem_ncv(sa, orc):
    """The basis-size autotuner picks, per size class, the candidate ncv with
    the least time per converged pair on sampled problems; solve(ncv=0) then
    runs every problem with its class's ncv, and each problem agrees with the
    oracle run alone with that ncv."""
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    cand = [16, 24, 32]
    B = Batch(csrs)
    B.set_true_residual(True)
    per = B.autotune_ncv(8, 1e-10, cand, which="LM", nclass=2, nsample=2, seed=3)
    tun = B.ncv_tuning()
    assert tun["scores"].shape == (2, 3) and np.all(np.isfinite(tun["scores"]))
    bests = {cand[int(np.argmin(tun["scores"][c]))] for c in range(2)}
    assert set(per.tolist()) == bests                # each problem takes its class's best
    ns = np.array([A.n for A in csrs])
    assert tun["class_hi"][0] <= tun["class_hi"][1] == ns.max()
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]
    B.solve(8, 0, 1e-10, "LM", v0=v0)
    assert np.array_equal(B.ncv(), per)
    err = B.errors()
    assert np.all(err <= 1e-10)
    for p, A in enumerate(csrs):
        r = orc.eps_solve(orc.CSR(A.n, A.rowptr, A.col, A.val), 8, int(per[p]), 1e-10, "LM", True, v0=v0[p],
                          trueres=True)
        vals = np.sort([B.eigenpair(p, i, False)[0] for i in range(8)])
        assert np.allclose(vals, np.sort(r["eigr"][:8]), rtol=1e-10, atol=0)
