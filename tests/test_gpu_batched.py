"""GPU tests of the batched solver: every problem of an uneven batch agrees
with the CPU oracle run on that problem alone (same start vector, same
algorithm), in FP64 and in mixed precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problems(sa, orc):
    from slepc_amd.batch import uneven_laplacians
    P = uneven_laplacians(P=7, seed=2, lo=20, hi=70)
    return P


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_matches_oracle(sa, orc, mixed):
    from slepc_amd.batch import Batch
    csrs = _problems(sa, orc)
    v0 = [orc.rand_uniform(orc.sub_seed(9, p), A.n) for p, A in enumerate(csrs)]
    B = Batch(csrs)
    B.solve(8, 24, 1e-10, "LM", mixed=mixed, v0=v0)
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
