"""Subspace angles of the returned eigenvectors (symmetric problems).

For a symmetric A, a unit vector x with residual r = A x - lambda x and an
eigenvalue cluster E of A whose other eigenvalues are at least gap away from
lambda, the Davis-Kahan sin theta theorem bounds the angle between x and the
cluster's eigenspace: sin angle(x, E) <= ||r|| / gap.  For the whole set of
returned vectors X (orthonormal) and the wanted invariant subspace U of the
nev extreme eigenvalues, sin theta_max(X, U) <= ||R||_2 / delta with R =
A X - X Lambda and delta the distance from the returned values to the
unwanted part of the spectrum.  These are the bounds a solve to relative
residual tol can meet (an angle of 1e-8 would need ||r|| / gap <= 1e-8, far
below tol |lambda| / gap on these spectra), and they are asserted here
against the exact eigenspaces (closed forms for the Laplacian, dense LAPACK
otherwise) for the configurations' operator families at reduced sizes, in
FP64 and mixed precision, single solves and the batched solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eigh(A):
    w, U = np.linalg.eigh(A.to_dense())
    return w, U


def _per_vector(vals, X, Ad, w, U, slack=1.05):
    """Every returned vector against the exact eigenspace of its value's
    cluster (exact multiplicity): sin angle <= ||r|| / gap."""
    scale = np.max(np.abs(w))
    worst = 0.0
    for i, lam in enumerate(vals):
        x = X[:, i] / np.linalg.norm(X[:, i])
        r = np.linalg.norm(Ad @ x - lam * x)
        cl = np.abs(w - lam) <= 1e-9 * scale
        assert cl.any()
        gap = np.min(np.abs(w[~cl] - lam))
        E = U[:, cl]
        s = np.linalg.norm(x - E @ (E.T @ x))
        assert s <= slack * r / gap + 1e-13, (i, lam, s, r, gap)
        worst = max(worst, s)
    return worst


def _whole(vals, X, Ad, w, U, wanted, slack=1.05):
    """The returned subspace against the wanted invariant subspace:
    sin theta_max <= ||A X - X Lambda||_2 / delta / sigma_min(X), delta the
    distance from the returned values to the unwanted eigenvalues (skipped
    when a cluster straddles the boundary and delta vanishes)."""
    scale = np.max(np.abs(w))
    rest = np.setdiff1d(np.arange(len(w)), wanted)
    delta = np.min(np.abs(w[rest][:, None] - np.asarray(vals)[None, :]))
    Q, _ = np.linalg.qr(X)
    Uw = U[:, wanted]
    sin_sub = np.linalg.norm(Q - Uw @ (Uw.T @ Q), 2)
    if delta > 1e-9 * scale:
        R = Ad @ X - X * np.asarray(vals)[None, :]
        smin = np.min(np.linalg.svd(X, compute_uv=False))
        assert sin_sub <= slack * np.linalg.norm(R, 2) / delta / smin + 1e-12, (sin_sub, delta)
    return sin_sub


def _solve(sa, Ah, nev, ncv, tol, which, v0, mixed=False):
    A = sa.Mat(Ah, sa.FP64)
    A32 = sa.Mat(Ah, sa.FP32) if mixed else None
    E = sa.EPS()
    E.setOperators(A, A32)
    E.setProblemType(True)
    E.setWhichEigenpairs(which)
    E.setDimensions(nev, ncv)
    E.setTolerances(tol, 20000)
    E.setTrueResidual(True)
    E.setInitialSpace(v0)
    if mixed:
        E.setMixedPrecision(True, 1e-5)
    E.solve()
    nc = E.getConverged()
    assert nc >= nev
    vals, vecs = [], []
    for i in range(nev):
        lam, _, x, _ = E.getEigenpair(i, True)
        vals.append(lam)
        vecs.append(x)
    return np.array(vals), np.stack(vecs, axis=1), E


def test_subspace_c1_laplacian_closed_form(sa, orc):
    """C1 at full size (2-D Laplacian 100x100, 10 largest magnitude, ncv 32,
    tol 1e-8): every returned vector lies in the exact eigenspace of its
    value (closed-form sine modes, double eigenvalues as 2-D spaces) within
    the Davis-Kahan bound ||r|| / gap."""
    m = 100
    Ah = sa.lap2d(m)
    v0 = orc.rand_uniform(0, Ah.n)
    vals, X, E = _solve(sa, Ah, 10, 32, 1e-8, "LM", v0)
    k = np.arange(1, m + 1)
    c = 2 * np.cos(k * np.pi / (m + 1))
    lamgrid = 4 - c[:, None] - c[None, :]
    wl = lamgrid.ravel()
    S = np.sqrt(2.0 / (m + 1)) * np.sin(np.outer(np.arange(1, m + 1), k) * np.pi / (m + 1))   # S[x, i]
    M = sa.Mat(Ah)
    for i, lam in enumerate(vals):
        x = X[:, i] / np.linalg.norm(X[:, i])
        r = np.linalg.norm(np.asarray(M.mult_host(x)) - lam * x)
        cl = np.abs(wl - lam) <= 1e-9 * 8
        gap = np.min(np.abs(wl[~cl] - lam))
        idx = np.argwhere(np.abs(lamgrid - lam) <= 1e-9 * 8)
        # eigenvectors of the grid index ii = iy*m + ix: S[:, i] (x) S[:, j]
        Eb = np.stack([np.kron(S[:, a], S[:, b]) for a, b in idx], axis=1)
        s = np.linalg.norm(x - Eb @ (Eb.T @ x))
        assert s <= 1.05 * r / gap + 1e-13, (i, lam, s, r, gap)


@pytest.mark.parametrize("mixed", [False, True])
def test_subspace_c4_family(sa, orc, mixed):
    """C4's operator (FV -div(k_theta grad u), sigma I - L, 16 largest real,
    ncv 64, tol 1e-10) on a 48x48 grid, held-out theta: per-vector and
    whole-subspace Davis-Kahan bounds against dense LAPACK, plain FP64 and
    mixed FP32 -> FP64."""
    Ah = sa.fv2d(48, 48, sa.fv_theta(5, 0))
    Ao = orc.CSR(Ah.n, Ah.rowptr, Ah.col, Ah.val)
    w, U = _eigh(Ao)
    vals, X, E = _solve(sa, Ah, 16, 64, 1e-10, "LR", sa.rand_uniform(0, Ah.n), mixed)
    Ad = Ao.to_dense()
    _per_vector(vals, X, Ad, w, U)
    _whole(vals, X, Ad, w, U, np.argsort(-w)[:16])


@pytest.mark.parametrize("mixed", [False, True])
def test_subspace_c3_stencil27(sa, orc, mixed):
    """C3's operator (27-point stencil + diagonal U[0,1) seed 1, 32 largest
    magnitude, ncv 96) at 12^3, tol 1e-10: the Davis-Kahan bounds against
    dense LAPACK, FP64 and mixed precision."""
    Ah = sa.stencil27(12, seed=1)
    Ao = orc.CSR(Ah.n, Ah.rowptr, Ah.col, Ah.val)
    w, U = _eigh(Ao)
    vals, X, E = _solve(sa, Ah, 32, 96, 1e-10, "LM", sa.rand_uniform(0, Ah.n), mixed)
    Ad = Ao.to_dense()
    _per_vector(vals, X, Ad, w, U)
    _whole(vals, X, Ad, w, U, np.argsort(-np.abs(w))[:32])


def test_subspace_c5_batched(sa, orc):
    """C5's family (Laplacians with a diagonal potential U[0,10), 8 largest
    magnitude, ncv 24, tol 1e-10) at reduced sizes through the batched
    solver: per-vector Davis-Kahan bounds against dense LAPACK for every
    problem."""
    from slepc_amd.batch import Batch, uneven_laplacians
    csrs = uneven_laplacians(P=6, seed=2, lo=16, hi=40)
    B = Batch(csrs)
    B.set_true_residual(True)
    B.solve(8, 24, 1e-10, "LM", seed=4)
    for p, A in enumerate(csrs):
        Ao = orc.CSR(A.n, A.rowptr, A.col, A.val)
        w, U = _eigh(Ao)
        vals, vecs = [], []
        for i in range(8):
            lam, x = B.eigenpair(p, i, True)
            vals.append(lam)
            vecs.append(x)
        vals, X, Ad = np.array(vals), np.stack(vecs, axis=1), Ao.to_dense()
        _per_vector(vals, X, Ad, w, U)
        _whole(vals, X, Ad, w, U, np.argsort(-np.abs(w))[:8])
