"""CPU pins of the workload generators that do not rest on the generators'
own formulas.  Each check is run on the oracle and on the product's host
generator, so a mistake shared by both (the two evaluate the same
definitions) still fails here.

- fv2d (finite-volume -div(k grad u), cell-centred, Dirichlet faces): at
  theta = 0 the spectrum of sigma I - L is 8 - mu_i - mu_j with
  mu_j = 2 - 2 cos(j pi / n), j = 1..n (the half-cell Dirichlet face gives the
  boundary coefficient 2k; eigenvectors sin(j pi (i - 1/2) / n)); sigma is
  the Gershgorin bound max_r sum_c |L_rc|; for any theta the cell
  conductivities recovered from the matrix through the harmonic-mean face
  rule are consistent on every face, and log k is the separable cosine
  expansion whose amplitudes are the Gaussian covariance's spectral density
  sqrt(l_p l_q), l_p = int exp(-r^2 / (2 ell^2)) cos(p pi r) dr, taken here by
  numerical quadrature.
- stencil27: A - diag(A) = -(J - I) with J the 27-point adjacency including
  the centre, J = T (x) T (x) T, T = tridiag(1, 1, 1), so its spectrum is
  1 - prod_d (1 + 2 cos(k_d pi / (N_d + 1))); the diagonal is 26 + u with u in
  [0, 1) drawn from splitmix64, whose first outputs are pinned by the
  published known answers of the generator.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate

ELL = 0.2
# the 8 leading non-constant separable modes (p, q) by sqrt(l_p l_q)
MODES = [(1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2), (2, 2)]


def _both(orc):
    import slepc_amd as sa
    return [("oracle", orc), ("product", sa)]


def _dense(A):
    D = np.zeros((A.n, A.n))
    for r in range(A.n):
        for p in range(A.rowptr[r], A.rowptr[r + 1]):
            D[r, A.col[p]] = A.val[p]
    return D


def _mu(n):
    return 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / n)


@pytest.mark.parametrize("nx,ny", [(12, 9), (7, 7), (1, 5)])
def test_fv2d_theta0_closed_form(orc, nx, ny):
    for name, g in _both(orc):
        A = g.fv2d(nx, ny, np.zeros(8))
        assert A.sigma == 8.0, name
        w = np.sort(np.linalg.eigvalsh(_dense(A)))
        ex = np.sort([8.0 - a - b for a in _mu(nx) for b in _mu(ny)])
        assert np.max(np.abs(w - ex)) < 1e-12, name


def test_fv2d_theta0_planted_boundary_factor_fails(orc):
    """A boundary face with coefficient k instead of 2k (planted in the
    oracle's output on the CPU) leaves the closed form: the check can fail."""
    A = orc.fv2d(6, 5, np.zeros(8))
    D = _dense(A)
    D[0, 0] += 1.0          # corner cell: one of its two boundary faces halved
    w = np.sort(np.linalg.eigvalsh(D))
    ex = np.sort([8.0 - a - b for a in _mu(6) for b in _mu(5)])
    assert np.max(np.abs(w - ex)) > 1e-3


def _lambda_quad(p):
    """Spectral density of exp(-r^2/(2 ell^2)) at frequency p pi, by quadrature."""
    f = lambda r: np.exp(-r * r / (2 * ELL * ELL)) * np.cos(p * np.pi * r)
    v, _ = integrate.quad(f, -12 * ELL, 12 * ELL, limit=200, epsabs=1e-14, epsrel=1e-12)
    return v


def _phi(p, x):
    return np.ones_like(x) if p == 0 else np.sqrt(2.0) * np.cos(p * np.pi * x)


def _recover_k(A, nx, ny):
    """Cell conductivities from sigma I - L: L = sigma I - A; interior faces
    carry t = 2 k0 k1 / (k0 + k1) (off-diagonal -t in L), a boundary face 2 k0
    (its share of the diagonal).  Boundary cells give k from their row; the
    others follow along faces by 1/k1 = 2/t - 1/k0.  Returns k and the
    largest relative inconsistency over all interior faces."""
    n = nx * ny
    D = A.sigma * np.eye(n) - _dense(A)
    k = np.full(n, np.nan)
    for r in range(n):
        j, i = divmod(r, nx)
        nb = (j == 0) + (i == 0) + (i == nx - 1) + (j == ny - 1)
        if nb:
            off = -(D[r].sum() - D[r, r])
            k[r] = (D[r, r] - off) / (2.0 * nb)
    changed = True
    while changed:
        changed = False
        for r in range(n):
            if not np.isnan(k[r]):
                continue
            j, i = divmod(r, nx)
            for c in (r - nx, r - 1, r + 1, r + nx):
                if 0 <= c < n and not np.isnan(k[c]) and D[r, c] != 0.0:
                    t = -D[r, c]
                    k[r] = 1.0 / (2.0 / t - 1.0 / k[c])
                    changed = True
                    break
    worst = 0.0
    for r in range(n):
        for c in (r + 1, r + nx):
            if c < n and D[r, c] != 0.0:
                t = -D[r, c]
                worst = max(worst, abs(2 * k[r] * k[c] / (k[r] + k[c]) - t) / t)
    return k, worst


@pytest.mark.parametrize("m", range(8))
def test_fv2d_kl_field_unit_modes(orc, m):
    nx, ny = 11, 8
    theta = np.zeros(8)
    theta[m] = 1.0
    p, q = MODES[m]
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    shape = np.outer(_phi(q, y), _phi(p, x)).ravel()
    amp = np.sqrt(_lambda_quad(p) * _lambda_quad(q))
    for name, g in _both(orc):
        A = g.fv2d(nx, ny, theta)
        k, worst = _recover_k(A, nx, ny)
        assert worst < 1e-12, name
        assert np.max(np.abs(np.log(k) - amp * shape)) < 1e-11, (name, m)
        # sigma: the Gershgorin bound of L
        L = A.sigma * np.eye(A.n) - _dense(A)
        assert abs(np.max(np.abs(L).sum(axis=1)) - A.sigma) <= 1e-13 * A.sigma, name


def test_fv2d_kl_field_random_theta(orc):
    nx, ny = 9, 13
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    lam = {p: _lambda_quad(p) for p in range(3)}
    for s in range(2):
        theta = orc.fv_theta(5, s)
        logk = sum(theta[m] * np.sqrt(lam[p] * lam[q]) * np.outer(_phi(q, y), _phi(p, x)).ravel()
                   for m, (p, q) in enumerate(MODES))
        for name, g in _both(orc):
            A = g.fv2d(nx, ny, theta)
            k, worst = _recover_k(A, nx, ny)
            assert worst < 1e-12, name
            assert np.max(np.abs(np.log(k) - logk)) < 1e-11, name


def test_fv2d_modes_are_the_leading_non_constant_ones():
    """The mode table is the 8 largest sqrt(l_p l_q) over (p, q) != (0, 0)."""
    lam = [_lambda_quad(p) for p in range(6)]
    pairs = sorted(((lam[p] * lam[q], p, q) for p in range(6) for q in range(6) if (p, q) != (0, 0)),
                   reverse=True)
    assert sorted((p, q) for _, p, q in pairs[:8]) == sorted(MODES)
    assert pairs[7][0] > pairs[8][0] * 1.05     # the eighth is not a tie with the ninth


def test_fv2d_harmonic_mean_planted_mistake_fails(orc):
    """Arithmetic instead of harmonic face means (planted on the CPU in a copy
    of the oracle's matrix) breaks the face consistency the recovery checks."""
    nx, ny = 6, 5
    A = orc.fv2d(nx, ny, orc.fv_theta(5, 0))
    k, worst = _recover_k(A, nx, ny)
    assert worst < 1e-12
    B = type(A)(A.n, A.rowptr.copy(), A.col.copy(), A.val.copy(), A.sigma)
    for r in range(A.n):
        for p in range(B.rowptr[r], B.rowptr[r + 1]):
            c = B.col[p]
            if c != r:
                B.val[p] = 0.5 * (k[r] + k[c])
    _, worst_b = _recover_k(B, nx, ny)
    assert worst_b > 1e-6


@pytest.mark.parametrize("dims", [(6, 5, 4), (3, 3, 3), (2, 7, 1)])
def test_stencil27_offdiagonal_spectrum(orc, dims):
    nx, ny, nz = dims
    ev = [1 + 2 * np.cos(np.arange(1, d + 1) * np.pi / (d + 1)) for d in dims]
    ex = np.sort([1.0 - a * b * c for a in ev[0] for b in ev[1] for c in ev[2]])
    for name, g in _both(orc):
        A = g.stencil27(nx, ny, nz, seed=1)
        D = _dense(A)
        dg = np.diag(D).copy()
        assert np.all(dg >= 26.0) and np.all(dg < 27.0), name
        assert np.unique(dg).size == A.n, name        # one draw per row
        O = D - np.diag(dg)
        assert np.array_equal(O, O.T), name
        assert set(np.unique(O)) <= {0.0, -1.0}, name
        w = np.sort(np.linalg.eigvalsh(O))
        assert np.max(np.abs(w - ex)) < 1e-12, name


def test_splitmix64_known_answers(orc):
    """splitmix64 (Steele, Lea, Flood 2014; Vigna's reference code): outputs
    of the generator seeded with 0 and with 1234567, as published."""
    L = orc.lib()
    f = L.orc_splitmix64
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64]
    gamma = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1
    for seed, expect in ((0, [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]),
                         (1234567, [6457827717110365317, 3203168211198807973, 9817491932198370423])):
        state = seed
        for e in expect:
            assert f(state) == e
            state = (state + gamma) & mask
