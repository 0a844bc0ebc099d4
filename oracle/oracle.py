"""ctypes wrapper over liboracle.so -- the CPU FP64 restatement of SLEPc's
Krylov-Schur hot path.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  Never imported by slepc_amd/.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LM, SM, LR, SR = 1, 2, 3, 4
WHICH = {"LM": LM, "SM": SM, "LR": LR, "SR": SR, "LI": 5, "SI": 6, "TM": 7, "TR": 8, "USER": 11}


class Opts(C.Structure):
    _fields_ = [("hermitian", C.c_int), ("which", C.c_int), ("nev", C.c_int),
                ("ncv", C.c_int), ("mpd", C.c_int), ("max_it", C.c_int),
                ("tol", C.c_double), ("keep", C.c_double), ("eta", C.c_double),
                ("lock", C.c_int), ("nthreads", C.c_int), ("solver", C.c_int),
                ("extraction", C.c_int), ("target", C.c_double),
                ("st", C.c_int), ("sigma", C.c_double), ("lanczos_local", C.c_int),
                ("arnoldi_delayed", C.c_int), ("balance", C.c_int), ("balance_its", C.c_int),
                ("balance_cutoff", C.c_double), ("lanczos_partial", C.c_int), ("trueres", C.c_int),
                ("refine", C.c_int), ("conv", C.c_int)]

REFINE = {"ifneeded": 0, "never": 1, "always": 2}


class Result(C.Structure):
    _fields_ = [("nconv", C.c_int), ("its", C.c_int), ("reason", C.c_int),
                ("ncv_used", C.c_int), ("eigr", C.POINTER(C.c_double)),
                ("eigi", C.POINTER(C.c_double)), ("errest", C.POINTER(C.c_double)),
                ("vecs", C.POINTER(C.c_double)), ("nvec", C.c_int),
                ("matvecs", C.c_int64), ("orth_passes", C.c_int64),
                ("seconds", C.c_double), ("trace_nconv", C.POINTER(C.c_int)),
                ("trace_eig0", C.POINTER(C.c_double)), ("trace_cap", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_gauss.restype = C.c_double
        L.orc_gauss.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_sub_seed.restype = C.c_uint64
        L.orc_sub_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_rand_uniform.argtypes = [C.c_uint64, C.c_int64, f64p]
        L.orc_lapack_status.argtypes = [C.c_char_p, C.c_int]
        for name in ("orc_gen_lap2d", "orc_gen_lap2d_pot", "orc_gen_convdiff3d",
                     "orc_gen_stencil27", "orc_gen_fv2d", "orc_gen_markov"):
            getattr(L, name).restype = C.c_int64
        L.orc_spmv.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p]
        L.orc_bv_dotvec.argtypes = [C.c_int64, C.c_int, f64p, C.c_int64, f64p, f64p]
        L.orc_bv_multvec.argtypes = [C.c_int64, C.c_int, C.c_double, f64p, C.c_int64,
                                     f64p, C.c_double, f64p]
        L.orc_bv_multinplace.argtypes = [C.c_int64, f64p, C.c_int64, C.c_int, C.c_int,
                                         C.c_int, C.c_int, f64p, C.c_int]
        L.orc_bv_norm2.restype = C.c_double
        L.orc_bv_norm2.argtypes = [C.c_int64, f64p]
        L.orc_bv_orthonormalize_column.argtypes = [C.c_int64, f64p, C.c_int64, C.c_int, f64p,
                                                   C.c_int, C.c_double,
                                                   C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_arrow_tridiag.argtypes = [C.c_int, f64p, f64p, f64p, C.c_int]
        L.orc_bv_dot.argtypes = [C.c_int64, f64p, C.c_int64, C.c_int, f64p, C.c_int64, C.c_int, f64p, C.c_int]
        L.orc_bv_matmult.argtypes = [C.c_int, i32p, i32p, f64p, f64p, C.c_int64, f64p, C.c_int64, C.c_int]
        L.orc_bv_orthogonalize.argtypes = [C.c_int64, f64p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_void_p, C.c_int]
        L.orc_eps_solve.argtypes = [C.c_int, i32p, i32p, f64p, C.POINTER(Opts),
                                    C.c_void_p, C.POINTER(Result)]
        L.orc_eps_solve_twosided.argtypes = [C.c_int, i32p, i32p, f64p, C.POINTER(Opts),
                                             C.c_void_p, C.POINTER(Result), C.c_void_p]
        L.orc_residual_rel.restype = C.c_double
        L.orc_residual_rel.argtypes = [C.c_int, i32p, i32p, f64p, C.c_double, C.c_double,
                                       f64p, C.c_void_p]
        _LIB = L
    return _LIB


def lapack_path():
    buf = C.create_string_buffer(512)
    ok = lib().orc_lapack_status(buf, 512)
    return buf.value.decode() if ok else None


# ---------------------------------------------------------------- generators
class CSR:
    def __init__(self, n, rowptr, col, val, sigma=None):
        self.n, self.rowptr, self.col, self.val = n, rowptr, col, val
        self.sigma = sigma

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def to_dense(self):
        A = np.zeros((self.n, self.n))
        for r in range(self.n):
            s, e = self.rowptr[r], self.rowptr[r + 1]
            A[r, self.col[s:e]] = self.val[s:e]
        return A

    def matvec(self, x):
        y = np.empty(self.n)
        lib().orc_spmv(self.n, self.rowptr, self.col, self.val, np.ascontiguousarray(x, np.float64), y)
        return y


def _gen(fn, n, args, sigma=False):
    L = lib()
    f = getattr(L, fn)
    nul = C.c_void_p(0)
    sig = C.c_double(0.0)
    extra = [C.byref(sig)] if sigma else []
    nnz = f(*args, *extra, nul, nul, nul)
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    f(*args, *extra, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
      v.ctypes.data_as(C.c_void_p))
    return CSR(n, rp, ci, v, sig.value if sigma else None)


def lap2d(nx, ny=None):
    ny = nx if ny is None else ny
    return _gen("orc_gen_lap2d", nx * ny, [C.c_int(nx), C.c_int(ny)])


def lap2d_pot(nx, ny, seed, scale=10.0):
    return _gen("orc_gen_lap2d_pot", nx * ny, [C.c_int(nx), C.c_int(ny), C.c_uint64(seed),
                                               C.c_double(scale)])


def convdiff3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _gen("orc_gen_convdiff3d", nx * ny * nz, [C.c_int(nx), C.c_int(ny), C.c_int(nz)])


def stencil27(nx, ny=None, nz=None, seed=1):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _gen("orc_gen_stencil27", nx * ny * nz, [C.c_int(nx), C.c_int(ny), C.c_int(nz),
                                                    C.c_uint64(seed)])


def fv_theta(seed, sample):
    th = np.zeros(8)
    lib().orc_fv2d_theta(C.c_uint64(seed), C.c_int(sample), th.ctypes.data_as(C.c_void_p))
    return th


def fv2d(nx, ny, theta):
    th = np.ascontiguousarray(theta, np.float64)
    return _gen("orc_gen_fv2d", nx * ny, [C.c_int(nx), C.c_int(ny), th.ctypes.data_as(C.c_void_p)],
                sigma=True)


def markov(m):
    return _gen("orc_gen_markov", m * (m + 1) // 2, [C.c_int(m)])


def rand_uniform(seed, n):
    out = np.empty(n)
    lib().orc_rand_uniform(seed, n, out)
    return out


def u01(seed, i):
    return lib().orc_u01(seed, i)


def sub_seed(seed, p):
    return lib().orc_sub_seed(seed, p)


# ---------------------------------------------------------------- BV ops
def bv_dotvec(V, x):
    n, k = V.shape
    Vf = np.asfortranarray(V)
    q = np.empty(k)
    lib().orc_bv_dotvec(n, k, Vf.ravel(order="F"), n, np.ascontiguousarray(x, np.float64), q)
    return q


def build_balance(A, twoside=False, its=5, cutoff=1e-8):
    """EPSBuildBalance_Krylov: the diagonal D of D A D^-1."""
    D = np.zeros(A.n)
    f = lib().orc_build_balance
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p]
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    va = np.ascontiguousarray(A.val, np.float64)
    f(A.n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 1 if twoside else 0, its, cutoff, D.ctypes.data)
    return D


def bv_orthogonalize_some_column(V, j, which, refine="ifneeded", eta=0.7071):
    """BVOrthogonalizeSomeColumn (MGS against the columns i < j with which[i])
    with a refinement type: returns (V, coefficients h[0:j+1], norm, lindep,
    passes)."""
    n, k = V.shape
    Vf = np.asfortranarray(V, np.float64).copy(order="F")
    m = k
    buf = np.zeros(m * m)
    w = np.zeros(max(k, 1), np.uint8)
    w[:len(which)] = np.asarray(which, np.uint8)
    nrm, lind = C.c_double(), C.c_int()
    f = lib().orc_bv_orthogonalize_some_column
    f.restype = C.c_int
    f.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
    passes = f(n, Vf.ctypes.data, n, j, w.ctypes.data, buf.ctypes.data, m, REFINE[refine], eta, C.byref(nrm),
               C.byref(lind))
    h = buf.reshape(m, m, order="F")[:j + 1, j].copy()
    return Vf, h, nrm.value, lind.value, passes


def bv_multvec(V, alpha, q, beta, y):
    n, k = V.shape
    Vf = np.asfortranarray(V).ravel(order="F")
    y = np.array(y, np.float64)
    lib().orc_bv_multvec(n, k, alpha, Vf, n, np.ascontiguousarray(q, np.float64), beta, y)
    return y


def bv_multinplace(V, Q, s, e, l=0, k=None):
    n, m = V.shape
    k = m if k is None else k
    flat = np.array(V, order="F").ravel(order="F")
    Qf = np.asfortranarray(Q)
    lib().orc_bv_multinplace(n, flat, n, l, k, s, e, Qf.ravel(order="F"), Qf.shape[0])
    return flat.reshape((n, m), order="F")


def bv_orthonormalize_column(V, j, eta=0.7071):
    """Returns (V', norm, lindep, h) -- h the coefficient column incl. norm."""
    n, m = V.shape
    flat = np.array(V, order="F").ravel(order="F")
    buf = np.zeros(m * m)
    nrm = C.c_double()
    lind = C.c_int()
    lib().orc_bv_orthonormalize_column(n, flat, n, j, buf, m, eta, C.byref(nrm), C.byref(lind))
    return flat.reshape((n, m), order="F"), nrm.value, lind.value, buf[j * m:j * m + j + 1].copy()


def bv_dot(X, Y):
    """M = Y^T X (BVDot, all columns)."""
    Xf, Yf = np.asfortranarray(X, np.float64), np.asfortranarray(Y, np.float64)
    n, mx = Xf.shape
    my = Yf.shape[1]
    M = np.zeros((my, mx), order="F")
    lib().orc_bv_dot(n, Xf.ravel(order="F"), n, mx, Yf.ravel(order="F"), n, my, M.ravel(order="K"), my)
    return M


def bv_matmult(A, X):
    """Y = A X column by column (BVMatMult)."""
    Xf = np.asfortranarray(X, np.float64)
    n, nc = Xf.shape
    Y = np.zeros((n, nc), order="F")
    lib().orc_bv_matmult(A.n, A.rowptr, A.col, A.val, Xf.ravel(order="F"), n, Y.ravel(order="K"), n, nc)
    return Y


def bv_orthogonalize(V, l, k, method):
    """BVOrthogonalize of columns l..k-1 (method 0 gs, 1 chol, 2 svqb,
    3 tsqr, 4 tsqrchol).
    Returns (V', R) with R k x k (columns l..k-1 filled), or raises on breakdown."""
    n, m = V.shape
    flat = np.array(V, order="F").ravel(order="F")
    R = np.zeros((k, k), order="F")
    rc = lib().orc_bv_orthogonalize(n, flat, n, m, l, k, method, R.ctypes.data, k)
    if rc:
        raise RuntimeError("BVOrthogonalize breakdown or LAPACK failure")
    return flat.reshape((n, m), order="F"), R


def arrow_tridiag(d, e):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64)
    Q = np.eye(n).ravel(order="F").copy()
    lib().orc_arrow_tridiag(n, d, e, Q, n)
    return d, e, Q.reshape((n, n), order="F")


# ---------------------------------------------------------------- solve
def lanczos_local_sweep(A, m, v0):
    """One EPSLocalLanczos sweep (MGS against v_{j-1}, v_j only) from v0:
    (V, alpha, beta) with V of shape (n, steps + 1)."""
    V = np.zeros((m + 1, A.n))
    al, be = np.zeros(m), np.zeros(m)
    f = lib().orc_lanczos_local_sweep
    f.restype = C.c_int
    x0 = np.ascontiguousarray(v0, np.float64)
    P = C.c_void_p
    steps = f(A.n, P(A.rowptr.ctypes.data), P(A.col.ctypes.data), P(A.val.ctypes.data), m, P(x0.ctypes.data),
              P(V.ctypes.data), P(al.ctypes.data), P(be.ctypes.data))
    return V[:steps + 1].T.copy(), al[:steps], be[:steps]


_CMPFN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int), C.c_void_p)


def eps_solve_user_order(A, nev, cmp, **kw):
    """eps_solve with which = USER: cmp(ar, ai, br, bi) returns < 0 when a goes
    first (EPSSetEigenvalueComparison)."""
    def tramp(ar, ai, br, bi, res, _ctx):
        res[0] = int(cmp(ar, ai, br, bi))
        return 0
    cb = _CMPFN(tramp)
    lib().orc_set_user_comparison.argtypes = [_CMPFN, C.c_void_p]
    lib().orc_set_user_comparison(cb, None)
    try:
        return eps_solve(A, nev, which="USER", **kw)
    finally:
        lib().orc_set_user_comparison(_CMPFN(), None)


def eps_solve(A, nev, ncv=0, tol=1e-8, which="LM", hermitian=True, v0=None, max_it=0,
              mpd=0, nvec=None, nthreads=0, keep=0.5, lock=1, solver="krylovschur",
              extraction="ritz", target=0.0, twosided=False, st="shift", sigma=None,
              reorthog="full", balance="none", trueres=False, trace=0, refine="ifneeded", conv="rel"):
    o = Opts(hermitian=1 if hermitian else 0, which=WHICH[which], nev=nev, ncv=ncv, mpd=mpd,
             extraction={"ritz": 0, "harmonic": 1}[extraction], target=float(target),
             st={"shift": 0, "sinvert": 1}[st], sigma=float(target if sigma is None else sigma),
             lanczos_local=1 if reorthog == "local" else 0,
             arnoldi_delayed=1 if (solver == "arnoldi" and reorthog == "delayed") else 0,
             balance={"none": 0, "oneside": 1, "twoside": 2}[balance],
             lanczos_partial={"full": 0, "local": 0, "periodic": 1, "partial": 2, "selective": 3,
                               "delayed": 4}[reorthog],
             max_it=max_it, tol=tol, keep=keep, eta=0.7071, lock=lock, nthreads=nthreads,
             trueres=1 if trueres else 0, refine=REFINE[refine], conv={"rel": 0, "abs": 1, "norm": 2}[conv],
             solver={"krylovschur": 0, "arnoldi": 1, "lanczos": 2}[solver])
    ncv_eff = ncv if ncv else min(A.n, max(2 * nev, nev + 15))
    eigr = np.zeros(ncv_eff + 1)
    eigi = np.zeros(ncv_eff + 1)
    err = np.zeros(ncv_eff + 1)
    nvec = ncv_eff if nvec is None else nvec
    vecs = np.zeros(A.n * max(nvec, 1))
    r = Result()
    r.eigr = eigr.ctypes.data_as(C.POINTER(C.c_double))
    r.eigi = eigi.ctypes.data_as(C.POINTER(C.c_double))
    r.errest = err.ctypes.data_as(C.POINTER(C.c_double))
    r.vecs = vecs.ctypes.data_as(C.POINTER(C.c_double)) if nvec else None
    r.nvec = nvec
    tr_n = np.zeros(max(trace, 1), np.int32)
    tr_e = np.zeros(max(trace, 1))
    if trace:   # per-iteration (converged count, first Ritz value), the monitor's view
        r.trace_nconv = tr_n.ctypes.data_as(C.POINTER(C.c_int))
        r.trace_eig0 = tr_e.ctypes.data_as(C.POINTER(C.c_double))
        r.trace_cap = trace
    v0p = None
    if v0 is not None:
        v0 = np.ascontiguousarray(v0, np.float64)
        v0p = v0.ctypes.data_as(C.c_void_p)
    lvecs = np.zeros(A.n * max(nvec, 1)) if twosided else None
    if twosided:
        rc = lib().orc_eps_solve_twosided(A.n, A.rowptr, A.col, A.val, C.byref(o), v0p, C.byref(r),
                                          lvecs.ctypes.data_as(C.c_void_p))
    else:
        rc = lib().orc_eps_solve(A.n, A.rowptr, A.col, A.val, C.byref(o), v0p, C.byref(r))
    if rc != 0:
        raise RuntimeError("oracle eps_solve failed rc=%d" % rc)
    nc = r.nconv
    X = vecs[:A.n * min(nc, nvec)].reshape((min(nc, nvec), A.n)).T.copy() if nvec else None
    out = dict(nconv=nc, its=r.its, reason=r.reason, eigr=eigr[:nc].copy(), eigi=eigi[:nc].copy(),
               errest=err[:nc].copy(), vecs=X, matvecs=r.matvecs, passes=r.orth_passes,
               seconds=r.seconds)
    if trace:
        m = min(r.its, trace)
        out["trace_nconv"] = tr_n[:m].copy()
        out["trace_eig0"] = tr_e[:m].copy()
    if twosided:
        out["lvecs"] = lvecs[:A.n * min(nc, nvec)].reshape((min(nc, nvec), A.n)).T.copy() if nvec else None
    return out


def residual_abs(A, lr, li, xr, xi=None):
    """||A x - lambda x|| (EPSComputeResidualNorm_Private, trans = FALSE)."""
    xr = np.ascontiguousarray(xr, np.float64)
    xip = None
    if xi is not None:
        xi = np.ascontiguousarray(xi, np.float64)
        xip = xi.ctypes.data_as(C.c_void_p)
    f = lib().orc_residual_abs
    f.restype = C.c_double
    f.argtypes = lib().orc_residual_rel.argtypes
    return f(A.n, A.rowptr, A.col, A.val, lr, li, xr, xip)


def residual_rel(A, lr, li, xr, xi=None):
    xr = np.ascontiguousarray(xr, np.float64)
    xip = None
    if xi is not None:
        xi = np.ascontiguousarray(xi, np.float64)
        xip = xi.ctypes.data_as(C.c_void_p)
    return lib().orc_residual_rel(A.n, A.rowptr, A.col, A.val, lr, li, xr, xip)
