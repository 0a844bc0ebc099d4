"""slepc_amd -- MI355X-native Krylov-Schur hot path.

Python host mirror of the reference's operator interface (names follow
SLEPc/slepc4py: ``Mat``, ``BV``, ``EPS`` with ``setOperators``,
``setDimensions``, ``setInitialSpace``, ``solve``, ``getEigenpair``,
``computeError``).  Every numeric operation runs in libslepc_amd.so on the GPU;
this module only moves arguments across the C ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (FP32, FP64, LARGEST_MAGNITUDE, LARGEST_REAL, SMALLEST_MAGNITUDE, SMALLEST_REAL,
                   ERROR_ABSOLUTE, ERROR_RELATIVE, ERROR_BACKWARD, SlepcAmdError, check, lib)

__all__ = ["Mat", "BV", "EPS", "CSR", "init", "lap2d", "convdiff3d", "stencil27", "fv2d", "markov",
           "fv_theta", "rand_uniform", "FP32", "FP64", "SlepcAmdError"]

WHICH = {"LM": LARGEST_MAGNITUDE, "SM": SMALLEST_MAGNITUDE, "LR": LARGEST_REAL, "SR": SMALLEST_REAL,
         "LI": 5, "SI": 6, "TM": 7, "TR": 8}


def init(device=0):
    check(lib().sa_init(device))


# ------------------------------------------------------------------ CSR host data
class CSR:
    """Host CSR arrays (int32 row pointers / columns, float64 values)."""

    def __init__(self, n, rowptr, col, val, sigma=None):
        self.n, self.rowptr, self.col, self.val, self.sigma = n, rowptr, col, val, sigma

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def _gen(fname, n, args, with_sigma=False):
    f = getattr(lib(), fname)
    sig = C.c_double(0.0)
    extra = [C.byref(sig)] if with_sigma else []
    nnz = f(*args, *extra, None, None, None)
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    f(*args, *extra, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
    return CSR(n, rp, ci, v, sig.value if with_sigma else None)


def lap2d(nx, ny=None, pot_scale=0.0, seed=0):
    ny = nx if ny is None else ny
    return _gen("sa_gen_lap2d", nx * ny, [nx, ny, pot_scale, seed])


def convdiff3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _gen("sa_gen_convdiff3d", nx * ny * nz, [nx, ny, nz])


def stencil27(nx, ny=None, nz=None, seed=1):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _gen("sa_gen_stencil27", nx * ny * nz, [nx, ny, nz, seed])


def fv2d(nx, ny, theta):
    th = np.ascontiguousarray(theta, np.float64)
    return _gen("sa_gen_fv2d", nx * ny, [nx, ny, th], with_sigma=True)


def markov(m):
    return _gen("sa_gen_markov", m * (m + 1) // 2, [m])


def fv_theta(seed, sample):
    th = np.zeros(8)
    lib().sa_fv2d_theta(seed, sample, th)
    return th


def rand_uniform(seed, n):
    out = np.empty(n)
    lib().sa_rand_uniform(seed, n, out)
    return out


def csr_assemble(n, rows, cols, vals):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    L = lib()
    nnz = L.sa_csr_assemble(n, len(rows), rows, cols, vals, None, None, None)
    if nnz < 0:
        raise SlepcAmdError("index out of range in csr_assemble")
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    L.sa_csr_assemble(n, len(rows), rows, cols, vals, rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
    return CSR(n, rp, ci, v)


# ------------------------------------------------------------------ file formats
def load_petsc_binary(path):
    """MatLoad of a PETSc binary AIJ file -> CSR (square matrices)."""
    L = lib()
    M, N, nz = C.c_int(), C.c_int(), C.c_int64()
    bp = path.encode()
    check(L.sa_mat_load_petsc_binary(bp, C.byref(M), C.byref(N), C.byref(nz), None, None, None))
    if M.value != N.value:
        raise SlepcAmdError("matrix in %s is %d x %d, not square" % (path, M.value, N.value))
    rp = np.zeros(M.value + 1, np.int32)
    ci = np.zeros(nz.value, np.int32)
    v = np.zeros(nz.value, np.float64)
    check(L.sa_mat_load_petsc_binary(bp, C.byref(M), C.byref(N), C.byref(nz), rp.ctypes.data, ci.ctypes.data,
                                     v.ctypes.data))
    return CSR(M.value, rp, ci, v)


def load_petsc_vec(path, index=0):
    """VecLoad of the index-th vector of a PETSc binary file."""
    L = lib()
    n = C.c_int64()
    bp = path.encode()
    check(L.sa_vec_load_petsc_binary(bp, index, C.byref(n), None))
    x = np.zeros(n.value)
    check(L.sa_vec_load_petsc_binary(bp, index, C.byref(n), x.ctypes.data))
    return x


def format_eigenvalues(re, im=None, fmt="ascii", name="eps"):
    """Text of EPSValuesView for the given eigenvalues ("ascii" or "matlab")."""
    re = np.ascontiguousarray(re, np.float64)
    im = np.zeros_like(re) if im is None else np.ascontiguousarray(im, np.float64)
    f = {"ascii": 0, "matlab": 1}[fmt]
    L = lib()
    need = L.sa_format_eigenvalues(len(re), re.ctypes.data, im.ctypes.data, f, name.encode(), None, 0)
    if need < 0:
        raise SlepcAmdError(L.sa_last_error().decode())
    buf = C.create_string_buffer(int(need) + 1)
    L.sa_format_eigenvalues(len(re), re.ctypes.data, im.ctypes.data, f, name.encode(), buf, need + 1)
    return buf.value.decode()


# ------------------------------------------------------------------ device objects
class DeviceArray:
    """A device buffer owned by the library allocator."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().sa_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes)
        check(lib().sa_memcpy_h2d(d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def to_host(self, dtype, count):
        out = np.empty(count, dtype)
        check(lib().sa_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().sa_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def set_spmv_stream_bytes(nbytes):
    """Operators whose values and indices exceed `nbytes` are created for the
    stream SpMV (row blocks staged in LDS); returns the previous threshold."""
    return int(lib().sa_set_spmv_stream_bytes(int(nbytes)))


def set_concurrent_solves(nsolves):
    """Partition hint: number of solves kept in flight at once (one host
    thread each); with 2 or more, cache-resident bases created afterwards use
    half as many Gram-Schmidt workgroups.  Returns the previous value."""
    return int(lib().sa_set_concurrent_solves(int(nsolves)))


class Mat:
    """CSR operator in HBM (MATAIJ analogue); MatMult runs the HIP SpMV."""

    def __init__(self, csr, prec=FP64):
        self.h = C.c_void_p()
        self.n, self.prec = csr.n, prec
        check(lib().sa_mat_create_csr(csr.n, csr.nnz, csr.rowptr, csr.col, csr.val, prec, C.byref(self.h)))

    def mult(self, x_dev, y_dev):
        check(lib().sa_mat_mult(self.h, x_dev.ptr if isinstance(x_dev, DeviceArray) else x_dev,
                                y_dev.ptr if isinstance(y_dev, DeviceArray) else y_dev))

    def mult_host(self, x):
        dt = np.float64 if self.prec == FP64 else np.float32
        xd = DeviceArray.from_host(np.asarray(x, dt))
        yd = DeviceArray(self.n * np.dtype(dt).itemsize)
        self.mult(xd, yd)
        return yd.to_host(dt, self.n)

    def build_balance(self, twoside=False, its=5, cutoff=1e-8):
        """EPSBuildBalance_Krylov: the diagonal D of D A D^-1 (host array)."""
        D = np.zeros(self.n)
        check(lib().sa_mat_build_balance(self.h, 1 if twoside else 0, its, cutoff, D.ctypes.data))
        return D

    def residual_norms(self, X, lr, li=None):
        """||A x_i - lambda_i x_i|| for the columns of X (host array n x k,
        uploaded; a complex pair occupies two columns, real then imaginary
        part, with li[i] != 0), as EPSComputeResidualNorm_Private computes
        them (sa_mat_residual_norms)."""
        dt = np.float64 if self.prec == FP64 else np.float32
        X = np.asarray(X, dt)
        if X.ndim == 1:
            X = X[:, None]
        k = X.shape[1]
        xd = DeviceArray.from_host(np.ascontiguousarray(X.T))
        lr = np.ascontiguousarray(lr, np.float64)
        li = None if li is None else np.ascontiguousarray(li, np.float64)
        res = np.zeros(k)
        check(lib().sa_mat_residual_norms(self.h, k, xd.ptr, self.n, lr.ctypes.data,
                                          None if li is None else li.ctypes.data, res.ctypes.data))
        return res

    def destroy(self):
        if self.h:
            lib().sa_mat_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BV:
    """Basis of m columns of length n in HBM (BVMAT analogue)."""

    def __init__(self, n, m, prec=FP64):
        self.h = C.c_void_p()
        self.n, self.m, self.prec = n, m, prec
        check(lib().sa_bv_create(n, m, prec, C.byref(self.h)))
        ld = C.c_int64()
        check(lib().sa_bv_get_info(self.h, None, None, C.byref(ld), None))
        self.ld = ld.value

    def set_active_columns(self, l, k):
        check(lib().sa_bv_set_active_columns(self.h, l, k))

    def get_active_columns(self):
        l, k = C.c_int(), C.c_int()
        check(lib().sa_bv_get_active_columns(self.h, C.byref(l), C.byref(k)))
        return l.value, k.value

    def set_column(self, j, x):
        check(lib().sa_bv_set_column(self.h, j, np.ascontiguousarray(x, np.float64)))

    def get_column(self, j):
        out = np.empty(self.n)
        check(lib().sa_bv_get_column(self.h, j, out))
        return out

    def set_matrix(self, X):
        for j in range(X.shape[1]):
            self.set_column(j, X[:, j])

    def get_matrix(self, ncols=None):
        ncols = self.m if ncols is None else ncols
        return np.stack([self.get_column(j) for j in range(ncols)], axis=1)

    def column_ptr(self, j):
        base = C.c_void_p()
        check(lib().sa_bv_get_array(self.h, C.byref(base)))
        es = 8 if self.prec == FP64 else 4
        return C.c_void_p(base.value + j * self.ld * es)

    def dotvec(self, x_dev):
        l, k = C.c_int(), C.c_int()
        check(lib().sa_bv_get_active_columns(self.h, C.byref(l), C.byref(k)))
        q = np.zeros(max(k.value - l.value, 1))
        check(lib().sa_bv_dotvec(self.h, x_dev, q.ctypes.data))
        return q[:k.value - l.value]

    def multvec(self, alpha, beta, y_dev, q):
        q = np.ascontiguousarray(q, np.float64)
        check(lib().sa_bv_multvec(self.h, alpha, beta, y_dev, q.ctypes.data))

    def multinplace(self, Q, s, e):
        Qf = np.asfortranarray(Q, np.float64)
        check(lib().sa_bv_multinplace(self.h, Qf.ravel(order="F"), Qf.shape[0], s, e))

    def dot(self, Y):
        """BVDot(self, Y, M): M = Y^T self over the active columns, returned as
        a Y.k x self.k array (rows below Y.l and columns below self.l zero)."""
        l, k = self.get_active_columns()
        ly, ky = Y.get_active_columns()
        M = np.zeros((ky, k), order="F")
        check(lib().sa_bv_dot(self.h, Y.h, M.ctypes.data, ky))
        return M

    def matmult(self, A, Y):
        """BVMatMult(self, A, Y): Y(:, Y.l+j) = A self(:, self.l+j)."""
        check(lib().sa_bv_matmult(self.h, A.h, Y.h))

    def orthogonalize(self, method="chol", with_R=True):
        """BVOrthogonalize of the active columns with block method gs, chol,
        svqb, tsqr or tsqrchol; returns the k x k factor R (columns l..k-1
        filled) or None."""
        m = {"gs": 0, "chol": 1, "svqb": 2, "tsqr": 3, "tsqrchol": 4}[method]
        l, k = self.get_active_columns()
        R = np.zeros((k, k), order="F") if with_R else None
        check(lib().sa_bv_orthogonalize(self.h, m, R.ctypes.data if with_R else None, k))
        return R

    def scale_column(self, j, alpha):
        check(lib().sa_bv_scale_column(self.h, j, alpha))

    def norm_column(self, j):
        v = C.c_double()
        check(lib().sa_bv_norm_column(self.h, j, C.byref(v)))
        return v.value

    def norm(self):
        v = C.c_double()
        check(lib().sa_bv_norm(self.h, C.byref(v)))
        return v.value

    def copy_column(self, src, dst):
        check(lib().sa_bv_copy_column(self.h, src, dst))

    def mat_mult_column(self, A, j):
        check(lib().sa_bv_mat_mult_column(self.h, A.h, j))

    def set_orthogonalization(self, refine="ifneeded"):
        """BVSetOrthogonalization's refinement: "ifneeded", "never", "always"."""
        check(lib().sa_bv_set_orthogonalization(self.h, REFINE[refine]))

    def orthogonalize_some_column(self, j, which, eta=0.7071):
        """BVOrthogonalizeSomeColumn (MGS against the columns i < j with
        which[i]); returns (coefficients with the norm at index j, norm, lindep)."""
        w = np.zeros(max(j, 1), np.int32)
        w[:j] = np.asarray(which[:j], np.int32)
        h = np.zeros(j + 1)
        nrm, lind = C.c_double(), C.c_int()
        check(lib().sa_bv_orthogonalize_some_column(self.h, j, w.ctypes.data, eta, h.ctypes.data, C.byref(nrm),
                                                    C.byref(lind)))
        return h, nrm.value, bool(lind.value)

    def orthonormalize_column(self, j, eta=0.7071):
        nrm, lind = C.c_double(), C.c_int()
        check(lib().sa_bv_orthonormalize_column(self.h, j, eta, C.byref(nrm), C.byref(lind)))
        return nrm.value, bool(lind.value)

    def coefficients(self, c0, c1):
        buf = np.zeros(self.m * max(c1 - c0, 1))
        check(lib().sa_bv_get_coefficients(self.h, c0, c1, buf))
        return buf[:self.m * (c1 - c0)].reshape((c1 - c0, self.m)).T

    def mat_arnoldi(self, A, k, m, eta=0.7071):
        mm, beta, bd = C.c_int(m), C.c_double(), C.c_int()
        check(lib().sa_bv_mat_arnoldi(self.h, A.h, k, C.byref(mm), eta, C.byref(beta), C.byref(bd)))
        return mm.value, beta.value, bool(bd.value)

    def destroy(self):
        if self.h:
            lib().sa_bv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# BVOrthogRefineType (BVSetOrthogonalization)
REFINE = {"ifneeded": 0, "never": 1, "always": 2}

# EPSConvergedReason (include/slepceps.h:179-185)
REASON_CONVERGED_TOL, REASON_CONVERGED_USER, REASON_ITERATING = 1, 2, 0
REASON_DIVERGED_ITS, REASON_DIVERGED_BREAKDOWN = -1, -2


class EPS:
    """Krylov-Schur eigensolver (EPSKRYLOVSCHUR default variant)."""

    def __init__(self):
        self.h = C.c_void_p()
        check(lib().sa_eps_create(C.byref(self.h)))
        self.A = self.A32 = None

    def setOperators(self, A, A32=None):
        self.A = A
        check(lib().sa_eps_set_operators(self.h, A.h))
        if A32 is not None:
            self.A32 = A32
            check(lib().sa_eps_set_operators_fp32(self.h, A32.h))

    def setProblemType(self, hermitian):
        check(lib().sa_eps_set_problem_type(self.h, 1 if hermitian else 0))

    def setType(self, eps_type):
        """EPSSetType: "krylovschur" (default), "arnoldi" or "lanczos" (full
        reorthogonalization)."""
        check(lib().sa_eps_set_type(self.h, {"krylovschur": 0, "arnoldi": 1, "lanczos": 2}[eps_type]))

    def setWhichEigenpairs(self, which):
        """EPSSetWhichEigenpairs: LM, SM, LR, SR, LI, SI, TM (closest to the
        target) or TR (closest real part)."""
        check(lib().sa_eps_set_which(self.h, WHICH[which] if isinstance(which, str) else which))

    def setTarget(self, target):
        check(lib().sa_eps_set_target(self.h, float(target)))

    def setST(self, st_type, shift=None):
        """STSetType ("shift" or "sinvert") and STSetShift (default: the target)."""
        check(lib().sa_eps_set_st(self.h, {"shift": 0, "sinvert": 1}[st_type], 0 if shift is None else 1,
                                  0.0 if shift is None else float(shift)))

    def setSTKSP(self, pc="lu", rtol=1e-12, max_it=10000):
        """Linear solver of the spectral transformation: "lu" (the default,
        KSPPREONLY + PCLU: banded LU on the device) or BiCGStab with a Jacobi
        ("jacobi") or no ("none") preconditioner."""
        check(lib().sa_eps_set_st_ksp(self.h, {"none": 0, "jacobi": 1, "lu": 2}[pc], float(rtol), int(max_it)))

    def setSTKSPType(self, ksp_type="bcgs", restart=30):
        """KSPSetType of the iterative solves: "bcgs" (BiCGStab) or "gmres"
        (GMRES(restart), right preconditioned, as KSPGMRES)."""
        check(lib().sa_eps_set_st_ksp_type(self.h, {"bcgs": 0, "gmres": 1}[ksp_type], int(restart)))

    def setBalance(self, balance="twoside", its=5, cutoff=1e-8):
        """EPSSetBalance: "none", "oneside" or "twoside" (non-Hermitian problems)."""
        check(lib().sa_eps_set_balance(self.h, {"none": 0, "oneside": 1, "twoside": 2}[balance], int(its),
                                       float(cutoff)))

    def setTrueResidual(self, trueres=True):
        """EPSSetTrueResidual: the convergence test takes explicit residuals
        of the Ritz vectors (Krylov-Schur, Arnoldi)."""
        check(lib().sa_eps_set_true_residual(self.h, 1 if trueres else 0))

    def getTrueResidual(self):
        v = C.c_int()
        check(lib().sa_eps_get_true_residual(self.h, C.byref(v)))
        return bool(v.value)

    def setKrylovSchurRestart(self, keep):
        """EPSKrylovSchurSetRestart: fraction of the basis kept at restart."""
        check(lib().sa_eps_krylovschur_set_restart(self.h, float(keep)))

    def getKrylovSchurRestart(self):
        v = C.c_double()
        check(lib().sa_eps_krylovschur_get_restart(self.h, C.byref(v)))
        return v.value

    def setKrylovSchurLocking(self, lock):
        """EPSKrylovSchurSetLocking: locking (True, default) or non-locking."""
        check(lib().sa_eps_krylovschur_set_locking(self.h, 1 if lock else 0))

    def getKrylovSchurLocking(self):
        v = C.c_int()
        check(lib().sa_eps_krylovschur_get_locking(self.h, C.byref(v)))
        return bool(v.value)

    def setArnoldiDelayed(self, delayed=True):
        """EPSArnoldiSetDelayed: delayed reorthogonalization in EPSARNOLDI."""
        check(lib().sa_eps_arnoldi_set_delayed(self.h, 1 if delayed else 0))

    def getArnoldiDelayed(self):
        """EPSArnoldiGetDelayed."""
        v = C.c_int()
        check(lib().sa_eps_arnoldi_get_delayed(self.h, C.byref(v)))
        return bool(v.value)

    def setLanczosReorthog(self, reorthog):
        """EPSLanczosSetReorthog: "local", "full" (the default here), "selective",
        "periodic", "partial" or "delayed"."""
        check(lib().sa_eps_lanczos_set_reorthog(self.h, {"local": 0, "full": 1, "selective": 2,
                                                         "periodic": 3, "partial": 4, "delayed": 5}[reorthog]))

    def getLanczosReorthog(self):
        """EPSLanczosGetReorthog: "local", "full", "selective", "periodic", "partial"
        or "delayed"."""
        v = C.c_int()
        check(lib().sa_eps_lanczos_get_reorthog(self.h, C.byref(v)))
        return {0: "local", 1: "full", 2: "selective", 3: "periodic", 4: "partial", 5: "delayed"}[v.value]

    def setOrthogonalization(self, refine="ifneeded", eta=0.0):
        """EPSGetBV + BVSetOrthogonalization: refinement and eta (0 keeps 0.7071)."""
        check(lib().sa_eps_set_orthogonalization(self.h, REFINE[refine], eta))

    def setTwoSided(self, twosided=True):
        """EPSSetTwoSided: also compute left eigenvectors (two-sided Krylov-Schur)."""
        check(lib().sa_eps_set_two_sided(self.h, 1 if twosided else 0))

    def getLeftEigenvector(self, i):
        """EPSGetLeftEigenvector: (yr, yi) of eigenpair i."""
        yr, yi = np.empty(self.A.n), np.empty(self.A.n)
        check(lib().sa_eps_get_left_eigenvector(self.h, i, yr.ctypes.data, yi.ctypes.data))
        return yr, yi

    def setExtraction(self, extraction):
        """EPSSetExtraction: "ritz" (default) or "harmonic"."""
        check(lib().sa_eps_set_extraction(self.h, {"ritz": 0, "harmonic": 1}[extraction]))

    def setDimensions(self, nev, ncv=0, mpd=0):
        check(lib().sa_eps_set_dimensions(self.h, nev, ncv, mpd))

    def setTolerances(self, tol=0.0, max_it=0):
        check(lib().sa_eps_set_tolerances(self.h, tol, max_it))

    def setInitialSpace(self, vecs):
        V = np.ascontiguousarray(np.atleast_2d(np.asarray(vecs, np.float64)))
        if V.shape[1] != self.A.n:
            V = np.ascontiguousarray(V.T)
        check(lib().sa_eps_set_initial_space(self.h, V.shape[0], V.ravel()))

    def setDeflationSpace(self, vecs):
        """EPSSetDeflationSpace (src/eps/interface/epssetup.c:544-583): the
        solve looks for eigenpairs orthogonal to these vectors (rows, or
        columns of an n x k array).  The space is removed by the next solve."""
        V = np.ascontiguousarray(np.atleast_2d(np.asarray(vecs, np.float64)))
        if V.shape[1] != self.A.n:
            V = np.ascontiguousarray(V.T)
        check(lib().sa_eps_set_deflation_space(self.h, V.shape[0], V.ravel()))

    def setMixedPrecision(self, mixed=True, tol32=1e-5):
        check(lib().sa_eps_set_mixed(self.h, 1 if mixed else 0, tol32))

    def solve(self):
        check(lib().sa_eps_solve(self.h))

    def getConverged(self):
        v = C.c_int()
        check(lib().sa_eps_get_converged(self.h, C.byref(v)))
        return v.value

    def getIterationNumber(self):
        v = C.c_int()
        check(lib().sa_eps_get_iteration_number(self.h, C.byref(v)))
        return v.value

    def getConvergedReason(self):
        v = C.c_int()
        check(lib().sa_eps_get_converged_reason(self.h, C.byref(v)))
        return v.value

    # EPSMonitorSet / EPSMonitorCancel (epsmon.c:96-135): fn(its, nconv, eigr,
    # eigi, errest, nest) with numpy copies of the first nest values; return
    # None or 0 to continue, anything else ends the solve with an error.
    _MONITOR = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.c_int, C.c_void_p)
    _STOPPING = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                            C.c_void_p)
    _CONVERGED = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                             C.c_void_p)
    CONV = {"abs": 0, "rel": 1, "norm": 2}

    def setConvergenceTest(self, conv):
        """EPSSetConvergenceTest: "abs" (||r||), "rel" (||r|| / |lambda|, the
        default) or "norm" (||r|| / (||A||_inf + |lambda| ||B||))."""
        check(lib().sa_eps_set_convergence_test(self.h, self.CONV[conv]))

    def setConvergenceTestFunction(self, fn):
        """EPSSetConvergenceTestFunction: fn(eigr, eigi, res) returns the error
        estimate; None restores the relative test."""
        if fn is None:
            self._conv_cb = None
            check(lib().sa_eps_set_convergence_test_function(self.h, None, None))
            return

        def tramp(_eps, er, ei, res, errest, _ctx):
            try:
                errest[0] = float(fn(er, ei, res))
                return 0
            except Exception:
                return 1
        self._conv_cb = self._CONVERGED(tramp)
        check(lib().sa_eps_set_convergence_test_function(self.h, C.cast(self._conv_cb, C.c_void_p), None))

    _COMPARE = C.CFUNCTYPE(C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int),
                           C.c_void_p)

    def setEigenvalueComparison(self, fn):
        """EPSSetEigenvalueComparison: fn(ar, ai, br, bi) < 0 when a goes first
        (which becomes EPS_WHICH_USER); None restores largest magnitude."""
        if fn is None:
            self._cmp_cb = None
            check(lib().sa_eps_set_eigenvalue_comparison(self.h, None, None))
            return

        def tramp(ar, ai, br, bi, res, _ctx):
            try:
                res[0] = int(fn(ar, ai, br, bi))
                return 0
            except Exception:
                return 1
        self._cmp_cb = self._COMPARE(tramp)
        check(lib().sa_eps_set_eigenvalue_comparison(self.h, C.cast(self._cmp_cb, C.c_void_p), None))

    def getMatrixNorms(self):
        """(||A||_inf, ||B||_inf) as computed for the norm test or a backward
        error; 0 until computed."""
        a, b = C.c_double(), C.c_double()
        check(lib().sa_eps_get_matrix_norms(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def setMonitor(self, fn):
        def tramp(_eps, its, nconv, er, ei, err, nest, _ctx):
            try:
                k = max(nest, 0)
                rc = fn(its, nconv, np.ctypeslib.as_array(er, (k,)).copy() if k else np.zeros(0),
                        np.ctypeslib.as_array(ei, (k,)).copy() if k else np.zeros(0),
                        np.ctypeslib.as_array(err, (k,)).copy() if k else np.zeros(0), nest)
                return int(rc or 0)
            except Exception:
                return 1
        cb = self._MONITOR(tramp)
        self._callbacks = getattr(self, "_callbacks", []) + [cb]    # kept alive with the EPS
        check(lib().sa_eps_monitor_set(self.h, C.cast(cb, C.c_void_p), None))

    def cancelMonitor(self):
        check(lib().sa_eps_monitor_cancel(self.h))

    # EPSSetStoppingTestFunction (epsopts.c:863-885): fn(its, max_it, nconv,
    # nev) returns the converged reason (REASON_ITERATING to continue); None
    # restores EPSStoppingBasic.
    def setStoppingTest(self, fn):
        if fn is None:
            self._stop_cb = None
            check(lib().sa_eps_set_stopping_test_function(self.h, None, None))
            return

        def tramp(_eps, its, max_it, nconv, nev, reason, _ctx):
            try:
                reason[0] = int(fn(its, max_it, nconv, nev))
                return 0
            except Exception:
                return 1
        self._stop_cb = self._STOPPING(tramp)
        check(lib().sa_eps_set_stopping_test_function(self.h, C.cast(self._stop_cb, C.c_void_p), None))

    def getEigenpair(self, i, vectors=True):
        re, im = C.c_double(), C.c_double()
        xr = np.empty(self.A.n) if vectors else None
        xi = np.empty(self.A.n) if vectors else None
        check(lib().sa_eps_get_eigenpair(self.h, i, C.byref(re), C.byref(im),
                                         xr.ctypes.data if vectors else None,
                                         xi.ctypes.data if vectors else None))
        return re.value, im.value, xr, xi

    def getEigenvalues(self):
        out = []
        for i in range(self.getConverged()):
            re, im, _, _ = self.getEigenpair(i, vectors=False)
            out.append(complex(re, im) if im else re)
        return out

    def valuesView(self, fmt="ascii", name="eps", path=None):
        """EPSValuesView (epsview.c:530-640): "ascii" or "matlab" text, or
        "binary" (the converged eigenvalues as big-endian complex doubles, as
        PetscViewerBinaryWrite of PETSC_COMPLEX writes them).  Returns the
        text (ascii/matlab) and writes it to `path` when given."""
        n = self.getConverged()
        re = np.zeros(n)
        im = np.zeros(n)
        for i in range(n):
            re[i], im[i], _, _ = self.getEigenpair(i, vectors=False)
        if fmt == "binary":
            if path is None:
                raise ValueError("binary view needs a path")
            np.column_stack([re, im]).astype(">f8").tofile(path)
            return None
        text = format_eigenvalues(re, im, fmt, name)
        if path is not None:
            with open(path, "w") as f:
                f.write(text)
        return text

    def vectorsView(self, path):
        """EPSVectorsView with a binary viewer: every converged eigenvector
        (real part; the imaginary part follows for a complex pair) appended
        to `path` in PETSc binary Vec format."""
        L = lib()
        first = True
        for i in range(self.getConverged()):
            _, im, xr, xi = self.getEigenpair(i)
            for x in ([xr, xi] if im else [xr]):
                x = np.ascontiguousarray(x)
                check(L.sa_vec_write_petsc_binary(path.encode(), len(x), x.ctypes.data, 0 if first else 1))
                first = False

    def getErrorEstimate(self, i):
        v = C.c_double()
        check(lib().sa_eps_get_error_estimate(self.h, i, C.byref(v)))
        return v.value

    def computeError(self, i, etype=ERROR_RELATIVE):
        v = C.c_double()
        check(lib().sa_eps_compute_error(self.h, i, etype, C.byref(v)))
        return v.value

    def getStats(self):
        s = _lib.EpsStats()
        check(lib().sa_eps_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def destroy(self):
        if self.h:
            lib().sa_eps_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
