"""Learned warm start for the parameterised finite-volume family.

A small network maps the 8 Karhunen-Loeve coefficients theta of the
conductivity field to 16 vectors on a coarse 64x64 grid that span the
eigenspace of the 16 lowest modes of -div(k_theta grad u).  The training
targets are the orthogonal projections X X^T w_i of 16 fixed random vectors w_i
onto that eigenspace, X being the eigenvectors computed by this package's own
Krylov-Schur solver on the coarse operators: unlike the eigenvectors
themselves, the projections do not depend on the sign or on the rotation of
eigenvectors inside (near-)degenerate clusters, so the map theta -> target is
smooth.  Forward, backward and Adam run in the library's BF16 MFMA kernels.  At solve time the predicted modes are
prolonged bilinearly to the fine grid and summed into the start vector passed
through EPSSetInitialSpace; Krylov-Schur's own convergence test still decides,
so accuracy never depends on the network, and a degenerate prediction falls
back to the random start.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check, lib

NMODES, COARSE, HIDDEN, DIN = 16, 64, 512, 8
THETA_SEED_TRAIN = 4


def _setup_signatures():
    L = lib()
    vp = C.c_void_p
    L.sa_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]
    L.sa_mlp_destroy.argtypes = [vp]
    L.sa_mlp_forward.argtypes = [vp, C.c_int, vp, vp]
    L.sa_mlp_train_step.argtypes = [vp, C.c_int, vp, vp, C.c_float, C.POINTER(C.c_float)]
    L.sa_mlp_gradients.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_float)]
    L.sa_mlp_get_params.argtypes = [vp] * 7
    L.sa_mlp_set_params.argtypes = [vp] * 7
    L.sa_prolong_bilinear.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
    return L


class MLP:
    """din -> hidden -> hidden -> dout ReLU network on the GPU (BF16 MFMA)."""

    def __init__(self, din=DIN, hidden=HIDDEN, dout=NMODES * COARSE * COARSE, seed=1):
        self.L = _setup_signatures()
        self.din, self.hidden, self.dout = din, hidden, dout
        self.h = C.c_void_p()
        check(self.L.sa_mlp_create(din, hidden, dout, seed, C.byref(self.h)))

    def shapes(self):
        h, o = self.hidden, self.dout
        return [(h, 32), (h,), (h, h), (h,), (o, h), (o,)]

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty((X.shape[0], self.dout), np.float32)
        check(self.L.sa_mlp_forward(self.h, X.shape[0], X.ctypes.data, Y.ctypes.data))
        return Y

    def train_step(self, X, T, lr):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = C.c_float()
        check(self.L.sa_mlp_train_step(self.h, X.shape[0], X.ctypes.data, T.ctypes.data, lr, C.byref(loss)))
        return loss.value

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = [np.zeros(s, np.float32) for s in self.shapes()]
        loss = C.c_float()
        check(self.L.sa_mlp_gradients(self.h, X.shape[0], X.ctypes.data, T.ctypes.data,
                                      *[a.ctypes.data for a in g], C.byref(loss)))
        return g, loss.value

    def get_params(self):
        p = [np.zeros(s, np.float32) for s in self.shapes()]
        check(self.L.sa_mlp_get_params(self.h, *[a.ctypes.data for a in p]))
        return p

    def set_params(self, p):
        p = [np.ascontiguousarray(a, np.float32) for a in p]
        check(self.L.sa_mlp_set_params(self.h, *[a.ctypes.data for a in p]))

    def destroy(self):
        if self.h:
            self.L.sa_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def prolong(modes, g, nx, ny):
    """Bilinear prolongation on the GPU: modes [nm, g*g] -> [nm, nx*ny] (float64)."""
    L = _setup_signatures()
    modes = np.ascontiguousarray(modes, np.float32)
    nm = modes.shape[0]
    out = np.empty((nm, nx * ny))
    check(L.sa_prolong_bilinear(nm, g, nx, ny, modes.ctypes.data, out.ctypes.data))
    return out


def coarse_modes(sa, theta, nmodes=NMODES, g=COARSE, tol=1e-6):
    """Lowest nmodes eigenvectors of the coarse operator from this package's
    Krylov-Schur (largest real of sigma I - L, FP64)."""
    Ah = sa.fv2d(g, g, theta)
    A = sa.Mat(Ah)
    E = sa.EPS()
    E.setOperators(A)
    E.setProblemType(True)
    E.setWhichEigenpairs("LR")
    E.setDimensions(nmodes, 64)
    E.setTolerances(tol, 2000)
    E.setInitialSpace(sa.rand_uniform(0, Ah.n))
    E.solve()
    nc = E.getConverged()
    if nc < nmodes:
        raise RuntimeError("coarse solve converged %d < %d modes" % (nc, nmodes))
    X = np.stack([E.getEigenpair(i)[2] for i in range(nmodes)])
    return X


def coarse_modes_batch(sa, thetas, nmodes=NMODES, g=COARSE, tol=1e-6, ncv=48):
    """The same eigenvectors for many parameter draws at once with the batched
    solver (one workgroup per coarse problem, one launch per restart phase)."""
    from .batch import Batch
    csrs = [sa.fv2d(g, g, th) for th in thetas]
    B = Batch(csrs)
    B.solve(nmodes, ncv, tol, "LR", seed=0, max_it=4000)
    out = []
    for p in range(len(csrs)):
        nconv, _, _ = B.converged(p)
        if nconv < nmodes:
            out.append(None)
            continue
        out.append(np.stack([B.eigenpair(p, i)[1] for i in range(nmodes)]))
    B.destroy()
    return out


def probe_vectors(n=COARSE * COARSE, k=NMODES, seed=77):
    """Fixed random vectors w_i (columns) that are projected onto the eigenspace."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, k))


def projection_targets(X, W):
    """Rows i = X^T X w_i / ||.||: projections of the probes onto span(X rows).
    X: [k, n] orthonormal rows (eigenvectors); W: [n, k]."""
    P = (X.T @ (X @ W)).T
    return P / np.linalg.norm(P, axis=1, keepdims=True)


class WarmStart:
    """Trained network + the rules to turn its output into a start vector."""

    def __init__(self, net, probes, scale):
        self.net, self.probes, self.scale = net, probes, scale

    @classmethod
    def train(cls, sa, n_train=2000, epochs=60, batch=64, lr=1e-3, seed=1, log=None):
        t0 = time.perf_counter()
        W = probe_vectors()
        thetas, targets = [], []
        draws = [sa.fv_theta(THETA_SEED_TRAIN, s) for s in range(n_train)]
        for th, X in zip(draws, coarse_modes_batch(sa, draws)):
            if X is None:
                continue
            thetas.append(th)
            targets.append(projection_targets(X, W).ravel())
        thetas = np.array(thetas, np.float32)
        scale = float(COARSE)               # unit-norm modes have entries ~ 1/64
        T = np.array(targets, np.float32) * scale
        t1 = time.perf_counter()
        net = MLP(seed=seed)
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(len(thetas))
            for b0 in range(0, len(order), batch):
                idx = order[b0:b0 + batch]
                losses.append(net.train_step(thetas[idx], T[idx], lr))
        t2 = time.perf_counter()
        ws_steps = len(losses)
        if log:
            log("learned warm start: %d coarse solves in %.1f s, %d epochs in %.1f s, loss %.4f -> %.4f"
                % (len(thetas), t1 - t0, epochs, t2 - t1, losses[0], losses[-1]))
        ws = cls(net, W, scale)
        ws.train_losses = losses
        ws.n_train = len(thetas)
        ws.seconds_solves, ws.seconds_epochs, ws.steps = t1 - t0, t2 - t1, ws_steps
        return ws

    _cached = None

    @classmethod
    def load_or_train(cls, grid, log=None, **kw):
        import slepc_amd as sa
        if cls._cached is None:
            cls._cached = cls.train(sa, log=log, **kw)
        return cls._cached

    def predict_modes(self, theta):
        Y = self.net.forward(np.asarray(theta, np.float32)[None, :])[0] / self.scale
        return Y.reshape(NMODES, COARSE * COARSE)

    def initial_vector(self, theta, nx=256, ny=256):
        """Sum of the predicted modes (each normalised) on the fine grid, or
        None when the prediction is degenerate (the caller then keeps the
        random start)."""
        M = self.predict_modes(theta)
        F = prolong(M, COARSE, nx, ny)
        nrm = np.linalg.norm(F, axis=1)
        if not np.all(np.isfinite(nrm)) or np.any(nrm == 0):
            return None
        v = (F / nrm[:, None]).sum(axis=0)
        if not np.isfinite(v).all() or np.linalg.norm(v) == 0:
            return None
        return v
