"""Batched Krylov-Schur over many independent symmetric problems (C ABI
sa_batch_*): one block-diagonal operator, one basis; each restart sweep is
either one launch with a workgroup per problem, or lock-step launches per
Gram-Schmidt slot over the row chunks of all problems (autotuned)."""
import ctypes as C

import numpy as np

from ._lib import check, lib


def _sig():
    L = lib()
    vp = C.c_void_p
    L.sa_batch_create.argtypes = [C.c_int, vp, vp, vp, vp, C.POINTER(vp)]
    L.sa_batch_destroy.argtypes = [vp]
    L.sa_batch_solve.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, vp,
                                 C.c_uint64]
    L.sa_batch_get_converged.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sa_batch_get_eigenpair.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), vp]
    L.sa_batch_compute_errors.argtypes = [vp, C.c_int, vp]
    L.sa_batch_get_stats.argtypes = [vp, vp, vp]
    L.sa_batch_set_sweep.argtypes = [vp, C.c_int]
    L.sa_batch_set_true_residual.argtypes = [vp, C.c_int]
    L.sa_batch_get_sweep.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.sa_batch_get_schedule.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
    L.sa_batch_set_graphs.argtypes = [vp, C.c_int]
    L.sa_batch_get_graph_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)]
    L.sa_batch_autotune_ncv.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, C.c_int, C.c_int, C.c_int,
                                        C.c_uint64, vp]
    L.sa_batch_get_ncv_tuning.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), vp, vp]
    L.sa_batch_get_ncv.argtypes = [vp, vp]
    return L


SWEEP_MODES = {"auto": 0, "persistent": 1, "wide": 2}


def uneven_laplacians(P=512, seed=2, lo=32, hi=256, pot=10.0):
    """The batched family: 2-D 5-point Laplacians on nx x ny rectangles with
    nx, ny uniform integers in [lo, hi] and a diagonal potential pot*U[0,1);
    problem p uses sub-seed sa_sub_seed(seed, p)."""
    import slepc_amd as sa
    L = lib()
    out = []
    span = hi - lo + 1
    for p in range(P):
        nx = lo + int(L.sa_u01(seed, 2 * p) * span)
        ny = lo + int(L.sa_u01(seed, 2 * p + 1) * span)
        out.append(sa.lap2d(nx, ny, pot, L.sa_sub_seed(seed, p)))
    return out


class Batch:
    def __init__(self, csrs):
        self.L = _sig()
        self.csrs = csrs
        self.P = len(csrs)
        self.n = np.array([A.n for A in csrs], np.int32)
        rp = np.concatenate([A.rowptr for A in csrs]).astype(np.int32)
        ci = np.concatenate([A.col for A in csrs]).astype(np.int32)
        va = np.concatenate([A.val for A in csrs]).astype(np.float64)
        self.h = C.c_void_p()
        check(self.L.sa_batch_create(self.P, self.n.ctypes.data, rp.ctypes.data, ci.ctypes.data, va.ctypes.data,
                                     C.byref(self.h)))
        self.nev = 0

    def set_true_residual(self, trueres=True):
        """EPSSetTrueResidual for every problem of the batch."""
        check(self.L.sa_batch_set_true_residual(self.h, 1 if trueres else 0))

    def solve(self, nev, ncv, tol, which="LM", mixed=False, tol32=1e-5, v0=None, seed=0, max_it=0):
        from . import WHICH
        self.nev = nev
        v = None
        if v0 is not None:
            v0 = np.ascontiguousarray(np.concatenate(v0) if isinstance(v0, (list, tuple)) else v0, np.float64)
            v = v0.ctypes.data
        check(self.L.sa_batch_solve(self.h, WHICH[which], nev, ncv, tol, max_it, 1 if mixed else 0, tol32, v, seed))

    def converged(self, p):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(self.L.sa_batch_get_converged(self.h, p, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def eigenpair(self, p, i, vector=True):
        re = C.c_double()
        x = np.empty(int(self.n[p])) if vector else None
        check(self.L.sa_batch_get_eigenpair(self.h, p, i, C.byref(re), x.ctypes.data if vector else None))
        return re.value, x

    def errors(self):
        e = np.zeros(self.P * self.nev)
        check(self.L.sa_batch_compute_errors(self.h, self.nev, e.ctypes.data))
        return e.reshape(self.P, self.nev)

    def stats(self):
        c = np.zeros(6, np.int64)
        s = np.zeros(3)
        check(self.L.sa_batch_get_stats(self.h, c.ctypes.data, s.ctypes.data))
        return dict(matvecs=int(c[0]), passes=int(c[1]), rounds=int(c[2]), matvecs32=int(c[3]),
                    passes32=int(c[4]), rounds32=int(c[5]), seconds=s[0], seconds32=s[1], seconds_ds=s[2])

    def set_sweep(self, mode):
        """"auto" (time both sweep kernels on the first restart round), "persistent"
        (a workgroup per problem, ncv < 64) or "wide" (lock-step over row chunks)."""
        check(self.L.sa_batch_set_sweep(self.h, SWEEP_MODES[mode]))

    def sweep_info(self):
        m, a, b = C.c_int(), C.c_double(), C.c_double()
        check(self.L.sa_batch_get_sweep(self.h, C.byref(m), C.byref(a), C.byref(b)))
        names = {v: k for k, v in SWEEP_MODES.items()}
        return dict(mode=names.get(m.value, "none"), ms_persistent=a.value, ms_wide=b.value)

    def schedule(self):
        """The last solve's sweep schedule and the autotuner's timings (ms) of the
        candidates [one workgroup per problem, lock-step, other tile loads,
        twice the SpMV lanes] (-1: not timed)."""
        k, nt, la = C.c_int(), C.c_int(), C.c_int()
        ms = np.zeros(4)
        check(self.L.sa_batch_get_schedule(self.h, C.byref(k), C.byref(nt), C.byref(la), ms.ctypes.data))
        return dict(kind={1: "persistent", 2: "wide"}.get(k.value, "none"), nt=nt.value, lanes=la.value,
                    ms=[float(x) for x in ms])

    def set_graphs(self, on=True):
        """Replay the lock-step restart sweeps from captured HIP graphs."""
        check(self.L.sa_batch_set_graphs(self.h, 1 if on else 0))

    def graph_stats(self):
        c, r, k = C.c_int64(), C.c_int64(), C.c_int()
        check(self.L.sa_batch_get_graph_stats(self.h, C.byref(c), C.byref(r), C.byref(k)))
        return dict(captures=c.value, replays=r.value, cached=k.value)

    def autotune_ncv(self, nev, tol, cand, which="LM", nclass=4, nsample=4, seed=0):
        """Basis-size autotuner per size class; returns the per-problem ncv
        (used by solve(ncv=0))."""
        from . import WHICH
        c = np.ascontiguousarray(cand, np.int32)
        out = np.zeros(self.P, np.int32)
        check(self.L.sa_batch_autotune_ncv(self.h, WHICH[which], nev, tol, c.ctypes.data, len(c), nclass, nsample,
                                           seed, out.ctypes.data))
        return out

    def ncv_tuning(self):
        nc, nq = C.c_int(), C.c_int()
        check(self.L.sa_batch_get_ncv_tuning(self.h, C.byref(nc), C.byref(nq), None, None))
        hi = np.zeros(nc.value, np.int32)
        sc = np.zeros(nc.value * nq.value)
        check(self.L.sa_batch_get_ncv_tuning(self.h, C.byref(nc), C.byref(nq), hi.ctypes.data, sc.ctypes.data))
        return dict(class_hi=hi, scores=sc.reshape(nc.value, nq.value))

    def ncv(self):
        out = np.zeros(self.P, np.int32)
        check(self.L.sa_batch_get_ncv(self.h, out.ctypes.data))
        return out

    def destroy(self):
        if self.h:
            self.L.sa_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
