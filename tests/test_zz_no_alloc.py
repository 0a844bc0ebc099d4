"""The allocation-free residual entry point (sa_mat_residual_norms_noalloc)
against the allocating one (sa_mat_residual_norms) through the C ABI: equal
bits on two calls in a row, the second on a smaller operator with fewer
vectors (the shared device scratch keeps the first call's contents), and
the size limit of the scratch rejected with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(sa, name, M, X, lr, li):
    L = sa.lib()
    f = getattr(L, name)
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    k = X.shape[1]
    xd = sa.DeviceArray.from_host(np.ascontiguousarray(X.T))
    lr = np.ascontiguousarray(lr, np.float64)
    li = np.ascontiguousarray(li, np.float64)
    res = np.full(k, -1.0)
    rc = f(M.h, k, xd.ptr, M.n, lr.ctypes.data, li.ctypes.data, res.ctypes.data)
    xd.free()
    return rc, res


def _case(orc, N, k, seed):
    A = orc.convdiff3d(N, N - 1, N - 2)
    X = np.stack([orc.rand_uniform(seed + j, A.n) for j in range(k)], axis=1)
    lr = np.linspace(1.0, 9.0, k)
    li = np.zeros(k)
    if k >= 4:
        li[1], li[2] = 0.5, -0.5
    return A, X, lr, li


def test_noalloc_matches_allocating_entry(sa, orc):
    for N, k, seed in ((40, 16, 10), (9, 3, 70)):      # second call: smaller input
        A, X, lr, li = _case(orc, N, k, seed)
        M = sa.Mat(sa.CSR(A.n, A.rowptr, A.col, A.val))
        rc0, ref = _call(sa, "sa_mat_residual_norms", M, X, lr, li)
        rc1, got = _call(sa, "sa_mat_residual_norms_noalloc", M, X, lr, li)
        assert rc0 == 0 and rc1 == 0, sa.lib().sa_last_error()
        assert np.array_equal(got, ref)
        # and the value itself against the oracle
        i = 0
        while i < k:
            if li[i] != 0.0:
                o = orc.residual_abs(A, lr[i], li[i], X[:, i], X[:, i + 1])
                assert abs(got[i] - o) <= 1e-12 * o and got[i + 1] == got[i]
                i += 2
            else:
                o = orc.residual_abs(A, lr[i], 0.0, X[:, i])
                assert abs(got[i] - o) <= 1e-12 * o
                i += 1
        M.destroy()


def test_noalloc_rejects_more_than_one_launch(sa, orc):
    A, X, lr, li = _case(orc, 8, 17, 5)
    M = sa.Mat(sa.CSR(A.n, A.rowptr, A.col, A.val))
    rc, res = _call(sa, "sa_mat_residual_norms_noalloc", M, X, lr, li)
    assert rc != 0
    sa.lib().sa_last_error.restype = C.c_char_p
    assert b"at most 16 vectors" in sa.lib().sa_last_error()
    assert np.all(res == -1.0)
    rc, res = _call(sa, "sa_mat_residual_norms", M, X, lr, li)    # the allocating entry takes any k
    assert rc == 0 and np.all(res > 0)
