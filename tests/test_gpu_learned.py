"""GPU tests of the learned warm start: BF16 MFMA forward/backward/Adam against
the CPU restatement, the prolongation, and the gain of the trained network
over the random start (Krylov-Schur still judges convergence)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _net(sa, din=8, hidden=64, dout=200, seed=3):
    from slepc_amd.learned import MLP
    return MLP(din=din, hidden=hidden, dout=dout, seed=seed)


def _relerr(a, b):
    return np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("dims", [(8, 64, 200, 5), (8, 512, 1000, 33)])
def test_mlp_forward_matches_oracle(sa, dims):
    from oracle import mlp_ref
    din, hid, dout, B = dims
    net = _net(sa, din, hid, dout)
    rng = np.random.default_rng(0)
    p = net.get_params()
    p[1][:] = rng.standard_normal(hid) * 0.1
    p[5][:] = rng.standard_normal(dout) * 0.1
    net.set_params(p)
    X = rng.standard_normal((B, din)).astype(np.float32)
    Y = net.forward(X)
    ref = mlp_ref.forward(net.get_params(), X)["Y"]
    assert _relerr(Y, ref) < 2e-3


def test_mlp_gradients_match_oracle(sa):
    from oracle import mlp_ref
    net = _net(sa, 8, 128, 300)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((17, 8)).astype(np.float32)
    T = rng.standard_normal((17, 300)).astype(np.float32)
    g, loss = net.gradients(X, T)
    gr, lr = mlp_ref.gradients(net.get_params(), X, T)
    assert abs(loss - lr) <= 1e-3 * lr
    for a, b in zip(g, gr):
        assert _relerr(a, b) < 1e-2


def test_mlp_adam_step(sa):
    from oracle import mlp_ref
    net = _net(sa, 8, 64, 100)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((9, 8)).astype(np.float32)
    T = rng.standard_normal((9, 100)).astype(np.float32)
    p0 = net.get_params()
    g, _ = net.gradients(X, T)
    net.train_step(X, T, 1e-2)
    p1 = net.get_params()
    for a0, a1, ga in zip(p0, p1, g):
        ref, _, _ = mlp_ref.adam(a0.astype(np.float64), ga.astype(np.float64), 0.0, 0.0, 1, 1e-2)
        assert np.max(np.abs(a1 - ref)) < 1e-5


def test_mlp_training_reduces_loss(sa):
    net = _net(sa, 8, 128, 256)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((64, 8)).astype(np.float32)
    Wt = rng.standard_normal((8, 256)).astype(np.float32)
    T = np.tanh(X @ Wt).astype(np.float32)
    l0 = net.train_step(X, T, 3e-3)
    for _ in range(300):
        l1 = net.train_step(X, T, 3e-3)
    assert l1 < 0.2 * l0


def test_prolongation_matches_oracle(sa):
    from oracle import mlp_ref
    from slepc_amd.learned import prolong
    rng = np.random.default_rng(5)
    M = rng.standard_normal((3, 16 * 16)).astype(np.float32)
    out = prolong(M, 16, 64, 48)
    ref = mlp_ref.prolong(M.astype(np.float64), 16, 64, 48)
    assert np.max(np.abs(out - ref)) < 1e-5
    # a bilinear field is reproduced exactly away from the walls
    g = 16
    yy, xx = np.meshgrid(np.arange(g), np.arange(g), indexing="ij")
    lin = (0.3 * xx + 0.7 * yy + 1.0).astype(np.float32).reshape(1, -1)
    out = prolong(lin, g, 64, 64).reshape(64, 64)
    ys = ((np.arange(64) + 0.5) / 64) * g - 0.5
    ex = 0.3 * ys[None, :] + 0.7 * ys[:, None] + 1.0
    assert np.max(np.abs(out[4:-4, 4:-4] - ex[4:-4, 4:-4])) < 1e-5


def test_learned_start_gain_over_random_start(sa):
    """What the trained network gains over the method without it, on held-out
    draws of the workload family at its full size (256 x 256, 16 pairs,
    ncv 64, tol 1e-10 on true residuals).  The network's start vector lies
    almost entirely in the wanted 16-dimensional eigenspace (measured against
    the converged FP64 eigenvectors of the same problem), where a random
    vector has a fraction ~16/n of its energy: that gain is asserted.
    Krylov-Schur starts from that one vector and still decides convergence,
    so every pair meets 1e-10 either way.  The time-to-solution ratio
    random/learned is measured on the same problems: on this family the
    matvec count is set by the relative gaps (~1e-6) at the edge of the
    wanted cluster, which one start vector does not change: the ratios are
    printed and asserted not to fall below 0.85 (no gain is claimed)."""
    import time
    from slepc_amd.learned import WarmStart
    ws = WarmStart.train(sa, n_train=256, epochs=40, batch=64, lr=1e-3)
    assert ws.train_losses[-1] < 0.5 * ws.train_losses[0]
    fracs, ratios, tratios = {"random": [], "learned": []}, [], []
    for s in range(4):
        th = sa.fv_theta(5, 20 + s)
        Ah = sa.fv2d(256, 256, th)
        A = sa.Mat(Ah)
        res, tt = {}, {}
        starts = (("random", lambda: sa.rand_uniform(0, Ah.n)), ("learned", lambda: ws.initial_vector(th, 256, 256)))
        vecs = {}
        for name, make in starts:
            t0 = time.perf_counter()
            v0 = make()
            assert v0 is not None
            E = sa.EPS()
            E.setOperators(A)
            E.setProblemType(True)
            E.setWhichEigenpairs("LR")
            E.setDimensions(16, 64)
            E.setTolerances(1e-10, 20000)      # the hardest draws need more than the default 2n/ncv restarts
            E.setTrueResidual(True)
            E.setInitialSpace(v0)
            E.solve()
            tt[name] = time.perf_counter() - t0
            vecs[name] = v0
            assert E.getConverged() >= 16
            for i in range(16):
                assert E.computeError(i) <= 1e-10
            res[name] = E.getStats()["matvecs"]
            if name == "random":
                X = np.stack([E.getEigenpair(i)[2] for i in range(16)])   # orthonormal rows
        for name in ("random", "learned"):
            v0 = vecs[name]
            fracs[name].append(np.linalg.norm(X @ v0) ** 2 / np.dot(v0, v0))
        ratios.append(res["random"] / res["learned"])
        tratios.append(tt["random"] / tt["learned"])
    print("energy fraction in the wanted eigenspace:", fracs, "matvec ratio random/learned:", ratios,
          "time ratio:", tratios)
    assert min(fracs["learned"]) > 0.5
    assert min(fracs["learned"]) > 100 * max(fracs["random"])
    assert np.mean(ratios) > 0.85, ratios
    assert np.mean(tratios) > 0.85, tratios
