"""CPU pins of the warm-start network's restatement (oracle/mlp_ref.py):
bfloat16 rounding against known bit patterns, gradients against central
finite differences (rounding switched off), Adam against its closed form after
one step, and bilinear prolongation against an exactly reproduced linear field."""
import numpy as np

from oracle import mlp_ref


def test_bf16_rounding_bit_patterns():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = mlp_ref.bf16(x)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e38) / 3.0e38 < 2 ** -8


def test_gradients_match_finite_differences(monkeypatch):
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, np.float64))
    rng = np.random.default_rng(0)
    din, h, o, B = 8, 12, 7, 5
    params = [rng.standard_normal((h, 32)) * 0.5, rng.standard_normal(h) * 0.1,
              rng.standard_normal((h, h)) * 0.3, rng.standard_normal(h) * 0.1,
              rng.standard_normal((o, h)) * 0.3, rng.standard_normal(o) * 0.1]
    params[0][:, din:] = 0.0
    X = rng.standard_normal((B, din))
    T = rng.standard_normal((B, o))
    g, _ = mlp_ref.gradients(params, X, T)

    def loss(pp):
        Y = mlp_ref.forward(pp, X)["Y"]
        return np.mean((Y - T) ** 2)

    for li, (pi, gi) in enumerate(zip(params, g)):
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in pi.shape)
            if li == 0 and idx[1] >= din:
                continue
            hstep = 1e-3
            pp = [q.copy() for q in params]
            pp[li][idx] += hstep
            lp = loss(pp)
            pp[li][idx] -= 2 * hstep
            lm = loss(pp)
            fd = (lp - lm) / (2 * hstep)
            assert abs(fd - gi[idx]) <= 2e-3 * max(1e-3, abs(fd)), (li, idx, fd, gi[idx])


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 0.0])
    out, m, v = mlp_ref.adam(p, g, 0.0, 0.0, 1, 1e-2)
    # after one step m/c1 = g and sqrt(v/c2) = |g|: the update is lr*sign(g)
    expect = p - 1e-2 * g / (np.abs(g) + 1e-8)
    assert np.allclose(out, expect, atol=1e-12)


def test_prolongation_reproduces_linear_fields():
    g = 8
    yy, xx = np.meshgrid(np.arange(g), np.arange(g), indexing="ij")
    M = (2.0 * xx - 1.5 * yy + 0.25).reshape(1, -1)
    out = mlp_ref.prolong(M, g, 32, 32).reshape(32, 32)
    c = ((np.arange(32) + 0.5) / 32) * g - 0.5
    ex = 2.0 * c[None, :] - 1.5 * c[:, None] + 0.25
    assert np.max(np.abs(out[2:-2, 2:-2] - ex[2:-2, 2:-2])) < 1e-12
