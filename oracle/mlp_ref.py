"""CPU restatement of the warm-start network (test infrastructure only).

Same arithmetic points as the GPU kernels: inputs, weights and activations are
rounded to bfloat16 (round to nearest even) where the GPU stores them in BF16;
products are accumulated in float64 here (FP32 on the GPU), so comparisons use
a relative tolerance.  Loss: mean squared error over batch x outputs.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, X):
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    B, din = X.shape
    Xp = np.zeros((B, W1.shape[1]))
    Xp[:, :din] = X
    Xb = bf16(Xp)
    H1 = np.maximum(Xb @ bf16(W1).T + b1, 0.0).astype(np.float32).astype(np.float64)
    H1b = bf16(H1)
    H2 = np.maximum(H1b @ bf16(W2).T + b2, 0.0).astype(np.float32).astype(np.float64)
    H2b = bf16(H2)
    Y = (H2b @ bf16(W3).T + b3).astype(np.float32).astype(np.float64)
    return dict(Xb=Xb, H1=H1, H1b=H1b, H2=H2, H2b=H2b, Y=Y)


def gradients(params, X, T):
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    f = forward(params, X)
    B, o = f["Y"].shape
    dY = (2.0 * (f["Y"] - T) / (B * o)).astype(np.float32).astype(np.float64)
    loss = np.mean((f["Y"] - T) ** 2)
    dYb = bf16(dY)
    gW3 = dYb.T @ f["H2b"]
    gb3 = dY.sum(0)
    dH2 = (dYb @ bf16(W3)) * (f["H2"] > 0)
    dZ2b = bf16(dH2)
    gb2 = dH2.sum(0)
    gW2 = dZ2b.T @ f["H1b"]
    dH1 = (dZ2b @ bf16(W2)) * (f["H1"] > 0)
    dZ1b = bf16(dH1)
    gb1 = dH1.sum(0)
    gW1 = dZ1b.T @ f["Xb"]
    return [gW1, gb1, gW2, gb2, gW3, gb3], loss


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def prolong(modes, g, nx, ny):
    """Bilinear interpolation of cell-centred g x g data to nx x ny cell centres
    (clamped at the walls)."""
    out = np.zeros((modes.shape[0], nx * ny))
    ys = ((np.arange(ny) + 0.5) / ny) * g - 0.5
    xs = ((np.arange(nx) + 0.5) / nx) * g - 0.5
    y0 = np.floor(ys).astype(int)
    x0 = np.floor(xs).astype(int)
    fy, fx = ys - y0, xs - x0
    y1, x1 = np.clip(y0 + 1, 0, g - 1), np.clip(x0 + 1, 0, g - 1)
    y0, x0 = np.clip(y0, 0, g - 1), np.clip(x0, 0, g - 1)
    for q in range(modes.shape[0]):
        M = modes[q].reshape(g, g)
        v = ((1 - fy)[:, None] * ((1 - fx)[None, :] * M[y0][:, x0] + fx[None, :] * M[y0][:, x1])
             + fy[:, None] * ((1 - fx)[None, :] * M[y1][:, x0] + fx[None, :] * M[y1][:, x1]))
        out[q] = v.ravel()
    return out
