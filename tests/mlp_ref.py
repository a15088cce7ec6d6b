"""CPU reference of the learned-closure network (test infrastructure).

Same parameter layout as the device network (fcg_mlp.h): flat FP32
W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[16][128]
b4[16]; ReLU hidden layers, linear output; loss 0.5 * mean_i |y_i - t_i|^2
over the two outputs.

``forward``/``gradient`` with ``bf16=True`` round the operands where the
device kernels round them (inputs, weights, the hidden activations after
ReLU and the errors that feed a matrix product; the bias gradients sum the
rounded errors) and accumulate in float64; with ``bf16=False`` they are the
plain float64 network, which ``finite_difference`` pins.
"""
import numpy as np

F, IN, H, OUT, OUTP = 5, 32, 128, 2, 16
W1 = 0
B1 = W1 + H * IN
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * H
W4 = B3 + H
B4 = W4 + OUTP * H
P = B4 + OUTP


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(H, IN), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, H), p[B3:W4],
            p[W4:B4].reshape(OUTP, H), p[B4:P])


def _r(a, on):
    """bfloat16-rounded (from float32) when on, else plain float64."""
    return bf16(np.asarray(a, dtype=np.float32)) if on else np.asarray(a, dtype=np.float64)


def forward(p, X, bf=True):
    """X: (n, 5) normalised features -> (n, 2) normalised outputs (+ activations)."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    n = X.shape[0]
    x = np.zeros((IN, n))
    x[:F] = np.asarray(X, dtype=np.float64).T
    xb = _r(x, bf)
    a1 = _r(np.maximum(_r(w1, bf) @ xb + b1[:, None], 0.0), bf)
    a2 = _r(np.maximum(_r(w2, bf) @ a1 + b2[:, None], 0.0), bf)
    a3 = _r(np.maximum(_r(w3, bf) @ a2 + b3[:, None], 0.0), bf)
    y = _r(w4, bf) @ a3 + b4[:, None]
    return y[:OUT].T, (xb, a1, a2, a3, y)


def loss(p, X, T, bf=False):
    y, _ = forward(p, X, bf)
    return 0.5 * np.mean(np.sum((y - T) ** 2, axis=1))


def gradient(p, X, T, bf=True):
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    n = X.shape[0]
    _, (xb, a1, a2, a3, y) = forward(p, X, bf)
    dy = np.zeros((OUTP, n))
    dy[:OUT] = (y[:OUT] - np.asarray(T, dtype=np.float64).T) / n
    dyb = _r(dy, bf)
    d3 = _r((_r(w4, bf).T @ dyb) * (a3 > 0), bf)
    d2 = _r((_r(w3, bf).T @ d3) * (a2 > 0), bf)
    d1 = _r((_r(w2, bf).T @ d2) * (a1 > 0), bf)
    g = np.zeros(P)
    g[W1:B1] = (d1 @ xb.T).reshape(-1)
    g[B1:W2] = d1.sum(axis=1)
    g[W2:B2] = (d2 @ a1.T).reshape(-1)
    g[B2:W3] = d2.sum(axis=1)
    g[W3:B3] = (d3 @ a2.T).reshape(-1)
    g[B3:W4] = d3.sum(axis=1)
    g[W4:B4] = (dyb @ a3.T).reshape(-1)
    g[B4:P] = dyb.sum(axis=1)
    g[~used_mask()] = 0.0   # padding carries no signal
    return g


def used_mask():
    m = np.ones(P, dtype=bool)
    m[W1:B1].reshape(H, IN)[:, F:] = False
    m[W4:B4].reshape(OUTP, H)[OUT:] = False
    m[B4 + OUT:P] = False
    return m


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    for off, rows, cols, stride, fi in ((W1, H, F, IN, F), (W2, H, H, H, H), (W3, H, H, H, H), (W4, OUT, H, H, H)):
        a = np.sqrt(6.0 / fi)
        blk = p[off:off + rows * stride].reshape(rows, stride)
        blk[:, :cols] = rng.uniform(-a, a, (rows, cols))
    for off in (B1, B2, B3):
        p[off:off + H] = rng.normal(0, 0.1, H)
    p[B4:B4 + OUT] = rng.normal(0, 0.1, OUT)
    return p


def finite_difference(p, X, T, idx, h=1e-6):
    """Central differences of the float64 loss at parameter indices idx."""
    out = []
    for i in idx:
        q = np.asarray(p, dtype=np.float64).copy()
        q[i] += h
        lp = loss(q, X, T)
        q[i] -= 2 * h
        lm = loss(q, X, T)
        out.append((lp - lm) / (2 * h))
    return np.array(out)
