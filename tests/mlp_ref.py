"""CPU reference of the learned-closure network (test infrastructure).

Same parameter layout as the device network (fcg_mlp.h): flat FP32
W1[64][16] b1[64] W2[64][64] b2[64] W3[16][64] b3[16], tanh hidden layers,
loss 0.5 * mean_i |y_i - t_i|^2 over the two outputs.

``forward``/``gradient`` with ``bf16=True`` round the operands where the
device kernels round them (inputs, weights and the activations / errors
that feed a matrix product) and accumulate in float64; with ``bf16=False``
they are the plain float64 network, which ``finite_difference`` pins.
"""
import numpy as np

F, IN, H, OUT = 5, 16, 64, 2
W1, B1 = 0, 64 * 16
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 16 * 64
P = B3 + 16


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(64, 16), p[B1:W2], p[W2:B2].reshape(64, 64), p[B2:W3], p[W3:B3].reshape(16, 64),
            p[B3:P])


def _r(a, on):
    """bfloat16-rounded (from float32) when on, else plain float64."""
    return bf16(np.asarray(a, dtype=np.float32)) if on else np.asarray(a, dtype=np.float64)


def forward(p, X, bf=True):
    """X: (n, 5) normalised features -> (n, 2) normalised outputs (+ activations)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    x = np.zeros((16, n))
    x[:F] = np.asarray(X, dtype=np.float64).T
    xb = _r(x, bf)
    t1 = np.tanh(_r(w1, bf) @ xb + b1[:, None])
    t1b = _r(t1, bf)
    t2 = np.tanh(_r(w2, bf) @ t1b + b2[:, None])
    t2b = _r(t2, bf)
    y = _r(w3, bf) @ t2b + b3[:, None]
    return y[:OUT].T, (xb, t1, t1b, t2, t2b, y)


def loss(p, X, T, bf=False):
    y, _ = forward(p, X, bf)
    return 0.5 * np.mean(np.sum((y - T) ** 2, axis=1))


def gradient(p, X, T, bf=True):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    _, (xb, t1, t1b, t2, t2b, y) = forward(p, X, bf)
    dy = np.zeros((16, n))
    dy[:OUT] = (y[:OUT] - np.asarray(T, dtype=np.float64).T) / n
    dyb = _r(dy, bf)
    d2 = (_r(w3, bf).T @ dyb) * (1.0 - t2 ** 2)
    d2b = _r(d2, bf)
    d1 = (_r(w2, bf).T @ d2b) * (1.0 - t1 ** 2)
    d1b = _r(d1, bf)
    g = np.zeros(P)
    g[W1:B1] = (d1b @ xb.T).reshape(-1)
    g[B1:W2] = d1.sum(axis=1)
    g[W2:B2] = (d2b @ t1b.T).reshape(-1)
    g[B2:W3] = d2.sum(axis=1)
    g[W3:B3] = (dyb @ t2b.T).reshape(-1)
    g[B3:P] = dy.sum(axis=1)
    # padding (W1 columns >= 5, W3 rows and b3 entries >= 2) carries no signal
    g[W1:B1].reshape(64, 16)[:, F:] = 0.0
    g[W3:B3].reshape(16, 64)[OUT:] = 0.0
    g[B3 + OUT:P] = 0.0
    return g


def used_mask():
    m = np.ones(P, dtype=bool)
    m[W1:B1].reshape(64, 16)[:, F:] = False
    m[W3:B3].reshape(16, 64)[OUT:] = False
    m[B3 + OUT:P] = False
    return m


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    for off, rows, cols, stride, fi, fo in ((W1, 64, F, 16, F, 64), (W2, 64, 64, 64, 64, 64),
                                            (W3, OUT, 64, 64, 64, OUT)):
        a = np.sqrt(6.0 / (fi + fo))
        blk = p[off:off + rows * stride].reshape(rows, stride)
        blk[:, :cols] = rng.uniform(-a, a, (rows, cols))
    p[B1:W2] = rng.normal(0, 0.1, 64)
    p[B2:W3] = rng.normal(0, 0.1, 64)
    p[B3:B3 + OUT] = rng.normal(0, 0.1, OUT)
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
