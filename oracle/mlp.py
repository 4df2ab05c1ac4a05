"""Forward/backward/Adam of the 3-layer GELU MLP used as a learned Davidson
preconditioner (no reference counterpart; the packing matches itg.h).

Y = W3 g(W2 g(W1 x + b1) + b2) + b3, g = exact GELU 0.5 x (1 + erf(x/sqrt2)),
loss = mean over all rows*out entries of (Y - T)^2.
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * INV_SQRT_2PI * np.exp(-0.5 * x * x)


def param_count(i, h, o):
    return h * i + h + h * h + h + o * h + o


def unpack(p, i, h, o):
    p = np.asarray(p, dtype=np.float64)
    sizes = [h * i, h, h * h, h, o * h, o]
    out, s = [], 0
    for n in sizes:
        out.append(p[s:s + n])
        s += n
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(h, i), b1, W2.reshape(h, h), b2, W3.reshape(o, h), b3


def forward(p, X, i, h, o):
    W1, b1, W2, b2, W3, b3 = unpack(p, i, h, o)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1.T + b1
    A1 = gelu(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = gelu(Z2)
    Y = A2 @ W3.T + b3
    return Y, (Z1, A1, Z2, A2)


def bf16_round(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even),
    returned as float32: the conversion the fused inference kernel makes
    (`(__bf16)float` in mlp.hip's mlp_infer_kernel). Finite inputs only."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16_hidden(p, X, i, h, o):
    """The fused inference kernel's arithmetic (itg_mlp_infer at h = 256):
    layer 1 in FP32, its GELU output rounded to bf16; the hidden product with
    W2 rounded to bf16 and exact (FP32-accumulated on the device) sums; bias,
    GELU and layer 3 as in `forward`. Evaluated in FP64 from the rounded
    operands."""
    W1, b1, W2, b2, W3, b3 = unpack(p, i, h, o)
    X = np.asarray(X, dtype=np.float64)
    Z1 = (X @ W1.T + b1).astype(np.float32)
    A1 = bf16_round(gelu(Z1.astype(np.float64)).astype(np.float32)).astype(np.float64)
    W2r = bf16_round(W2.astype(np.float32)).astype(np.float64)
    Z2 = A1 @ W2r.T + b2
    return gelu(Z2) @ W3.T + b3


def loss_and_grad(p, X, T, i, h, o):
    W1, b1, W2, b2, W3, b3 = unpack(p, i, h, o)
    Y, (Z1, A1, Z2, A2) = forward(p, X, i, h, o)
    R = Y.shape[0]
    diff = Y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(diff ** 2))
    dY = 2.0 * diff / (R * o)
    gW3 = dY.T @ A2
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3) * gelu_grad(Z2)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * gelu_grad(Z1)
    gW1 = dZ1.T @ np.asarray(X, dtype=np.float64)
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------ residual feature rows
def _rows(chiL, chiR, row0, step, nrows):
    if nrows is None:
        nrows = (chiL * chiR - row0 + step - 1) // step
    return row0 + step * np.arange(nrows)


def _clip(den, floor):
    return np.where(np.abs(den) < floor, np.where(den < 0, -floor, floor), den)


def features(r, dg, lam, chiL, d, chiR, floor=1e-8, row0=0, step=1, nrows=None):
    """Rows (a + chiL*b) of [r/rms(r), delta/rms(delta), (r/delta)/rms(r/delta)]
    over each d*d physical block; delta = dg - lam clipped to |delta| >= floor."""
    r = np.asarray(r, dtype=np.float64).reshape(chiL, d * d, chiR, order="F")
    de = np.asarray(dg, dtype=np.float64).reshape(chiL, d * d, chiR, order="F") - lam
    p = r / _clip(de, floor)

    def inv_rms(x):
        s = np.sqrt(np.mean(x * x))
        return 1.0 / s if s > 0 else 1.0

    rows = _rows(chiL, chiR, row0, step, nrows)
    a, b = rows % chiL, rows // chiL
    return np.concatenate([r[a, :, b] * inv_rms(r), de[a, :, b] * inv_rms(de),
                           p[a, :, b] * inv_rms(p)], axis=1)


def gather(v, scale, chiL, d, chiR, row0=0, step=1, nrows=None):
    v = np.asarray(v, dtype=np.float64).reshape(chiL, d * d, chiR, order="F")
    rows = _rows(chiL, chiR, row0, step, nrows)
    return v[rows % chiL, :, rows // chiL] * scale


def scatter(Y, chiL, d, chiR, row0, nrows, v):
    v = np.asarray(v, dtype=np.float64).reshape(chiL, d * d, chiR, order="F").copy()
    rows = row0 + np.arange(nrows)
    v[rows % chiL, :, rows // chiL] = np.asarray(Y).reshape(nrows, d * d)
    return v.ravel(order="F")
