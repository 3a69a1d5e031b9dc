"""CPU restatement of the active-set MLP (test infrastructure only).

The MLP has no counterpart in the reference; this module is the checker of
the engine's matrix-core kernels (grappolo_amd/csrc/engine/mlp.hip): the
same 8 -> 64 -> 64 -> 1 network, the same parameter layout, the same
weighted logistic loss and the same BF16 rounding points, evaluated in
float64.  With ``rounding=False`` every rounding is skipped and the gradient
is the exact derivative of the loss, which the tests pin against central
finite differences.

Only tests, the smoke check and the benchmark's CPU baseline may import it.
"""
import numpy as np

IN, HID = 8, 64
W1, B1 = 0, 512
W2, B2 = 576, 4672
W3, B3 = 4736, 4800
NPARAMS = 4801


def bf16(x):
    """float64 value of round-to-nearest-even float32 -> bfloat16 of x."""
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64).reshape(x32.shape)


def bf16_bits_to_float(bits):
    """float64 values of BF16 bit patterns (uint16 / int16 arrays)."""
    b = np.asarray(bits).astype(np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def float_to_bf16_bits(x):
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[W1:B1].reshape(IN, HID), P[B1:W2], P[W2:B2].reshape(HID, HID), P[B2:W3], P[W3:B3],
            P[B3])


def init(seed=0):
    """He-normal weights, zero biases (float32, NPARAMS)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAMS, np.float32)
    P[W1:B1] = rng.normal(0, np.sqrt(2.0 / IN), IN * HID)
    P[W2:B2] = rng.normal(0, np.sqrt(2.0 / HID), HID * HID)
    P[W3:B3] = rng.normal(0, np.sqrt(1.0 / HID), HID)
    return P


def forward(P, X, rounding=True):
    """Logits of the rows of X (n x 8); returns (z, cache)."""
    w1, b1, w2, b2, w3, b3 = unpack(P)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    Xr = r(X)
    w1r, w2r, w3r = r(w1), r(w2), r(w3)
    z1 = Xr @ w1r + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ w2r + b2
    h2 = np.maximum(z2, 0.0)
    z = h2 @ w3r + b3
    return z, (Xr, z1, h1, z2, h2, w2r, w3r)


def loss_grad(P, X, y, pos_weight=1.0, rounding=True):
    """Mean weighted logistic loss over the rows and its gradient (NPARAMS)."""
    y = np.asarray(y, np.float64)
    B = len(y)
    z, (Xr, z1, h1, z2, h2, w2r, w3r) = forward(P, X, rounding)
    r = bf16 if rounding else (lambda a: a)
    c = np.where(y > 0.5, pos_weight, 1.0) / B
    p = 1.0 / (1.0 + np.exp(-z))
    dz = c * (p - y)
    loss = float(np.sum(c * (np.logaddexp(0.0, z) - y * z)))
    g = np.zeros(NPARAMS)
    g[B3] = dz.sum()
    g[W3:B3] = h2.T @ dz
    dZ2 = (z2 > 0) * (w3r[None, :] * dz[:, None])
    g[B2:W3] = dZ2.sum(0)
    dZ2r = r(dZ2)
    g[W2:B2] = (h1.T @ dZ2r).ravel()
    dH1 = dZ2r @ w2r.T
    dZ1 = (z1 > 0) * dH1
    g[B1:W2] = dZ1.sum(0)
    g[W1:B1] = (Xr.T @ r(dZ1)).ravel()
    return loss, g


def adam_step(P, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = 1 for the first), float64; returns (P, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v
