"""CPU oracle of the learned L3 operator's convolutions (FP64 NumPy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
convolution kernels.  No reference counterpart (the reference has no learned
part); pinned in tests/test_oracle_pins.py by an identity kernel, a shift
kernel (zero padding), the adjoint identities of the two gradients and
finite differences of a two-layer loss.

Layouts as the kernels: X[nb, D, H, W, C] (channels last), weights
Wk[27, cout, cin] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); stride 1, zero
padding 1.
"""
import numpy as np


def _taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shifted(X, dz, dy, dx):
    """S[b, z, y, x] = X[b, z+dz, y+dy, x+dx], zero outside."""
    nb, D, H, W, C = X.shape
    S = np.zeros_like(X)
    zs, ys, xs = (slice(max(0, -d), n - max(0, d)) for d, n in ((dz, D), (dy, H), (dx, W)))
    zd, yd, xd = (slice(max(0, d), n + min(0, d)) for d, n in ((dz, D), (dy, H), (dx, W)))
    S[:, zs, ys, xs] = X[:, zd, yd, xd]
    return S


def conv3d(X, Wk, bias=None):
    """Y[v, o] = bias[o] + sum_{tap, i} X[v + s(tap), i] Wk[tap, o, i]."""
    X = np.asarray(X, np.float64)
    Wk = np.asarray(Wk, np.float64)
    Y = np.zeros(X.shape[:4] + (Wk.shape[1],))
    for t, (dz, dy, dx) in enumerate(_taps()):
        Y += _shifted(X, dz, dy, dx) @ Wk[t].T
    if bias is not None:
        Y += np.asarray(bias, np.float64)
    return Y


def conv3d_bwd_data(dY, Wk):
    """dX[v, i] = sum_{tap, o} dY[v - s(tap), o] Wk[tap, o, i]."""
    dY = np.asarray(dY, np.float64)
    Wk = np.asarray(Wk, np.float64)
    dX = np.zeros(dY.shape[:4] + (Wk.shape[2],))
    for t, (dz, dy, dx) in enumerate(_taps()):
        dX += _shifted(dY, -dz, -dy, -dx) @ Wk[t]
    return dX


def conv3d_bwd_weight(X, dY):
    """dW[tap, o, i] = sum_v dY[v, o] X[v + s(tap), i]."""
    X = np.asarray(X, np.float64)
    dY = np.asarray(dY, np.float64)
    co, ci = dY.shape[-1], X.shape[-1]
    dW = np.zeros((27, co, ci))
    for t, (dz, dy, dx) in enumerate(_taps()):
        dW[t] = dY.reshape(-1, co).T @ _shifted(X, dz, dy, dx).reshape(-1, ci)
    return dW


def two_layer_loss(X, W1, b1, W2, b2, target):
    """L = 0.5 sum (conv(relu(conv(X; W1) + b1); W2) + b2 - target)^2, and its
    gradients (dW1, db1, dW2, db2) by the chain rule."""
    Z1 = conv3d(X, W1, b1)
    A1 = np.maximum(Z1, 0.0)
    Y = conv3d(A1, W2, b2)
    R = Y - target
    L = 0.5 * float((R * R).sum())
    dW2 = conv3d_bwd_weight(A1, R)
    db2 = R.reshape(-1, R.shape[-1]).sum(axis=0)
    dA1 = conv3d_bwd_data(R, W2)
    dZ1 = dA1 * (Z1 > 0)
    dW1 = conv3d_bwd_weight(X, dZ1)
    db1 = dZ1.reshape(-1, dZ1.shape[-1]).sum(axis=0)
    return L, (dW1, db1, dW2, db2), (Z1, A1, Y, R, dZ1)


def bf16_round(a):
    """Round FP32 values to the nearest bf16 (ties to even), as FP32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def bf16_bits(a):
    """uint16 bit patterns of bf16_round(a)."""
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)
