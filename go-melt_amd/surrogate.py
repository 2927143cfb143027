"""Kernels of the learned L3 operator: 3x3x3 convolutions with BF16 operands
on the MFMA matrix cores (FP32 accumulation) and Adam, called through the C
ABI (gm_conv3d_*, gm_adam in include/gomelt.h).  PyTorch only holds the
tensors (device memory and the current stream); every computation here is
one of the library's kernels, and there is no fallback without them.

Layouts: activations (nb, D, H, W, C) channels last, weights (27, cout, cin)
with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); stride 1, zero padding 1; channel
counts multiples of 16.
"""

from __future__ import annotations

import ctypes as C

import torch

import _native as N


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _shape(X, co):
    nb, D, H, W, ci = X.shape
    return (C.c_int32 * 6)(nb, D, H, W, ci, co)


def _check(t, dtype, name):
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise ValueError(f"{name}: need a contiguous {dtype} tensor on the GPU")


def conv3d(X, Wk, bias=None, relu=False, want_bf16=True):
    """Y = conv(X; Wk) + bias (then max(0, .) if relu).  X bf16 (nb,D,H,W,ci),
    Wk bf16 (27, co, ci), bias FP32 (co,) or None.  Returns (Y FP32, Y bf16
    or None)."""
    _check(X, torch.bfloat16, "X")
    _check(Wk, torch.bfloat16, "Wk")
    co = Wk.shape[1]
    if Wk.shape != (27, co, X.shape[4]):
        raise ValueError("Wk must be (27, cout, cin)")
    Y = torch.empty(X.shape[:4] + (co,), dtype=torch.float32, device=X.device)
    Yb = torch.empty_like(Y, dtype=torch.bfloat16) if want_bf16 else None
    if bias is not None:
        _check(bias, torch.float32, "bias")
    N.check(N.lib().gm_conv3d_fwd(_shape(X, co), _ptr(X), _ptr(Wk), _ptr(bias), int(relu),
                                  _ptr(Y), _ptr(Yb), _stream()))
    return Y, Yb


def conv3d_bwd_data(dY, Wk, want_bf16=True):
    """dX of the convolution with weights Wk, from dY bf16 (nb,D,H,W,co)."""
    _check(dY, torch.bfloat16, "dY")
    _check(Wk, torch.bfloat16, "Wk")
    co, ci = Wk.shape[1], Wk.shape[2]
    work = torch.empty_like(Wk)
    dX = torch.empty(dY.shape[:4] + (ci,), dtype=torch.float32, device=dY.device)
    dXb = torch.empty_like(dX, dtype=torch.bfloat16) if want_bf16 else None
    shape = (C.c_int32 * 6)(*dY.shape[:4], ci, co)
    N.check(N.lib().gm_conv3d_bwd_data(shape, _ptr(dY), _ptr(Wk), _ptr(work), _ptr(dX), _ptr(dXb),
                                       _stream()))
    return dX, dXb


def conv3d_bwd_weight(X, dY, dW):
    """dW (FP32, (27, co, ci)) += the weight gradient from X and dY (bf16)."""
    _check(X, torch.bfloat16, "X")
    _check(dY, torch.bfloat16, "dY")
    _check(dW, torch.float32, "dW")
    co = dY.shape[4]
    N.check(N.lib().gm_conv3d_bwd_weight(_shape(X, co), _ptr(X), _ptr(dY), _ptr(dW), _stream()))
    return dW


def adam(w, wbf, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on FP32 master weights w (and their bf16 copy wbf)."""
    for t, name in ((w, "w"), (g, "g"), (m, "m"), (v, "v")):
        _check(t, torch.float32, name)
    N.check(N.lib().gm_adam(_ptr(w), _ptr(wbf), _ptr(g), _ptr(m), _ptr(v), w.numel(), float(lr),
                            float(beta1), float(beta2), float(eps), int(step), _stream()))


def to_bf16(a):
    """bf16 copy (round to nearest even) of an FP32 tensor, by the library."""
    _check(a, torch.float32, "a")
    b = torch.empty_like(a, dtype=torch.bfloat16)
    N.check(N.lib().gm_f32_to_bf16(_ptr(a), _ptr(b), a.numel(), _stream()))
    return b
