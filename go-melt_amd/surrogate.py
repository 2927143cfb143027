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

import numpy as np
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
    shape = _shape(X, co)
    nb = C.c_int64()
    N.check(N.lib().gm_conv3d_fwd_ws_bytes(shape, C.byref(nb)))
    if nb.value > 0:  # few output tiles: the K sum split over blocks
        ws = torch.empty(nb.value, dtype=torch.uint8, device=X.device)
        N.check(N.lib().gm_conv3d_fwd_ws(shape, _ptr(X), _ptr(Wk), _ptr(bias), int(relu),
                                         _ptr(Y), _ptr(Yb), _ptr(ws), nb.value, _stream()))
    else:
        N.check(N.lib().gm_conv3d_fwd(shape, _ptr(X), _ptr(Wk), _ptr(bias), int(relu),
                                      _ptr(Y), _ptr(Yb), _stream()))
    return Y, Yb


def conv3d_bwd_data(dY, Wk, want_bf16=True):
    """dX of the convolution with weights Wk, from dY bf16 (nb,D,H,W,co)."""
    _check(dY, torch.bfloat16, "dY")
    _check(Wk, torch.bfloat16, "Wk")
    co, ci = Wk.shape[1], Wk.shape[2]
    if dY.shape[4] != co:
        raise ValueError("conv3d_bwd_data: dY channels must equal the weights' cout")
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
    if tuple(X.shape[:4]) != tuple(dY.shape[:4]) or tuple(dW.shape) != (27, co, X.shape[4]):
        raise ValueError("conv3d_bwd_weight: X, dY voxels and dW (27, co, ci) must agree")
    N.check(N.lib().gm_conv3d_bwd_weight(_shape(X, co), _ptr(X), _ptr(dY), _ptr(dW), _stream()))
    return dW


def adam(w, wbf, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on FP32 master weights w (and their bf16 copy wbf)."""
    for t, name in ((w, "w"), (g, "g"), (m, "m"), (v, "v")):
        _check(t, torch.float32, name)
    N.check(N.lib().gm_adam(_ptr(w), _ptr(wbf), _ptr(g), _ptr(m), _ptr(v), w.numel(), float(lr),
                            float(beta1), float(beta2), float(eps), int(step), _stream()))


def zero(t):
    """t[:] = 0 by the library (hipMemsetAsync on the current stream)."""
    N.check(N.lib().gm_memset(_ptr(t), 0, t.numel() * t.element_size(), _stream()))


def to_bf16(a):
    """bf16 copy (round to nearest even) of an FP32 tensor, by the library."""
    _check(a, torch.float32, "a")
    b = torch.empty_like(a, dtype=torch.bfloat16)
    N.check(N.lib().gm_f32_to_bf16(_ptr(a), _ptr(b), a.numel(), _stream()))
    return b


# ---------------------------------------------------------------------------
# Network helpers (each one kernel of the library)
# ---------------------------------------------------------------------------

def pool2(X):
    """2x2x2 average pooling of a bf16 (nb, 2D, 2H, 2W, C) tensor."""
    nb, D2, H2, W2, Cc = X.shape
    if D2 % 2 or H2 % 2 or W2 % 2:
        raise ValueError("pool2 needs even extents")
    Y = torch.empty((nb, D2 // 2, H2 // 2, W2 // 2, Cc), dtype=torch.bfloat16, device=X.device)
    sh = (C.c_int32 * 5)(nb, D2 // 2, H2 // 2, W2 // 2, Cc)
    N.check(N.lib().gm_nn_pool2(0, sh, _ptr(X), _ptr(Y), _stream()))
    return Y


def pool2_bwd(dY):
    nb, D, H, W, Cc = dY.shape
    dX = torch.empty((nb, 2 * D, 2 * H, 2 * W, Cc), dtype=torch.float32, device=dY.device)
    sh = (C.c_int32 * 5)(nb, D, H, W, Cc)
    N.check(N.lib().gm_nn_pool2(1, sh, _ptr(dY), _ptr(dX), _stream()))
    return dX


def up2(X):
    nb, D, H, W, Cc = X.shape
    Y = torch.empty((nb, 2 * D, 2 * H, 2 * W, Cc), dtype=torch.bfloat16, device=X.device)
    sh = (C.c_int32 * 5)(nb, D, H, W, Cc)
    N.check(N.lib().gm_nn_up2(0, sh, _ptr(X), _ptr(Y), _stream()))
    return Y


def up2_bwd(dY):
    nb, D2, H2, W2, Cc = dY.shape
    dX = torch.empty((nb, D2 // 2, H2 // 2, W2 // 2, Cc), dtype=torch.float32, device=dY.device)
    sh = (C.c_int32 * 5)(nb, D2 // 2, H2 // 2, W2 // 2, Cc)
    N.check(N.lib().gm_nn_up2(1, sh, _ptr(dY), _ptr(dX), _stream()))
    return dX


def concat(A, B):
    if tuple(A.shape[:4]) != tuple(B.shape[:4]):
        raise ValueError(f"concat: voxel shapes {tuple(A.shape)} and {tuple(B.shape)} differ")
    nv = A.numel() // A.shape[-1]
    Y = torch.empty(A.shape[:4] + (A.shape[4] + B.shape[4],), dtype=torch.bfloat16,
                    device=A.device)
    N.check(N.lib().gm_nn_concat(nv, A.shape[4], B.shape[4], _ptr(A), _ptr(B), _ptr(Y),
                                 _stream()))
    return Y


def split(dY, ca):
    cb = dY.shape[4] - ca
    nv = dY.numel() // dY.shape[-1]
    dA = torch.empty(dY.shape[:4] + (ca,), dtype=torch.float32, device=dY.device)
    dB = torch.empty(dY.shape[:4] + (cb,), dtype=torch.float32, device=dY.device)
    N.check(N.lib().gm_nn_split(nv, ca, cb, _ptr(dY), _ptr(dA), _ptr(dB), _stream()))
    return dA, dB


def _same(a, b, what):
    if b is not None and tuple(a.shape) != tuple(b.shape):
        raise ValueError(f"{what}: shapes {tuple(a.shape)} and {tuple(b.shape)} differ")


def relu_bwd(dA, A, dA2=None):
    _same(dA, A, "relu_bwd")
    _same(dA, dA2, "relu_bwd")
    dZ = torch.empty_like(dA)
    dZb = torch.empty_like(dA, dtype=torch.bfloat16)
    N.check(N.lib().gm_nn_relu_bwd(dA.numel(), _ptr(dA), _ptr(dA2), _ptr(A), _ptr(dZ), _ptr(dZb),
                                   _stream()))
    return dZ, dZb


def bias_grad(dZ, db):
    if db.numel() != dZ.shape[-1]:
        raise ValueError("bias_grad: one bias per channel")
    N.check(N.lib().gm_nn_bias_grad(dZ.numel() // dZ.shape[-1], dZ.shape[-1], _ptr(dZ), _ptr(db),
                                    _stream()))


def relu_bias_bwd(dA, A, db, dA2=None):
    """ReLU backward and bias gradient in one pass: returns the bf16 dZ and
    adds its per-channel sums (of the FP32 values) to db.  Channel counts the
    fused kernel does not take go through relu_bwd and bias_grad."""
    _same(dA, A, "relu_bias_bwd")
    _same(dA, dA2, "relu_bias_bwd")
    nc = dA.shape[-1]
    if db.numel() != nc:
        raise ValueError("relu_bias_bwd: one bias per channel")
    aligned = all(t is None or t.data_ptr() % 16 == 0 for t in (dA, dA2, A))
    if nc % 4 or 256 % nc or not aligned:
        dZ, dZb = relu_bwd(dA, A, dA2)
        bias_grad(dZ, db)
        return dZb
    dZb = torch.empty_like(dA, dtype=torch.bfloat16)
    N.check(N.lib().gm_nn_relu_bias_bwd(dA.numel() // nc, nc, _ptr(dA), _ptr(dA2), _ptr(A), None,
                                        _ptr(dZb), _ptr(db), _stream()))
    return dZb


def mse(Y, target, loss):
    """loss (FP64 device scalar) += 0.5 mean (Y[..., 0] - target)^2; returns
    the FP32 gradient wrt Y."""
    if target.numel() != Y.numel() // Y.shape[-1] or loss.numel() < 1:
        raise ValueError("mse: target needs one value per voxel of Y")
    dY = torch.empty_like(Y)
    N.check(N.lib().gm_nn_mse(Y.numel() // Y.shape[-1], Y.shape[-1], _ptr(Y), _ptr(target),
                              _ptr(dY), _ptr(loss), _stream()))
    return dY


def mse_energy(Y, target, loss, lam):
    """loss += 0.5 mean (Y[..., 0] - target)^2 + 0.5 lam mean over the windows
    of m_w^2, m_w the sum of the error over window w's free voxels (off the
    Dirichlet faces) over its voxel count; Y is (nb, D, H, W, C).  Returns
    dloss/dY."""
    nb, D, H, Wd = Y.shape[:4]
    nvox = nb * D * H * Wd
    if target.numel() != nvox or loss.numel() < 1:
        raise ValueError("mse_energy: target needs one value per voxel of Y")
    dY = torch.empty_like(Y)
    ws = torch.empty(nb, dtype=torch.float64, device=Y.device)
    zero(ws)
    N.check(N.lib().gm_nn_mse_energy(D * H * Wd, nb, Y.shape[-1], Wd, H, float(lam), _ptr(Y),
                                     _ptr(target), _ptr(ws), _ptr(dY), _ptr(loss), _stream()))
    return dY


# ---------------------------------------------------------------------------
# The operator: a 3-D U-Net of 3x3x3 convolutions
# ---------------------------------------------------------------------------

class Conv:
    """One 3x3x3 convolution layer: FP32 master weights, their bf16 copy,
    FP32 gradients and Adam moments (all device tensors)."""

    def __init__(self, cin, cout, relu, gen):
        std = (2.0 / (27 * cin)) ** 0.5 if relu else (1.0 / (27 * cin)) ** 0.5
        w = torch.randn((27, cout, cin), generator=gen, dtype=torch.float32) * std
        self.w = w.cuda()
        self.b = _zeros(cout)
        self.wb = to_bf16(self.w)
        self.relu = relu
        self.params = [self.w, self.b]
        self.grads = [_zeros(self.w.shape), _zeros(self.b.shape)]
        self.moments = [(_zeros(p.shape), _zeros(p.shape)) for p in self.params]

    def forward(self, Xb):
        self.Xb = Xb
        self.A, Ab = conv3d(Xb, self.wb, self.b, relu=self.relu)
        return Ab

    def backward(self, dA, dA2=None, need_input_grad=True):
        """dA: FP32 gradient wrt this layer's output; returns the FP32
        gradient wrt its input (or None)."""
        dZb = relu_bias_bwd(dA, self.A if self.relu else None, self.grads[1], dA2)
        conv3d_bwd_weight(self.Xb, dZb, self.grads[0])
        if not need_input_grad:
            return None
        dX, _ = conv3d_bwd_data(dZb, self.wb, want_bf16=False)
        return dX

    def step(self, t, lr):
        for p, g, (m, v), bf in zip(self.params, self.grads, self.moments, (self.wb, None)):
            adam(p, bf, g, m, v, t, lr)
        for g in self.grads:
            zero(g)


class UNet:
    """3-D U-Net operator: encoder levels of two convolutions (ReLU) with
    2x2x2 average pooling between them, decoder levels of nearest 2x
    upsampling, a skip concatenation and two convolutions, and a linear
    head whose channel 0 is the predicted (T_next - T) / T_scale.  Input: 16
    channels from gm_nn_window_input."""

    def __init__(self, channels=(32, 64, 128, 128), cin=16, cout=16, seed=0):
        gen = torch.Generator().manual_seed(seed)
        self.ch = list(channels)
        self.enc = []
        c = cin
        for ch in self.ch:
            self.enc.append((Conv(c, ch, True, gen), Conv(ch, ch, True, gen)))
            c = ch
        self.dec = []
        for l in range(len(self.ch) - 2, -1, -1):
            self.dec.append((Conv(self.ch[l] + self.ch[l + 1], self.ch[l], True, gen),
                             Conv(self.ch[l], self.ch[l], True, gen)))
        self.head = Conv(self.ch[0], cout, False, gen)
        self.t = 0

    def layers(self):
        out = [c for pair in self.enc for c in pair] + [c for pair in self.dec for c in pair]
        return out + [self.head]

    def forward(self, Xb):
        x = Xb
        self.skips = []
        for l, (c1, c2) in enumerate(self.enc):
            x = c2.forward(c1.forward(x))
            self.skips.append(x)
            if l < len(self.enc) - 1:
                x = pool2(x)
        for k, (c1, c2) in enumerate(self.dec):
            l = len(self.ch) - 2 - k
            x = concat(self.skips[l], up2(x))
            x = c2.forward(c1.forward(x))
        self.head.forward(x)
        return self.head.A

    def backward(self, dY):
        """dY: FP32 gradient wrt the head output (from mse)."""
        g = self.head.backward(dY)
        skip_grads = [None] * len(self.ch)
        # decoder levels in the reverse of the forward order (finest first)
        for k in reversed(range(len(self.dec))):
            c1, c2 = self.dec[k]
            l = len(self.ch) - 2 - k
            g = c1.backward(c2.backward(g))
            ds, du = split(g, self.ch[l])
            skip_grads[l] = ds
            g = up2_bwd(du)
        for l in range(len(self.ch) - 1, -1, -1):
            c1, c2 = self.enc[l]
            g2 = c2.backward(g, skip_grads[l]) if l < len(self.ch) - 1 else c2.backward(g)
            g = c1.backward(g2, need_input_grad=(l > 0))
            if l > 0:
                g = pool2_bwd(g)

    def step(self, lr=1e-3):
        self.t += 1
        for c in self.layers():
            c.step(self.t, lr)

    def train_step(self, Xb, target, lr=1e-3, energy=0.0):
        """One Adam step on a batch; returns the loss (FP64 device scalar).
        energy > 0 adds the penalty on each window's mean error (mse_energy)."""
        Y = self.forward(Xb)
        loss = _zeros(1, torch.float64)
        dY = mse_energy(Y, target, loss, energy) if energy > 0 else \
            mse(Y, target, loss)
        self.backward(dY)
        self.step(lr)
        return loss

    def flops_per_voxel_fwd(self, D, H, W):
        """Multiply-adds x 2 of the convolutions for one input window."""
        tot = 0
        shapes = [(D, H, W)]
        for _ in range(len(self.ch) - 1):
            d, h, w = shapes[-1]
            shapes.append((d // 2, h // 2, w // 2))
        for l, (c1, c2) in enumerate(self.enc):
            n = shapes[l][0] * shapes[l][1] * shapes[l][2]
            tot += n * 27 * (c1.w.shape[1] * c1.w.shape[2] + c2.w.shape[1] * c2.w.shape[2])
        for k, (c1, c2) in enumerate(self.dec):
            l = len(self.ch) - 2 - k
            n = shapes[l][0] * shapes[l][1] * shapes[l][2]
            tot += n * 27 * (c1.w.shape[1] * c1.w.shape[2] + c2.w.shape[1] * c2.w.shape[2])
        n = D * H * W
        tot += n * 27 * self.head.w.shape[1] * self.head.w.shape[2]
        return 2 * tot


# ---------------------------------------------------------------------------
# Training windows from the plain solver (batched L3 instances)
# ---------------------------------------------------------------------------

T_SCALE = 1000.0
P_SCALE = 400.0


def _copy_field(dev_array):
    """Device copy (library memcpy) of a DeviceArray view as an FP32 tensor."""
    t = torch.empty(dev_array.n, dtype=torch.float32, device="cuda")
    N.check(N.lib().gm_memcpy_d2d(_ptr(t), C.c_void_p(dev_array.ptr), dev_array.nbytes,
                                  _stream()))
    return t


def window_input(inst, grid, T, Tface, row, out):
    """Fill out (bf16, (D, H, W, 16) view of one window) from the fields."""
    n = (C.c_int32 * 3)(grid.n[0], grid.n[1], grid.n[2])
    geo = (C.c_float * 4)(grid.x0[0], grid.x0[1], grid.x0[2], grid.h[0])
    p = inst["properties"]
    prm = (C.c_float * 9)(p.get("T_amb", 300.0), T_SCALE, float(row[0]), float(row[1]),
                          float(row[2]), p["laser_radius"], p.get("laser_depth", 0.05),
                          float(row[6]), P_SCALE)
    N.check(N.lib().gm_nn_window_input(n, geo, _ptr(T), _ptr(Tface), prm, _ptr(out), _stream()))


def make_windows(instances, K=10, warm=20, n_snap=4, gap=10):
    """Run the plain batched L3 step on the instances and cut training pairs:
    input (bf16, (n, D, H, W, 16)) at a step t and target (T(t+K) - T(t)) /
    T_SCALE (FP32, (n, D*H*W)), n = len(instances) * n_snap.  Also returns the
    fields T(t), T(t+K), the face values and the laser rows, for checks."""
    import workloads as W
    import batch as B

    class _TorchStream:
        """The batch runs on torch's current stream, so the copies below are
        ordered after the steps without host synchronisation."""
        ptr = torch.cuda.current_stream().cuda_stream

        @staticmethod
        def sync():
            torch.cuda.current_stream().synchronize()

    check_one_mesh_size(instances)
    bt = B.Batch.from_instances(instances, stream=_TorchStream())
    g0 = bt.grids[0]
    D, H, Wd = g0.n[2], g0.n[1], g0.n[0]
    nw = len(instances) * n_snap
    X = torch.empty((nw, D, H, Wd, 16), dtype=torch.bfloat16, device="cuda")
    Y = torch.empty((nw, D * H * Wd), dtype=torch.float32, device="cuda")
    keep = []
    s = 0
    for _ in range(warm):
        bt.step(W.batch_rows(instances, s))
        s += 1
    w = 0
    for _ in range(n_snap):
        rows = W.batch_rows(instances, s)
        T0 = [_copy_field(bt._field(i, 0)) for i in range(len(instances))]
        faces = [_copy_field(bt._field(i, 2)) for i in range(len(instances))]
        for _k in range(K):
            bt.step(W.batch_rows(instances, s))
            s += 1
        for i in range(len(instances)):
            T1 = _copy_field(bt._field(i, 0))
            window_input(instances[i], bt.grids[i], T0[i], faces[i], rows[i], X[w])
            zero(Y[w])
            N.check(N.lib().gm_nn_axpy(Y[w].numel(), 1.0 / T_SCALE, _ptr(T1), _ptr(Y[w]),
                                       _stream()))
            N.check(N.lib().gm_nn_axpy(Y[w].numel(), -1.0 / T_SCALE, _ptr(T0[i]), _ptr(Y[w]),
                                       _stream()))
            keep.append((T0[i], T1, faces[i], rows[i].copy(), i))
            w += 1
        for _g in range(gap):
            bt.step(W.batch_rows(instances, s))
            s += 1
    torch.cuda.synchronize()
    return X, Y, keep, bt


def apply_jump(net, X, T, Tface, grid, T_floor=300.0):
    """T after one jump of the operator from one window's input X (bf16,
    (1, D, H, W, 16)): T + T_SCALE * prediction inside, faces held."""
    Yh = net.forward(X)
    Tn = torch.empty_like(T)
    n = (C.c_int32 * 3)(grid.n[0], grid.n[1], grid.n[2])
    N.check(N.lib().gm_nn_window_apply(n, _ptr(T), _ptr(Tface), _ptr(Yh), Yh.shape[-1],
                                       T_SCALE, float(T_floor), _ptr(Tn), _stream()))
    return Tn


def _zeros(shape, dtype=torch.float32):
    t = torch.empty(shape, dtype=dtype, device="cuda")
    zero(t)
    return t


def _gather(src, idx, out):
    """out[k] = src[idx[k]] by library copies (rows of a contiguous tensor)."""
    row = src[0].numel() * src.element_size()
    for k, j in enumerate(idx):
        N.check(N.lib().gm_memcpy_d2d(C.c_void_p(out.data_ptr() + k * row),
                                      C.c_void_p(src.data_ptr() + int(j) * row), row, _stream()))
    return out


def rel_error(net, X, Y, batch=4):
    """||prediction - target|| / ||target|| over the windows (the zero
    predictor, i.e. T(t+K) = T(t), scores 1)."""
    num = den = 0.0
    for b0 in range(0, X.shape[0], batch):
        xb = X[b0:b0 + batch]
        yb = Y[b0:b0 + batch].reshape(-1)
        pred = net.forward(xb)
        loss = _zeros(1, torch.float64)
        mse(pred, yb, loss)
        ref = _zeros(1, torch.float64)
        mse(_zeros(pred.shape), yb, ref)
        num += float(loss.item())
        den += float(ref.item())
    return (num / den) ** 0.5 if den > 0 else 0.0


def train(net, X, Y, iters=100, batch=4, lr=1e-3, seed=0, log=None, energy=0.0):
    """Adam on mini-batches drawn without replacement (reshuffled each pass);
    returns the loss history.  energy: weight of the windows' mean-error
    penalty (mse_energy), 0 for the plain MSE."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    if n < batch:
        raise ValueError(f"train: {n} windows for batches of {batch}")
    xb = torch.empty((batch,) + tuple(X.shape[1:]), dtype=X.dtype, device="cuda")
    yb = torch.empty((batch,) + tuple(Y.shape[1:]), dtype=Y.dtype, device="cuda")
    order = []
    hist = []
    for it in range(iters):
        if len(order) < batch:
            order = list(rng.permutation(n)) + order
        idx = [order.pop() for _ in range(batch)]
        _gather(X, idx, xb)
        _gather(Y, idx, yb)
        hist.append(net.train_step(xb, yb.reshape(-1), lr, energy))
        if log and (it % 20 == 0 or it == iters - 1):
            log(f"iter {it} loss {float(hist[-1].item()):.4e}")
    return [float(h.item()) for h in hist]


# ---------------------------------------------------------------------------
# Operator in the loop (batched windows) against the plain solver
# ---------------------------------------------------------------------------

def _jump_audit(inst, grid, T0, Tn, rows, K):
    """Energy audit of a K-step jump T0 -> Tn of one window (gm_energy_balance
    with dt_total = K dt): properties of T0, the source averaged over the
    jump's K laser rows plus the surface flux of T0, Dirichlet faces and the
    T_amb floor held.  Returns the relative residual.  The identity is exact
    for one explicit step; over a jump it carries the change of the
    conduction term across the jump, so the plain solver's own K steps score
    a small nonzero residual that this is compared against."""
    import computeFunctions as cf
    p = cf.SetupProperties(inst["properties"])
    mat = cf._material(p)
    nn = T0.numel()
    st = _stream()
    S1 = _zeros(nn)
    S1o = torch.empty(nn, dtype=torch.float32, device="cuda")
    S2o = torch.empty(nn, dtype=torch.uint8, device="cuda")
    k = torch.empty(nn, dtype=torch.float32, device="cuda")
    rc = torch.empty(nn, dtype=torch.float32, device="cuda")
    N.check(N.lib().gm_state_properties(C.byref(mat), _ptr(T0), _ptr(S1), nn, 0, _ptr(S1o),
                                        _ptr(S2o), _ptr(k), _ptr(rc), st))
    F = _zeros(nn)
    Fk = torch.empty(nn, dtype=torch.float32, device="cuda")
    for r in rows:
        v = (C.c_float * 3)(float(r[0]), float(r[1]), float(r[2]))
        N.check(N.lib().gm_source_l3(C.byref(grid), C.byref(mat), v, float(r[6]), _ptr(Fk), st))
        N.check(N.lib().gm_nn_axpy(nn, 1.0 / len(rows), _ptr(Fk), _ptr(F), st))
    ne = (grid.n[0] - 1) * (grid.n[1] - 1) * (grid.n[2] - 1)
    N.check(N.lib().gm_conv_rad_bc(C.byref(grid), C.byref(mat), _ptr(T0), ne, _ptr(F), st))
    sums = _zeros(8, torch.float64)
    dt = float(rows[0][5]) * K
    N.check(N.lib().gm_energy_balance(C.byref(grid), ne, _ptr(k), _ptr(rc), dt, _ptr(T0), _ptr(F),
                                      None, _ptr(Tn), 31, float(p["T_amb"]), _ptr(sums), st))
    return float(sums[6].item())


def check_one_mesh_size(instances):
    """Raise ValueError unless every instance has the same element counts: the
    operator's window tensors and the per-window kernels (gm_nn_window_input,
    gm_nn_window_apply) take one (D, H, W) for the whole batch."""
    if not instances:
        raise ValueError("no instances")
    el0 = list(instances[0]["elements"])
    for i, b in enumerate(instances):
        if list(b["elements"]) != el0:
            raise ValueError(f"instance {i} has elements {list(b['elements'])}, instance 0 "
                             f"{el0}: the learned operator needs windows of one mesh size")


def run_in_loop(net, instances, n_jumps, K=10, warm=20, compare_every=2, residual_limit=0.01):
    """Advance batched windows by operator jumps of K steps next to the plain
    batched solver from the same start.  Each jump is audited
    (gm_energy_balance over the jump); when any window's relative residual
    exceeds residual_limit the jump is discarded and the windows take K plain
    steps instead (the fallback).  Every compare_every jumps the melt-pool
    depth and width of the two runs are compared (relative error, averaged
    over the windows that have a pool).  Returns a dict of the errors,
    residuals (operator jumps and, for calibration, the plain K steps),
    fallbacks and times (operator jump vs K plain steps, all windows at
    once)."""
    import numpy as np
    import workloads as W
    import batch as B
    import computeFunctions as cf

    class _TorchStream:
        ptr = torch.cuda.current_stream().cuda_stream

        @staticmethod
        def sync():
            torch.cuda.current_stream().synchronize()

    # the network input and output buffers below have one window shape
    check_one_mesh_size(instances)
    plain = B.Batch.from_instances(instances, stream=_TorchStream())
    sur = B.Batch.from_instances(instances, stream=_TorchStream())
    n = len(instances)
    g = plain.grids[0]
    D, H, Wd = g.n[2], g.n[1], g.n[0]
    lv = {"node_coords": g.axes(), "nodes": [g.n[0], g.n[1], g.n[2]], "h": list(g.h)}
    Tl = instances[0]["properties"]["T_liquidus"]
    s = 0
    for _ in range(warm):
        rows = W.batch_rows(instances, s)
        plain.step(rows)
        sur.step(rows)
        s += 1
    X = torch.empty((n, D, H, Wd, 16), dtype=torch.bfloat16, device="cuda")
    t_net = t_plain = 0.0
    fallbacks = 0
    derr, werr, r_sur, r_plain = [], [], [], []
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for j in range(n_jumps):
        rows = W.batch_rows(instances, s)
        T0 = [_copy_field(sur._field(i, 0)) for i in range(n)]
        Tp0 = [_copy_field(plain._field(i, 0)) for i in range(n)]
        faces = [_copy_field(sur._field(i, 2)) for i in range(n)]
        e[0].record()
        for i in range(n):
            window_input(instances[i], sur.grids[i], T0[i], faces[i], rows[i], X[i])
        Yh = net.forward(X)
        Tn = []
        nn3 = (C.c_int32 * 3)(g.n[0], g.n[1], g.n[2])
        for i in range(n):
            t = torch.empty_like(T0[i])
            N.check(N.lib().gm_nn_window_apply(nn3, _ptr(T0[i]), _ptr(faces[i]),
                                               C.c_void_p(Yh.data_ptr() + i * Yh[0].numel() * 4),
                                               Yh.shape[-1], T_SCALE,
                                               float(instances[i]["properties"]["T_amb"]),
                                               _ptr(t), _stream()))
            Tn.append(t)
            d = sur._field(i, 0)
            N.check(N.lib().gm_memcpy_d2d(C.c_void_p(d.ptr), _ptr(t), d.nbytes, _stream()))
            # nodes the jump took over the liquidus keep their melted flag
            N.check(N.lib().gm_batch_jump_state(sur.h, i, _ptr(T0[i])))
        N.check(N.lib().gm_batch_refresh(sur.h))
        e[1].record()
        jrows = [W.batch_rows(instances, s + k) for k in range(K)]
        e[2].record()
        for k in range(K):
            plain.step(jrows[k])
        e[3].record()
        torch.cuda.synchronize()
        t_net += e[0].elapsed_time(e[1])
        t_plain += e[2].elapsed_time(e[3])
        worst = 0.0
        for i in range(n):
            rws = [jr[i] for jr in jrows]
            r = _jump_audit(instances[i], sur.grids[i], T0[i], Tn[i], rws, K)
            r_sur.append(r)
            worst = max(worst, abs(r))
            Tp1 = _copy_field(plain._field(i, 0))
            r_plain.append(_jump_audit(instances[i], plain.grids[i], Tp0[i], Tp1, rws, K))
        if worst > residual_limit:
            # fallback: back to the jump's start, K plain steps
            fallbacks += 1
            for i in range(n):
                d = sur._field(i, 0)
                N.check(N.lib().gm_memcpy_d2d(C.c_void_p(d.ptr), _ptr(T0[i]), d.nbytes,
                                              _stream()))
            N.check(N.lib().gm_batch_refresh(sur.h))
            for k in range(K):
                sur.step(jrows[k])
        s += K
        if (j + 1) % compare_every == 0:
            for i in range(n):
                dp = cf.meltPoolDimensions(lv, _copy_field(plain._field(i, 0)), Tl)
                ds = cf.meltPoolDimensions(lv, _copy_field(sur._field(i, 0)), Tl)
                if dp[0] > 0 and dp[1] > 0:
                    derr.append(abs(ds[0] - dp[0]) / dp[0])
                    werr.append(abs(ds[1] - dp[1]) / dp[1])
    return {
        "jumps": n_jumps, "windows": n, "K": K,
        "melt_pool_rel_err_depth": float(np.mean(derr)) if derr else None,
        "melt_pool_rel_err_width": float(np.mean(werr)) if werr else None,
        "residual_limit": residual_limit,
        "fallbacks": fallbacks,
        "energy_residual_operator_mean_abs": float(np.mean(np.abs(r_sur))) if r_sur else None,
        "energy_residual_plain_mean_abs": float(np.mean(np.abs(r_plain))) if r_plain else None,
        "ms_per_jump_operator": t_net / n_jumps,
        "ms_per_K_plain_steps": t_plain / n_jumps,
        "final_T": ([_copy_field(sur._field(i, 0)) for i in range(n)],
                    [_copy_field(plain._field(i, 0)) for i in range(n)]),
    }


# ---------------------------------------------------------------------------
# Operator in the three-level engine's moving L3 window
# ---------------------------------------------------------------------------

class _Raw:
    """A device pointer with a tensor's data_ptr(), for the wrappers above."""

    def __init__(self, ptr):
        self.p = int(ptr)

    def data_ptr(self):
        return self.p


def _stage_field(e, which):
    p = C.c_void_p()
    n = C.c_int64()
    N.check(N.lib().gm_engine_stage_field(e.h, int(which), C.byref(p), C.byref(n)))
    return _Raw(p.value), int(n.value)


class EngineOperator:
    """The learned operator inside the engine's moving L3 window.

    Attached to a facade engine (``computeFunctions.attach_l3_operator``),
    it runs subcycleGOMELT (cF:3224-3632) stage by stage and replaces the N3
    L3 solves of every L2 substep (cF:3367-3412 in Part 1, cF:3530-3590 in
    Part 2) by one jump: the network sees the window's field at the start of
    the substep, the substep's new L3 face values (the L2-interpolated
    Dirichlet values the last of the N3 solves uses), the laser position and
    power of the substep's first row, and returns the field at its end.

    Each jump is audited with gm_energy_balance over the substep (properties
    and surface flux of the start field, the source averaged over the N3
    rows, faces and the T_amb floor held); when the relative residual exceeds
    ``residual_limit`` the jump is dropped and the N3 plain solves run
    (the fallback).  Every ``judge_every`` jumps the plain solves run from the
    same start as well, and the jump is judged against them (peak T,
    melt-pool depth and width of the L3 window); the run continues on the
    plain field there.  ``record`` (a callable) receives (input field, face
    values, rows of the substep, plain output field) at every substep whose
    plain solves ran: the windows the operator is trained on.
    ``net`` None runs the plain solves everywhere (with ``record``, the
    training-window generator)."""

    def __init__(self, net, properties, residual_limit=0.01, judge_every=200, record=None,
                 graph=True):
        self.net = net
        self.p = properties
        # the network's forward pass recorded once as a graph and replayed
        # on every jump (its launches then cost no host time)
        self.graph = bool(graph)
        self._g = None
        self._gout = None
        self._gx = None
        self.limit = float(residual_limit)
        self.judge_every = int(judge_every) if judge_every else 0
        self.record = record
        self.X = None
        self.stats = {"jumps": 0, "fallbacks": 0, "residuals": [], "judged": []}
        self._tmp = {}

    def _buf(self, name, n, dtype=torch.float32):
        t = self._tmp.get(name)
        if t is None or t.numel() != n or t.dtype != dtype:
            t = torch.empty(n, dtype=dtype, device="cuda")
            self._tmp[name] = t
        return t

    def _audit(self, e, g, T0, Tn, S1u8, rows):
        import computeFunctions as cf
        lib = N.lib()
        nn = g.n[0] * g.n[1] * g.n[2]
        st = _stream()
        mat = cf._material(self.p)
        S1 = self._buf("S1", nn)
        N.check(lib.gm_u8_to_f32(nn, _ptr(S1u8), _ptr(S1), st))
        S1o, k, rc = self._buf("S1o", nn), self._buf("k", nn), self._buf("rc", nn)
        S2o = self._buf("S2o", nn, torch.uint8)
        N.check(lib.gm_state_properties(C.byref(mat), _ptr(T0), _ptr(S1), nn, 0, _ptr(S1o),
                                        _ptr(S2o), _ptr(k), _ptr(rc), st))
        F = self._buf("F", nn)
        r32 = np.ascontiguousarray(rows, dtype=np.float32)
        N.check(lib.gm_source_l3_mean(C.byref(g), C.byref(mat), N.fptr(r32), r32.shape[0],
                                      _ptr(F), st))
        ne = (g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1)
        N.check(lib.gm_conv_rad_bc(C.byref(g), C.byref(mat), _ptr(T0), ne, _ptr(F), st))
        sums = self._buf("sums", 8, torch.float64)
        zero(sums)
        dt = float(np.float32(rows[:, 5]).sum())
        N.check(lib.gm_energy_balance(C.byref(g), ne, _ptr(k), _ptr(rc), dt, _ptr(T0), _ptr(F),
                                      None, _ptr(Tn), 31, float(self.p["T_amb"]), _ptr(sums), st))
        return float(sums[6].item())

    def _pool(self, g, T):
        import computeFunctions as cf
        lv = {"node_coords": g.axes(), "nodes": [g.n[0], g.n[1], g.n[2]], "h": list(g.h)}
        return cf.meltPoolDimensions(lv, T, self.p["T_liquidus"])

    def subcycle(self, e, rows):
        """One L1 step (N2*N3 rows) with the operator in the L3 window."""
        lib = N.lib()
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        with torch.cuda.stream(torch.cuda.ExternalStream(e.stream.ptr)):
            N.check(lib.gm_engine_stage_begin(e.h, N.fptr(rows), rows.shape[0]))
            part, k = C.c_int32(), C.c_int32()
            for _ in range(2 * e.N2):
                N.check(lib.gm_engine_stage_l2(e.h, C.byref(part), C.byref(k)))
                self._substep(e, rows[k.value * e.N3:(k.value + 1) * e.N3])

    def _substep(self, e, rows):
        lib = N.lib()
        T0, n3 = _stage_field(e, N.GM_STAGE_L3_IN)
        out, _ = _stage_field(e, N.GM_STAGE_L3_OUT)
        fa, _ = _stage_field(e, N.GM_STAGE_FACE_NEW)
        s1, _ = _stage_field(e, N.GM_STAGE_S1_IN)
        g = e.grid(3)
        if self.net is None:
            if self.record is not None:
                T0c = self._copy(T0, n3)
                fac = self._copy(fa, n3)
                N.check(lib.gm_engine_stage_l3(e.h, 0))
                last, _ = _stage_field(e, N.GM_STAGE_L3_LAST)
                self.record(T0c, fac, rows.copy(), self._copy(last, n3))
            else:
                N.check(lib.gm_engine_stage_l3(e.h, 0))
            return
        shape = (1, g.n[2], g.n[1], g.n[0], 16)
        if self.X is None or tuple(self.X.shape) != shape:
            self.X = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
        window_input({"properties": self.p}, g, T0, fa, rows[0], self.X[0])
        Yh = self._forward()
        nn3 = (C.c_int32 * 3)(g.n[0], g.n[1], g.n[2])
        N.check(lib.gm_nn_window_apply(nn3, _ptr(T0), _ptr(fa), _ptr(Yh), Yh.shape[-1], T_SCALE,
                                       float(self.p["T_amb"]), _ptr(out), _stream()))
        res = self._audit(e, g, T0, out, s1, rows)
        self.stats["jumps"] += 1
        self.stats["residuals"].append(res)
        if self.judge_every and self.stats["jumps"] % self.judge_every == 0:
            Tj = self._copy(out, n3)
            T0c = self._copy(T0, n3)  # the N3 solves may reuse the start field's buffer
            N.check(lib.gm_engine_stage_l3(e.h, 0))
            last, _ = _stage_field(e, N.GM_STAGE_L3_LAST)
            Tp = self._copy(last, n3)
            # the audit of the plain solves over the same substep, for scale
            res_plain = self._audit(e, g, T0c, Tp, s1, rows)
            e.sync()
            pj, pp = self._pool(g, Tj), self._pool(g, Tp)
            # peak temperatures on the host (a judged jump is rare)
            peak_j = float(np.max(N.DeviceArray(n3, ptr=Tj.data_ptr(), owner=False).numpy()))
            peak_p = float(np.max(N.DeviceArray(n3, ptr=Tp.data_ptr(), owner=False).numpy()))
            rel = lambda a, b: abs(a - b) / b if b > 0 else (0.0 if a == 0 else float("inf"))
            self.stats["judged"].append({
                "jump": self.stats["jumps"], "residual": res, "residual_plain": res_plain,
                "peak_T_rel_err": rel(peak_j, peak_p),
                "depth_rel_err": rel(pj[0], pp[0]), "width_rel_err": rel(pj[1], pp[1]),
                "depth_mm": [pj[0], pp[0]], "width_mm": [pj[1], pp[1]]})
            return
        if not (abs(res) <= self.limit):
            self.stats["fallbacks"] += 1
            N.check(lib.gm_engine_stage_l3(e.h, 0))
            return
        N.check(lib.gm_engine_stage_l3(e.h, 1))

    def _copy(self, raw, n):
        t = torch.empty(n, dtype=torch.float32, device="cuda")
        N.check(N.lib().gm_memcpy_d2d(_ptr(t), C.c_void_p(raw.p), 4 * n, _stream()))
        return t

    def _forward(self):
        """The network on self.X: eager, or the replay of its recorded graph
        (recorded on first use; the network's kernels allocate through
        PyTorch and clear nothing with memset, so the graph holds kernels
        only).  The output tensor is the graph's own, reused by every
        replay."""
        if not self.graph:
            return self.net.forward(self.X)
        if self._g is None or self._gx is not self.X:
            cur = torch.cuda.current_stream()
            side = torch.cuda.Stream()
            side.wait_stream(cur)
            with torch.cuda.stream(side):  # warm-up outside the capture
                self.net.forward(self.X)
            cur.wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                out = self.net.forward(self.X)
            self._g, self._gout, self._gx = g, out, self.X
        self._g.replay()
        return self._gout


def engine_windows(n_windows, seed0=0, n_l1=(80, 80, 40), length=4.0, per_run=1000,
                   max_runs=400, log=None):
    """Training pairs from the plain three-level engine: runs of
    workloads.window_deck(seed) for seed = seed0, seed0 + 1, ... (P, speed and
    path drawn from the seed), and at every L2 substep of every subcycled L1
    step the window's field at the substep's start, its new face values and
    first laser row (the input, as EngineOperator builds it) and the field
    after the N3 plain L3 solves (target (T1 - T0) / T_SCALE), at most
    `per_run` windows from each run.  Returns
    X (bf16, (n, 32, 64, 64, 16)), Y (FP32, (n, 64*64*32)) and the seed of each
    window."""
    import json
    import tempfile
    import go_melt as G
    import computeFunctions as cf
    import workloads as W
    e3 = W.OP_L3_ELEMENTS
    D, H, Wd = e3[2] + 1, e3[1] + 1, e3[0] + 1
    X = torch.empty((n_windows, D, H, Wd, 16), dtype=torch.bfloat16, device="cuda")
    Y = torch.empty((n_windows, D * H * Wd), dtype=torch.float32, device="cuda")
    seeds = []
    state = {"w": 0}
    seed = seed0
    while state["w"] < n_windows and seed - seed0 < max_runs:
        inp, g = W.window_deck(seed, n_l1=n_l1, length=length)
        d = tempfile.mkdtemp(prefix="gomelt_win_")
        with open(W.write_deck(inp, g, d)) as fh:
            deck = json.load(fh)
        run = G.TimeLoop(deck, quiet=True, write_outputs=False)
        p = run.Properties

        state["run_w"] = 0

        def rec(T0, face, rows, T1, run=run, p=p, seed=seed):
            w = state["w"]
            if w >= n_windows or state["run_w"] >= per_run:
                return
            state["run_w"] += 1
            window_input({"properties": p}, run.eng.grid(3), T0, face, rows[0], X[w])
            zero(Y[w])
            N.check(N.lib().gm_nn_axpy(Y[w].numel(), 1.0 / T_SCALE, _ptr(T1), _ptr(Y[w]),
                                       _stream()))
            N.check(N.lib().gm_nn_axpy(Y[w].numel(), -1.0 / T_SCALE, _ptr(T0), _ptr(Y[w]),
                                       _stream()))
            seeds.append(seed)
            state["w"] = w + 1

        cf.attach_l3_operator(run.Levels, EngineOperator(None, p, record=rec))
        for rows in run.chunks():
            if state["w"] >= n_windows or state["run_w"] >= per_run or \
                    not run.step_chunk(rows):
                break
        run.eng.sync()
        if log:
            log(f"run {seed}: {state['w']} windows")
        del run
        seed += 1
    w = state["w"]
    torch.cuda.synchronize()
    return X[:w], Y[:w], seeds
