"""GPU convolution kernels of the learned L3 operator (BF16 MFMA, FP32
accumulation) against the FP64 oracle (oracle/conv_oracle.py) on the same
bf16-rounded inputs; tolerances stated per test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch
    import _native as N
    N.init_device()
    import surrogate as S
    return S


def _dev_bf16(a):
    import torch
    from oracle import conv_oracle as CO
    bits = CO.bf16_bits(np.ascontiguousarray(a, np.float32))
    return torch.from_numpy(bits.view(np.int16)).view(torch.bfloat16).cuda()


def _rel(g, r):
    return float(np.max(np.abs(g - r)) / max(np.max(np.abs(r)), 1e-30))


# the kernels chosen by channel counts: 16-channel K stages (cin 16, 48), the
# pipelined 32-channel stages with 128 x 32 tiles (cout <= 32) and 64 x 64
# tiles, the weight gradient's 32 x 32, 32-multiple and general kernels
SHAPES = [(2, 5, 6, 7, 16, 32), (1, 8, 8, 4, 32, 16), (1, 3, 3, 3, 16, 48), (1, 9, 7, 11, 48, 80),
          (2, 4, 6, 8, 32, 32), (1, 5, 6, 7, 64, 96), (1, 7, 5, 9, 96, 32),
          (1, 6, 8, 10, 16, 32), (1, 5, 7, 9, 32, 16), (2, 3, 4, 5, 16, 16),
          # 128-voxel blocks of whole rows, W from 2 to 64
          (2, 3, 4, 32, 32, 32), (1, 3, 2, 64, 64, 16), (1, 2, 64, 2, 96, 32),
          # a larger volume: 32768 voxels, 128 outputs
          (2, 8, 32, 64, 32, 128),
          # 2^18 voxels: 4096-voxel parts in the weight gradient, and the
          # forward kernel with one sub-stage per barrier pair (the smaller
          # volumes above take three): 256 x 32 blocks of 64 x 32 waves (up to
          # 32 outputs, also with 16 input channels), 64 x 64 blocks beyond
          (1, 8, 128, 256, 32, 32), (1, 8, 128, 256, 16, 32), (1, 8, 128, 256, 64, 64),
          (2, 8, 64, 128, 128, 128)]


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_forward(S, shape):
    """Y = conv(X; W) + b, relu off and on: 1e-5 of max |Y| (FP32 sums of
    exact bf16 products)."""
    import torch
    from oracle import conv_oracle as CO
    nb, D, H, W, ci, co = shape
    rng = np.random.default_rng(sum(shape))
    X = CO.bf16_round(rng.standard_normal((nb, D, H, W, ci)).astype(np.float32))
    Wk = CO.bf16_round((0.2 * rng.standard_normal((27, co, ci))).astype(np.float32))
    b = rng.standard_normal(co).astype(np.float32)
    ref = CO.conv3d(X, Wk, b)
    for relu in (False, True):
        Y, Yb = S.conv3d(_dev_bf16(X), _dev_bf16(Wk), torch.from_numpy(b).cuda(), relu=relu)
        torch.cuda.synchronize()
        r = np.maximum(ref, 0) if relu else ref
        assert _rel(Y.cpu().numpy(), r) < 1e-5
        # the bf16 copy is the rounded FP32 output
        yb = Yb.float().cpu().numpy()
        np.testing.assert_array_equal(yb, CO.bf16_round(Y.cpu().numpy()))


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_gradients(S, shape):
    """Input and weight gradients against the oracle: 1e-5 of the max; the
    weight gradient is added to what dW holds."""
    import torch
    from oracle import conv_oracle as CO
    nb, D, H, W, ci, co = shape
    rng = np.random.default_rng(7 + sum(shape))
    X = CO.bf16_round(rng.standard_normal((nb, D, H, W, ci)).astype(np.float32))
    Wk = CO.bf16_round((0.2 * rng.standard_normal((27, co, ci))).astype(np.float32))
    dY = CO.bf16_round(rng.standard_normal((nb, D, H, W, co)).astype(np.float32))
    dX, _ = S.conv3d_bwd_data(_dev_bf16(dY), _dev_bf16(Wk))
    torch.cuda.synchronize()
    assert _rel(dX.cpu().numpy(), CO.conv3d_bwd_data(dY, Wk)) < 1e-5
    base = rng.standard_normal((27, co, ci)).astype(np.float32)
    dW = torch.from_numpy(base.copy()).cuda()
    S.conv3d_bwd_weight(_dev_bf16(X), _dev_bf16(dY), dW)
    torch.cuda.synchronize()
    ref = CO.conv3d_bwd_weight(X, dY)
    assert _rel(dW.cpu().numpy() - base, ref) < 1e-5


def test_two_layer_training_gradients(S):
    """Gradients of L = 0.5 |conv(relu(conv(X)+b1))+b2 - T|^2 through the
    kernels: with the oracle fed the same bf16-rounded intermediates, 1e-5;
    against the oracle's all-FP64 gradients, 3e-2 (bf16 activations)."""
    import torch
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(5)
    nb, D, H, W, ci, ch, co = 2, 6, 5, 4, 16, 32, 16
    X = CO.bf16_round(rng.standard_normal((nb, D, H, W, ci)).astype(np.float32))
    W1 = CO.bf16_round((0.1 * rng.standard_normal((27, ch, ci))).astype(np.float32))
    W2 = CO.bf16_round((0.1 * rng.standard_normal((27, co, ch))).astype(np.float32))
    b1 = (0.1 * rng.standard_normal(ch)).astype(np.float32)
    b2 = (0.1 * rng.standard_normal(co)).astype(np.float32)
    T = rng.standard_normal((nb, D, H, W, co)).astype(np.float32)
    Z1, A1b = S.conv3d(_dev_bf16(X), _dev_bf16(W1), torch.from_numpy(b1).cuda(), relu=True)
    Y, _ = S.conv3d(A1b, _dev_bf16(W2), torch.from_numpy(b2).cuda())
    R = Y - torch.from_numpy(T).cuda()
    Rb = S.to_bf16(R.contiguous())
    dW2 = torch.zeros((27, co, ch), dtype=torch.float32, device="cuda")
    S.conv3d_bwd_weight(A1b, Rb, dW2)
    dA1, _ = S.conv3d_bwd_data(Rb, _dev_bf16(W2))
    dZ1 = dA1 * (Z1 > 0)
    dZ1b = S.to_bf16(dZ1.contiguous())
    dW1 = torch.zeros((27, ch, ci), dtype=torch.float32, device="cuda")
    S.conv3d_bwd_weight(_dev_bf16(X), dZ1b, dW1)
    torch.cuda.synchronize()
    # matched: the oracle on the GPU's own bf16 intermediates
    a1 = A1b.float().cpu().numpy()
    rb = Rb.float().cpu().numpy()
    dz1b = dZ1b.float().cpu().numpy()
    assert _rel(dW2.cpu().numpy(), CO.conv3d_bwd_weight(a1, rb)) < 1e-5
    assert _rel(dA1.cpu().numpy(), CO.conv3d_bwd_data(rb, W2)) < 1e-5
    assert _rel(dW1.cpu().numpy(), CO.conv3d_bwd_weight(X, dz1b)) < 1e-5
    # unmatched: all-FP64 chain rule of the oracle
    _, (gW1, gb1, gW2, gb2), _ = CO.two_layer_loss(X, W1, b1, W2, b2, T)
    assert _rel(dW2.cpu().numpy(), gW2) < 3e-2
    assert _rel(dW1.cpu().numpy(), gW1) < 3e-2


def test_adam(S):
    import torch
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(9)
    n = 5000
    w = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    tw, tm, tv = (torch.from_numpy(a.copy()).cuda() for a in (w, m, v))
    twb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    w64, m64, v64 = w.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in (1, 2, 3):
        g = rng.standard_normal(n).astype(np.float32)
        S.adam(tw, twb, torch.from_numpy(g).cuda(), tm, tv, step, lr, b1, b2, eps)
        m64 = b1 * m64 + (1 - b1) * g
        v64 = b2 * v64 + (1 - b2) * g.astype(np.float64) ** 2
        w64 = w64 - lr * (m64 / (1 - b1 ** step)) / (np.sqrt(v64 / (1 - b2 ** step)) + eps)
    torch.cuda.synchronize()
    assert float(np.max(np.abs(tw.cpu().numpy() - w64))) < 1e-5
    np.testing.assert_array_equal(twb.float().cpu().numpy(), CO.bf16_round(tw.cpu().numpy()))
