"""CPU trace of the operator network's forward and backward passes: every
kernel wrapper is replaced by a shape-checking stand-in, so a wiring mistake
(a tensor handed to the wrong level) fails here instead of on the GPU."""
import pytest


class _Shapes:
    """Stand-ins of the surrogate module's kernel calls (CPU tensors)."""

    def __init__(self, torch):
        self.t = torch

    def conv3d(self, X, Wk, bias=None, relu=False, want_bf16=True):
        assert X.dim() == 5 and Wk.shape == (27, Wk.shape[1], X.shape[4]), (X.shape, Wk.shape)
        assert X.shape[4] % 16 == 0 and Wk.shape[1] % 16 == 0
        Y = self.t.empty(X.shape[:4] + (Wk.shape[1],))
        return Y, Y.to(self.t.bfloat16)

    def conv3d_bwd_data(self, dY, Wk, want_bf16=True):
        assert dY.shape[4] == Wk.shape[1]
        return self.t.empty(dY.shape[:4] + (Wk.shape[2],)), None

    def conv3d_bwd_weight(self, X, dY, dW):
        assert X.shape[:4] == dY.shape[:4] and dW.shape == (27, dY.shape[4], X.shape[4])
        return dW

    def relu_bwd(self, dA, A, dA2=None):
        assert A is None or A.shape == dA.shape, (A.shape, dA.shape)
        assert dA2 is None or dA2.shape == dA.shape, (dA2.shape, dA.shape)
        return self.t.empty(dA.shape), self.t.empty(dA.shape)

    def bias_grad(self, dZ, db):
        assert db.numel() == dZ.shape[-1]

    def relu_bias_bwd(self, dA, A, db, dA2=None):
        self.bias_grad(dA, db)
        return self.relu_bwd(dA, A, dA2)[1]

    def pool2(self, X):
        assert all(s % 2 == 0 for s in X.shape[1:4])
        return self.t.empty(X.shape[:1] + tuple(s // 2 for s in X.shape[1:4]) + X.shape[4:])

    def pool2_bwd(self, dY):
        return self.t.empty(dY.shape[:1] + tuple(2 * s for s in dY.shape[1:4]) + dY.shape[4:])

    up2 = pool2_bwd

    def up2_bwd(self, dY):
        return self.pool2(dY)

    def concat(self, A, B):
        assert A.shape[:4] == B.shape[:4], (A.shape, B.shape)
        return self.t.empty(A.shape[:4] + (A.shape[4] + B.shape[4],))

    def split(self, dY, ca):
        return (self.t.empty(dY.shape[:4] + (ca,)),
                self.t.empty(dY.shape[:4] + (dY.shape[4] - ca,)))

    def mse(self, Y, target, loss):
        assert target.numel() == Y.numel() // Y.shape[-1]
        return self.t.empty(Y.shape)

    def to_bf16(self, a):
        return a.to(self.t.bfloat16)

    def zeros(self, shape, dtype=None):
        return self.t.zeros(shape)


@pytest.mark.parametrize("channels,window", [((16, 32), (8, 8, 8)), ((16, 32, 32), (16, 32, 32)),
                                             ((32, 64, 128, 128), (32, 64, 64))])
def test_unet_pass_shapes(monkeypatch, channels, window):
    torch = pytest.importorskip("torch")
    import surrogate as S
    f = _Shapes(torch)
    for name in ("conv3d", "conv3d_bwd_data", "conv3d_bwd_weight", "relu_bwd", "bias_grad",
                 "relu_bias_bwd", "pool2", "pool2_bwd", "up2", "up2_bwd", "concat", "split", "mse",
                 "to_bf16"):
        monkeypatch.setattr(S, name, getattr(f, name))
    monkeypatch.setattr(S, "_zeros", f.zeros)
    monkeypatch.setattr(torch.Tensor, "cuda", lambda self, *a, **k: self)
    net = S.UNet(channels=channels, seed=0)
    X = torch.empty((2,) + window + (16,), dtype=torch.bfloat16)
    Y = net.forward(X)
    assert Y.shape == (2,) + window + (16,)
    net.backward(torch.empty(Y.shape))
