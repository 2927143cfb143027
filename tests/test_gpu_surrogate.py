"""The learned L3 operator's network kernels against NumPy, and a short
training run on windows of the plain solver (the operator must beat the
persistence predictor T(t+K) = T(t) on held-out windows)."""
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


def _bf(a):
    import torch
    from oracle import conv_oracle as CO
    bits = CO.bf16_bits(np.ascontiguousarray(a, np.float32))
    return torch.from_numpy(bits.view(np.int16)).view(torch.bfloat16).cuda()


def _np(t):
    import torch
    if t.dtype == torch.bfloat16:
        return (t.view(torch.int16).cpu().numpy().astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
    return t.cpu().numpy()


def test_network_elementwise_kernels(S):
    """pool/unpool, upsample, concat/split, ReLU backward, bias gradient and
    the loss against NumPy (exact or FP32 rounding)."""
    import torch
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(0)
    X = CO.bf16_round(rng.standard_normal((2, 4, 6, 8, 16)).astype(np.float32))
    P = S.pool2(_bf(X))
    ref = X.reshape(2, 2, 2, 3, 2, 4, 2, 16).mean(axis=(2, 4, 6))
    np.testing.assert_allclose(_np(P), CO.bf16_round(ref), rtol=0, atol=0)
    g = rng.standard_normal((2, 2, 3, 4, 16)).astype(np.float32)
    dX = _np(S.pool2_bwd(torch.from_numpy(g).cuda()))
    np.testing.assert_array_equal(dX, np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8)
    U = _np(S.up2(_bf(ref)))
    np.testing.assert_array_equal(U, np.repeat(np.repeat(np.repeat(CO.bf16_round(ref), 2, 1), 2, 2), 2, 3))
    gu = rng.standard_normal((2, 4, 6, 8, 16)).astype(np.float32)
    du = _np(S.up2_bwd(torch.from_numpy(gu).cuda()))
    np.testing.assert_allclose(du, gu.reshape(2, 2, 2, 3, 2, 4, 2, 16).sum(axis=(2, 4, 6)),
                               rtol=1e-6, atol=1e-6)
    B = CO.bf16_round(rng.standard_normal((2, 4, 6, 8, 32)).astype(np.float32))
    Cc = _np(S.concat(_bf(X), _bf(B)))
    np.testing.assert_array_equal(Cc, np.concatenate([X, B], axis=-1))
    gc = rng.standard_normal((2, 4, 6, 8, 48)).astype(np.float32)
    da, db = S.split(torch.from_numpy(gc).cuda(), 16)
    np.testing.assert_array_equal(_np(da), gc[..., :16])
    np.testing.assert_array_equal(_np(db), gc[..., 16:])
    A = np.maximum(rng.standard_normal((5000, 16)).astype(np.float32), 0)
    dA = rng.standard_normal((5000, 16)).astype(np.float32)
    dA2 = rng.standard_normal((5000, 16)).astype(np.float32)
    dZ, dZb = S.relu_bwd(torch.from_numpy(dA).cuda(), torch.from_numpy(A).cuda(),
                         torch.from_numpy(dA2).cuda())
    np.testing.assert_array_equal(_np(dZ), np.where(A > 0, dA + dA2, 0.0).astype(np.float32))
    np.testing.assert_array_equal(_np(dZb), CO.bf16_round(_np(dZ)))
    dbias = torch.from_numpy(np.ones(16, np.float32)).cuda()
    S.bias_grad(dZ, dbias)
    np.testing.assert_allclose(_np(dbias), 1 + _np(dZ).astype(np.float64).sum(0), rtol=1e-5)
    Y = rng.standard_normal((3000, 16)).astype(np.float32)
    T = rng.standard_normal(3000).astype(np.float32)
    loss = torch.from_numpy(np.zeros(1)).cuda()
    dY = _np(S.mse(torch.from_numpy(Y).cuda(), torch.from_numpy(T).cuda(), loss))
    r = Y[:, 0].astype(np.float64) - T
    # FP32 residuals, FP64 sum
    assert float(loss.item()) == pytest.approx(0.5 * np.mean(r * r), rel=1e-6)
    np.testing.assert_allclose(dY[:, 0], r / 3000, rtol=1e-6)
    assert np.all(dY[:, 1:] == 0)


def test_network_copies_odd_channels(S):
    """Channel counts that the 16-byte copy kernels do not take (not a
    multiple of 8 bf16 or 4 floats) run the one-value kernels: same results."""
    import torch
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(3)
    X = CO.bf16_round(rng.standard_normal((2, 2, 3, 4, 6)).astype(np.float32))
    U = _np(S.up2(_bf(X)))
    np.testing.assert_array_equal(U, np.repeat(np.repeat(np.repeat(X, 2, 1), 2, 2), 2, 3))
    g = rng.standard_normal((2, 2, 3, 4, 6)).astype(np.float32)
    dX = _np(S.pool2_bwd(torch.from_numpy(g).cuda()))
    np.testing.assert_array_equal(dX, np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8)
    B = CO.bf16_round(rng.standard_normal((2, 2, 3, 4, 5)).astype(np.float32))
    np.testing.assert_array_equal(_np(S.concat(_bf(X), _bf(B))), np.concatenate([X, B], axis=-1))
    gc = rng.standard_normal((2, 2, 3, 4, 11)).astype(np.float32)
    da, db = S.split(torch.from_numpy(gc).cuda(), 6)
    np.testing.assert_array_equal(_np(da), gc[..., :6])
    np.testing.assert_array_equal(_np(db), gc[..., 6:])
    # a large loss: many blocks, each adding once
    Y = rng.standard_normal((300000, 16)).astype(np.float32)
    T = rng.standard_normal(300000).astype(np.float32)
    loss = torch.from_numpy(np.zeros(1)).cuda()
    dY = _np(S.mse(torch.from_numpy(Y).cuda(), torch.from_numpy(T).cuda(), loss))
    r = Y[:, 0].astype(np.float64) - T
    assert float(loss.item()) == pytest.approx(0.5 * np.mean(r * r), rel=1e-6)
    np.testing.assert_allclose(dY[:, 0], r / 300000, rtol=1e-6)
    assert np.all(dY[:, 1:] == 0)


@pytest.mark.parametrize("nc,nvox", [(16, 5000), (32, 70000), (128, 3000), (256, 1100), (4, 777)])
@pytest.mark.parametrize("with_a,with_a2", [(True, True), (True, False), (False, False)])
def test_relu_bias_fused(S, nc, nvox, with_a, with_a2):
    """The fused ReLU backward + bias gradient against the two separate
    kernels: identical bf16 and FP32 dZ, bias sums equal to FP32 rounding
    (they add in a different order, and atomically) and to the FP64 sum of
    dZ."""
    import ctypes
    import torch
    import _native as N
    rng = np.random.default_rng(nc + nvox)
    dA = torch.from_numpy(rng.standard_normal((nvox, nc)).astype(np.float32)).cuda()
    dA2 = torch.from_numpy(rng.standard_normal((nvox, nc)).astype(np.float32)).cuda() if with_a2 else None
    A = (torch.from_numpy(np.maximum(rng.standard_normal((nvox, nc)), 0).astype(np.float32)).cuda()
         if with_a else None)
    db0 = torch.from_numpy(np.linspace(-1, 1, nc).astype(np.float32)).cuda()
    dZ_ref, dZb_ref = S.relu_bwd(dA, A, dA2)
    db_ref = db0.clone()
    S.bias_grad(dZ_ref, db_ref)
    db = db0.clone()
    dZb = S.relu_bias_bwd(dA, A, db, dA2)
    np.testing.assert_array_equal(_np(dZb), _np(dZb_ref))
    # FP32 sums of nvox terms of unit size whose running partial sums reach
    # ~sqrt(nvox); up to 256 atomic adds in no fixed order each round at half
    # an ulp of a partial sum: |error| <~ 256 * 2^-24 * 4 sqrt(nvox), which is
    # 1.6e-2 at 70,000 voxels (measured up to 2.2e-4 where a channel's sum
    # cancels to 1e-2)
    atol = 256 * 2.0 ** -24 * 4 * np.sqrt(nvox)
    np.testing.assert_allclose(_np(db), _np(db_ref), rtol=2e-5, atol=atol)
    exact = _np(db0).astype(np.float64) + _np(dZ_ref).astype(np.float64).sum(0)
    np.testing.assert_allclose(_np(db), exact, rtol=2e-5, atol=atol)
    # the FP32 dZ output of the C entry point, called directly
    dZ = torch.empty_like(dA)
    db2 = db0.clone()
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N.check(N.lib().gm_nn_relu_bias_bwd(nvox, nc, p(dA), p(dA2), p(A), p(dZ), None, p(db2), st))
    np.testing.assert_array_equal(_np(dZ), _np(dZ_ref))
    # up to 256 blocks add their partial sums atomically, in no fixed order:
    # each add rounds at half an ulp of the running sum
    np.testing.assert_allclose(_np(db2), _np(db), rtol=2e-5, atol=atol)


def test_relu_bias_fused_rejects(S):
    """Channel counts outside the fused kernel's layout are refused by the
    C entry point; the Python wrapper routes them through the two kernels."""
    import ctypes
    import torch
    import _native as N
    dA = torch.ones((10, 48), dtype=torch.float32).cuda()
    db = torch.zeros(48, dtype=torch.float32).cuda()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert N.lib().gm_nn_relu_bias_bwd(10, 48, p(dA), None, None, None, None, p(db), st) != 0
    dZb = S.relu_bias_bwd(dA, None, db)
    np.testing.assert_array_equal(_np(db), np.full(48, 10.0, np.float32))
    np.testing.assert_array_equal(_np(dZb), np.ones((10, 48), np.float32))


def test_unet_gradient_direction(S):
    """Directional derivative of the loss through the whole U-Net: the
    kernels' gradient g predicts L(w + e d) - L(w - e d) ~ 2 e <g, d> for a
    random weight direction d (BF16 weights and activations: 10%)."""
    import torch
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(3)
    net = S.UNet(channels=(16, 32), seed=1)
    X = _bf(0.5 * rng.standard_normal((2, 8, 8, 8, 16)).astype(np.float32))
    T = torch.from_numpy(0.1 * rng.standard_normal(2 * 8 * 8 * 8).astype(np.float32)).cuda()

    def loss():
        out = net.forward(X)
        l = S._zeros(1, torch.float64)
        S.mse(out, T, l)
        return float(l.item())

    out = net.forward(X)
    l0 = S._zeros(1, torch.float64)
    dY = S.mse(out, T, l0)
    net.backward(dY)
    layers = net.layers()
    dirs = [rng.standard_normal(tuple(c.w.shape)).astype(np.float32) for c in layers]
    w0 = [_np(c.w).copy() for c in layers]
    eps = 1e-1
    vals = []
    for sgn in (1, -1):
        for c, w, d in zip(layers, w0, dirs):
            c.w.copy_(torch.from_numpy(w + sgn * eps * d * np.abs(w).mean()))
            c.wb = S.to_bf16(c.w)
        vals.append(loss())
    scale = [np.abs(w).mean() for w in w0]
    gd = sum(float((_np(c.grads[0]).astype(np.float64) * d * s).sum())
             for c, d, s in zip(layers, dirs, scale))
    fd = (vals[0] - vals[1]) / (2 * eps)
    assert fd == pytest.approx(gd, rel=1e-1), (fd, gd)


def test_operator_training_beats_persistence(S):
    """Windows of the plain batched solver (32 x 32 x 16 nodes, K = 10
    steps); a small U-Net trained for 150 steps predicts T(t+K) better than
    T(t) on windows of instances it did not see."""
    import workloads as W
    inst_tr = W.batch_instances(n=8, seed=11, sizes=[(31, 31, 15)])
    inst_te = W.batch_instances(n=4, seed=12, sizes=[(31, 31, 15)])
    Xtr, Ytr, _, _ = S.make_windows(inst_tr, K=10, warm=10, n_snap=3, gap=5)
    Xte, Yte, _, _ = S.make_windows(inst_te, K=10, warm=10, n_snap=2, gap=5)
    net = S.UNet(channels=(16, 32, 32), seed=0)
    e0 = S.rel_error(net, Xte, Yte)
    hist = S.train(net, Xtr, Ytr, iters=150, batch=4, lr=2e-3)
    e1 = S.rel_error(net, Xte, Yte)
    print("held-out relative error: untrained", e0, "trained", e1, "loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.5 * hist[0]
    assert e1 < 0.7


def test_operator_in_the_loop_fallback(S):
    """In the loop, the energy audit separates the plain solver's K steps
    (residual 1e-7 to 1e-4) from an untrained operator's jump (residual far above
    1%); every such jump falls back to K plain steps, so the run ends
    bitwise on the plain solver's fields."""
    import numpy as np
    import workloads as W
    inst = W.batch_instances(n=3, seed=13, sizes=[(31, 31, 15)])
    net = S.UNet(channels=(16, 32), seed=4)
    out = S.run_in_loop(net, inst, n_jumps=3, K=10, warm=10, compare_every=3,
                        residual_limit=0.01)
    print({k: v for k, v in out.items() if k != "final_T"})
    assert out["energy_residual_plain_mean_abs"] < 1e-3
    assert out["energy_residual_operator_mean_abs"] > 0.01
    assert out["fallbacks"] == 3
    sur, plain = out["final_T"]
    for a, b in zip(sur, plain):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_energy_consistent_loss(S):
    """mse_energy against NumPy: loss 0.5 mean r^2 + 0.5 lam mean_w m_w^2 and
    gradient (r + lam m_w [free]) / nvox on channel 0 (r = Y[..., 0] - target,
    m_w the sum of r over window w's voxels off the faces x = 0, x = max,
    y = 0, y = max, z = 0, over the window's voxel count); a planted uniform
    offset on one window's free voxels moves m_w by the offset times their
    share, an offset on its faces leaves m_w alone."""
    import torch
    rng = np.random.default_rng(21)
    nb, D, H, Wd, C, lam = 3, 10, 20, 25, 16, 250.0
    nw = D * H * Wd
    Y = rng.standard_normal((nb, D, H, Wd, C)).astype(np.float32)
    T = rng.standard_normal(nb * nw).astype(np.float32)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(Wd), indexing="ij")
    free = ~((x == 0) | (x == Wd - 1) | (y == 0) | (y == H - 1) | (z == 0))
    free = free.reshape(-1)

    def ref(Yn):
        r = Yn[..., 0].astype(np.float64).reshape(nb, nw) - T.astype(np.float64).reshape(nb, nw)
        m = (r * free).sum(axis=1) / nw
        return r, m, 0.5 * np.mean(r ** 2) + 0.5 * lam * np.mean(m ** 2)

    loss = S._zeros(1, torch.float64)
    dY = _np(S.mse_energy(torch.from_numpy(Y).cuda(), torch.from_numpy(T).cuda(), loss, lam))
    r, m, want = ref(Y)
    # FP32 residuals, FP64 sums
    assert float(loss.item()) == pytest.approx(want, rel=1e-6)
    g = (r + lam * m[:, None] * free[None, :]) / (nb * nw)
    # FP32 sum of two terms of size <~ 4 over nvox: a few ulps of 4 / 15000
    np.testing.assert_allclose(dY[..., 0].reshape(nb, nw), g, rtol=2e-5, atol=1e-10)
    assert not np.any(dY[..., 1:])
    for where, moves in ((free, True), (~free, False)):
        Y2 = Y.copy()
        v = Y2[1, ..., 0].reshape(-1).copy()
        v[where] += 0.01
        Y2[1, ..., 0] = v.reshape(D, H, Wd)
        l2 = S._zeros(1, torch.float64)
        S.mse_energy(torch.from_numpy(Y2).cuda(), torch.from_numpy(T).cuda(), l2, lam)
        r2, m2, want2 = ref(Y2)
        assert float(l2.item()) == pytest.approx(want2, rel=1e-6)
        assert (abs(m2[1] - m[1]) > 1e-3) == moves
