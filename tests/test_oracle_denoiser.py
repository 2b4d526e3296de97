"""Pins of the denoiser's numpy restatement (oracle/denoiser_ref.py).

Each function is checked against something independent of its own formula:
a direct loop or scipy for the convolution, adjointness for pooling and
upsampling, central finite differences in float64 for every backward,
closed forms for the kernel prediction and Adam's first step, and known bit
patterns for the BF16 rounding. A planted error (a transposed kernel, a
dropped softmax term) must fail.
"""
import numpy as np
import pytest
from scipy import signal

from oracle import denoiser_ref as D


def _rng(s=0):
    return np.random.default_rng(s)


def test_bf16_rounding_bits():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1.0 + 2 ** -7], dtype=np.float32)
    r = D.bf16_round(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r.tolist() == [1.0, 1.0, 1.0 + 2 ** -6, -2.5, np.float32(D.bf16_round(np.float32(3.0e38))), 1.0 + 2 ** -7]
    assert np.isnan(D.bf16_round(np.array([np.nan], np.float32)))[0]
    # every result is representable: low 16 bits zero
    y = D.bf16_round(_rng().normal(size=1000).astype(np.float32))
    assert not np.any(y.view(np.uint32) & 0xFFFF)


def test_package_bf16_helpers_agree():
    from vulkanhybridraytracer_amd import denoise
    x = _rng(1).normal(size=4096).astype(np.float32) * 100
    assert np.array_equal(denoise.from_bf16_bits(denoise.to_bf16_bits(x)), D.bf16_round(x))


def _conv_loop(x, w, b):
    N, H, W, C = x.shape
    co = w.shape[0]
    y = np.zeros((N, H, W, co))
    for n in range(N):
        for i in range(H):
            for j in range(W):
                for dy in range(3):
                    for dx in range(3):
                        yy, xx = i + dy - 1, j + dx - 1
                        if 0 <= yy < H and 0 <= xx < W:
                            y[n, i, j] += w[:, dy, dx, :] @ x[n, yy, xx]
    return y + b


def test_conv3x3_direct_loop_and_scipy():
    r = _rng(2)
    x = r.normal(size=(2, 5, 7, 3))
    w = r.normal(size=(4, 3, 3, 3))
    b = r.normal(size=4)
    y = D.conv3x3(x, w, b)
    assert np.allclose(y, _conv_loop(x, w, b), atol=1e-12)
    # scipy: cross-correlation with zero padding ('same')
    for co in range(4):
        ref = sum(signal.correlate(x[0, :, :, c], w[co, :, :, c], mode="same") for c in range(3)) + b[co]
        assert np.allclose(y[0, :, :, co], ref, atol=1e-12)
    # a transposed kernel is a different convolution
    assert not np.allclose(D.conv3x3(x, w.transpose(0, 2, 1, 3), b), y)
    assert np.array_equal(D.conv3x3(x, w, b, relu=True), np.maximum(y, 0))


def _fd(f, a, idx, h=1e-6):
    a0 = a[idx]
    a[idx] = a0 + h
    fp = f()
    a[idx] = a0 - h
    fm = f()
    a[idx] = a0
    return (fp - fm) / (2 * h)


def test_conv3x3_backward_finite_differences():
    r = _rng(3)
    x = r.normal(size=(1, 4, 5, 3))
    w = r.normal(size=(2, 3, 3, 3))
    b = r.normal(size=2)
    g = r.normal(size=(1, 4, 5, 2))
    dx, dw, db = D.conv3x3_backward(x, w, g)
    f = lambda: float((D.conv3x3(x, w, b) * g).sum())
    for idx in [(0, 0, 0, 0), (0, 3, 4, 2), (0, 1, 2, 1), (0, 2, 0, 0)]:
        assert abs(_fd(f, x, idx) - dx[idx]) < 1e-6
    for idx in [(0, 0, 0, 0), (1, 2, 1, 2), (0, 1, 1, 1), (1, 0, 2, 0)]:
        assert abs(_fd(f, w, idx) - dw[idx]) < 1e-6
    for i in range(2):
        assert abs(_fd(f, b, (i,)) - db[i]) < 1e-6


def test_pool_up_adjoint_and_closed_form():
    r = _rng(4)
    x = r.normal(size=(2, 8, 6, 3))
    g = r.normal(size=(2, 4, 3, 3))
    assert np.isclose((D.pool2(x) * g).sum(), (x * D.pool2_backward(g)).sum())
    y = r.normal(size=(2, 4, 3, 3))
    G = r.normal(size=(2, 8, 6, 3))
    assert np.isclose((D.up2(y) * G).sum(), (y * D.up2_backward(G)).sum())
    assert np.isclose(D.pool2(x)[1, 2, 1, 0], x[1, 4:6, 2:4, 0].mean())
    assert np.array_equal(D.up2(y)[0, 5, 3], y[0, 2, 1])


def test_kpcn_closed_forms():
    r = _rng(5)
    img = r.uniform(0, 2, size=(6, 7, 4))
    # uniform logits: 5x5 box filter with clamped coordinates
    out = D.kpcn_apply(np.zeros((6, 7, 25)), img)
    box = np.zeros((6, 7, 3))
    for i in range(6):
        for j in range(7):
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    box[i, j] += img[min(max(i + dy, 0), 5), min(max(j + dx, 0), 6), :3] / 25
    assert np.allclose(out, box, atol=1e-12)
    # a dominant logit at k = 5 * (dy + 2) + dx + 2: the shifted image
    lg = np.full((6, 7, 25), -200.0)
    lg[..., 5 * 3 + 4] = 0.0  # (dx, dy) = (2, 1)
    out = D.kpcn_apply(lg, img)
    sh = img[np.clip(np.arange(6) + 1, 0, 5)][:, np.clip(np.arange(7) + 2, 0, 6), :3]
    assert np.allclose(out, sh, atol=1e-12)


def test_kpcn_loss_backward_finite_differences():
    r = _rng(6)
    lg = r.normal(size=(4, 5, 25))
    noisy = r.uniform(0, 3, size=(4, 5, 4))
    tgt = r.uniform(0, 3, size=(4, 5, 4))
    loss, dl = D.kpcn_loss_backward(lg, noisy, tgt, 1.0)
    f = lambda: D.kpcn_loss_backward(lg, noisy, tgt, 1.0)[0]
    for idx in [(0, 0, 0), (3, 4, 24), (2, 2, 12), (1, 3, 7)]:
        assert abs(_fd(f, lg, idx) - dl[idx]) < 1e-6
    # the loss is the tone-mapped L1 distance
    o = D.kpcn_apply(lg, noisy)
    assert np.isclose(loss, np.abs(o / (1 + o) - tgt[..., :3] / (1 + tgt[..., :3])).sum())
    # relative MSE
    loss, dl = D.kpcn_loss_backward(lg, noisy, tgt, 1.0, loss_kind=1)
    assert np.isclose(loss, ((o - tgt[..., :3]) ** 2 / (tgt[..., :3] ** 2 + 0.01)).sum())
    f = lambda: D.kpcn_loss_backward(lg, noisy, tgt, 1.0, loss_kind=1)[0]
    for idx in [(0, 0, 0), (3, 4, 24), (2, 2, 12)]:
        num = _fd(f, lg, idx)
        assert abs(num - dl[idx]) < 1e-6 * max(1.0, abs(num)), idx


def _small_problem(seed, n=1, h=16, w=16):
    r = _rng(seed)
    params = D.he_init(seed)
    # non-zero biases so every bias gradient is exercised
    params = [(wt, r.normal(0, 0.05, size=b.shape)) for wt, b in params]
    x0 = r.uniform(0, 1, size=(n, h, w, 9))
    noisy = r.uniform(0, 2, size=(n, h, w, 4))
    target = r.uniform(0, 2, size=(n, h, w, 4))
    return params, x0, noisy, target


def test_network_gradients_finite_differences():
    params, x0, noisy, target = _small_problem(7)
    loss, grads = D.loss_and_grads(x0, noisy, target, params)
    assert loss > 0

    def f():
        return D.loss_and_grads(x0, noisy, target, params)[0]

    r = _rng(8)
    checked = 0
    for layer in (0, 1, 5, 9, 10, 14, 16, 18):
        w, b = params[layer]
        idx = tuple(int(r.integers(s)) for s in w.shape)
        num = _fd(f, w, idx)
        ana = grads[layer][0][idx]
        assert abs(num - ana) <= 1e-6 + 1e-4 * abs(ana), (layer, idx, num, ana)
        j = (int(r.integers(b.shape[0])),)
        num = _fd(f, b, j)
        assert abs(num - grads[layer][1][j]) <= 1e-6 + 1e-4 * abs(num), (layer, "bias", num)
        checked += 1
    assert checked == 8


def test_planted_error_in_backward_fails():
    """Dropping the softmax's mean term from the logits gradient breaks the
    finite-difference agreement."""
    r = _rng(9)
    lg = r.normal(size=(3, 3, 25))
    noisy = r.uniform(0, 3, size=(3, 3, 4))
    tgt = r.uniform(0, 3, size=(3, 3, 4))
    _, dl = D.kpcn_loss_backward(lg, noisy, tgt, 1.0)
    w = D.softmax25(lg)
    c = D._taps(noisy)
    o = (w[..., None] * c).sum(axis=2)
    d = D.tonemap(o) - D.tonemap(tgt[..., :3])
    go = np.sign(d) / (1.0 + o) ** 2
    planted = w * (go[:, :, None, :] * c).sum(-1)  # softmax Jacobian without its mean term
    f = lambda: D.kpcn_loss_backward(lg, noisy, tgt, 1.0)[0]
    num = _fd(f, lg, (1, 1, 3))
    assert abs(num - dl[1, 1, 3]) < 1e-6
    assert abs(num - planted[1, 1, 3]) > 1e-6


def test_adam_first_step_closed_form():
    r = _rng(10)
    w, g = r.normal(size=50), r.normal(size=50)
    w1, m, v = D.adam_step(w, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    assert np.allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_features_closed_form():
    H, W = 2, 3
    pos = np.zeros((H, W, 4))
    nrm = np.zeros((H, W, 4))
    alb = np.zeros((H, W, 4))
    mat = np.zeros((H, W, 4))
    noisy = np.zeros((H, W, 4))
    pos[0, 1] = (3, 4, 12, 0)   # material 0 is a surface: distance 13
    nrm[0, 1] = (0, 1, 0, 1)
    alb[0, 1, :3] = (0.1, 0.2, 0.3)
    mat[0, 1, 1] = 0.7
    noisy[0, 1, :3] = (1.0, 1.0, 1.0)  # luminance 1
    f = D.features(pos, nrm, alb, mat, noisy, (0, 0, 0))
    assert np.allclose(f[0, 1], [0.1, 0.2, 0.3, 0, 1, 0, 0.13, 0.7, np.log(2.0)])
    assert np.allclose(f[1, 2], 0.0)  # background


def test_param_count():
    assert D.param_count() == 589081
    pairs = D.he_init(0)
    flat = D.join_params(pairs)
    assert flat.size == D.param_count()
    back = D.split_params(flat)
    assert all(np.array_equal(a[0], b[0]) for a, b in zip(pairs, back))
