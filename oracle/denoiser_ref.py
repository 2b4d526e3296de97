"""Numpy restatement of the learned denoiser (test infrastructure only: the
checker of the HIP kernels, never part of the product path).

The denoiser has no counterpart in the reference (its only "denoise" is the
temporal shadow filter blur.comp:35-81); this module states the network of
the HIP implementation in float64 numpy:

- 3x3 convolution with zero padding, W[cout][3][3][cin], tap (dy+1)*3+(dx+1);
- 2x2 average pooling, nearest 2x upsampling, channel concatenation;
- the 5-level kernel-predicting U-Net (layer table LAYERS);
- softmax kernel prediction over the clamped 5x5 neighbourhood;
- loss = mean |T(out) - T(target)|, T(v) = max(v,0) / (1 + max(v,0));
- Adam (Kingma & Ba 2015).

With ``bf16=True`` values are rounded to bfloat16 (round to nearest even) at
the points where the HIP path stores BF16 (features, activations, weights,
activation gradients), so GPU results can be compared with a tolerance set by
FP32 accumulation order only.

Pins (tests/test_oracle_denoiser.py): convolution against a direct loop and
against scipy.signal.correlate; pooling / upsampling adjointness; every
backward against central finite differences in float64; the kernel
prediction against closed forms (uniform logits = clamped box filter, a
one-hot logit = a shift); Adam's first step against its closed form.
"""
import numpy as np

# (cin, cout, level) in exchange order
LAYERS = [(9, 32, 0), (32, 32, 0), (32, 48, 1), (48, 48, 1), (48, 64, 2), (64, 64, 2), (64, 64, 3), (64, 64, 3),
          (64, 64, 4), (64, 64, 4), (128, 64, 3), (64, 64, 3), (128, 64, 2), (64, 64, 2), (112, 48, 1), (48, 48, 1),
          (80, 32, 0), (32, 32, 0), (32, 25, 0)]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "e4a", "e4b", "d3a", "d3b", "d2a", "d2b", "d1a",
         "d1b", "d0a", "d0b", "head"]
N_FEATURES = 9


def param_count():
    return sum(9 * ci * co + co for ci, co, _ in LAYERS)


def bf16_round(x):
    """float -> bfloat16 -> float, round to nearest even (NaN kept)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(a), a, out)
    return out.astype(np.float64) if np.asarray(x).dtype == np.float64 else out


def _q(x, bf16):
    return bf16_round(x).astype(np.float64) if bf16 else x


def split_params(flat):
    """exchange layout -> list of (W[cout,3,3,cin], b[cout])"""
    out, o = [], 0
    for ci, co, _ in LAYERS:
        n = 9 * ci * co
        w = np.asarray(flat[o:o + n], dtype=np.float64).reshape(co, 3, 3, ci)
        o += n
        b = np.asarray(flat[o:o + co], dtype=np.float64)
        o += co
        out.append((w, b))
    assert o == len(flat)
    return out


def join_params(pairs):
    return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in pairs])


# ---- layers ------------------------------------------------------------------------
def _im2col(x):
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    cols = np.empty((N, H, W, 9, C), dtype=x.dtype)
    for dy in range(3):
        for dx in range(3):
            cols[:, :, :, dy * 3 + dx] = xp[:, dy:dy + H, dx:dx + W]
    return cols.reshape(N * H * W, 9 * C)


def conv3x3(x, w, b=None, relu=False):
    """y[n,y,x,co] = sum_{dy,dx,ci} x[n, y+dy-1, x+dx-1, ci] w[co,dy,dx,ci] + b[co]"""
    N, H, W, C = x.shape
    co = w.shape[0]
    y = (_im2col(x) @ w.reshape(co, 9 * C).T).reshape(N, H, W, co)
    if b is not None:
        y = y + b
    return np.maximum(y, 0.0) if relu else y


def conv3x3_backward(x, w, dy):
    """gradients of sum(dy * conv3x3(x, w, b)) w.r.t. x, w, b"""
    N, H, W, C = x.shape
    co = w.shape[0]
    g = dy.reshape(N * H * W, co)
    dw = (g.T @ _im2col(x)).reshape(co, 3, 3, C)
    db = g.sum(axis=0)
    dcols = (g @ w.reshape(co, 9 * C)).reshape(N, H, W, 9, C)
    dxp = np.zeros((N, H + 2, W + 2, C), dtype=dy.dtype)
    for ddy in range(3):
        for ddx in range(3):
            dxp[:, ddy:ddy + H, ddx:ddx + W] += dcols[:, :, :, ddy * 3 + ddx]
    return dxp[:, 1:-1, 1:-1], dw, db


def pool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_backward(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_backward(g):
    N, H, W, C = g.shape
    return g.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


# ---- input features ---------------------------------------------------------------
def features(position, normal, albedo, material, noisy, eye):
    """9 features per pixel from (H, W, 4) planes: albedo, normal, distance to
    the eye / 100 (0 on background, normal.w == 0), roughness (material.y),
    log(1 + luminance of the noisy colour)."""
    H, W = position.shape[:2]
    f = np.zeros((H, W, N_FEATURES), dtype=np.float64)
    f[..., 0:3] = albedo[..., :3]
    f[..., 3:6] = normal[..., :3]
    d = np.sqrt(((position[..., :3] - np.asarray(eye, dtype=np.float64)) ** 2).sum(-1))
    f[..., 6] = np.where(normal[..., 3] == 0, 0.0, d * 0.01)
    f[..., 7] = material[..., 1]
    lum = 0.2126 * noisy[..., 0] + 0.7152 * noisy[..., 1] + 0.0722 * noisy[..., 2]
    f[..., 8] = np.log1p(np.maximum(lum, 0.0))
    return f


# ---- kernel prediction and loss -------------------------------------------------------
def _taps(img):
    """(H, W, 25, 3): img at clamp(p + (k % 5 - 2, k // 5 - 2))"""
    H, W = img.shape[:2]
    ys, xs = np.arange(H), np.arange(W)
    out = np.empty((H, W, 25, 3), dtype=np.float64)
    for k in range(25):
        yy = np.clip(ys + k // 5 - 2, 0, H - 1)
        xx = np.clip(xs + k % 5 - 2, 0, W - 1)
        out[:, :, k] = img[yy][:, xx, :3]
    return out


def softmax25(logits):
    m = logits.max(axis=-1, keepdims=True)
    e = np.exp(logits - m)
    return e / e.sum(axis=-1, keepdims=True)


def kpcn_apply(logits, noisy):
    """logits (H, W, 25), noisy (H, W, >=3) -> (H, W, 3)"""
    return (softmax25(logits)[..., None] * _taps(noisy)).sum(axis=2)


def tonemap(v):
    v = np.maximum(v, 0.0)
    return v / (1.0 + v)


def kpcn_loss_backward(logits, noisy, target, scale, loss_kind=0):
    """loss_kind 0: sum |T(out) - T(target)| over pixels and channels;
    loss_kind 1: sum (out - target)^2 / (target^2 + 0.01) (relative MSE);
    and its gradient w.r.t. the logits times scale"""
    w = softmax25(logits)
    c = _taps(noisy)
    o = (w[..., None] * c).sum(axis=2)
    if loss_kind == 1:
        t = target[..., :3]
        den = t * t + 0.01
        loss = ((o - t) ** 2 / den).sum()
        go = 2.0 * (o - t) / den * scale
    else:
        d = tonemap(o) - tonemap(target[..., :3])
        loss = np.abs(d).sum()
        go = np.where(o > 0, np.sign(d) * scale / (1.0 + np.maximum(o, 0.0)) ** 2, 0.0)
    dw = (go[:, :, None, :] * c).sum(-1)
    dl = w * (dw - (w * dw).sum(-1, keepdims=True))
    return loss, dl


# ---- the network ----------------------------------------------------------------------
def forward(x0, params, bf16=False):
    """x0 (N, H, W, 9) features (H, W multiples of 16) -> logits (N, H, W, 25)
    and the activations the backward pass needs"""
    P = [(_q(w, bf16), b) for w, b in params]
    acts = {}

    def conv(i, x, relu=True):
        w, b = P[i]
        return _q(conv3x3(x, w, b, relu), bf16) if relu else conv3x3(x, w, b, False)

    x = _q(x0, bf16)
    acts["x0"] = x
    skips = []
    for lvl in range(5):
        a = conv(2 * lvl, x)
        acts["e%da" % lvl] = a
        e = conv(2 * lvl + 1, a)
        acts["e%db" % lvl] = e
        skips.append(e)
        if lvl < 4:
            x = _q(pool2(e), bf16)
            acts["p%d" % (lvl + 1)] = x
    y = skips[4]
    for lvl in (3, 2, 1, 0):
        cat = np.concatenate([skips[lvl], up2(y)], axis=-1)
        acts["cat%d" % lvl] = cat
        i = 10 + 2 * (3 - lvl)
        a = conv(i, cat)
        acts["d%da" % lvl] = a
        y = conv(i + 1, a)
        acts["d%db" % lvl] = y
    logits = conv(18, y, relu=False)
    return logits, acts


def backward(dlogits, params, acts, bf16=False):
    """gradient of the loss w.r.t. every parameter (list of (dW, db)) for the
    logits gradient dlogits"""
    P = [(_q(w, bf16), b) for w, b in params]
    grads = [None] * len(P)

    def conv_bwd(i, x, y, g, relu=True, need_dx=True):
        if relu:
            g = np.where(y > 0, g, 0.0)
        dx, dw, db = conv3x3_backward(x, P[i][0], g)
        grads[i] = (dw, db)
        return _q(dx, bf16) if need_dx else None

    g = _q(dlogits, bf16)
    g = conv_bwd(18, acts["d0b"], None, g, relu=False)
    skip_grads = {}
    for lvl in (0, 1, 2, 3):
        i = 10 + 2 * (3 - lvl)
        g = conv_bwd(i + 1, acts["d%da" % lvl], acts["d%db" % lvl], g)
        gcat = conv_bwd(i, acts["cat%d" % lvl], acts["d%da" % lvl], g)
        cs = LAYERS[2 * lvl + 1][1]
        skip_grads[lvl] = gcat[..., :cs]
        g = _q(up2_backward(gcat[..., cs:]), bf16)
    # g: gradient of e4b
    for lvl in (4, 3, 2, 1, 0):
        if lvl < 4:
            g = _q(skip_grads[lvl] + _q(pool2_backward(gp), bf16), bf16) if bf16 else skip_grads[lvl] + pool2_backward(gp)
        ga = conv_bwd(2 * lvl + 1, acts["e%da" % lvl], acts["e%db" % lvl], g)
        xin = acts["x0"] if lvl == 0 else acts["p%d" % lvl]
        gp = conv_bwd(2 * lvl, xin, acts["e%da" % lvl], ga, need_dx=lvl > 0)
    return grads


def loss_and_grads(x0, noisy, target, params, bf16=False, loss_kind=0):
    """mean loss (kpcn_loss_backward's loss_kind) of a batch and the
    parameter gradients. x0 (N, H, W, 9); noisy, target (N, H, W, >=3)"""
    N, H, W, _ = x0.shape
    logits, acts = forward(x0, params, bf16)
    scale = 1.0 / (3.0 * N * H * W)
    total, dl = 0.0, np.empty_like(logits)
    for n in range(N):
        l, dl[n] = kpcn_loss_backward(logits[n], noisy[n], target[n], scale, loss_kind)
        total += l
    return total * scale, backward(dl, params, acts, bf16)


def denoise(x0, noisy, params, bf16=False):
    logits, _ = forward(x0, params, bf16)
    return np.stack([kpcn_apply(logits[n], noisy[n]) for n in range(x0.shape[0])])


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def he_init(seed):
    """He-normal weights (head scaled by 0.1), zero biases (a numpy stream, not
    the library's own initialiser)"""
    rng = np.random.default_rng(seed)
    pairs = []
    for i, (ci, co, _) in enumerate(LAYERS):
        sd = np.sqrt(2.0 / (9 * ci)) * (0.1 if i == 18 else 1.0)
        pairs.append((rng.normal(0, sd, (co, 3, 3, ci)), np.zeros(co)))
    return pairs


def flops(H, W):
    """multiply-adds x 2 of the convolutions of one forward pass at H x W
    (padded to multiples of 16), plus the 5x5 kernel application"""
    Hp, Wp = -(-H // 16) * 16, -(-W // 16) * 16
    f = sum(2 * 9 * ci * co * (Hp >> lvl) * (Wp >> lvl) for ci, co, lvl in LAYERS)
    return f + 2 * 25 * 3 * H * W


def rmse(x, ref, eps=0.01):
    """relative MSE: mean over pixels and colour channels of (x - ref)^2 / (ref^2 + eps)"""
    x, ref = np.asarray(x, np.float64)[..., :3], np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))
