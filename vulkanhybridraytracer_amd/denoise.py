"""Learned denoiser of the low-sample hybrid pass (Python host side of the
hr_denoiser_* / hr_dn_* entry points of include/hr/hr.h).

The denoiser has no counterpart in the reference, whose only "denoise" is
the temporal shadow filter (blur.comp:35-81). It is a kernel-predicting
U-Net run entirely by the library's HIP kernels (BF16 MFMA convolutions):
  Denoiser        weights, inference (hr_denoise), training steps and
                  gradients (hr_denoiser_train_step / _gradients), resumable
                  checkpoints (hr_denoiser_get_state / _set_state)
  PairRenderer    renders training / evaluation pairs with the tracer: the
                  same view at low (4 spp) and high (64 spp) sample counts,
                  buffers kept in device memory
  conv3x3_forward / conv3x3_backward   one layer, for kernel-level tests
"""
import ctypes as C
import json
import os

import numpy as np

from . import abi, hr, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "weights", "kpcn_bf16.npy")

LAYERS = [(9, 32, 0), (32, 32, 0), (32, 48, 1), (48, 48, 1), (48, 64, 2), (64, 64, 2), (64, 64, 3), (64, 64, 3),
          (64, 64, 4), (64, 64, 4), (128, 64, 3), (64, 64, 3), (128, 64, 2), (64, 64, 2), (112, 48, 1), (48, 48, 1),
          (80, 32, 0), (32, 32, 0), (32, 25, 0)]


# ---- BF16 helpers (bit patterns as uint16) ----------------------------------------------
def to_bf16_bits(x):
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return np.where(np.isnan(a), np.uint16(0x7FC0), r)


def from_bf16_bits(b):
    return (np.ascontiguousarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def flops(height, width):
    """FLOPs (2 x multiply-adds) of one forward pass: the convolutions at the
    padded size (multiples of 16) plus the 5x5 kernel application."""
    hp, wp = -(-height // 16) * 16, -(-width // 16) * 16
    conv = sum(2 * 9 * ci * co * (hp >> lvl) * (wp >> lvl) for ci, co, lvl in LAYERS)
    return conv + 2 * 25 * 3 * height * width


class Denoiser:
    """The network's FP32 master weights and Adam state live in device memory."""

    def __init__(self, ctx, weights=None, seed=0):
        self.ctx = ctx
        L = ctx.L
        self.count = int(L.hr_denoiser_param_count())
        h = C.c_void_p()
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float32)
            if w.size != self.count:
                raise ValueError("expected %d weights, got %d" % (self.count, w.size))
        ctx.check(L.hr_denoiser_create(ctx.h, None if w is None else w.ctypes.data, self.count if w is not None
                                       else 0, seed, C.byref(h)))
        self.h = h
        ctx.own(self)

    @staticmethod
    def param_count():
        return sum(9 * ci * co + co for ci, co, _ in LAYERS)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.hr_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        w = np.empty(self.count, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_denoiser_get_weights(self.h, w.ctypes.data, self.count))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_denoiser_set_weights(self.h, w.ctypes.data, w.size))

    def get_state(self):
        """(weights, adam_m, adam_v, step): the FP32 training state"""
        w, m, v = (np.empty(self.count, dtype=np.float32) for _ in range(3))
        step = C.c_uint64(0)
        self.ctx.check(self.ctx.L.hr_denoiser_get_state(self.h, w.ctypes.data, m.ctypes.data, v.ctypes.data,
                                                        self.count, C.byref(step)))
        return w, m, v, step.value

    def set_state(self, weights, adam_m, adam_v, step):
        w, m, v = (np.ascontiguousarray(a, dtype=np.float32) for a in (weights, adam_m, adam_v))
        for a in (w, m, v):
            if a.size != self.count:
                raise ValueError("expected %d values, got %d" % (self.count, a.size))
        self.ctx.check(self.ctx.L.hr_denoiser_set_state(self.h, w.ctypes.data, m.ctypes.data, v.ctypes.data,
                                                        self.count, int(step)))

    def save_checkpoint(self, path, meta=None):
        """Resumable training state as .npz (arrays only, no pickles): FP32
        master weights, Adam moments, step count, loss kind, and `meta` (a
        JSON-serialisable dict, e.g. the batch sampler's state)."""
        w, m, v, step = self.get_state()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp.npz"
        np.savez(tmp, weights=w, adam_m=m, adam_v=v, step=np.uint64(step), loss=np.uint32(self.loss),
                 meta=np.array(json.dumps(meta or {})))
        os.replace(tmp, path)

    @classmethod
    def load_checkpoint(cls, ctx, path):
        """(Denoiser, meta) restored from save_checkpoint's file"""
        with np.load(path, allow_pickle=False) as z:
            dn = cls(ctx)
            dn.set_loss(int(z["loss"]))
            dn.set_state(z["weights"], z["adam_m"], z["adam_v"], int(z["step"]))
            return dn, json.loads(str(z["meta"]))

    def save(self, path):
        """BF16 bit patterns (the precision the convolutions use)."""
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.save(path, to_bf16_bits(self.get_weights()))

    @classmethod
    def load(cls, ctx, path=WEIGHTS_PATH):
        return cls(ctx, weights=load_weights(path))

    @staticmethod
    def image(gbuffer, noisy, target=None, out=None, eye=(0.0, 0.0, 0.0)):
        """hr_dn_image from a GBuffer and device pointers (ints)."""
        im = abi.hr_dn_image()
        im.gbuffer = gbuffer.c if isinstance(gbuffer, hr.GBuffer) else gbuffer
        im.noisy = noisy
        im.target = target
        im.out = out
        im.eye[:] = [float(v) for v in eye]
        return im

    def denoise(self, image, timing=False):
        ms = C.c_float(0.0)
        self.ctx.check(self.ctx.L.hr_denoise(self.h, C.byref(image), C.byref(ms) if timing else None))
        return ms.value if timing else None

    LOSS_L1_TONEMAPPED, LOSS_REL_MSE = 0, 1

    loss = LOSS_L1_TONEMAPPED

    def set_loss(self, kind):
        self.ctx.check(self.ctx.L.hr_denoiser_set_loss(self.h, kind))
        self.loss = kind

    def train_step(self, images, lr=1e-3, want_loss=True):
        """One Adam step; without want_loss nothing waits for the device."""
        arr = (abi.hr_dn_image * len(images))(*images)
        loss = C.c_float(0.0)
        self.ctx.check(self.ctx.L.hr_denoiser_train_step(self.h, arr, len(images), lr,
                                                         C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def gradients(self, images):
        arr = (abi.hr_dn_image * len(images))(*images)
        loss = C.c_float(0.0)
        g = np.empty(self.count, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_denoiser_gradients(self.h, arr, len(images), C.byref(loss), g.ctypes.data,
                                                        self.count))
        return loss.value, g


def load_weights(path=WEIGHTS_PATH):
    b = np.load(path, allow_pickle=False)
    return from_bf16_bits(b) if b.dtype == np.uint16 else b.astype(np.float32)


# ---- one layer -------------------------------------------------------------------------------
def conv3x3_forward(ctx, x_bits, w, b=None, relu=False):
    """x_bits: uint16 BF16 [N, H, W, Cin]; w float32 [Cout, 3, 3, Cin] -> uint16 [N, H, W, Cout]"""
    n, h, wd, cin = x_bits.shape
    cout = w.shape[0]
    xb, yb = ctx.alloc(x_bits.nbytes), ctx.alloc(n * h * wd * cout * 2)
    xb.upload(x_bits)
    wf = np.ascontiguousarray(w, dtype=np.float32)
    bf = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    ctx.check(ctx.L.hr_dn_conv3x3_forward(ctx.h, xb.ptr, n, h, wd, cin, wf.ctypes.data,
                                          None if bf is None else bf.ctypes.data, cout, 1 if relu else 0, yb.ptr))
    return yb.download((n, h, wd, cout), np.uint16)


def conv3x3_backward(ctx, x_bits, dy_bits, w):
    """-> dx uint16 [N, H, W, Cin], dw float32 [Cout, 3, 3, Cin], db float32 [Cout]"""
    n, h, wd, cin = x_bits.shape
    cout = w.shape[0]
    xb, gb, dxb = ctx.alloc(x_bits.nbytes), ctx.alloc(dy_bits.nbytes), ctx.alloc(x_bits.nbytes)
    xb.upload(x_bits)
    gb.upload(dy_bits)
    wf = np.ascontiguousarray(w, dtype=np.float32)
    dw = np.empty_like(wf)
    db = np.empty(cout, dtype=np.float32)
    ctx.check(ctx.L.hr_dn_conv3x3_backward(ctx.h, xb.ptr, gb.ptr, n, h, wd, cin, wf.ctypes.data, cout, dxb.ptr,
                                           dw.ctypes.data, db.ctypes.data))
    return dxb.download((n, h, wd, cin), np.uint16), dw, db


# ---- training / evaluation pairs rendered by the tracer -------------------------------------
class Sample:
    """One view: its G-buffer, the low-sample and high-sample lit colours
    (device buffers) and the eye position."""

    def __init__(self, gb, noisy, target, eye):
        self.gb, self.noisy, self.target, self.eye = gb, noisy, target, eye

    def image(self, out=None, with_target=True):
        return Denoiser.image(self.gb, self.noisy.ptr, self.target.ptr if with_target else None,
                              None if out is None else out.ptr, self.eye)


def random_view(rng, extent):
    """An eye above the terrain looking at a point on it."""
    r = rng.uniform(0.15, 0.75) * extent
    a = rng.uniform(0, 2 * np.pi)
    eye = (r * np.cos(a), rng.uniform(0.08, 0.45) * extent, r * np.sin(a))
    tgt = (rng.uniform(-0.2, 0.2) * extent, 0.0, rng.uniform(-0.2, 0.2) * extent)
    return eye, tgt


class PairRenderer:
    """Renders (noisy, target) pairs of one scene at any views: the same
    G-buffer traced at low_spp and high_spp shadow samples per light and
    reflection paths per pixel (non-adaptive)."""

    def __init__(self, ctx, scene, width, height, low_spp=4, high_spp=64):
        self.ctx, self.scene, self.width, self.height = ctx, scene, width, height
        self.hs = hr.HybridScene(ctx, scene)
        self.rd = hr.HybridRenderer(ctx, self.hs, width, height)
        self.low = dict(scene.params, shadow_samples=low_spp, refl_spp=low_spp, flags=0)
        self.high = dict(scene.params, shadow_samples=high_spp, refl_spp=high_spp, flags=0)

    def render(self, eye, target, frame=0.0):
        sc = self.scene
        cam = scenes.rt_camera(eye, target, self.width, self.height, frame=frame)
        rd = self.rd
        rd.gb = hr.GBuffer(self.ctx, self.width, self.height)
        npx = self.width * self.height * 16
        noisy, ref = self.ctx.alloc(npx), self.ctx.alloc(npx)
        rd.set_frame(cam, sc.lights, sc.shapes, self.low)
        rd.gbuffer_primary()
        rd.lit = noisy
        rd.trace_shadows()
        rd.shade_hybrid()
        rd.set_frame(cam, sc.lights, sc.shapes, self.high)
        rd.lit = ref
        rd.trace_shadows()
        rd.shade_hybrid()
        return Sample(rd.gb, noisy, ref, tuple(float(v) for v in eye))

    def close(self):
        self.hs.close()


def training_scene(seed, quads=None, instances=None, extent=None):
    """A smaller scene of the terrain generator (393,216 triangles, roughness
    U[0,1]; configs/denoise_1080p_4m.json "training"); its views are 256 x 256."""
    t = scenes.config("denoise_1080p_4m")["training"]
    g = t["scene_generator"]
    w, h = t["crop"]
    return scenes.make_terrain_scene(seed=seed, width=w, height=h,
                                     quads=g["terrain_quads"] if quads is None else quads,
                                     instances=g["instances"] if instances is None else instances,
                                     extent=g["extent"] if extent is None else extent,
                                     rough_max=g["roughness"][1], name="train%04d" % seed)


def make_dataset(ctx, seeds, views, size=256, low_spp=4, high_spp=64, log=None, scene_kw=None):
    """views x len(seeds) samples: random views of seeded training scenes."""
    out = []
    for s in seeds:
        sc = training_scene(s, **(scene_kw or {}))
        pr = PairRenderer(ctx, sc, size, size, low_spp, high_spp)
        rng = np.random.default_rng(50_000 + s)
        extent = (scene_kw or {}).get("extent", 120.0)
        for v in range(views):
            eye, tgt = random_view(rng, extent)
            out.append(pr.render(eye, tgt, frame=float(v)))
        ctx.sync()
        pr.close()
        if log:
            log("scene %d: %d views" % (s, views))
    return out


def rmse(x, ref, eps=0.01):
    """relative MSE over pixels and colour channels: (x - ref)^2 / (ref^2 + eps)"""
    x, ref = np.asarray(x, np.float64)[..., :3], np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


# ---- the method without the network: fixed 5x5 filters (host-side evaluation baselines) ----
def _taps(img):
    H, W = img.shape[:2]
    out = np.empty((H, W, 25) + img.shape[2:], dtype=np.float64)
    for k in range(25):
        yy = np.clip(np.arange(H) + k // 5 - 2, 0, H - 1)
        xx = np.clip(np.arange(W) + k % 5 - 2, 0, W - 1)
        out[:, :, k] = img[yy][:, xx]
    return out


def box_filter(noisy):
    """5x5 box (uniform kernel weights) with clamped coordinates."""
    return _taps(np.asarray(noisy, np.float64)[..., :3]).mean(axis=2)


def bilateral_filter(noisy, planes, eye, sigma_s=1.5, sigma_n=0.25, sigma_a=0.15, sigma_d=0.03):
    """5x5 joint bilateral filter guided by the G-buffer: spatial Gaussian x
    normal, albedo and relative-depth similarity."""
    n = np.asarray(planes["normal"], np.float64)[..., :3]
    a = np.asarray(planes["albedo"], np.float64)[..., :3]
    d = np.sqrt(((np.asarray(planes["position"], np.float64)[..., :3] - np.asarray(eye)) ** 2).sum(-1))
    tn, ta, td, tc = _taps(n), _taps(a), _taps(d), _taps(np.asarray(noisy, np.float64)[..., :3])
    k = np.arange(25)
    ws = np.exp(-((k % 5 - 2) ** 2 + (k // 5 - 2) ** 2) / (2 * sigma_s ** 2))
    w = ws * np.exp(-((tn - n[:, :, None]) ** 2).sum(-1) / sigma_n ** 2) \
        * np.exp(-((ta - a[:, :, None]) ** 2).sum(-1) / sigma_a ** 2) \
        * np.exp(-((td - d[:, :, None]) / np.maximum(d[:, :, None], 1e-3)) ** 2 / sigma_d ** 2)
    return (w[..., None] * tc).sum(2) / w.sum(2)[..., None]


def evaluate(dn, samples, with_filters=True):
    """Mean rMSE against the high-sample colour of the raw low-sample colour,
    the network's output and (optionally) the fixed filters; max abs error of
    the network's output."""
    ctx = dn.ctx
    acc = {"noisy": [], "network": [], "box": [], "bilateral": [], "max_abs": []}
    for s in samples:
        h, w = s.gb.height, s.gb.width
        out = ctx.alloc(h * w * 16)
        dn.denoise(s.image(out=out, with_target=False))
        o = out.download((h, w, 4))
        noisy = s.noisy.download((h, w, 4))
        ref = s.target.download((h, w, 4))
        acc["noisy"].append(rmse(noisy, ref))
        acc["network"].append(rmse(o, ref))
        acc["max_abs"].append(float(np.abs(o[..., :3] - ref[..., :3]).max()))
        if with_filters:
            acc["box"].append(rmse(box_filter(noisy), ref))
            acc["bilateral"].append(rmse(bilateral_filter(noisy, s.gb.download(), s.eye), ref))
    return {k: (float(np.mean(v)) if k != "max_abs" else float(np.max(v))) for k, v in acc.items() if v}
