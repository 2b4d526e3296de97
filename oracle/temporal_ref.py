"""Float32 numpy restatement of the temporal shadow filter and its motion
vectors (test infrastructure only: the checker of temporal.hip).

Follows blur.comp:35-81 (dispatch renderer.cpp:2661-2675, frame counter
vk_engine.cpp:669-686, 722) generalised to any size and light count, and the
motion vectors of geometry_shader.frag:72-75 / basic.vert:38-46, with the
decisions the HIP path takes:
  - every pixel is filtered (the reference dispatches W/16 x H/16 groups);
  - imageLoad outside the frame reads 0;
  - the motion vector is read at the pixel's own texel;
  - the y offset of the reprojection uses the x component, as blur.comp:52
    evaluates vec2(x + r.x * W, y + r * H);
  - reprojected reads see the previous output;
  - mix(x, y, a) = x * (1 - a) + y * a.
Every operation is a float32 operation in the order the kernel evaluates it,
so the GPU result is bit-exact against this restatement.

Pins (tests/test_oracle_temporal.py): the running mean equals the arithmetic
mean of the frames; zero motion equals a clamp against scipy's 3x3 min/max
filters; a one-pixel motion is a shift; background motion reads 0.
"""
import numpy as np

F = np.float32


def _mix(x, y, a):
    a = F(a)
    return (x * (F(1.0) - a)).astype(F) + (y * a).astype(F)


def filter_shadows(shadow, history, frame, motion=None):
    """shadow, history: float32 (nl, H, W); motion: float32 (H, W, 2).
    Returns the new history."""
    shadow = np.asarray(shadow, dtype=F)
    prev = np.asarray(history, dtype=F)
    nl, H, W = shadow.shape
    if frame > 0:
        a = F(1.0) / (F(frame) + F(1.0))
        return _mix(prev, shadow, a)
    m = np.asarray(motion, dtype=F)
    rx = (m[..., 0] * F(2.0)) - F(1.0)
    xs = np.arange(W, dtype=F)[None, :]
    ys = np.arange(H, dtype=F)[:, None]
    lx = xs + (rx * F(W))
    ly = ys + (rx * F(H))
    px = np.trunc(lx).astype(np.int64)
    py = np.trunc(ly).astype(np.int64)
    inside = (px >= 0) & (px < W) & (py >= 0) & (py < H)
    pxc, pyc = np.clip(px, 0, W - 1), np.clip(py, 0, H - 1)
    out = np.empty_like(shadow)
    for l in range(nl):
        s = shadow[l]
        pad = np.zeros((H + 2, W + 2), dtype=F)
        pad[1:-1, 1:-1] = s
        mn = s.copy()
        mx = s.copy()
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == 0 and dy == 0:
                    continue
                c = pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
                mn = np.minimum(mn, c)
                mx = np.maximum(mx, c)
        old = np.where(inside, prev[l][pyc, pxc], F(0.0))
        old = np.maximum(mn, old)
        old = np.minimum(mx, old)
        out[l] = _mix(old, s, F(0.4))
    return out


def motion_vectors(position, normal, view_proj, prev_view_proj):
    """(H, W, 2) float32: (ndc - ndc_prev).xy * 0.5 + 0.5 of each pixel's
    world position; 0 for background (position.w == 0 and normal.w == 0)."""
    P = np.asarray(position, dtype=F)
    N = np.asarray(normal, dtype=F)
    vp = np.asarray(view_proj, dtype=F).reshape(16)
    pv = np.asarray(prev_view_proj, dtype=F).reshape(16)
    v = [P[..., 0], P[..., 1], P[..., 2], np.ones(P.shape[:2], dtype=F)]

    def clip(m):
        return [(((m[r] * v[0]) + (m[4 + r] * v[1])) + (m[8 + r] * v[2])) + (m[12 + r] * v[3]) for r in range(4)]

    c, q = clip(vp), clip(pv)
    with np.errstate(divide="ignore", invalid="ignore"):
        nx, ny = c[0] / c[3], c[1] / c[3]
        px, py = q[0] / q[3], q[1] / q[3]
        out = np.stack([((nx - px) * F(0.5)) + F(0.5), ((ny - py) * F(0.5)) + F(0.5)], -1).astype(F)
    bg = (P[..., 3] == 0) & (N[..., 3] == 0)
    out[bg] = 0.0
    return out
