"""Pins of the temporal shadow filter's restatement (oracle/temporal_ref.py)
against closed forms and independent implementations."""
import numpy as np
from scipy import ndimage

from oracle import temporal_ref as T


def _rng(s):
    return np.random.default_rng(s)


def test_running_mean_is_the_arithmetic_mean():
    """frame 1, 2, 3, ...: mix(h, s, 1 / (frame + 1)) accumulates the mean of
    the first output and every later frame (blur.comp:40-47)."""
    r = _rng(0)
    h0 = r.uniform(0, 1, size=(2, 5, 7)).astype(np.float32)
    frames = [r.uniform(0, 1, size=(2, 5, 7)).astype(np.float32) for _ in range(6)]
    h = h0
    for k, s in enumerate(frames, start=1):
        h = T.filter_shadows(s, h, k)
        mean = np.mean(np.stack([h0] + frames[:k]).astype(np.float64), axis=0)
        assert np.allclose(h, mean, atol=2e-6)


def test_zero_motion_is_a_neighbourhood_clamp():
    """motion 0.5 means no motion: the previous output at the pixel, clamped
    to the 3x3 min / max of the new factors (0 outside the frame), blended
    0.6 / 0.4. Checked against scipy's min / max filters."""
    r = _rng(1)
    s = r.uniform(0, 1, size=(3, 9, 11)).astype(np.float32)
    h = r.uniform(0, 1, size=(3, 9, 11)).astype(np.float32)
    m = np.full((9, 11, 2), 0.5, dtype=np.float32)
    for frame in (0, -1, -5):
        out = T.filter_shadows(s, h, frame, m)
        for l in range(3):
            mn = ndimage.minimum_filter(s[l], size=3, mode="constant", cval=0.0)
            mx = ndimage.maximum_filter(s[l], size=3, mode="constant", cval=0.0)
            ref = np.clip(h[l].astype(np.float64), mn, mx) * 0.6 + s[l] * 0.4
            assert np.allclose(out[l], ref, atol=1e-6)


def test_one_pixel_motion_is_a_shift_and_y_uses_x():
    """W = 64, H = 32, motion.x = 0.5 + 1/128: r.x = 1/64, so the history is
    read one pixel to the right and half a pixel down (truncated: the same
    row) - the y offset comes from r.x; motion.y is ignored."""
    W, H = 64, 32
    r = _rng(2)
    s = np.full((1, H, W), 0.5, dtype=np.float32)
    s[0, :, ::2] = 0.25
    h = r.uniform(0, 1, size=(1, H, W)).astype(np.float32)
    m = np.zeros((H, W, 2), dtype=np.float32)
    m[..., 0] = 0.5 + 1.0 / 128
    m[..., 1] = r.uniform(0, 1, size=(H, W))  # not read
    out = T.filter_shadows(s, h, 0, m)
    shifted = np.zeros_like(h[0])
    shifted[:, :-1] = h[0][:, 1:]
    mn = ndimage.minimum_filter(s[0], size=3, mode="constant", cval=0.0)
    mx = ndimage.maximum_filter(s[0], size=3, mode="constant", cval=0.0)
    ref = np.clip(shifted.astype(np.float64), mn, mx) * 0.6 + s[0] * 0.4
    assert np.allclose(out[0], ref, atol=1e-6)


def test_background_motion_reads_outside():
    """motion 0 (background) gives r = -1: the reprojected pixel is outside
    the frame and reads 0, which the clamp raises to the neighbourhood min."""
    r = _rng(3)
    s = r.uniform(0.2, 1, size=(1, 6, 6)).astype(np.float32)
    h = r.uniform(0, 1, size=(1, 6, 6)).astype(np.float32)
    out = T.filter_shadows(s, h, 0, np.zeros((6, 6, 2), np.float32))
    mn = ndimage.minimum_filter(s[0], size=3, mode="constant", cval=0.0)
    assert np.allclose(out[0], mn * 0.6 + s[0] * 0.4, atol=1e-6)


def test_motion_vectors_closed_form():
    """A static camera gives 0.5 everywhere; a camera translated along x by
    d in NDC units gives 0.5 + d / 2 in x; background pixels give 0."""
    H, W = 4, 5
    r = _rng(4)
    P = np.concatenate([r.uniform(-10, 10, size=(H, W, 3)), np.ones((H, W, 1))], -1).astype(np.float32)
    N = np.zeros((H, W, 4), np.float32)
    N[..., 3] = 1
    P[0, 0, 3] = 0
    N[0, 0, 3] = 0
    vp = np.eye(4, dtype=np.float32)
    vp[3, 2] = 0.0
    vp[2, 3] = 1.0  # w = z: a projective matrix (column-major element (row 3, col 2) -> index 11)
    flat = vp.T.reshape(16)  # column-major storage of the row/column matrix above
    m = T.motion_vectors(P, N, flat, flat)
    valid = ~((P[..., 3] == 0) & (N[..., 3] == 0))
    assert np.allclose(m[valid], 0.5) and np.all(m[0, 0] == 0)
    shift = np.eye(4, dtype=np.float32)
    shift[0, 3] = 0.25  # previous clip x = x + 0.25 w
    prev = (shift @ vp).T.reshape(16)
    m2 = T.motion_vectors(P, N, flat, prev)
    assert np.allclose(m2[valid][:, 0], 0.5 - 0.125, atol=1e-6)
    assert np.allclose(m2[valid][:, 1], 0.5, atol=1e-6)
