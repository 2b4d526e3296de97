"""Oracle: material textures, the normal-map frame and the post pass, pinned by
known answers and closed forms that do not restate the oracle's code:

  - NEAREST / REPEAT texel selection: a texture whose texels encode their own
    coordinates, on a square whose uv at each pixel is computed here in
    float64 from the camera model (ray / plane), positive and negative uv;
  - geometry_shader.frag:54: the G-buffer albedo is texel * vertex colour,
    untextured materials ignore the vertex colour;
  - a flat normal-map texel (128, 128, 255) leaves the normal unchanged;
  - a camera-facing square: the perturbed normal has the closed form
    normalize(-n.x, -n.y, n.z) (derivation in the test), which fails if the
    frame used the uv derivatives (the square's uv is rotated against the
    screen) or swapped dFdx / dFdy;
  - a constant texture is the same as the untextured material with that
    value (albedo, metallic-roughness), in the G-buffer, the hybrid hit
    shader and the pure ray tracing hit shader;
  - an emissive texel seen directly in pure ray tracing mode with black
    lights and environment is the pixel value;
  - postFragment.frag against float64 pow.
"""
import numpy as np
import pytest

from vulkanhybridraytracer_amd import scenes


def _pixel_uv(sc, size=8.0, uv_scale=3.0, eye_z=10.0, fov=60.0):
    """float64 uv of the facing square at each pixel centre (rt_camera model)."""
    W, H = sc.width, sc.height
    x = (np.arange(W) + 0.5) / W * 2 - 1
    y = (np.arange(H) + 0.5) / H * 2 - 1
    t = np.tan(np.radians(fov) / 2)
    aspect = W / H
    # camera at (0, 0, eye_z) looking down -z, up +y; projection y flipped
    X = x[None, :] * t * aspect * eye_z
    Y = -y[:, None] * t * eye_z
    X, Y = np.broadcast_arrays(X, Y)
    h = size / 2
    u = (Y + h) / size * uv_scale
    v = (X + h) / size * uv_scale
    inside = (np.abs(X) < h) & (np.abs(Y) < h)
    return u, v, inside


def _coord_texture(w=16, h=8):
    t = np.zeros((h, w, 4), dtype=np.uint8)
    t[..., 0] = np.arange(w)[None, :] * 7
    t[..., 1] = np.arange(h)[:, None] * 11
    t[..., 2] = 77
    t[..., 3] = 255
    return t


@pytest.mark.parametrize("shift", [0.0, -5.3])
def test_nearest_repeat_texel_selection(oracle_lib, shift):
    w, h = 16, 8
    sc = scenes.make_facing_quad((0, -1, -1, -1), width=96, height=64)
    sc.meshes[0][0][:, 9] += shift
    sc.textures = [_coord_texture(w, h)]
    os_ = oracle_lib.OracleScene(sc)
    gb, _ = os_.gbuffer()
    u, v, inside = _pixel_uv(sc)
    u = u + shift
    # away from texel edges (float32 vs float64 rounding)
    fu, fv = u * w, v * h
    safe = inside & (np.abs(fu - np.rint(fu)) > 1e-3) & (np.abs(fv - np.rint(fv)) > 1e-3)
    assert safe.sum() > 1000
    alb = gb["albedo"]
    got_x = np.rint(alb[..., 0] * 255 / 7).astype(int)
    got_y = np.rint(alb[..., 1] * 255 / 11).astype(int)
    exp_x = np.floor(fu).astype(int) % w
    exp_y = np.floor(fv).astype(int) % h
    assert np.array_equal(got_x[safe], exp_x[safe])
    assert np.array_equal(got_y[safe], exp_y[safe])
    assert len(np.unique(exp_x[safe])) == w and len(np.unique(exp_y[safe])) == h
    # UNORM decode c / 255 in float32, times the interpolated white vertex
    # colour ((b0 + b1) + b2 = 1 within rounding)
    assert np.abs(alb[..., 2][inside] - np.float32(77) / np.float32(255)).max() <= 1.2e-7


def test_vertex_colour_multiplies_gbuffer_albedo(oracle_lib):
    col = (0.5, 0.25, 1.0)
    t = np.zeros((4, 4, 4), dtype=np.uint8)
    t[...] = (200, 100, 50, 255)
    sc = scenes.make_facing_quad((0, -1, -1, -1), color=col)
    sc.textures = [t]
    gb, _ = oracle_lib.OracleScene(sc).gbuffer()
    _, _, inside = _pixel_uv(sc)
    exp = np.array([200, 100, 50], dtype=np.float32) / np.float32(255) * np.array(col, dtype=np.float32)
    assert np.abs(gb["albedo"][inside][:, :3] - exp).max() <= 2.5e-7
    # untextured: the material colour, no vertex colour (pushC.color.xyz)
    sc2 = scenes.make_facing_quad((-1, -1, -1, -1), color=col, albedo=(0.3, 0.6, 0.9))
    gb2, _ = oracle_lib.OracleScene(sc2).gbuffer()
    assert np.all(gb2["albedo"][inside][:, :3] == np.array((0.3, 0.6, 0.9), dtype=np.float32))


def _const_tex(c, n=4):
    t = np.zeros((n, n, 4), dtype=np.uint8)
    t[...] = c
    return t


def test_flat_normal_map_keeps_the_normal(oracle_lib):
    sc = scenes.make_textured(width=96, height=54)
    flat = _const_tex((128, 128, 255, 255))
    sc.textures[1] = flat
    gb, _ = oracle_lib.OracleScene(sc).gbuffer()
    sc_plain = scenes.make_textured(width=96, height=54)
    sc_plain.materials[:, 5] = -1.0  # no normal maps
    gb0, _ = oracle_lib.OracleScene(sc_plain).gbuffer()
    hit = gb0["normal"][..., 3] == 1.0
    assert hit.mean() > 0.5
    d = np.abs(gb["normal"][hit][:, :3] - gb0["normal"][hit][:, :3])
    assert d.max() < 2e-6


@pytest.mark.parametrize("texel", [(255, 128, 128), (128, 255, 128), (255, 60, 200), (10, 200, 180)])
def test_facing_square_normal_map_closed_form(oracle_lib, texel):
    """Camera at +z looking down -z, square in z = 0 with N = (0, 0, 1).
    Screen columns move world x by a > 0 and rows move world y by -b < 0, so
    dFdx(p) = (a, 0, 0), dFdy(p) = (0, -b, 0). The shader's frame
    (geometry_shader.frag:28-42, with duv = dFdx(p).xy, dFdy(p).xy):
      dp2perp = cross(dp2, N) = (-b, 0, 0), dp1perp = cross(N, dp1) = (0, a, 0)
      T = dp2perp * a + dp1perp * 0 = (-ab, 0, 0)
      B = dp2perp * 0 + dp1perp * (-b) = (0, -ab, 0)
    scaled by 1 / ab: T = -x, B = -y, so N' = normalize(-n.x, -n.y, n.z) with
    n = (c - 128) / 127."""
    sc = scenes.make_facing_quad((-1, 0, -1, -1), width=64, height=64)
    sc.textures = [_const_tex(texel + (255,))]
    gb, _ = oracle_lib.OracleScene(sc).gbuffer()
    _, _, inside = _pixel_uv(sc)
    n = (np.array(texel, dtype=np.float64) - 128.0) / 127.0
    exp = np.array([-n[0], -n[1], n[2]])
    exp /= np.linalg.norm(exp)
    got = gb["normal"][inside][:, :3]
    assert np.abs(got - exp).max() < 2e-5


def _equiv_scenes(slot, texel, attr):
    """make_textured with every material using `slot` set to a constant
    texture, and the same scene with the slot off and the material factor set
    to the texel value instead."""
    a = scenes.make_textured(width=96, height=54, spheres=12)
    b = scenes.make_textured(width=96, height=54, spheres=12)
    # texture 5 = the constant texel; every material uses it in `slot`
    a.textures.append(_const_tex(texel))
    a.materials[:, 4 + slot] = 5.0
    b.materials[:, 4 + slot] = -1.0
    c = np.array(texel[:3], dtype=np.float32) / np.float32(255)
    attr(b.materials, c)
    return a, b


def _all_outputs(oracle_lib, sc, gb_in=None):
    """G-buffer of sc, then shadows and hybrid shading of gb_in (default: that
    G-buffer) and the pure ray traced frame"""
    os_ = oracle_lib.OracleScene(sc)
    gb, _ = os_.gbuffer()
    g = gb if gb_in is None else gb_in
    sh, _ = os_.trace_shadows(g)
    lit, refl, _ = os_.shade_hybrid(g, sh)
    rt, _ = os_.render_rt()
    return gb, lit, refl, rt


def test_constant_textures_equal_material_factors(oracle_lib):
    def set_rm(m, c):
        m[:, 9] = c[2]   # metallic = b
        m[:, 10] = c[1]  # roughness = g

    def set_albedo(m, c):
        m[:, 0:3] = c

    for slot, texel, attr in ((3, (0, 140, 255, 255), set_rm), (3, (0, 30, 0, 255), set_rm),
                              (0, (230, 120, 40, 255), set_albedo)):
        a, b = _equiv_scenes(slot, texel, attr)
        if slot == 0:
            # the G-buffer albedo is texel * interpolated vertex colour: white
            # vertices give the material colour within rounding of (b0 + b1) + b2;
            # the hit shaders take no vertex colour, so shading both scenes from
            # one G-buffer must agree exactly
            for V, _ in a.meshes + b.meshes:
                V[:, 6:9] = 1.0
        gb_, lb, rb, tb = _all_outputs(oracle_lib, b)
        ga, la, ra, ta = _all_outputs(oracle_lib, a, gb_in=gb_ if slot == 0 else None)
        for k in ("position", "normal", "albedo", "material", "emissive"):
            if slot == 0 and k == "albedo":
                assert np.abs(ga[k] - gb_[k]).max() <= 1.2e-7
                continue
            assert np.array_equal(ga[k], gb_[k], equal_nan=True), (slot, k)
        assert np.array_equal(la, lb, equal_nan=True), slot
        assert np.array_equal(ra, rb, equal_nan=True), slot
        assert np.array_equal(ta, tb, equal_nan=True), slot


def test_emissive_texel_seen_directly(oracle_lib):
    w, h = 16, 8
    tex = _coord_texture(w, h)
    sc = scenes.make_facing_quad((-1, -1, 0, -1), width=96, height=64)
    sc.textures = [tex]
    sc.lights = np.stack([scenes.point_light((0, 0, 50), color=(0, 0, 0), intensity=0.0, max_distance=1.0)])
    rt, _ = oracle_lib.OracleScene(sc).render_rt(camera=_frame(sc.camera, -1))
    u, v, inside = _pixel_uv(sc)
    fu, fv = u * w, v * h
    safe = inside & (np.abs(fu - np.rint(fu)) > 1e-3) & (np.abs(fv - np.rint(fv)) > 1e-3)
    ex = tex[np.floor(fv).astype(int) % h, np.floor(fu).astype(int) % w, :3].astype(np.float32) / np.float32(255)
    assert safe.sum() > 1000
    assert np.array_equal(rt[..., :3][safe], ex[safe])


def _frame(cam, f):
    c = np.asarray(cam, dtype=np.float32).copy()
    c[32] = f
    return c


def test_post_gamma_closed_form(oracle_lib):
    rng = np.random.default_rng(0)
    img = rng.uniform(-0.5, 4.0, size=(64, 64, 4)).astype(np.float32)
    img[0, 0, :3], img[0, 1, :3], img[0, 2, :3] = 0.0, 0.001, 1.0
    out, q = oracle_lib.post_gamma(img)
    ref = np.power(np.maximum(img[..., :3].astype(np.float64), np.float64(np.float32(0.001))),
                   np.float64(np.float32(1.0) / np.float32(2.2)))
    assert np.abs(out[..., :3] - ref).max() <= 4e-7 * max(1.0, ref.max())
    assert np.array_equal(out[..., 3], img[..., 3])
    assert out[0, 2, 0] == 1.0
    exp_q = np.rint(np.clip(out, 0, 1) * np.float32(255)).astype(np.uint8)
    assert np.array_equal(q, exp_q)
