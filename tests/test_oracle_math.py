"""The oracle's shading mathematics against closed forms and libm.

helpers.glsl:33-93 (GGX, Smith, Fresnel, sampleDisk) and the Taylor-series
atan2 / asin / sincos(2 pi u) used for texel selection and sample
directions."""
import numpy as np
import pytest


def test_atan2_asin_sincos_accuracy(oracle_lib):
    rng = np.random.default_rng(0)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    x[:100] = 0.0
    y[100:200] = 0.0
    got = oracle_lib.atan2(y, x).astype(np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    want[(x == 0) & (y == 0)] = 0.0
    assert np.max(np.abs(got - want)) < 5e-7
    s = rng.uniform(-1, 1, size=200000).astype(np.float32)
    s[:4] = [-1.0, 1.0, 0.0, 0.5]
    assert np.max(np.abs(oracle_lib.asin(s) - np.arcsin(s.astype(np.float64)))) < 5e-7
    u = rng.uniform(0, 1, size=200000).astype(np.float32)
    u[:5] = [0.0, 0.25, 0.5, 0.75, np.nextafter(np.float32(1), np.float32(0))]
    sn, cs = oracle_lib.sincos2pi(u)
    ang = 2 * np.pi * u.astype(np.float64)
    assert np.max(np.abs(sn - np.sin(ang))) < 5e-7
    assert np.max(np.abs(cs - np.cos(ang))) < 5e-7
    # quadrant boundaries are exact
    assert sn[1] == 1.0 and cs[0] == 1.0 and cs[2] == -1.0 and sn[3] == -1.0


def test_ggx_closed_forms(oracle_lib):
    N = np.array([0, 1, 0], np.float32)
    for r in (0.25, 0.5, 1.0):
        # NdotH = 1: D = a^2 / (pi (a^2)^2) = 1 / (pi r^4)
        assert oracle_lib.distribution_ggx(N, N, r) == pytest.approx(1.0 / (np.pi * r ** 4), rel=1e-6)
    H = np.array([0, np.cos(0.3), np.sin(0.3)], np.float32)
    r = 0.5
    a2 = r ** 4
    nh = float(np.float32(H[1]))
    want = a2 / (np.pi * (nh * nh * (a2 - 1) + 1) ** 2)
    assert oracle_lib.distribution_ggx(N, H, r) == pytest.approx(want, rel=1e-5)
    # roughness 0 with NdotH exactly 1 is 0/0 (helpers.glsl:33-44 as written)
    assert np.isnan(oracle_lib.distribution_ggx(N, N, 0.0))


def test_smith_and_fresnel(oracle_lib):
    N = np.array([0, 0, 1], np.float32)
    for r in (0.0, 0.3, 0.9):
        assert oracle_lib.geometry_smith(N, N, N, r) == pytest.approx(1.0, abs=1e-7)
    V = np.array([0, np.sin(1.0), np.cos(1.0)], np.float32)
    r = 0.4
    k = (r + 1) ** 2 / 8
    nv = float(np.cos(1.0))
    want = (nv / (nv * (1 - k) + k)) * 1.0
    assert oracle_lib.geometry_smith(N, V, N, r) == pytest.approx(want, rel=1e-6)
    F0 = np.array([0.04, 0.5, 0.9], np.float32)
    assert np.allclose(oracle_lib.fresnel_schlick(1.0, F0), F0)
    assert np.allclose(oracle_lib.fresnel_schlick(0.0, F0), 1.0)
    c = 0.3
    assert np.allclose(oracle_lib.fresnel_schlick(c, F0), F0 + (1 - F0) * (1 - c) ** 5, rtol=1e-6)


def test_sample_disk(oracle_lib):
    light = np.zeros(12, np.float32)
    light[8] = 0.0  # radius 0: every sample is the light direction
    L = np.array([0.3, 0.8, -0.52], np.float32)
    L /= np.linalg.norm(L)
    pos = np.array([1.0, 2.0, 3.0], np.float32)
    d = oracle_lib.sample_disk(light, pos, L, 0.7, 0.2)
    assert np.allclose(d, L, atol=1e-6)
    # radius r at unit distance: the sample lies inside the cone of half-angle atan(r)
    light[8] = 0.5
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = oracle_lib.sample_disk(light, pos, L, rng.uniform(), rng.uniform())
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        assert np.dot(d, L) >= np.cos(np.arctan(0.5)) - 1e-6
    # L == +Y exactly takes the special tangent (helpers.glsl:87-88)
    d = oracle_lib.sample_disk(light, pos, np.array([0, 1, 0], np.float32), 1.0 - 2 ** -24, 0.0)
    assert np.all(np.isfinite(d)) and d[2] > 0.4
