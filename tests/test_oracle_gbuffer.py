"""Pins of the raster G-buffer restatement (oracle/gbuffer_ref.py)."""
import numpy as np

from oracle import gbuffer_ref as G


def _raster_from_codes(h, w, rng):
    halves = rng.integers(0, 0x7C00, size=(h, w, 4), dtype=np.uint16)  # finite positive halves
    halves[..., 0] |= rng.integers(0, 2, size=(h, w), dtype=np.uint16) << 15  # signs for x
    return {"position": halves.view(np.float16),
            "normal": rng.integers(0, 0x3C01, size=(h, w, 4), dtype=np.uint16).view(np.float16),  # [0, 1]
            "albedo": rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8),
            "material": rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8),
            "emissive": rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8),
            "motion": rng.integers(0, 0x3C01, size=(h, w, 2), dtype=np.uint16).view(np.float16)}


def test_unorm8_decode_and_roundtrip():
    codes = np.arange(256, dtype=np.uint8).reshape(16, 16)
    rgba = np.stack([codes] * 4, -1)
    planes, _ = G.ingest({"position": np.zeros((16, 16, 4), np.float16), "normal": np.zeros((16, 16, 4), np.float16),
                          "albedo": rgba, "material": rgba, "emissive": rgba})
    assert np.all(planes["albedo"][..., 0] == (codes.astype(np.float32) / np.float32(255)))
    assert planes["albedo"][15, 15, 0] == 1.0 and planes["albedo"][0, 0, 0] == 0.0
    back = G.quantize(planes)
    assert np.array_equal(back["albedo"], rgba) and np.array_equal(back["material"], rgba)


def test_half_known_bits_and_roundtrip():
    bits = np.array([0x3C00, 0xC000, 0x7BFF, 0x0001, 0x3555], np.uint16)
    vals = bits.view(np.float16).astype(np.float32)
    assert vals.tolist() == [1.0, -2.0, 65504.0, 2.0 ** -24, np.float32(np.float16(0.333251953125))]
    rng = np.random.default_rng(0)
    r = _raster_from_codes(8, 9, rng)
    planes, mot = G.ingest(r)
    back = G.quantize(planes, mot)
    assert np.array_equal(back["position"].view(np.uint16), r["position"].view(np.uint16))
    assert np.array_equal(back["motion"].view(np.uint16), r["motion"].view(np.uint16))


def test_normal_encoding():
    """0.5 n + 0.5 maps -1 / 0 / 1 to 0 / 0.5 / 1 and decodes back exactly."""
    N = np.zeros((1, 3, 4), np.float32)
    N[0, :, 0] = [-1.0, 0.0, 1.0]
    N[..., 3] = 1.0
    planes = {"position": np.ones((1, 3, 4), np.float32), "normal": N, "albedo": np.zeros((1, 3, 4), np.float32),
              "material": np.zeros((1, 3, 4), np.float32), "emissive": np.zeros((1, 3, 4), np.float32)}
    q = G.quantize(planes)
    assert q["normal"][0, :, 0].astype(np.float32).tolist() == [0.0, 0.5, 1.0]
    back, _ = G.ingest(q)
    assert back["normal"][0, :, 0].tolist() == [-1.0, 0.0, 1.0]
    # background normals are written as 0 and read back with w == 0
    planes["normal"][0, 1] = 0.0
    q = G.quantize(planes)
    assert np.all(q["normal"][0, 1].astype(np.float32) == 0)
