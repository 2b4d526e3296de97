"""Numpy restatement of the raster G-buffer formats (test infrastructure only:
the checker of gbuffer_io.hip).

Attachments of renderer.cpp:252-261; encoding as geometry_shader.frag:66-75
and skybox.frag:16-21 write them and the attachments store them (half by
round to nearest even, UNORM8 = rint(clamp(v, 0, 1) * 255)); decoding as
shadowRaygen.rgen:36-41 and hybridRaygen.rgen:40-47 read them (half to
float, normal * 2 - 1, c / 255). All float32 operations in the order the
kernels evaluate them.

Pins (tests/test_oracle_gbuffer.py): every UNORM8 code decodes to c / 255
and re-encodes to itself; every finite half round-trips; known half bit
patterns; the normal encoding maps [-1, 1] onto [0, 1].
"""
import numpy as np

F = np.float32


def _u8(x):
    return (np.asarray(x, dtype=np.uint8).astype(F) / F(255.0)).astype(F)


def _to_u8(v):
    c = (np.minimum(np.maximum(np.asarray(v, F), F(0.0)), F(1.0)) * F(255.0)).astype(F)
    return np.rint(c).astype(np.uint8)


def ingest(raster):
    """raster: dict of position / normal (float16 H x W x 4), albedo /
    material / emissive (uint8 H x W x 4), motion (float16 H x W x 2, optional).
    Returns (planes of float32 H x W x 4, motion float32 H x W x 2 or None)."""
    p = np.asarray(raster["position"], np.float16).astype(F)
    n = np.asarray(raster["normal"], np.float16).astype(F)
    nrm = np.empty_like(n)
    nrm[..., :3] = (n[..., :3] * F(2.0)) - F(1.0)
    nrm[..., 3] = n[..., 3]
    emi = _u8(raster["emissive"])
    emi[..., 3] = 0.0
    planes = {"position": p, "normal": nrm, "albedo": _u8(raster["albedo"]), "material": _u8(raster["material"]),
              "emissive": emi}
    mot = raster.get("motion")
    return planes, (None if mot is None else np.asarray(mot, np.float16).astype(F))


def quantize(planes, motion=None):
    P = np.asarray(planes["position"], F)
    N = np.asarray(planes["normal"], F)
    out = {"position": P.astype(np.float16)}
    enc = np.empty_like(N)
    enc[..., :3] = (N[..., :3] * F(0.5)) + F(0.5)
    enc[..., 3] = N[..., 3]
    enc[N[..., 3] == 0] = 0.0
    out["normal"] = enc.astype(np.float16)
    out["albedo"] = _to_u8(planes["albedo"])
    out["material"] = _to_u8(planes["material"])
    e = np.asarray(planes["emissive"], F).copy()
    e[..., 3] = 1.0
    out["emissive"] = _to_u8(e)
    H, W = P.shape[:2]
    out["motion"] = (np.zeros((H, W, 2), F) if motion is None else np.asarray(motion, F)).astype(np.float16)
    return out
