"""Asset ingest (host only): the native image and model loaders of libhr
against the Python restatement in oracle/asset_ref.py and, for images,
against Pillow's decoders (libpng / libjpeg):

  - PNG: every colour type and bit depth, tRNS, all five row filters and
    Adam7 - bit-exact against the restatement and, for the 8-bit types,
    against Pillow;
  - JPEG (Pillow-encoded baseline 4:4:4 / 4:2:2 / 4:2:0, grey, progressive,
    restart intervals): within 3 levels of Pillow's decode, mean error below
    0.6 (JPEG decoders are not bit-identical across libraries);
  - Radiance HDR, flat and run-length: bit-exact against the restatement;
  - OBJ: known answers for a quad-faced cube (24 unique vertices, 36 indices,
    v flipped) and a fanned pentagon; equality with the restatement;
  - glTF / GLB with external, data-URI and embedded images: equality with the
    restatement (NaN normals where NORMAL is missing, as the loader
    normalises a zero vector), node matrices against float64 products,
    material deduplication, the texture-index quirk and its flag;
  - scene composition: entity x node matrices, index offsets, and a textured
    asset rendered by the CPU oracle.
"""
import io
import json
import os

import numpy as np
import pytest

import asset_fixtures as fx
from oracle import asset_ref
from vulkanhybridraytracer_amd import assets, scenes

PIL = pytest.importorskip("PIL.Image")


def _lib_ok():
    from vulkanhybridraytracer_amd import hr
    try:
        hr.load()
        return True
    except Exception:
        return False


pytestmark = pytest.mark.skipif(not _lib_ok(), reason="libhr.so is not built")


# ---------------------------------------------------------------- PNG
@pytest.mark.parametrize("case", fx.PNG_CASES, ids=[c[0] for c in fx.PNG_CASES])
def test_png_against_restatement(case):
    for seed, (w, h) in enumerate([(13, 11), (1, 1), (9, 17)]):
        data = fx.make_png(case, w, h, seed)
        got = assets.decode_image(data)
        ref = asset_ref.decode_png(data)
        assert got.shape == (h, w, 4)
        assert np.array_equal(got, ref), (case[0], w, h)


@pytest.mark.parametrize("case", [c for c in fx.PNG_CASES if c[2] == 8 and c[5] is None],
                         ids=[c[0] for c in fx.PNG_CASES if c[2] == 8 and c[5] is None])
def test_png_against_pillow(case):
    data = fx.make_png(case, 23, 19, 5)
    ref = np.asarray(PIL.open(io.BytesIO(data)).convert("RGBA"))
    assert np.array_equal(assets.decode_image(data), ref)


def test_png_grey_low_depth_scaling():
    # 1/2/4-bit grey is scaled to 8 bits by 255 / 85 / 17 (no palette scaling)
    for depth, mul in ((1, 255), (2, 85), (4, 17)):
        s = np.arange((1 << depth)).reshape(1, -1, 1)
        got = assets.decode_image(fx.png_bytes(s, 0, depth))
        assert np.array_equal(got[0, :, 0], (np.arange(1 << depth) * mul).astype(np.uint8))
        assert np.all(got[..., 3] == 255)


# ---------------------------------------------------------------- JPEG
def _photo(w, h, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([128 + 100 * np.sin(x / 7.0 + y / 11.0), 128 + 90 * np.cos(x / 5.0 - y / 9.0),
                    (x * 255 // max(1, w - 1) + y) % 256], -1)
    img = img + rng.normal(0, 6, size=img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


JPEG_CASES = [("444", dict(subsampling=0)), ("422", dict(subsampling=1)), ("420", dict(subsampling=2)),
              ("420_q50", dict(subsampling=2, quality=50)), ("progressive", dict(subsampling=2, progressive=True)),
              ("progressive_444", dict(subsampling=0, progressive=True)), ("grey", dict(mode="L")),
              ("restart", dict(subsampling=2, restart_marker_blocks=3))]


@pytest.mark.parametrize("name,opts", JPEG_CASES, ids=[c[0] for c in JPEG_CASES])
def test_jpeg_against_pillow(name, opts):
    opts = dict(opts)
    mode = opts.pop("mode", "RGB")
    for w, h in ((67, 45), (16, 16), (5, 3)):
        im = PIL.fromarray(_photo(w, h, 1))
        if mode == "L":
            im = im.convert("L")
        buf = io.BytesIO()
        try:
            im.save(buf, "JPEG", quality=opts.pop("quality", 90) if "quality" in opts else 90, **opts)
        except (TypeError, ValueError, OSError) as e:  # an encoder option this Pillow lacks
            pytest.skip("Pillow cannot write this variant: %s" % e)
        data = buf.getvalue()
        ref = np.asarray(PIL.open(io.BytesIO(data)).convert("RGBA")).astype(int)
        got = assets.decode_image(data).astype(int)
        d = np.abs(got - ref)
        assert got.shape == ref.shape
        assert d.max() <= 3 and d.mean() < 0.6, (name, w, h, d.max(), d.mean())


# ---------------------------------------------------------------- HDR
@pytest.mark.parametrize("w,h,rle", [(5, 4, True), (37, 6, True), (37, 6, False), (8, 3, True)])
def test_hdr_against_restatement(w, h, rle):
    data = fx.hdr_bytes(fx.make_rgbe(w, h, seed=w + h), rle=rle)
    got = assets.decode_image(data)
    assert np.array_equal(got, asset_ref.decode_hdr(data))
    assert np.all(got[..., 3] == 255)


def test_hdr_known_values():
    # RGBE (128, 64, 0, 129): v = c * 2^(129 - 136) = (1, 0.5, 0);
    # 8-bit = (uint8)(pow(v, 1/2.2) * 255 + 0.5) = (255, 186, 0)
    rgbe = np.array([[[128, 64, 0, 129]]], dtype=np.uint8)
    got = assets.decode_image(fx.hdr_bytes(rgbe, rle=False))
    assert got[0, 0].tolist() == [255, int(0.5 ** (1 / 2.2) * 255 + 0.5), 0, 255]


# ---------------------------------------------------------------- OBJ
def _check_same(a, ref):
    assert np.array_equal(a.vertices, ref["vertices"], equal_nan=True)
    assert np.array_equal(a.indices, ref["indices"])
    assert len(a.prims) == len(ref["prims"])
    for p, (fi, cnt, mi, m) in zip(a.prims, ref["prims"]):
        assert (p["first_index"], p["index_count"], p["material"]) == (fi, cnt, mi)
        assert np.array_equal(p["model"], np.asarray(m, dtype=np.float32).reshape(16))
    assert np.array_equal(a.materials, ref["materials"])
    assert len(a.textures) == len(ref["textures"])
    for t, r in zip(a.textures, ref["textures"]):
        assert np.array_equal(t, r)


def test_obj_cube_known_answers(tmp_path):
    p = tmp_path / "cube.obj"
    p.write_text(fx.OBJ_CUBE)
    a = assets.load_asset(str(p))
    assert a.vertices.shape == (24, 11) and a.indices.shape == (36,)
    # face 1 (f 1/1/1 4/4/1 3/3/1 2/2/1) is fanned from its first vertex
    assert a.indices[:6].tolist() == [0, 1, 2, 0, 2, 3]
    # uv = (u, 1 - v); colour 1
    assert a.vertices[1, 9:11].tolist() == [0.0, 0.0] and a.vertices[0, 9:11].tolist() == [0.0, 1.0]
    assert np.all(a.vertices[:, 6:9] == 1.0)
    # negative indices of the fourth face name the same positions as 4 8 7 3
    assert a.vertices[a.indices[18], :3].tolist() == [-1, 1, -1]
    assert a.materials.tolist() == [[1, 1, 1, 1, -1, -1, -1, -1, 0, 1, 1, 0]]
    _check_same(a, asset_ref.load_obj(str(p)))


def test_obj_polygon_without_uv(tmp_path):
    p = tmp_path / "poly.obj"
    p.write_bytes(fx.OBJ_POLY_NO_UV.encode())
    a = assets.load_asset(str(p))
    assert a.indices.tolist() == [0, 1, 2, 0, 2, 3, 0, 3, 4, 0, 2, 3]
    assert np.all(a.vertices[:, 9:11] == 0)
    _check_same(a, asset_ref.load_obj(str(p)))


def test_obj_errors(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(assets.AssetError, match="normal"):
        assets.load_asset(str(p))
    p.write_text("v 0 0 0\nvn 0 0 1\nf 1//1 2//1 5//1\n")
    with pytest.raises(assets.AssetError, match="out of range"):
        assets.load_asset(str(p))
    with pytest.raises(assets.AssetError, match="cannot open"):
        assets.load_asset(str(tmp_path / "missing.obj"))


# ---------------------------------------------------------------- glTF
@pytest.mark.parametrize("glb,images", [(False, "files"), (False, "data"), (True, "embedded"), (True, "files")])
def test_gltf_against_restatement(tmp_path, glb, images):
    path = fx.make_gltf(str(tmp_path), glb=glb, images=images)
    for invert in (False, True):
        a = assets.load_asset(path, invert_normals=invert)
        _check_same(a, asset_ref.load_gltf(path, invert_normals=invert))
    # NaN normals for the primitive without NORMAL (normalize of a zero vector)
    assert np.isnan(a.vertices[:, 3:6]).any()
    # instance order: root 0 (mesh 0), its child (mesh 1: 2 prims), grandchild (mesh 2), root 3 (mesh 1: 2 prims)
    assert len(a.prims) == 6
    # Loading order: node 0's children first (node 2: no material -> the
    # default; node 1: file materials 1 and 2), then node 0 (file material 0),
    # then node 3 (file materials 1, 2). loadMaterial compares a new material's
    # raw texture indices with stored materials whose indices were already
    # replaced by texture ids (first-use order: image 2 first), so the file's
    # identical materials 0 and 2 and the repeats of node 3 do not all merge:
    # 6 materials, as the reference ends up with.
    assert len(a.materials) == 6
    assert a.materials[0, 4:8].tolist() == [-1, -1, -1, -1]
    assert a.materials[1, 4:8].tolist() == [-1, -1, 0, -1]   # emissive = images[2], the first texture id
    assert a.materials[2, 4:8].tolist() == [1, 2, -1, 0]     # diffuse, normal, emissive, metallic-roughness
    assert len(a.textures) == 3
    assert np.all(a.materials[:, 4:8] < 3)


def test_gltf_node_matrices(tmp_path):
    path = fx.make_gltf(str(tmp_path))
    doc = json.load(open(path))
    a = assets.load_asset(path)

    def local(n):
        M = np.eye(4)
        if "matrix" in n:
            return np.array(n["matrix"], dtype=np.float64).reshape(4, 4).T
        if "translation" in n:
            M[:3, 3] = n["translation"]
        if "rotation" in n:
            x, y, z, w = n["rotation"]
            R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                          [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                          [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
            M[:3, :3] = R
        if "scale" in n:
            M = M @ np.diag(n["scale"] + [1.0])
        return M
    N = doc["nodes"]
    g0 = local(N[0])
    g1 = g0 @ local(N[1])
    g2 = g1 @ local(N[2])
    g3 = local(N[3])
    expect = [g0, g1, g1, g2, g3, g3]
    for p, m in zip(a.prims, expect):
        got = np.asarray(p["model"], dtype=np.float64).reshape(4, 4).T
        assert np.abs(got - m).max() < 1e-5


def test_gltf_texture_index_quirk_and_flag(tmp_path):
    path = fx.make_gltf(str(tmp_path))
    doc = json.load(open(path))
    doc["textures"] = [{"source": 2}, {"source": 0}, {"source": 1}]
    json.dump(doc, open(path, "w"))
    quirk = assets.load_asset(path)
    spec = assets.load_asset(path, texture_source=True)
    _check_same(quirk, asset_ref.load_gltf(path))
    _check_same(spec, asset_ref.load_gltf(path, texture_source=True))
    # the diffuse texture of material 0 is images[0] with the reference's
    # indexing and images[2] with textures[0].source
    imgs = [asset_ref.decode_png(open(os.path.join(str(tmp_path), "tex %d.png" % i), "rb").read()) for i in range(3)]
    # the file's material 0 (base colour 0.8, 0.5, 0.25) is the first loaded with a diffuse texture
    q0 = next(m for m in quirk.materials if m[0] == np.float32(0.8))
    s0 = next(m for m in spec.materials if m[0] == np.float32(0.8))
    assert np.array_equal(quirk.textures[int(q0[4])], imgs[0])
    assert np.array_equal(spec.textures[int(s0[4])], imgs[2])


def test_gltf_errors(tmp_path):
    p = tmp_path / "bad.gltf"
    p.write_text('{"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}], ')
    with pytest.raises(assets.AssetError, match="JSON"):
        assets.load_asset(str(p))
    p.write_text(json.dumps({"asset": {"version": "2.0"}}))
    with pytest.raises(assets.AssetError, match="no scene"):
        assets.load_asset(str(p))
    q = tmp_path / "x.fbx"
    q.write_text("")
    with pytest.raises(assets.AssetError, match="not an"):
        assets.load_asset(str(q))
    with pytest.raises(assets.AssetError, match="format"):
        assets.decode_image(b"GIF89a....")


# ---------------------------------------------------------------- corrupt files
def _mutants(data, rng, n):
    """n seeded corruptions of `data`: random byte flips, flips in the header
    (the first 40 bytes, where sizes live), truncations, and splices of a
    piece of the file into itself"""
    out = []
    for _ in range(n):
        b = bytearray(data)
        kind = int(rng.integers(4))
        if kind in (0, 1):
            span = len(b) if kind == 0 else min(len(b), 40)
            for _ in range(int(rng.integers(1, 9))):
                b[int(rng.integers(span))] = int(rng.integers(256))
        elif kind == 2:
            b = b[:int(rng.integers(len(b)))]
        else:
            i, j = sorted(int(v) for v in rng.integers(len(b), size=2))
            k = int(rng.integers(len(b)))
            b = b[:k] + b[i:j] + b[k:]
        out.append(bytes(b))
    return out


def _jpeg(w, h, **opts):
    buf = io.BytesIO()
    PIL.fromarray(_photo(w, h, 2)).save(buf, "JPEG", quality=85, **opts)
    return buf.getvalue()


CORRUPT_CASES = {
    "png": lambda: fx.make_png(fx.PNG_CASES[0], 23, 19, 1),
    "png_interlaced": lambda: fx.make_png(next(c for c in fx.PNG_CASES if c[6]), 23, 19, 1),
    "jpeg": lambda: _jpeg(41, 29, subsampling=2),
    "jpeg_progressive": lambda: _jpeg(41, 29, subsampling=2, progressive=True),
    "hdr_rle": lambda: fx.hdr_bytes(fx.make_rgbe(37, 6, seed=3), rle=True),
    "hdr_flat": lambda: fx.hdr_bytes(fx.make_rgbe(37, 6, seed=3), rle=False),
}


@pytest.mark.parametrize("fmt", sorted(CORRUPT_CASES))
def test_decoders_survive_corrupt_files(fmt):
    """Seeded corruptions of valid files each decode to an RGBA8 image within
    the size limit or raise AssetError: no crash, no hang, no oversized
    allocation (the sanitizer run of this suite also checks every access)."""
    data = CORRUPT_CASES[fmt]()
    assets.decode_image(data)  # the original decodes
    rng = np.random.default_rng(sum(map(ord, fmt)))
    ok = bad = 0
    for m in _mutants(data, rng, 300):
        try:
            img = assets.decode_image(m)
        except assets.AssetError:
            bad += 1
            continue
        assert img.ndim == 3 and img.shape[2] == 4 and img.shape[0] * img.shape[1] <= 1 << 28
        ok += 1
    assert bad > 0


def test_decoders_refuse_oversized_headers():
    """A header claiming more than 2^28 pixels is refused before any buffer
    of that size exists; so is a claim the remaining bytes cannot back."""
    with pytest.raises(assets.AssetError, match="limit"):
        assets.decode_image(fx.png_bytes(np.zeros((1, 1, 3), np.uint8), 2, 8)[:16]
                            + (40000).to_bytes(4, "big") + (40000).to_bytes(4, "big")
                            + fx.png_bytes(np.zeros((1, 1, 3), np.uint8), 2, 8)[24:])
    hdr = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 30000 +X 30000\n" + bytes(64)
    with pytest.raises(assets.AssetError, match="limit"):
        assets.decode_image(hdr)
    hdr = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 9000 +X 9000\n" + bytes(64)
    with pytest.raises(assets.AssetError, match="truncated"):
        assets.decode_image(hdr)
    jpg = bytearray(_jpeg(16, 16))
    k = jpg.index(b"\xff\xc0")
    jpg[k + 5:k + 9] = (12000).to_bytes(2, "big") + (12000).to_bytes(2, "big")
    with pytest.raises(assets.AssetError, match="truncated"):
        assets.decode_image(bytes(jpg))


def test_gltf_loader_survives_corrupt_files(tmp_path):
    """Seeded corruptions of a .glb and of a .gltf's JSON numbers load or raise
    AssetError; offsets, counts and strides that would reach outside their
    buffer (negative, huge, or wrapping around when multiplied) are refused."""
    (tmp_path / "g").mkdir()
    (tmp_path / "t").mkdir()
    glb = fx.make_gltf(str(tmp_path / "g"), glb=True, images="embedded")
    data = open(glb, "rb").read()
    rng = np.random.default_rng(7)
    p = tmp_path / "m.glb"
    for m in _mutants(data, rng, 150):
        p.write_bytes(m)
        try:
            assets.load_asset(str(p))
        except assets.AssetError:
            pass
    gltf = fx.make_gltf(str(tmp_path / "t"), glb=False, images="data")
    doc = json.load(open(gltf))
    acc = doc["accessors"][0]
    for field, value in (("count", 10 ** 15), ("count", -3), ("byteOffset", -16), ("byteOffset", 10 ** 15)):
        bad = json.loads(json.dumps(doc))
        bad["accessors"][0] = dict(acc, **{field: value})
        q = tmp_path / "t" / ("bad_%s_%d.gltf" % (field, value))
        q.write_text(json.dumps(bad))
        with pytest.raises(assets.AssetError):
            assets.load_asset(str(q))
    view = dict(doc["bufferViews"][acc["bufferView"]], byteStride=10 ** 15)
    bad = json.loads(json.dumps(doc))
    bad["bufferViews"][acc["bufferView"]] = view
    bad["accessors"][0] = dict(acc, count=2 ** 49 + 1)
    q = tmp_path / "t" / "bad_stride.gltf"
    q.write_text(json.dumps(bad))
    with pytest.raises(assets.AssetError, match="exceeds"):
        assets.load_asset(str(q))
    # a subtree shared at every level would multiply without bound
    bad = json.loads(json.dumps(doc))
    bad["nodes"] = [{"children": [k + 1, k + 1]} for k in range(40)] + [{"mesh": 0}]
    bad["scenes"] = [{"nodes": [0]}]
    q = tmp_path / "t" / "bad_nodes.gltf"
    q.write_text(json.dumps(bad))
    with pytest.raises(assets.AssetError, match="node visits"):
        assets.load_asset(str(q))


# ---------------------------------------------------------------- composition
def test_compose_scene_and_render(tmp_path, oracle_lib):
    path = fx.make_gltf(str(tmp_path))
    a = assets.load_asset(path)
    cube = assets.cube().with_material(assets.material((0.9, 0.2, 0.2), metallic=0.0, roughness=0.5))
    E = [assets.Entity(a, assets.translate((0, 1, -6))),
         assets.Entity(a, assets.translate((3, 0, -8)) @ assets.scale(0.5)),
         assets.Entity(cube, assets.translate((0, -2, -6)) @ assets.scale((6, 0.2, 6)))]
    sky, irr = scenes.make_env(64, 32, 16, 8)
    lights = np.stack([scenes.point_light((0, 10, -2), intensity=800.0, radius=0.5)])
    sc = assets.compose_scene(E, lights, assets.reference_camera((0, 2, 4), 96, 64), sky, irr, 96, 64)
    assert len(sc.meshes) == 2 and len(sc.prims) == 2 * len(a.prims) + 1
    assert sc.prims["mesh"].tolist() == [0] * 6 + [0] * 6 + [1]
    # the cube's materials (default, override) follow the asset's; its texture slots stay -1
    na = len(a.materials)
    assert sc.prims["material"][-1] == na + 1 and np.all(sc.materials[na:, 4:8] == -1)
    assert sc.materials[:, 11].tolist() == list(range(len(sc.materials)))
    # entity x node matrix
    for k, p in enumerate(sc.prims[6:12]):
        m = np.asarray(p["model"], dtype=np.float64).reshape(4, 4).T
        ref = (assets.translate((3, 0, -8)) @ assets.scale(0.5)) @ np.asarray(a.prims[k]["model"],
                                                                              dtype=np.float64).reshape(4, 4).T
        assert np.abs(m - ref).max() < 1e-5
    gb, _ = oracle_lib.OracleScene(sc).gbuffer()
    hit = gb["normal"][..., 3] == 1.0
    assert hit.mean() > 0.1
    # textured materials reach the G-buffer: the checker texels differ across pixels
    assert len(np.unique(gb["albedo"][hit][:, 0])) > 10
