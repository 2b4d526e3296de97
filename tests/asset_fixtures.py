"""Writers of small asset files for the loader tests: PNG (every colour type,
bit depth, row filter and Adam7), Radiance HDR (flat and run-length
scanlines), OBJ and glTF / GLB. Everything is generated from seeds; nothing
is read from the reference tree."""
import base64
import json
import os
import struct
import zlib

import numpy as np


# ---------------------------------------------------------------- PNG
def _chunk(typ, data):
    return struct.pack(">I", len(data)) + typ + data + struct.pack(">I", zlib.crc32(typ + data) & 0xFFFFFFFF)


def _filter_rows(px_rows, bpp, filters):
    """px_rows: list of bytes per row; returns filtered scanlines."""
    out = b""
    prev = bytes(len(px_rows[0]))
    for y, row in enumerate(px_rows):
        ft = filters[y % len(filters)]
        f = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                p = 0
            elif ft == 1:
                p = a
            elif ft == 2:
                p = b
            elif ft == 3:
                p = (a + b) >> 1
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            f[i] = (row[i] - p) & 255
        out += bytes([ft]) + bytes(f)
        prev = row
    return out


def _pack_row(samples, depth):
    if depth == 8:
        return bytes(int(s) for s in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(s)) for s in samples)
    per = 8 // depth
    out = bytearray((len(samples) + per - 1) // per)
    for i, s in enumerate(samples):
        out[i // per] |= (int(s) & ((1 << depth) - 1)) << (8 - depth * (i % per + 1))
    return bytes(out)


def png_bytes(samples, ctype, depth, palette=None, trns=None, interlace=False, filters=(0, 1, 2, 3, 4)):
    """samples: (h, w, channels) integer array of raw sample values."""
    h, w, ch = samples.shape
    bits_pp = ch * depth
    bpp = max(1, bits_pp // 8)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0)
    data = b""
    if not interlace:
        rows = [_pack_row(samples[y].reshape(-1), depth) for y in range(h)]
        data = _filter_rows(rows, bpp, filters)
    else:
        passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
        for xo, yo, xs, ys in passes:
            sub = samples[yo::ys, xo::xs]
            if sub.shape[0] == 0 or sub.shape[1] == 0:
                continue
            rows = [_pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
            data += _filter_rows(rows, bpp, filters)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, dtype=np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    # split the image data over two IDAT chunks
    z = zlib.compress(data, 6)
    out += _chunk(b"IDAT", z[:len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])
    return out + _chunk(b"IEND", b"")


PNG_CASES = [
    # (name, ctype, depth, channels, palette size, tRNS kind, interlace)
    ("rgba8", 6, 8, 4, 0, None, False),
    ("rgb8", 2, 8, 3, 0, None, False),
    ("rgb8_key", 2, 8, 3, 0, "key", False),
    ("grey8", 0, 8, 1, 0, None, False),
    ("grey8_key", 0, 8, 1, 0, "key", False),
    ("grey1", 0, 1, 1, 0, None, False),
    ("grey2", 0, 2, 1, 0, None, False),
    ("grey4", 0, 4, 1, 0, "key", False),
    ("grey16", 0, 16, 1, 0, None, False),
    ("greya8", 4, 8, 2, 0, None, False),
    ("greya16", 4, 16, 2, 0, None, False),
    ("rgb16", 2, 16, 3, 0, None, False),
    ("rgba16", 6, 16, 4, 0, None, False),
    ("pal1", 3, 1, 1, 2, None, False),
    ("pal2", 3, 2, 1, 4, "alpha", False),
    ("pal4", 3, 4, 1, 16, None, False),
    ("pal8", 3, 8, 1, 200, "alpha", False),
    ("rgba8_adam7", 6, 8, 4, 0, None, True),
    ("pal4_adam7", 3, 4, 1, 16, "alpha", True),
    ("grey2_adam7", 0, 2, 1, 0, None, True),
]


def make_png(case, w=13, h=11, seed=0):
    name, ctype, depth, ch, npal, trns, inter = case
    rng = np.random.default_rng(seed)
    hi = npal if ctype == 3 else (1 << depth)
    s = rng.integers(0, hi, size=(h, w, ch))
    palette = rng.integers(0, 256, size=(npal, 3)) if ctype == 3 else None
    tr = None
    if trns == "key":
        key = [int(v) for v in s[0, 0, :3 if ctype == 2 else 1]]
        tr = b"".join(struct.pack(">H", k) for k in key)
    elif trns == "alpha":
        tr = bytes(rng.integers(0, 256, size=npal // 2).astype(np.uint8))
    return png_bytes(s, ctype, depth, palette, tr, inter)


# ---------------------------------------------------------------- Radiance HDR
def hdr_bytes(rgbe, rle=True):
    """rgbe: (h, w, 4) uint8; run-length scanlines when rle and 8 <= w < 32768."""
    h, w, _ = rgbe.shape
    out = b"#?RADIANCE\n# made by the tests\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w)
    if not rle or w < 8:
        return out + rgbe.astype(np.uint8).tobytes()
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            comp = rgbe[y, :, k]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and comp[x + run] == comp[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, int(comp[x])])
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and comp[x + n] == comp[x + n + 1] == comp[x + n + 2]):
                        n += 1
                    out += bytes([n]) + bytes(int(v) for v in comp[x:x + n])
                    x += n
    return out


def make_rgbe(w, h, seed=0):
    rng = np.random.default_rng(seed)
    rgbe = rng.integers(0, 256, size=(h, w, 4)).astype(np.uint8)
    rgbe[..., 3] = rng.integers(120, 140, size=(h, w))
    rgbe[:, : w // 3] = rgbe[:, :1]           # long runs
    rgbe[h // 2, :, 3] = 0                    # exponent 0 -> black
    return rgbe


# ---------------------------------------------------------------- OBJ
OBJ_CUBE = """# unit cube, quads, shared positions, per-face normals
o cube
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
vn 0 0 -1
vn 0 0 1
vn 0 -1 0
vn 0 1 0
vn -1 0 0
vn 1 0 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
g faces
usemtl whatever
f 1/1/1 4/4/1 3/3/1 2/2/1
f 5/1/2 6/2/2 7/3/2 8/4/2
f 1/1/3 2/2/3 6/3/3 5/4/3
f -5/1/4 -1/2/4 -2/3/4 -6/4/4
f 1/1/5 5/2/5 8/3/5 4/4/5
f 2/1/6 3/2/6 7/3/6 6/4/6  # trailing comment
"""

OBJ_POLY_NO_UV = "v 0 0 0\nv 1 0 0\nv 1.5 1 0\nv 0.5 1.7 0\nv -0.5 1 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1 5//1\r\nf 1//1 3//1 4//1\r\n"


# ---------------------------------------------------------------- glTF
def _pad4(b):
    return b + b"\0" * (-len(b) % 4)


def make_gltf(dirname, glb=False, seed=0, images="files"):
    """A small glTF 2.0 asset: root (TRS) -> child (matrix) -> grandchild,
    plus a second root; three meshes with u8 / u16 / u32 indices, a
    primitive without NORMAL, one without TEXCOORD_0, one non-indexed, an
    interleaved (byteStride) position/normal view; materials with repeated
    factors (deduplicated) and four textures. images: "files" (PNG files
    next to the .gltf), "data" (data URIs) or "embedded" (buffer views)."""
    rng = np.random.default_rng(seed)
    buf = b""
    views, accessors = [], []

    def add_view(data, stride=None):
        nonlocal buf
        buf = _pad4(buf)
        v = {"buffer": 0, "byteOffset": len(buf), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        buf += data
        views.append(v)
        return len(views) - 1

    def add_acc(view, ctype, count, typ, off=0):
        a = {"bufferView": view, "componentType": ctype, "count": count, "type": typ}
        if off:
            a["byteOffset"] = off
        accessors.append(a)
        return len(accessors) - 1

    def grid_mesh(n):
        P = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        N = rng.normal(size=(n, 3)).astype(np.float32)
        T = rng.uniform(-0.5, 2.5, size=(n, 2)).astype(np.float32)
        return P, N, T

    prims = []
    # mesh 0 prim 0: u16 indices, all attributes, material 0
    P, N, T = grid_mesh(9)
    pa = add_acc(add_view(P.tobytes()), 5126, 9, "VEC3")
    na = add_acc(add_view(N.tobytes()), 5126, 9, "VEC3")
    ta = add_acc(add_view(T.tobytes()), 5126, 9, "VEC2")
    ia = add_acc(add_view(np.array([0, 1, 2, 2, 3, 4, 5, 6, 7, 8, 0, 4], np.uint16).tobytes()), 5123, 12, "SCALAR")
    prims.append([{"attributes": {"POSITION": pa, "NORMAL": na, "TEXCOORD_0": ta}, "indices": ia, "material": 0}])
    # mesh 1: prim 0 u8 indices, no NORMAL, material 1; prim 1 u32, no TEXCOORD_0, material 2 (same as 0)
    P, N, T = grid_mesh(6)
    pa = add_acc(add_view(P.tobytes()), 5126, 6, "VEC3")
    ta = add_acc(add_view(T.tobytes()), 5126, 6, "VEC2")
    ia = add_acc(add_view(np.array([0, 1, 2, 3, 4, 5], np.uint8).tobytes()), 5121, 6, "SCALAR")
    P2, N2, _ = grid_mesh(7)
    # interleaved positions and normals, stride 24, accessor offsets 0 and 12
    inter = np.concatenate([P2, N2], 1).astype(np.float32).tobytes()
    iv = add_view(inter, stride=24)
    pa2 = add_acc(iv, 5126, 7, "VEC3")
    na2 = add_acc(iv, 5126, 7, "VEC3", off=12)
    ia2 = add_acc(add_view(np.array([6, 5, 4, 3, 2, 1, 0, 1, 2], np.uint32).tobytes()), 5125, 9, "SCALAR")
    prims.append([{"attributes": {"POSITION": pa, "TEXCOORD_0": ta}, "indices": ia, "material": 1},
                  {"attributes": {"POSITION": pa2, "NORMAL": na2}, "indices": ia2, "material": 2}])
    # mesh 2: non-indexed, no material
    P, N, T = grid_mesh(6)
    pa = add_acc(add_view(P.tobytes()), 5126, 6, "VEC3")
    na = add_acc(add_view(N.tobytes()), 5126, 6, "VEC3")
    prims.append([{"attributes": {"POSITION": pa, "NORMAL": na}}])

    def tex_png(i):
        s = rng.integers(0, 256, size=(5 + i, 7, 4))
        return png_bytes(s, 6, 8)

    pngs = [tex_png(i) for i in range(3)]
    images_json = []
    for i, data in enumerate(pngs):
        if images == "files":
            fn = "tex %d.png" % i  # a space, written %20 in the uri
            open(os.path.join(dirname, fn), "wb").write(data)
            images_json.append({"uri": "tex%%20%d.png" % i})
        elif images == "data":
            images_json.append({"uri": "data:image/png;base64," + base64.b64encode(data).decode()})
        else:
            images_json.append({"bufferView": add_view(data), "mimeType": "image/png"})
    doc = {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": [0, 3]}],
        "nodes": [
            {"translation": [1.0, 2.0, -3.0], "rotation": [0.2, 0.3, 0.1, 0.927361849549570], "scale": [2.0, 0.5, 1.5],
             "children": [1], "mesh": 0},
            {"matrix": [0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 5.0, 0.0, 0.0, 1.0],
             "children": [2], "mesh": 1},
            {"translation": [0.0, 0.0, 1.0], "mesh": 2},
            {"scale": [3.0, 3.0, 3.0], "mesh": 1},
        ],
        "meshes": [{"primitives": p} for p in prims],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.5, 0.25, 1.0], "metallicFactor": 0.3,
                                      "roughnessFactor": 0.7, "baseColorTexture": {"index": 0},
                                      "metallicRoughnessTexture": {"index": 2}},
             "normalTexture": {"index": 1}},
            {"pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.9, 0.2, 0.5]}, "emissiveTexture": {"index": 2}},
            {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.5, 0.25, 1.0], "metallicFactor": 0.3,
                                      "roughnessFactor": 0.7, "baseColorTexture": {"index": 0},
                                      "metallicRoughnessTexture": {"index": 2}},
             "normalTexture": {"index": 1}},
        ],
        "textures": [{"source": 0}, {"source": 1}, {"source": 2}],
        "images": images_json,
        "accessors": accessors,
        "bufferViews": views,
    }
    buf = _pad4(buf)
    if glb:
        doc["buffers"] = [{"byteLength": len(buf)}]
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        out = b"glTF" + struct.pack("<II", 2, 12 + 8 + len(js) + 8 + len(buf))
        out += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(buf), 0x004E4942) + buf
        path = os.path.join(dirname, "model.glb")
        open(path, "wb").write(out)
    else:
        open(os.path.join(dirname, "model.bin"), "wb").write(buf)
        doc["buffers"] = [{"uri": "model.bin", "byteLength": len(buf)}]
        path = os.path.join(dirname, "model.gltf")
        open(path, "w").write(json.dumps(doc, indent=1))
    return path
