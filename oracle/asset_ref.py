"""TEST INFRASTRUCTURE ONLY - restatement of the reference's asset loaders in
Python, the checker of the library's hr_asset_load / hr_image_load.

The reference loads models with tinyobjloader and tinygltf and images with
stb_image (dependencies its tree does not vendor; their versions are not
pinned there). This module restates what the reference's calls return:

  load_obj    Mesh::load_from_obj (vk_mesh.cpp:96-150): tinyobjloader
              LoadObj with triangulation (faces fanned from their first
              vertex), per face vertex (position, normal, colour 1, uv =
              (u, 1 - v) when the file has texture coordinates), deduplicated
              by value in first-seen order; one primitive, default material
              (Prefab::createOBJprefab, vk_mesh.cpp:832-841).
  load_gltf   Prefab::GET / loadNode / loadMaterial / loadTextures
              (vk_mesh.cpp:646-901), primitives in Node::node_to_instance
              order (vk_mesh.cpp:494-515) with the node's global matrix
              (vk_mesh.cpp:416-422).
  decode_png  stbi_load(..., STBI_rgb_alpha) for PNG: Python's zlib for the
              inflate, the five PNG row filters, 16-bit samples -> high byte,
              1/2/4-bit grey scaled by 255/85/17, palettes, tRNS, Adam7.
  decode_hdr  stbi_load for Radiance HDR: RGBE (flat or run-length
              scanlines) -> float -> (uint8) clamp(pow(v, 1/2.2) * 255 + 0.5).
JPEG is checked against libjpeg (Pillow) within a tolerance instead: stb's
decoder is absent, and JPEG decoders are not bit-identical across libraries.
"""
import json
import math
import os
import struct
import zlib

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------- images
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    if pb <= pc:
        return b
    return c


def _unfilter(raw, w, h, bits_pp):
    stride = (w * bits_pp + 7) // 8
    bpp = max(1, bits_pp // 8)
    out = bytearray(h * stride)
    pos = 0
    for y in range(h):
        ft = raw[pos]
        src = raw[pos + 1:pos + 1 + stride]
        pos += 1 + stride
        row = bytearray(stride)
        prev = out[(y - 1) * stride:y * stride] if y else bytes(stride)
        for i in range(stride):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if (i >= bpp) else 0
            x = src[i]
            if ft == 0:
                v = x
            elif ft == 1:
                v = x + a
            elif ft == 2:
                v = x + b
            elif ft == 3:
                v = x + ((a + b) >> 1)
            elif ft == 4:
                v = x + _paeth(a, b, c)
            else:
                raise ValueError("bad filter")
            row[i] = v & 255
        out[y * stride:(y + 1) * stride] = row
    return bytes(out), pos


def _samples(row, n, depth):
    if depth == 8:
        return list(row[:n])
    if depth == 16:
        return [(row[2 * i] << 8) | row[2 * i + 1] for i in range(n)]
    per = 8 // depth
    out = []
    for i in range(n):
        byte = row[i // per]
        shift = 8 - depth * (i % per + 1)
        out.append((byte >> shift) & ((1 << depth) - 1))
    return out


def decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p = 8
    idat = b""
    pal = [[0, 0, 0, 255] for _ in range(256)]
    key = None
    while p + 8 <= len(data):
        ln, typ = struct.unpack(">I4s", data[p:p + 8])
        c = data[p + 8:p + 8 + ln]
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, inter = struct.unpack(">IIBBBBB", c[:13])
        elif typ == b"PLTE":
            for i in range(ln // 3):
                pal[i][:3] = list(c[3 * i:3 * i + 3])
        elif typ == b"tRNS":
            if ctype == 3:
                for i, a in enumerate(c):
                    pal[i][3] = a
            elif ctype == 0:
                key = (struct.unpack(">H", c[:2])[0],)
            elif ctype == 2:
                key = struct.unpack(">HHH", c[:6])
        elif typ == b"IDAT":
            idat += c
        elif typ == b"IEND":
            break
        p += 12 + ln
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bits_pp = ch * depth
    raw = zlib.decompress(idat)
    out = np.zeros((h, w, 4), dtype=np.uint8)
    scale = {1: 255, 2: 85, 4: 17, 8: 1, 16: 1}[depth]

    def to8(v):
        return v >> 8 if depth == 16 else (v * scale if depth < 8 else v)

    def emit(px, pw, ph, x0, y0, dx, dy):
        stride = (pw * bits_pp + 7) // 8
        for y in range(ph):
            s = _samples(px[y * stride:(y + 1) * stride], pw * ch, depth)
            for x in range(pw):
                v = s[x * ch:(x + 1) * ch]
                o = out[y0 + y * dy, x0 + x * dx]
                if ctype == 0:
                    g = to8(v[0])
                    t = key is not None and (v[0] == key[0] if depth == 16 else g == (key[0] * scale) & 255)
                    o[:] = (g, g, g, 0 if t else 255)
                elif ctype == 2:
                    rgb = [to8(q) for q in v]
                    t = key is not None and all((v[k] == key[k]) if depth == 16 else (rgb[k] == key[k] & 255)
                                                for k in range(3))
                    o[:] = rgb + [0 if t else 255]
                elif ctype == 3:
                    o[:] = pal[v[0]]
                elif ctype == 4:
                    g = to8(v[0])
                    o[:] = (g, g, g, to8(v[1]))
                else:
                    o[:] = [to8(q) for q in v]

    if not inter:
        px, _ = _unfilter(raw, w, h, bits_pp)
        emit(px, w, h, 0, 0, 1, 1)
        return out
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    off = 0
    for xo, yo, xs, ys in passes:
        pw = (w - xo + xs - 1) // xs if w > xo else 0
        ph = (h - yo + ys - 1) // ys if h > yo else 0
        if not pw or not ph:
            continue
        px, used = _unfilter(raw[off:], pw, ph, bits_pp)
        emit(px, pw, ph, xo, yo, xs, ys)
        off += used
    return out


def decode_hdr(data):
    lines = data.split(b"\n")
    assert lines[0] in (b"#?RADIANCE", b"#?RGBE")
    i = 1
    while lines[i] != b"":
        i += 1
    res = lines[i + 1].split()
    h, w = int(res[1]), int(res[3])
    p = sum(len(l) + 1 for l in lines[:i + 2])
    rgbe = np.zeros((h, w, 4), dtype=np.uint8)
    if w < 8 or w >= 32768:
        rgbe[...] = np.frombuffer(data[p:p + w * h * 4], dtype=np.uint8).reshape(h, w, 4)
    else:
        for y in range(h):
            if data[p] != 2 or data[p + 1] != 2 or data[p + 2] & 0x80:
                rest = np.frombuffer(data[p:p + (h - y) * w * 4], dtype=np.uint8).reshape(h - y, w, 4)
                rgbe[y:] = rest
                break
            p += 4
            for k in range(4):
                x = 0
                while x < w:
                    c = data[p]
                    p += 1
                    if c > 128:
                        c -= 128
                        rgbe[y, x:x + c, k] = data[p]
                        p += 1
                    else:
                        rgbe[y, x:x + c, k] = np.frombuffer(data[p:p + c], dtype=np.uint8)
                        p += c
                    x += c
    out = np.full((h, w, 4), 255, dtype=np.uint8)
    gamma = float(F32(1.0) / F32(2.2))
    for y in range(h):
        for x in range(w):
            r, g, b, e = (int(v) for v in rgbe[y, x])
            f = [0.0, 0.0, 0.0]
            if e:
                m = F32(math.ldexp(1.0, e - 136))
                f = [F32(F32(r) * m), F32(F32(g) * m), F32(F32(b) * m)]
            for k in range(3):
                z = F32(F32(math.pow(float(f[k]), gamma)) * F32(255)) + F32(0.5)
                z = min(max(float(z), 0.0), 255.0)
                out[y, x, k] = int(z)
    return out


# ---------------------------------------------------------------- OBJ
DEFAULT_MATERIAL = [1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 0.0, 1.0, 1.0, 0.0]


def load_obj(path):
    V, N, T = [], [], []
    corners = []
    for raw in open(path, "rb").read().split(b"\n"):
        line = raw.rstrip(b"\r").strip().decode()
        if not line or line.startswith("#"):
            continue
        tok = line.split()
        if tok[0] == "v":
            V.append([F32(float(t)) for t in tok[1:4]])
        elif tok[0] == "vn":
            N.append([F32(float(t)) for t in tok[1:4]])
        elif tok[0] == "vt":
            T.append([F32(float(tok[1])), F32(float(tok[2])) if len(tok) > 2 else F32(0)])
        elif tok[0] == "f":
            face = []
            for t in tok[1:]:
                if t.startswith("#"):
                    break
                parts = t.split("/")

                def fix(s, n):
                    i = int(s)
                    return i - 1 if i > 0 else n + i
                v = fix(parts[0], len(V))
                vt = fix(parts[1], len(T)) if len(parts) > 1 and parts[1] else -1
                vn = fix(parts[2], len(N)) if len(parts) > 2 and parts[2] else -1
                face.append((v, vt, vn))
            for k in range(2, len(face)):
                corners += [face[0], face[k - 1], face[k]]
    has_uv = len(T) > 0
    verts, idx, seen = [], [], {}
    for v, vt, vn in corners:
        rec = tuple(float(x) for x in V[v]) + tuple(float(x) for x in N[vn]) + (1.0, 1.0, 1.0)
        rec += (float(T[vt][0]), float(F32(1) - T[vt][1])) if has_uv else (0.0, 0.0)
        key = tuple(0.0 if x == 0.0 else x for x in rec)
        if key not in seen:
            seen[key] = len(verts)
            verts.append(rec)
        idx.append(seen[key])
    return dict(vertices=np.array(verts, dtype=np.float32).reshape(-1, 11), indices=np.array(idx, dtype=np.uint32),
                prims=[(0, len(idx), 0, np.eye(4, dtype=np.float32))],
                materials=np.array([DEFAULT_MATERIAL], dtype=np.float32), textures=[])


# ---------------------------------------------------------------- glTF
def _mat_mul(A, B):
    """glm float mat4 product, column i = ((A0 b0 + A1 b1) + A2 b2) + A3 b3."""
    R = np.zeros((4, 4), dtype=np.float32)  # R[col][row]
    for i in range(4):
        for r in range(4):
            s = F32(A[0][r] * B[i][0])
            s = F32(s + F32(A[1][r] * B[i][1]))
            s = F32(s + F32(A[2][r] * B[i][2]))
            s = F32(s + F32(A[3][r] * B[i][3]))
            R[i][r] = s
    return R


def _local(n):
    m = np.eye(4, dtype=np.float32)  # m[col][row]
    if len(n.get("translation", [])) == 3:
        m[3][:3] = [F32(v) for v in n["translation"]]
    if len(n.get("rotation", [])) == 4:
        x, y, z, w = n["rotation"]
        R = np.eye(4, dtype=np.float32)
        R[0][:3] = [1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)]
        R[1][:3] = [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)]
        R[2][:3] = [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]
        m = _mat_mul(m, R)
    if len(n.get("scale", [])) == 3:
        for c in range(3):
            m[c] = m[c] * F32(n["scale"][c])
    if len(n.get("matrix", [])) == 16:
        m = np.array(n["matrix"], dtype=np.float32).reshape(4, 4)
    return m


def load_gltf(path, invert_normals=False, texture_source=False):
    data = open(path, "rb").read()
    base = os.path.dirname(path)
    bin_chunk = None
    if data[:4] == b"glTF":
        o = 12
        doc = None
        while o + 8 <= len(data):
            ln, typ = struct.unpack("<II", data[o:o + 8])
            if typ == 0x4E4F534A:
                doc = json.loads(data[o + 8:o + 8 + ln])
            elif typ == 0x004E4942 and bin_chunk is None:
                bin_chunk = data[o + 8:o + 8 + ln]
            o += 8 + ln
    else:
        doc = json.loads(data)

    def uri_bytes(uri):
        if uri.startswith("data:"):
            import base64
            return base64.b64decode(uri[uri.index(";base64,") + 8:])
        from urllib.parse import unquote
        return open(os.path.join(base, unquote(uri)), "rb").read()

    buffers = [uri_bytes(b["uri"]) if "uri" in b else bin_chunk for b in doc.get("buffers", [])]

    def accessor(ai, comps):
        a = doc["accessors"][ai]
        v = doc["bufferViews"][a["bufferView"]]
        off = v.get("byteOffset", 0) + a.get("byteOffset", 0)
        stride = v.get("byteStride", 0) or 4 * comps
        buf = buffers[v["buffer"]]
        return [np.frombuffer(buf, dtype="<f4", count=comps, offset=off + e * stride) for e in range(a["count"])]

    verts, idx, mats, tex_keys, textures = [], [], [], [], []
    nodes = []  # dicts: local, parent, children, prims
    prims = []

    def same(m1, m2):
        return all(m1[k] == m2[k] for k in ("shading", "diffuse", "metallic", "roughness", "ior", "tex"))

    def load_material(index):
        m = dict(shading=0, diffuse=[F32(1)] * 4, metallic=F32(1), roughness=F32(1), ior=F32(1), tex=[-1] * 4)
        if index > -1:
            tm = doc["materials"][index]
            pbr = tm.get("pbrMetallicRoughness", {})
            if len(pbr.get("baseColorFactor", [])) == 4:
                m["diffuse"] = [F32(v) for v in pbr["baseColorFactor"]]
            m["metallic"] = F32(pbr.get("metallicFactor", 1.0))
            m["roughness"] = F32(pbr.get("roughnessFactor", 1.0))
            m["tex"] = [pbr.get("baseColorTexture", {}).get("index", -1),
                        pbr.get("metallicRoughnessTexture", {}).get("index", -1),
                        tm.get("emissiveTexture", {}).get("index", -1),
                        tm.get("normalTexture", {}).get("index", -1)]
        for i, o in enumerate(mats):
            if same(o, m):
                return i
        mats.append(m)
        return len(mats) - 1

    def texture_id(t):
        image = doc["textures"][t]["source"] if texture_source else t
        im = doc["images"][image]
        from urllib.parse import unquote
        if "uri" in im:
            key = ("#data%d" % image) if im["uri"].startswith("data:") else os.path.join(base, unquote(im["uri"]))
        else:
            key = "#image%d" % image
        if key in tex_keys:
            return tex_keys.index(key)
        if "uri" in im:
            raw = uri_bytes(im["uri"])
        else:
            v = doc["bufferViews"][im["bufferView"]]
            o = v.get("byteOffset", 0)
            raw = buffers[v["buffer"]][o:o + v["byteLength"]]
        textures.append(decode_image(raw))
        tex_keys.append(key)
        return len(tex_keys) - 1

    def load_textures(mi):
        m = mats[mi]
        for k in (0, 3, 2, 1):
            if m["tex"][k] > -1:
                m["tex"][k] = texture_id(m["tex"][k])

    def load_node(ni, parent, invert):
        tn = doc["nodes"][ni]
        me = dict(local=_local(tn), parent=parent, children=[], prims=[])
        nid = len(nodes)
        nodes.append(me)
        for c in tn.get("children", []):
            me["children"].append(load_node(c, nid, False))
        if tn.get("mesh", -1) > -1:
            for tp in doc["meshes"][tn["mesh"]]["primitives"]:
                first_index, first_vertex = len(idx), len(verts)
                at = tp.get("attributes", {})
                P = accessor(at["POSITION"], 3) if "POSITION" in at else []
                Nn = accessor(at["NORMAL"], 3) if "NORMAL" in at else None
                Tc = accessor(at["TEXCOORD_0"], 2) if "TEXCOORD_0" in at else None
                for v in range(len(P)):
                    n = Nn[v].astype(np.float32) if Nn is not None else np.zeros(3, np.float32)
                    with np.errstate(divide="ignore", invalid="ignore"):
                        d = F32(F32(F32(n[0] * n[0]) + F32(n[1] * n[1])) + F32(n[2] * n[2]))
                        inv = F32(F32(1) / F32(np.sqrt(d)))
                        nn = (n * inv).astype(np.float32)
                    if invert:
                        nn = (nn * F32(-1)).astype(np.float32)
                    uv = Tc[v] if Tc is not None else (0.0, 0.0)
                    verts.append(list(P[v]) + list(nn) + [1.0, 1.0, 1.0] + list(uv))
                if "indices" in tp:
                    a = doc["accessors"][tp["indices"]]
                    vw = doc["bufferViews"][a["bufferView"]]
                    dt = {5125: "<u4", 5123: "<u2", 5121: "u1"}[a["componentType"]]
                    arr = np.frombuffer(buffers[vw["buffer"]], dtype=dt, count=a["count"],
                                        offset=vw.get("byteOffset", 0) + a.get("byteOffset", 0))
                    idx.extend(int(i) + first_vertex for i in arr)
                    count = a["count"]
                else:
                    idx.extend(range(first_vertex, first_vertex + len(P)))
                    count = len(P)
                mi = load_material(tp.get("material", -1))
                load_textures(mi)
                me["prims"].append(len(prims))
                prims.append((first_index, count, mi))
        return nid

    def global_m(n):
        g = nodes[n]
        return g["local"] if g["parent"] < 0 else _mat_mul(global_m(g["parent"]), g["local"])

    out_prims = []

    def instance(n):
        if nodes[n]["prims"]:
            m = global_m(n)
            for pi in nodes[n]["prims"]:
                fi, cnt, mi = prims[pi]
                out_prims.append((fi, cnt, mi, m.copy()))
        for c in nodes[n]["children"]:
            instance(c)

    roots = [load_node(r, -1, invert_normals) for r in doc["scenes"][0].get("nodes", [])]
    for r in roots:
        instance(r)
    gm = []
    for i, m in enumerate(mats):
        t = m["tex"]
        gm.append([*m["diffuse"][:3], m["ior"], t[0], t[3], t[2], t[1], m["shading"], m["metallic"], m["roughness"], i])
    if not gm:
        gm = [DEFAULT_MATERIAL]
    return dict(vertices=np.array(verts, dtype=np.float32).reshape(-1, 11), indices=np.array(idx, dtype=np.uint32),
                prims=out_prims, materials=np.array(gm, dtype=np.float32), textures=textures)


def decode_image(data):
    if data[:8] == b"\x89PNG\r\n\x1a\n":
        return decode_png(data)
    if data[:2] == b"#?":
        return decode_hdr(data)
    raise ValueError("the oracle restates PNG and HDR only")
