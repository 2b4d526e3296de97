"""Asset ingest: models and images through libhr's native loaders
(hr_asset_load, hr_image_load), composed into a scenes.Scene the way the
reference's Scene / Renderer compose entities (an entity = a prefab placed by
a matrix, scene.cpp:34-272; its primitives instanced with entity matrix x node
global matrix, vk_mesh.cpp:494-515).

    img = load_image("brick.png")                 # (h, w, 4) uint8 RGBA8
    helmet = load_asset("DamagedHelmet.gltf")     # Asset
    sc = compose_scene([Entity(helmet, translate((0, 2.5, -5)))], lights, camera, sky, irr, 1920, 1080)
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi, scenes

HR_ASSET_INVERT_NORMALS = 1
HR_ASSET_TEXTURE_SOURCE = 2


class AssetError(RuntimeError):
    pass


def _L():
    from . import hr
    return hr.load()


def _image(rc, im, err):
    if rc != 0:
        raise AssetError(err.value.decode(errors="replace") or "image decode failed (%d)" % rc)
    try:
        n = im.width * im.height * 4
        out = np.ctypeslib.as_array(im.rgba8, shape=(n,)).copy().reshape(im.height, im.width, 4)
    finally:
        _L().hr_image_free(C.byref(im))
    return out


def load_image(path):
    """RGBA8 (height, width, 4) uint8, as stbi_load(..., STBI_rgb_alpha)
    returns PNG / JPEG / Radiance HDR files to Texture::GET (vk_textures.cpp:15)."""
    im, err = abi.hr_image(), C.create_string_buffer(512)
    return _image(_L().hr_image_load(os.fsencode(path), C.byref(im), err, 512), im, err)


def decode_image(data):
    """as load_image, from the file's bytes"""
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    im, err = abi.hr_image(), C.create_string_buffer(512)
    return _image(_L().hr_image_decode(C.cast(buf, C.c_void_p), len(data), C.byref(im), err, 512), im, err)


@dataclass
class Asset:
    """One prefab: a shared vertex / index buffer, its primitives in instance
    order (model = node global matrix, column-major), materials (GPUMaterial
    rows of 12 floats, index = position in this asset) and RGBA8 textures."""
    name: str
    vertices: np.ndarray
    indices: np.ndarray
    prims: np.ndarray  # scenes.PRIM_DTYPE (mesh = 0)
    materials: np.ndarray
    textures: list = field(default_factory=list)

    def with_material(self, material):
        """Node::addMaterial on the first root (vk_mesh.cpp:533-544): the first
        primitive takes `material` (12 floats); returns a new prefab that
        shares the mesh."""
        m = np.asarray(material, dtype=np.float32).reshape(12).copy()
        mats = np.concatenate([self.materials, m[None]], 0)
        mats[-1, 11] = len(mats) - 1
        prims = self.prims.copy()
        if len(prims):
            prims["material"][0] = len(mats) - 1
        return Asset(self.name, self.vertices, self.indices, prims, mats, self.textures)


def load_asset(path, invert_normals=False, texture_source=False):
    """Prefab::GET(filename, invertNormals) for .obj / .gltf / .glb."""
    L = _L()
    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    flags = (HR_ASSET_INVERT_NORMALS if invert_normals else 0) | (HR_ASSET_TEXTURE_SOURCE if texture_source else 0)
    rc = L.hr_asset_load(os.fsencode(path), flags, C.byref(h), err, 1024)
    if rc != 0:
        raise AssetError(err.value.decode(errors="replace") or "asset load failed (%d)" % rc)
    try:
        v = abi.hr_asset_view()
        if L.hr_asset_view_get(h, C.byref(v)) != 0:
            raise AssetError("hr_asset_view_get failed")
        nv, ni = v.mesh.vertex_count, v.mesh.index_count
        verts = np.zeros((nv, 11), dtype=np.float32)
        if nv:
            C.memmove(verts.ctypes.data, v.mesh.vertices, nv * 44)
        idx = np.zeros(ni, dtype=np.uint32)
        if ni:
            C.memmove(idx.ctypes.data, v.mesh.indices, ni * 4)
        prims = np.zeros(v.primitive_count, dtype=scenes.PRIM_DTYPE)
        for i in range(v.primitive_count):
            p = v.primitives[i]
            prims[i] = (p.mesh, p.first_index, p.index_count, p.material, np.array(p.model[:], dtype=np.float32))
        mats = np.zeros((v.material_count, 12), dtype=np.float32)
        if v.material_count:
            C.memmove(mats.ctypes.data, v.materials, v.material_count * 48)
        texs = []
        for i in range(v.texture_count):
            t = v.textures[i]
            a = np.zeros((t.height, t.width, 4), dtype=np.uint8)
            C.memmove(a.ctypes.data, t.rgba8, a.nbytes)
            texs.append(a)
    finally:
        L.hr_asset_destroy(h)
    return Asset(os.path.basename(path), verts, idx, prims, mats, texs)


def _glm_mul(A, B):
    """float32 column-major 4x4 product as glm evaluates it (column i =
    ((A0 b0 + A1 b1) + A2 b2) + A3 b3); A, B: 16 floats, column-major."""
    A = np.asarray(A, dtype=np.float32).reshape(4, 4)  # [col][row]
    B = np.asarray(B, dtype=np.float32).reshape(4, 4)
    R = np.zeros((4, 4), dtype=np.float32)
    for i in range(4):
        s = A[0] * B[i][0]
        s = s + A[1] * B[i][1]
        s = s + A[2] * B[i][2]
        s = s + A[3] * B[i][3]
        R[i] = s
    return R.reshape(16)


def colmajor(m):
    """4x4 math matrix (m[row, col]) -> 16 float32, column-major"""
    return np.asarray(m, dtype=np.float32).T.reshape(16)


@dataclass
class Entity:
    """Object (entity.h): a prefab placed by m_matrix (4x4 math matrix)."""
    asset: Asset
    matrix: np.ndarray = field(default_factory=lambda: np.eye(4))


def compose_scene(entities, lights, camera, env_sky, env_irr, width, height, params=None, name="assets",
                  shapes=None):
    """A scenes.Scene from entities: every prefab's mesh once, its materials
    and textures appended with their indices offset, and one primitive per
    (entity, prefab primitive) with model = entity matrix x node global
    matrix, in entity order (Renderer's TLAS instance order)."""
    meshes, mesh_of, mat_base, tex_base = [], {}, {}, {}
    mats, texs, rows = [], [], []
    for e in entities:
        a = e.asset
        key = id(a)
        if key not in mesh_of:
            mesh_of[key] = len(meshes)
            meshes.append((a.vertices, a.indices))
            mat_base[key] = len(mats)
            tex_base[key] = len(texs)
            for m in a.materials:
                m = m.copy()
                for k in range(4, 8):
                    if m[k] > -1:
                        m[k] += tex_base[key]
                m[11] = len(mats)
                mats.append(m)
            texs.extend(a.textures)
        em = colmajor(e.matrix)
        for p in a.prims:
            model = _glm_mul(em, p["model"])
            rows.append((mesh_of[key], p["first_index"], p["index_count"], p["material"] + mat_base[key], model))
    prims = np.zeros(len(rows), dtype=scenes.PRIM_DTYPE)
    for i, (m, f, c, mat, model) in enumerate(rows):
        prims[i] = (m, f, c, mat, model)
    p = dict(shadow_samples=1, refl_spp=1, max_depth=8)
    if params:
        p.update(params)
    return scenes.Scene(name, meshes, prims, np.stack(mats).astype(np.float32),
                        np.asarray(lights, dtype=np.float32).reshape(-1, 12), shapes,
                        np.asarray(camera, dtype=np.float32), np.asarray(env_sky, dtype=np.float32),
                        np.asarray(env_irr, dtype=np.float32), width, height, p, textures=texs)


# ---------------------------------------------------------------- built-in meshes
def _prefab(name, V, idx):
    V = np.asarray(V, dtype=np.float32).reshape(-1, 11)
    idx = np.asarray(idx, dtype=np.uint32)
    prims = np.zeros(1, dtype=scenes.PRIM_DTYPE)
    prims[0] = (0, 0, len(idx), 0, np.eye(4, dtype=np.float32).reshape(16))
    mats = np.array([material()], dtype=np.float32)
    return Asset(name, V, idx, prims, mats, [])


def quad():
    """Mesh::get_quad (vk_mesh.cpp:161-184): unit quad in z = 0, normal +z."""
    v = [[1, 1, 0, 0, 0, 1, 1, 1, 1, 1, 1], [-1, 1, 0, 0, 0, 1, 1, 1, 1, 0, 1],
         [-1, -1, 0, 0, 0, 1, 1, 1, 1, 0, 0], [1, -1, 0, 0, 0, 1, 1, 1, 1, 1, 0]]
    return _prefab("quad", v, [0, 1, 2, 2, 3, 0])


def cube():
    """Mesh::get_cube (vk_mesh.cpp:208-263): 24 vertices, 6 faces, uv (1, 1)."""
    faces = [((0, 0, -1), [(1, 1, -1), (-1, 1, -1), (-1, -1, -1), (1, -1, -1)]),
             ((0, 0, 1), [(1, 1, 1), (-1, 1, 1), (-1, -1, 1), (1, -1, 1)]),
             ((0, 1, 0), [(1, 1, 1), (-1, 1, 1), (-1, 1, -1), (1, 1, -1)]),
             ((0, -1, 0), [(1, -1, 1), (-1, -1, 1), (-1, -1, -1), (1, -1, -1)]),
             ((-1, 0, 0), [(-1, 1, 1), (-1, -1, 1), (-1, -1, -1), (-1, 1, -1)]),
             ((1, 0, 0), [(1, 1, 1), (1, -1, 1), (1, -1, -1), (1, 1, -1)])]
    v = [list(p) + list(n) + [1, 1, 1, 1, 1] for n, ps in faces for p in ps]
    idx = []
    for f in range(6):
        b = 4 * f
        idx += [b, b + 1, b + 2, b + 2, b + 3, b]
    return _prefab("cube", v, idx)


def material(diffuse=(1.0, 1.0, 1.0), shading=0, metallic=1.0, roughness=1.0, ior=1.0):
    """Material defaults (material.h:17-31) as materialToShader packs them."""
    m = np.zeros(12, dtype=np.float32)
    m[0:3] = diffuse
    m[3] = ior
    m[4:8] = -1
    m[8:11] = (shading, metallic, roughness)
    return m


def translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


def scale(s):
    return np.diag([*np.broadcast_to(np.asarray(s, dtype=np.float64), (3,)), 1.0])


def rotate_x(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])


def reference_camera(position, width, height, fov=60.0, yaw=-90.0, pitch=0.0, frame=0.0):
    """Camera(position) of camera.cpp:3-11, 49-69 (yaw -90, pitch 0: looking
    down -z) as the RTCameraData scenes.rt_camera builds."""
    y, p = np.float32(np.radians(np.float32(yaw))), np.float32(np.radians(np.float32(pitch)))
    front = np.array([np.cos(y) * np.cos(p), np.sin(p), np.sin(y) * np.cos(p)], dtype=np.float32)
    front /= np.linalg.norm(front)
    pos = np.asarray(position, dtype=np.float64)
    return scenes.rt_camera(pos, pos + front, width, height, fov=fov, frame=frame)


def _env(asset_dir):
    """skybox (renderer.cpp:3073) and irradiance (renderer.cpp:3074) maps as
    the shaders sample them: RGBA8 UNORM -> c / 255."""
    sky = load_image(os.path.join(asset_dir, "LA_Downtown_Helipad_GoldenHour_8k.jpg"))
    irr = load_image(os.path.join(asset_dir, "LA_Downtown_Helipad_GoldenHour_Env.hdr"))
    return (sky.astype(np.float32) / np.float32(255)), (irr.astype(np.float32) / np.float32(255))


def reference_scene(which, asset_dir, width=1700, height=900, params=None):
    """Scene::default_scene (which = 0, scene.cpp:34-158) or
    Scene::cornell_scene (which = 1, scene.cpp:160-271) from the reference's
    asset files in asset_dir (sphere.obj, lucy.obj, DamagedHelmet.gltf and its
    textures, the LA_Downtown_Helipad_GoldenHour skybox / environment maps)."""
    f = lambda n: os.path.join(asset_dir, n)  # noqa: E731
    sphere = load_asset(f("sphere.obj"))
    helmet = load_asset(f("DamagedHelmet.gltf"))
    mirror = material(shading=3, metallic=1.0)
    glass = material((0.7, 0.7, 1.0), shading=4, metallic=1.0, ior=1.125)
    if which == 0:
        lights = [scenes.point_light((10, 12, -5), (1.0, 0.8, 0.5), 500.0, 500.0, 0.1),
                  scenes.point_light((-10, 15, -5), (0.5, 1.0, 1.0), 250.0, 500.0, 0.2),
                  scenes.point_light((0, 15, 5), (1, 1, 1), 500.0, 500.0, 0.09)]
        gold = material((1.0, 0.71, 0.29), metallic=0.5, roughness=0.1)
        floor_m = material(metallic=0.1)
        lucy = load_asset(f("lucy.obj")).with_material(gold)
        ents = [Entity(quad().with_material(floor_m), translate((0, 0, -5)) @ rotate_x(-90) @ scale(15)),
                Entity(sphere.with_material(mirror), translate((5, 1, -5))),
                Entity(sphere.with_material(glass), translate((0, 1, -5))),
                Entity(cube().with_material(mirror), translate((0, 4, -10)) @ scale((4, 4, 1))),
                Entity(lucy, scale(0.01)),
                Entity(helmet, translate((-5, 1, -5)))]
    elif which == 1:
        lights = [scenes.point_light((0, 8.5, -5), (1.0, 0.8, 0.5), 250.0, 500.0, 0.1)]
        white, red, green = material((1, 1, 1)), material((1, 0, 0)), material((0, 1, 0))
        wq, rq, gq = cube().with_material(white), cube().with_material(red), cube().with_material(green)
        ents = [Entity(wq, translate((0, 0, -5)) @ scale((5, 0.1, 5))),
                Entity(wq, translate((0, 5, -10)) @ scale((5, 5, 0.1))),
                Entity(rq, translate((-5, 5, -5)) @ scale((0.1, 5, 5))),
                Entity(gq, translate((5, 5, -5)) @ scale((0.1, 5, 5))),
                Entity(wq, translate((0, 10, -5)) @ scale((5, 0.1, 5))),
                Entity(sphere.with_material(glass), translate((-3.0, 2.2, -2.5)) @ scale(2)),
                Entity(sphere.with_material(mirror), translate((3.5, 1.0, -2.5))),
                Entity(helmet, translate((0, 2.5, -5)))]
    else:
        raise ValueError("scene 0 (default) or 1 (cornell)")
    sky, irr = _env(asset_dir)
    return compose_scene(ents, np.stack(lights), reference_camera((0, 5, 10), width, height), sky, irr,
                         width, height, params, name=("default_scene", "cornell_scene")[which])
