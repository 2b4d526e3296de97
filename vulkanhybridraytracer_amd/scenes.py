"""Seeded synthetic scenes for the hybrid pass.

The reference renders hand-made OBJ/glTF scenes (scene.cpp:34-272) whose
assets are not shipped, so the workloads here are procedural. Every scene is
returned in the reference's host-side data model:
  - meshes: (vertices float32[n, 11] = Vertex{position, normal, color, uv}
    (vk_mesh.h:21-27), indices uint32[m]) (Mesh::_vertices/_indices);
  - prims: one Primitive per drawable (vk_mesh.h:60-70) with the world matrix
    entity.m_matrix * node.global (vk_mesh.cpp:476), column-major;
  - materials: GPUMaterial rows (material.cpp:19-26);
  - lights: uboLight rows packed as vk_engine.cpp:146-164;
  - camera: RTCameraData (vk_engine.cpp:730-740) from glm-style lookAt and
    perspective (camera.cpp:49-58) with projection[1][1] *= -1 (vk_engine.cpp:695).

Workloads (sizes in the names):
  hybrid_640x360_20k   ground plane + 64 icospheres, 1 point + 1 directional light,
                       1 shadow ray per light, mirror bounces (depth 8)
  hybrid_1080p_1m      512x512-quad value-noise terrain + 256 triangle soups of 2048
                       triangles, 4 rectangular 4x4 area lights, 16 stratified shadow
                       samples per light, GGX glossy reflections at 4 spp, 1 bounce
  hybrid_1080p_4m      the same generator with a 1024x1024-quad terrain and 1024 soups,
                       roughness U[0,1] and adaptive sampling (1-64 shadow rays per
                       light, 1-64 reflection paths per pixel)
  hybrid_batch32       32 scenes of the terrain generator, 50k-2M triangles, 256x256
                       to 1920x1080, adaptive sampling up to 16

The sizes, camera, lights, material ranges and sampling of every workload are
read from configs/<workload>.json at the repository root (the terrain
generator's constants from configs/terrain_generator.json); the functions'
keyword arguments override them for tests.
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

CONFIG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")
_configs = {}


def config(name):
    """configs/<name>.json (cached; a workload name or terrain_generator)"""
    if name not in _configs:
        with open(os.path.join(CONFIG_DIR, name + ".json")) as f:
            _configs[name] = json.load(f)
    return _configs[name]


def _params(p):
    """the trace parameters of a config's "params" block"""
    return dict(shadow_samples=p["shadow_samples"], refl_spp=p["refl_spp"], max_depth=p["max_depth"],
                flags=abi.HR_FLAG_ADAPTIVE if p["adaptive"] else 0)

PRIM_DTYPE = np.dtype([("mesh", "<u4"), ("first_index", "<u4"), ("index_count", "<u4"),
                       ("material", "<i4"), ("model", "<f4", 16)])


@dataclass
class Scene:
    name: str
    meshes: list
    prims: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    shapes: object
    camera: np.ndarray
    env_sky: np.ndarray
    env_irr: np.ndarray
    width: int
    height: int
    params: dict = field(default_factory=dict)
    textures: list = field(default_factory=list)  # RGBA8 (height, width, 4) uint8 arrays

    @property
    def tri_count(self):
        return int(self.prims["index_count"].sum() // 3)

    @property
    def prim_tri_offsets(self):
        c = np.concatenate([[0], np.cumsum(self.prims["index_count"] // 3)])
        return c.astype(np.int64)


# ---------------------------------------------------------------- matrices
def _colmajor(a):
    """4x4 row/col math matrix (a[row, col]) -> glm column-major 16 floats."""
    return np.asarray(a, dtype=np.float64).T.reshape(16)


def look_at(eye, center, up=(0.0, 1.0, 0.0)):
    """glm::lookAt (right-handed), camera.cpp:49-52."""
    eye, center, up = (np.asarray(v, dtype=np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = -f
    m[0, 3] = -s @ eye
    m[1, 3] = -u @ eye
    m[2, 3] = f @ eye
    return m


def perspective(fovy_deg, aspect, near=0.1, far=1000.0):
    """glm::perspective (depth -1..1), camera.cpp:54-58."""
    t = np.tan(np.radians(fovy_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = -(far + near) / (far - near)
    m[3, 2] = -1.0
    m[2, 3] = -(2.0 * far * near) / (far - near)
    return m


def rt_camera(eye, center, width, height, fov=60.0, frame=0.0, near=0.1, far=1000.0):
    """RTCameraData as vk_engine.cpp:688-740: inverse(view), inverse(proj) with
    projection[1][1] *= -1, frame."""
    view = look_at(eye, center)
    proj = perspective(fov, width / height, near, far)
    proj[1, 1] *= -1.0
    cam = np.zeros(36, dtype=np.float32)
    cam[0:16] = _colmajor(np.linalg.inv(view))
    cam[16:32] = _colmajor(np.linalg.inv(proj))
    cam[32:36] = frame
    return cam


def translate_scale(t, s):
    m = np.eye(4)
    m[:3, :3] *= s
    m[:3, 3] = t
    return m


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------- materials / lights
def material(i, albedo, mode, metallic, roughness, ior=1.0):
    """GPUMaterial as Material::materialToShader (material.cpp:19-26)."""
    m = np.zeros(12, dtype=np.float32)
    m[0:3] = albedo
    m[3] = ior
    m[4:8] = -1.0
    m[8:12] = (mode, metallic, roughness, i)
    return m


def point_light(pos, color=(1, 1, 1), intensity=1000.0, max_distance=500.0, radius=1.0):
    """uboLight for a point light (vk_engine.cpp:157-161); defaults of entity.h:64."""
    lt = np.zeros(12, dtype=np.float32)
    lt[0:3] = pos
    lt[3] = max_distance
    lt[4:7] = color
    lt[7] = intensity
    lt[8] = radius
    return lt


def directional_light(direction, color=(1, 1, 1), intensity=1.0, radius=0.0):
    """uboLight for a directional light (vk_engine.cpp:152-155): pos.w = -1."""
    lt = np.zeros(12, dtype=np.float32)
    lt[0:3] = direction
    lt[3] = -1.0
    lt[4:7] = color
    lt[7] = intensity
    lt[8] = radius
    return lt


# ---------------------------------------------------------------- environment
def make_env(sky_w=512, sky_h=256, irr_w=64, irr_h=32):
    """Analytic sky as equirect float RGBA maps (row 0 = zenith, v = 0.5 - asin(y)/pi)."""
    def sky_color(w, h, blur):
        v = (np.arange(h) + 0.5) / h
        u = (np.arange(w) + 0.5) / w
        el = (0.5 - v) * np.pi                                  # elevation
        az = (u - 0.5) * 2 * np.pi
        el, az = np.meshgrid(el, az, indexing="ij")
        y = np.sin(el)
        horizon = np.array([0.80, 0.85, 0.90])
        zenith = np.array([0.25, 0.45, 0.85])
        ground = np.array([0.30, 0.27, 0.22])
        t = np.clip(y, 0, 1)[..., None]
        col = np.where(y[..., None] >= 0, horizon * (1 - t) + zenith * t,
                       ground + (horizon - ground) * np.clip(1 + y * 4, 0, 1)[..., None])
        if not blur:
            sd = np.array([np.cos(0.6) * np.sin(0.8), np.sin(0.6), np.cos(0.6) * np.cos(0.8)])
            d = np.stack([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)], -1)
            sun = np.clip((d @ sd - 0.995) / 0.005, 0, 1)[..., None]
            col = col * (1 - sun) + np.array([1.0, 0.95, 0.8]) * sun
        else:
            col = 0.5 * col + 0.25
        out = np.ones((h, w, 4), dtype=np.float32)
        out[..., :3] = col
        return out
    return sky_color(sky_w, sky_h, False), sky_color(irr_w, irr_h, True)


# ---------------------------------------------------------------- meshes
def icosphere(subdiv=2):
    """Unit icosphere: 20 * 4^subdiv triangles, normals = positions."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(verts)
    V = np.zeros((len(P), 11), dtype=np.float32)
    V[:, 0:3] = P
    V[:, 3:6] = P
    V[:, 6:9] = 1.0
    V[:, 9] = 0.5 + np.arctan2(P[:, 0], P[:, 2]) / (2 * np.pi)
    V[:, 10] = 0.5 - np.arcsin(np.clip(P[:, 1], -1, 1)) / np.pi
    return V, np.array(faces, dtype=np.uint32).reshape(-1)


def quad_mesh(size=100.0):
    h = size / 2
    V = np.zeros((4, 11), dtype=np.float32)
    V[:, 0:3] = [(-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h)]
    V[:, 3:6] = (0, 1, 0)
    V[:, 6:9] = 1
    V[:, 9:11] = [(0, 0), (1, 0), (1, 1), (0, 1)]
    return V, np.array([0, 2, 1, 0, 3, 2], dtype=np.uint32)


def value_noise(x, z, rng, octaves=6, base_cells=4):
    """Sum of octaves of smoothstep-interpolated lattice noise in [-1, 1]."""
    h = np.zeros(np.broadcast(x, z).shape)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = base_cells * 2 ** o
        lat = rng.uniform(-1, 1, size=(cells + 2, cells + 2))
        gx, gz = x * cells, z * cells
        ix, iz = np.floor(gx).astype(int), np.floor(gz).astype(int)
        fx, fz = gx - ix, gz - iz
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        ix, iz = np.clip(ix, 0, cells), np.clip(iz, 0, cells)
        a = lat[ix, iz] * (1 - sx) + lat[ix + 1, iz] * sx
        b = lat[ix, iz + 1] * (1 - sx) + lat[ix + 1, iz + 1] * sx
        h += amp * (a * (1 - sz) + b * sz)
        total += amp
        amp *= 0.5
    return h / total


def terrain_mesh(quads, extent, amplitude, rng, octaves=6, base_cells=4):
    n = quads + 1
    t = np.linspace(0.0, 1.0, n)
    X, Z = np.meshgrid(t, t, indexing="ij")
    Hh = amplitude * value_noise(X, Z, rng, octaves, base_cells)
    px = (X - 0.5) * extent
    pz = (Z - 0.5) * extent
    step = extent / quads
    gx = np.gradient(Hh, step, axis=0)
    gz = np.gradient(Hh, step, axis=1)
    nrm = np.stack([-gx, np.ones_like(gx), -gz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    V = np.zeros((n * n, 11), dtype=np.float32)
    V[:, 0] = px.reshape(-1)
    V[:, 1] = Hh.reshape(-1)
    V[:, 2] = pz.reshape(-1)
    V[:, 3:6] = nrm.reshape(-1, 3)
    V[:, 6:9] = 1.0
    V[:, 9] = X.reshape(-1)
    V[:, 10] = Z.reshape(-1)
    i, j = np.meshgrid(np.arange(quads), np.arange(quads), indexing="ij")
    a = (i * n + j).reshape(-1)
    b = ((i + 1) * n + j).reshape(-1)
    c = ((i + 1) * n + j + 1).reshape(-1)
    d = (i * n + j + 1).reshape(-1)
    idx = np.stack([a, c, b, a, d, c], -1).reshape(-1).astype(np.uint32)
    height_at = lambda x, z: float(Hh[int(round((x / extent + 0.5) * quads)),
                                     int(round((z / extent + 0.5) * quads))])
    return V, idx, height_at


def soup_mesh(ntri, rng, tri_size=0.12):
    """Random triangle soup in the unit cube, face normals, no shared vertices."""
    c = rng.uniform(-0.5, 0.5, size=(ntri, 1, 3))
    off = rng.normal(scale=tri_size, size=(ntri, 3, 3))
    P = (c + off).reshape(-1, 3)
    e1 = off[:, 1] - off[:, 0]
    e2 = off[:, 2] - off[:, 0]
    fn = np.cross(e1, e2)
    fn /= np.maximum(np.linalg.norm(fn, axis=-1, keepdims=True), 1e-12)
    V = np.zeros((ntri * 3, 11), dtype=np.float32)
    V[:, 0:3] = P
    V[:, 3:6] = np.repeat(fn, 3, axis=0)
    V[:, 6:9] = 1.0
    V[:, 9:11] = rng.uniform(0, 1, size=(ntri * 3, 2))
    return V, np.arange(ntri * 3, dtype=np.uint32)


def _prims(rows):
    p = np.zeros(len(rows), dtype=PRIM_DTYPE)
    for i, (mesh, first, count, mat, model) in enumerate(rows):
        p[i] = (mesh, first, count, mat, _colmajor(model).astype(np.float32))
    return p


# ---------------------------------------------------------------- workloads
def _light(d):
    """uboLight row of a config's light entry"""
    if d["type"] == "directional":
        return directional_light(d["direction"], intensity=d["intensity"], radius=d["radius"])
    return point_light(d["position"], intensity=d["intensity"], max_distance=d["max_distance"], radius=d["radius"])


def make_small(seed=None, width=None, height=None, spheres=None, frame=0.0):
    """hybrid_640x360_20k (configs/hybrid_640x360_20k.json): 2-triangle 100x100
    ground + 64 subdivision-2 icospheres (20,482 triangles), centres
    U([-40,40]x[1,10]x[-40,40]), radii U[0.5,3]; albedo U[0,1]^3, metallic 1
    (shading model 3) with p = 0.3 else 0 (model 0), roughness 0; one point
    light at (0,50,0) (radius 0: hard shadows) and one directional light; 1
    shadow ray per light, mirror chains up to depth 8."""
    cfg = config("hybrid_640x360_20k")
    seed = cfg["seed"] if seed is None else seed
    width = cfg["width"] if width is None else width
    height = cfg["height"] if height is None else height
    sp = cfg["spheres"]
    spheres = sp["count"] if spheres is None else spheres
    rng = np.random.default_rng(seed)
    ground = quad_mesh(cfg["ground"]["size"])
    sphere = icosphere(sp["subdivision"])
    mats, rows = [], []
    n_obj = 1 + spheres
    for i in range(n_obj):
        metallic = 1.0 if rng.uniform() < sp["p_metallic"] else 0.0
        albedo = rng.uniform(0, 1, size=3)
        mats.append(material(i, albedo, 3 if metallic else 0, metallic, 0.0))
    rows.append((0, 0, len(ground[1]), 0, np.eye(4)))
    for i in range(spheres):
        c = rng.uniform(sp["centre_min"], sp["centre_max"])
        r = rng.uniform(*sp["radius"])
        rows.append((1, 0, len(sphere[1]), 1 + i, translate_scale(c, r)))
    lights = np.stack([_light(d) for d in cfg["lights"]])
    sky, irr = make_env(*cfg["env"]["sky"], *cfg["env"]["irradiance"])
    cam = cfg["camera"]
    return Scene(cfg["workload"], [ground, sphere], _prims(rows), np.stack(mats), lights, None,
                 rt_camera(cam["eye"], cam["target"], width, height, fov=cam["fov_deg"], frame=frame,
                           near=cam["near"], far=cam["far"]), sky, irr, width, height,
                 {k: v for k, v in _params(cfg["params"]).items() if k != "flags"})


def make_terrain_scene(seed=0, width=1920, height=1080, quads=512, instances=256, soup_tris=None,
                       extent=200.0, samples=16, spp=4, max_depth=1, frame=0.0, name=None, rough_max=0.6,
                       adaptive=False):
    """The terrain generator (configs/terrain_generator.json) of
    hybrid_1080p_1m (quads=512, instances=256) / hybrid_1080p_4m (1024, 1024):
    value-noise terrain (6 octaves, amplitude 20) with a glossy material
    (shading model 3, metallic 1, roughness U[0,0.6]); triangle-soup instances
    with uniform positions, uniformly random rotations and scales U[1,5],
    materials glossy (model 3) with p = 0.5 else PBR (model 0), roughness
    U[0,0.6]; four 4x4 rectangular area lights at heights U[30,60]."""
    g = config("terrain_generator")
    gt, gi, gl, gc = g["terrain"], g["instances"], g["lights"], g["camera"]
    soup_tris = gi["triangles"] if soup_tris is None else soup_tris
    rng = np.random.default_rng(seed)
    tV, tI, height_at = terrain_mesh(quads, extent, gt["amplitude"], rng, gt["octaves"], gt["base_cells"])
    meshes = [(tV, tI)]
    mats = [material(0, rng.uniform(*gt["albedo"], size=3), gt["shading_model"], gt["metallic"],
                     rng.uniform(0.0, rough_max))]
    rows = [(0, 0, len(tI), 0, np.eye(4))]
    half = extent * gi["placement_frac"]
    for i in range(instances):
        meshes.append(soup_mesh(soup_tris, rng, gi["triangle_size"]))
        glossy = rng.uniform() < gi["p_glossy"]
        mats.append(material(1 + i, rng.uniform(0, 1, size=3), 3 if glossy else 0,
                             1.0 if glossy else 0.0, rng.uniform(0.0, rough_max)))
        s = rng.uniform(*gi["scale"])
        x, z = rng.uniform(-half, half, size=2)
        y = height_at(x, z) + 0.5 * s + gi["lift"]
        m = translate_scale((x, y, z), s)
        m[:3, :3] = random_rotation(rng) * s
        rows.append((1 + i, 0, soup_tris * 3, 1 + i, m))
    lights, shapes = [], []
    assert gl["type"] == "rect"
    for _ in range(gl["count"]):
        x, z = rng.uniform(-gl["xz_frac"] * extent, gl["xz_frac"] * extent, size=2)
        y = rng.uniform(*gl["height"])
        lights.append(point_light((x, y, z), color=rng.uniform(*gl["color"], size=3), intensity=gl["intensity"],
                                  max_distance=gl["max_distance"], radius=gl["radius"]))
        su, sv = gl["size"]
        shapes.append(np.array([1, su, 0, 0, 0, 0, sv, 0], dtype=np.float32))
    sky, irr = make_env(*g["env"]["sky"], *g["env"]["irradiance"])
    cam = rt_camera(tuple(f * extent for f in gc["eye_frac"]), tuple(f * extent for f in gc["target_frac"]),
                    width, height, fov=gc["fov_deg"], frame=frame, near=gc["near"], far=gc["far"])
    nm = name or ("hybrid_%dx%d_%dtri" % (width, height, quads * quads * 2 + instances * soup_tris))
    return Scene(nm, meshes, _prims(rows), np.stack(mats), np.stack(lights), np.stack(shapes), cam, sky,
                 irr, width, height, dict(shadow_samples=samples, refl_spp=spp, max_depth=max_depth,
                                          flags=abi.HR_FLAG_ADAPTIVE if adaptive else 0))


def _terrain_workload(name, seed, kw):
    cfg = config(name)
    p = cfg["params"]
    kw.setdefault("width", cfg["width"])
    kw.setdefault("height", cfg["height"])
    kw.setdefault("samples", p["shadow_samples"])
    kw.setdefault("spp", p["refl_spp"])
    kw.setdefault("max_depth", p["max_depth"])
    kw.setdefault("adaptive", p["adaptive"])
    kw.setdefault("rough_max", cfg["roughness"][1])
    return make_terrain_scene(cfg["seed"] if seed is None else seed, quads=cfg["terrain_quads"],
                              instances=cfg["instances"], extent=cfg["extent"], name=cfg["workload"], **kw)


def make_1080p_1m(seed=None, **kw):
    """hybrid_1080p_1m (configs/hybrid_1080p_1m.json)"""
    return _terrain_workload("hybrid_1080p_1m", seed, kw)


def make_1080p_4m(seed=None, **kw):
    """hybrid_1080p_4m (configs/hybrid_1080p_4m.json): the 4.2M-triangle scene
    with adaptive sampling: shadows 1 ray per hard light, else stages of 4,
    16, 32, 64 rays per light until the penumbra estimate settles;
    reflections 1 + floor(63 * roughness) paths, roughness U[0,1]."""
    return _terrain_workload("hybrid_1080p_4m", seed, kw)


BATCH_RESOLUTIONS = tuple(tuple(r) for r in config("hybrid_batch32")["resolutions"])


def make_batch(count=None, seed_base=None, tri_range=None, resolutions=BATCH_RESOLUTIONS, spp=None):
    """hybrid_batch32 (configs/hybrid_batch32.json): `count` scenes of the
    terrain generator with seeds seed_base.., triangle counts log-uniform in
    tri_range (half terrain, half 2048-triangle soups), resolutions drawn from
    `resolutions`, adaptive sampling up to `spp` (shadow penumbra samples and
    glossy paths), roughness U[0,1]."""
    cfg = config("hybrid_batch32")
    count = cfg["count"] if count is None else count
    seed_base = cfg["seeds"][0] if seed_base is None else seed_base
    tri_range = cfg["triangles_log_uniform"] if tri_range is None else tri_range
    spp = cfg["params"]["shadow_samples"] if spp is None else spp
    soup = config("terrain_generator")["instances"]["triangles"]
    out = []
    for k in range(count):
        rng = np.random.default_rng(cfg["rng_seed_base"] + seed_base + k)
        tris = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        quads = max(8, int(round(np.sqrt(tris / 4.0))))
        instances = max(1, int(round(tris / 2.0 / soup)))
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        extent = 200.0 * quads / 512.0 + 40.0
        out.append(make_terrain_scene(seed=seed_base + k, width=w, height=h, quads=quads, instances=instances,
                                      extent=extent, samples=spp, spp=spp, rough_max=cfg["roughness"][1],
                                      adaptive=cfg["params"]["adaptive"], name="batch%02d" % k))
    return out


# ---------------------------------------------------------------- textures
def texture_checker(n=64, cells=8, rng=None):
    """RGBA8 checkerboard of random cell colours, (n, n, 4) uint8."""
    rng = rng if rng is not None else np.random.default_rng(0)
    col = rng.integers(24, 256, size=(cells, cells, 3))
    idx = np.arange(n) * cells // n
    t = np.full((n, n, 4), 255, dtype=np.uint8)
    t[..., :3] = col[idx[:, None], idx[None, :]]
    return t


def texture_normal_map(n=64, bumps=6, strength=3.0, rng=None):
    """Tangent-space normal map of a sum of Gaussian bumps, encoded as the
    shader decodes it: c = n * 127 + 128 (geometry_shader.frag:46)."""
    rng = rng if rng is not None else np.random.default_rng(1)
    y, x = np.mgrid[0:n, 0:n] / n
    h = np.zeros((n, n))
    for _ in range(bumps):
        cx, cy, r, a = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.05, 0.2), rng.uniform(-1, 1)
        h += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * r * r))
    gy, gx = np.gradient(h, 1.0 / n)
    nv = np.stack([-gx * strength / n * 8, -gy * strength / n * 8, np.ones_like(h)], -1)
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    t = np.full((n, n, 4), 255, dtype=np.uint8)
    t[..., :3] = np.clip(np.rint(nv * 127 + 128), 0, 255)
    return t


def texture_emissive(n=32, rng=None):
    """Emissive stripes: dim coloured bands on black."""
    rng = rng if rng is not None else np.random.default_rng(2)
    t = np.zeros((n, n, 4), dtype=np.uint8)
    t[..., 3] = 255
    band = (np.arange(n) // 4) % 2 == 0
    t[band, :, :3] = rng.integers(40, 200, size=3)
    return t


def texture_metal_rough(n=32, rng=None):
    """Metallic-roughness map as glTF lays it out: g = roughness, b = metallic."""
    rng = rng if rng is not None else np.random.default_rng(3)
    cells = 4
    idx = np.arange(n) * cells // n
    rough = rng.integers(0, 256, size=(cells, cells))
    metal = rng.integers(0, 2, size=(cells, cells)) * 255
    t = np.zeros((n, n, 4), dtype=np.uint8)
    t[..., 1] = rough[idx[:, None], idx[None, :]]
    t[..., 2] = metal[idx[:, None], idx[None, :]]
    t[..., 3] = 255
    return t


def spherical_uv(V, shift=(0.0, 0.0), scale=1.0):
    """Copy of a unit-sphere mesh's vertex array with uv = (0.5 + atan2(z, x) /
    2pi, acos(y) / pi) * scale + shift."""
    V = V.copy()
    x, y, z = V[:, 0].astype(np.float64), V[:, 1].astype(np.float64), V[:, 2].astype(np.float64)
    V[:, 9] = (0.5 + np.arctan2(z, x) / (2 * np.pi)) * scale + shift[0]
    V[:, 10] = np.arccos(np.clip(y, -1, 1)) / np.pi * scale + shift[1]
    return V


def _tex_material(i, albedo, mode, metallic, roughness, tex, ior=1.0):
    m = material(i, albedo, mode, metallic, roughness, ior)
    m[4:8] = tex
    return m


def make_textured(seed=0, width=320, height=180, spheres=24, frame=0.0):
    """Textured variant of the small scene: a ground quad with an albedo
    checker (uv x6, so REPEAT wraps), a normal map, a metallic-roughness map
    and vertex colours; spheres with spherical uv (some shifted to negative
    uv) carrying albedo + normal maps, emissive stripes, glossy mirrors whose
    roughness comes from the metallic-roughness map, glass and plain
    materials. Textures: 0 checker 64^2, 1 normal map 64^2, 2 emissive 32^2,
    3 metallic-roughness 32^2, 4 checker 16^2."""
    rng = np.random.default_rng(seed)
    ground = quad_mesh(100.0)
    gV = ground[0].copy()
    gV[:, 9:11] *= 6.0
    gV[:, 6:9] = [(1.0, 0.8, 0.6), (0.6, 1.0, 0.8), (0.8, 0.6, 1.0), (1.0, 1.0, 1.0)]
    sph = icosphere(2)
    s1 = (spherical_uv(sph[0]), sph[1])
    s2 = (spherical_uv(sph[0], shift=(-1.37, -2.25), scale=2.0), sph[1])
    textures = [texture_checker(64, 8, rng), texture_normal_map(64, 6, 3.0, rng), texture_emissive(32, rng),
                texture_metal_rough(32, rng), texture_checker(16, 4, rng)]
    mats = [_tex_material(0, (1, 1, 1), 0, 0.0, 0.5, (0, 1, -1, 3))]
    rows = [(0, 0, len(ground[1]), 0, np.eye(4))]
    kinds = ["albedo_normal", "emissive", "glossy_tex", "glass", "plain", "albedo_neg"]
    for i in range(spheres):
        kind = kinds[i % len(kinds)]
        alb = rng.uniform(0.2, 1.0, size=3)
        if kind == "albedo_normal":
            m = _tex_material(1 + i, alb, 0, 0.0, 0.5, (4, 1, -1, -1))
        elif kind == "emissive":
            m = _tex_material(1 + i, alb, 0, 0.0, 0.5, (-1, -1, 2, -1))
        elif kind == "glossy_tex":
            m = _tex_material(1 + i, alb, 3, 1.0, 0.0, (0, -1, -1, 3))
        elif kind == "glass":
            m = material(1 + i, alb, 4, 0.0, 0.0, ior=1.5)
        elif kind == "plain":
            m = material(1 + i, alb, 0, float(rng.uniform() < 0.5), float(rng.uniform(0, 1)))
        else:
            m = _tex_material(1 + i, alb, 0, 0.0, 0.5, (0, 1, 2, 3))
        mats.append(m)
        c = rng.uniform([-40, 1, -40], [40, 10, 40])
        r = rng.uniform(1.0, 4.0)
        rows.append((2 if kind == "albedo_neg" else 1, 0, len(sph[1]), 1 + i, translate_scale(c, r)))
    lights = np.stack([point_light((0, 50, 10), intensity=1500.0, max_distance=500.0, radius=2.0),
                       directional_light((0.4, 1.0, 0.3), intensity=1.0)])
    sky, irr = make_env()
    return Scene("textured_%dx%d" % (width, height), [(gV, ground[1]), s1, s2], _prims(rows), np.stack(mats),
                 lights, None, rt_camera((0, 25, 70), (0, 0, 0), width, height, frame=frame), sky, irr, width,
                 height, dict(shadow_samples=2, refl_spp=2, max_depth=3), textures=textures)


def make_facing_quad(tex, width=64, height=64, size=8.0, uv_scale=3.0, eye_z=10.0, color=(1.0, 1.0, 1.0),
                     albedo=(1.0, 1.0, 1.0), mode=0, metallic=0.0, roughness=0.5):
    """One square of side `size` in the plane z = 0 facing +z, seen by a camera
    at (0, 0, eye_z) looking at the origin, with uv rotated against the screen
    axes: uv = ((y + size/2), (x + size/2)) / size * uv_scale. tex: the four
    material texture slots; textures are set by the caller. Black environment,
    no lights."""
    h = size / 2.0
    V = np.zeros((4, 11), dtype=np.float32)
    V[:, 0:3] = [(-h, -h, 0), (h, -h, 0), (h, h, 0), (-h, h, 0)]
    V[:, 3:6] = (0, 0, 1)
    V[:, 6:9] = color
    V[:, 9] = (V[:, 1] + h) / size * uv_scale
    V[:, 10] = (V[:, 0] + h) / size * uv_scale
    idx = np.array([0, 1, 2, 0, 2, 3], dtype=np.uint32)
    mats = np.stack([_tex_material(0, albedo, mode, metallic, roughness, tex)])
    sky = np.zeros((2, 4, 4), dtype=np.float32)
    irr = np.zeros((2, 4, 4), dtype=np.float32)
    return Scene("facing_quad", [(V, idx)], _prims([(0, 0, 6, 0, np.eye(4))]), mats, np.zeros((0, 12), np.float32),
                 None, rt_camera((0, 0, eye_z), (0, 0, 0), width, height), sky, irr, width, height,
                 dict(shadow_samples=1, refl_spp=1, max_depth=1))


WORKLOADS = {
    "hybrid_640x360_20k": make_small,
    "hybrid_1080p_1m": make_1080p_1m,
    "hybrid_1080p_4m": make_1080p_4m,
}
