"""ctypes mirrors of the POD structs in include/hr/hr.h.

Each struct mirrors a reference layout byte for byte (sizes are asserted at
import):
  hr_vertex    Vertex          vk_mesh.h:21-27       44 B
  hr_material  GPUMaterial     material.h:5-10       48 B
  hr_light     uboLight        vk_engine.h:76-80     48 B
  hr_camera    RTCameraData    vk_engine.h:56-62     144 B
"""
import ctypes as C

import numpy as np

HR_OK, HR_E_INVALID, HR_E_OOM, HR_E_DEVICE, HR_E_UNSUPPORTED, HR_E_OVERFLOW = range(6)
STATUS_NAMES = {0: "HR_OK", 1: "HR_E_INVALID", 2: "HR_E_OOM", 3: "HR_E_DEVICE",
                4: "HR_E_UNSUPPORTED", 5: "HR_E_OVERFLOW"}
HR_MAX_LIGHTS = 16
HR_LIGHT_DISK, HR_LIGHT_RECT = 0, 1
HR_FLAG_COUNT_VISITS, HR_FLAG_TIMING, HR_FLAG_ADAPTIVE = 1, 2, 4
HR_FLAG_SHADOW_R8, HR_FLAG_LIT_BGRA8 = 8, 16  # the reference's R8_UNORM shadow / B8G8R8A8_UNORM lit images
HR_TRACE_ANY_HIT, HR_TRACE_COUNT_VISITS = 1, 2  # hr_trace_rays flags
HR_BVH_NO_TREELETS = 1  # hr_tuning.bvh_flags: the plain Karras tree


class hr_vertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_float * 3),
                ("color", C.c_float * 3), ("uv", C.c_float * 2)]


class hr_material(C.Structure):
    _fields_ = [("diffuse_ior", C.c_float * 4), ("textures", C.c_float * 4),
                ("shading_metal_rough_idx", C.c_float * 4)]


class hr_light(C.Structure):
    _fields_ = [("pos_maxdist", C.c_float * 4), ("color_intensity", C.c_float * 4),
                ("radius", C.c_float), ("_pad", C.c_float * 3)]


class hr_light_shape(C.Structure):
    _fields_ = [("type", C.c_uint32), ("edge_u", C.c_float * 3), ("edge_v", C.c_float * 3),
                ("_pad", C.c_uint32)]


class hr_camera(C.Structure):
    _fields_ = [("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16), ("frame", C.c_float * 4)]


class hr_mesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32),
                ("indices", C.c_void_p), ("index_count", C.c_uint32)]


class hr_primitive(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("first_index", C.c_uint32), ("index_count", C.c_uint32),
                ("material", C.c_int32), ("model", C.c_float * 16)]


class hr_env(C.Structure):
    _fields_ = [("skybox", C.c_void_p), ("sky_width", C.c_uint32), ("sky_height", C.c_uint32),
                ("irradiance", C.c_void_p), ("irr_width", C.c_uint32), ("irr_height", C.c_uint32)]


class hr_texture(C.Structure):
    _fields_ = [("rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class hr_image(C.Structure):
    _fields_ = [("rgba8", C.POINTER(C.c_uint8)), ("width", C.c_uint32), ("height", C.c_uint32)]


class hr_asset_view(C.Structure):
    _fields_ = [("mesh", hr_mesh), ("primitives", C.POINTER(hr_primitive)), ("primitive_count", C.c_uint32),
                ("materials", C.POINTER(hr_material)), ("material_count", C.c_uint32),
                ("textures", C.POINTER(hr_texture)), ("texture_count", C.c_uint32)]


class hr_gbuffer(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("position", C.c_void_p),
                ("normal", C.c_void_p), ("albedo", C.c_void_p), ("material", C.c_void_p),
                ("emissive", C.c_void_p)]


class hr_raster_gbuffer(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("position", C.c_void_p), ("normal", C.c_void_p),
                ("albedo", C.c_void_p), ("motion", C.c_void_p), ("material", C.c_void_p), ("emissive", C.c_void_p)]


class hr_dn_image(C.Structure):
    _fields_ = [("gbuffer", hr_gbuffer), ("noisy", C.c_void_p), ("target", C.c_void_p), ("out", C.c_void_p),
                ("eye", C.c_float * 3), ("_reserved", C.c_float)]


class hr_trace_params(C.Structure):
    _fields_ = [("shadow_samples", C.c_uint32), ("refl_spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("flags", C.c_uint32)]


class hr_ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class hr_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("tri", C.c_uint32)]


class hr_stats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("reflection_rays", C.c_uint64), ("hit_shadow_rays", C.c_uint64),
                ("occluded_rays", C.c_uint64), ("node_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("ms_gbuffer", C.c_float), ("ms_shadow", C.c_float),
                ("ms_shade", C.c_float), ("ms_build", C.c_float)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class hr_build_params(C.Structure):
    _fields_ = [("leaf_max", C.c_uint32), ("_reserved", C.c_uint32 * 3)]


class hr_frame_desc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("gbuffer", hr_gbuffer), ("lights", C.c_void_p), ("shapes", C.c_void_p),
                ("light_count", C.c_uint32), ("camera", hr_camera), ("params", hr_trace_params),
                ("shadow", C.c_void_p), ("lit", C.c_void_p), ("refl", C.c_void_p)]


class hr_tuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("tile_w", C.c_uint32), ("wave_budget_mb", C.c_uint32),
                ("bvh_flags", C.c_uint32)]


for _s, _n in ((hr_vertex, 44), (hr_material, 48), (hr_light, 48), (hr_light_shape, 32),
               (hr_camera, 144), (hr_primitive, 80), (hr_ray, 32), (hr_hit, 16)):
    assert C.sizeof(_s) == _n, (_s.__name__, C.sizeof(_s))

# numpy views with the same layouts
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<u4")])
NODE_DTYPE = np.dtype([("lo0", "<f4", 3), ("c0", "<u4"), ("hi0", "<f4", 3), ("c1", "<u4"),
                       ("lo1", "<f4", 3), ("pad0", "<u4"), ("hi1", "<f4", 3), ("pad1", "<u4")])
STRI_DTYPE = np.dtype([("v0", "<f4", 3), ("id", "<u4"), ("v1", "<f4", 3), ("pad0", "<u4"),
                       ("v2", "<f4", 3), ("pad1", "<u4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
assert NODE_DTYPE.itemsize == 64 and STRI_DTYPE.itemsize == 48
LEAF_BIT = 0x80000000


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class PackedScene:
    """ctypes arrays for hr_scene_build / or_scene_build built from a Scene.

    Holds references to the numpy buffers so the pointers stay valid."""

    def __init__(self, scene):
        self._keep = []
        self.meshes = (hr_mesh * max(1, len(scene.meshes)))()
        for i, (v, idx) in enumerate(scene.meshes):
            v = np.ascontiguousarray(v, dtype=np.float32)
            idx = np.ascontiguousarray(idx, dtype=np.uint32)
            assert v.ndim == 2 and v.shape[1] == 11
            self._keep += [v, idx]
            self.meshes[i].vertices = ptr(v)
            self.meshes[i].vertex_count = v.shape[0]
            self.meshes[i].indices = ptr(idx)
            self.meshes[i].index_count = idx.shape[0]
        self.mesh_count = len(scene.meshes)
        p = scene.prims
        self.prims = (hr_primitive * max(1, len(p)))()
        for i in range(len(p)):
            self.prims[i].mesh = int(p["mesh"][i])
            self.prims[i].first_index = int(p["first_index"][i])
            self.prims[i].index_count = int(p["index_count"][i])
            self.prims[i].material = int(p["material"][i])
            self.prims[i].model[:] = [float(x) for x in p["model"][i]]
        self.prim_count = len(p)
        mats = np.ascontiguousarray(scene.materials, dtype=np.float32).reshape(-1, 12)
        self._keep.append(mats)
        self.materials = C.cast(mats.ctypes.data, C.POINTER(hr_material))
        self.material_count = mats.shape[0]
        sky = np.ascontiguousarray(scene.env_sky, dtype=np.float32)
        irr = np.ascontiguousarray(scene.env_irr, dtype=np.float32)
        self._keep += [sky, irr]
        self.env = hr_env(ptr(sky), sky.shape[1], sky.shape[0], ptr(irr), irr.shape[1], irr.shape[0])
        texs = list(getattr(scene, "textures", None) or [])
        self.texture_count = len(texs)
        self.textures = (hr_texture * max(1, len(texs)))()
        for i, t in enumerate(texs):
            t = np.ascontiguousarray(t, dtype=np.uint8)
            assert t.ndim == 3 and t.shape[2] == 4, "textures are (height, width, 4) uint8"
            self._keep.append(t)
            self.textures[i] = hr_texture(ptr(t), t.shape[1], t.shape[0])


def pack_lights(lights, shapes=None):
    lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 12)
    la = (hr_light * len(lights))()
    C.memmove(la, lights.ctypes.data, lights.nbytes)
    sa = None
    if shapes is not None:
        shapes = np.ascontiguousarray(shapes, dtype=np.float32).reshape(-1, 8)
        sa = (hr_light_shape * len(shapes))()
        for i, row in enumerate(shapes):
            sa[i].type = int(row[0])
            sa[i].edge_u[:] = [float(x) for x in row[1:4]]
            sa[i].edge_v[:] = [float(x) for x in row[4:7]]
    return la, sa


def pack_camera(cam):
    cam = np.ascontiguousarray(cam, dtype=np.float32).reshape(36)
    c = hr_camera()
    C.memmove(C.addressof(c), cam.ctypes.data, 144)
    return c
