"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle_hr.so).

Used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
as the checker of the GPU library; never by the library itself.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from vulkanhybridraytracer_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# HR_ORACLE_LIB selects another build of the same source (the sanitizer run's)
LIB_PATH = os.environ.get("HR_ORACLE_LIB") or os.path.join(HERE, "liboracle_hr.so")


# A one-GPU box gives a process a share of 16 host threads, whatever the
# machine has; the oracle's row workers never exceed that share.
HOST_THREAD_SHARE = 16


def default_threads():
    """Worker threads for the oracle: the CPUs this process may run on, at
    most HOST_THREAD_SHARE."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_THREAD_SHARE, n))


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (None when unreadable)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def build(force=False):
    """Compile the oracle with its Makefile (g++)."""
    cmd = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.check_call(cmd, stdout=subprocess.DEVNULL)


class or_counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("primary_rays", "shadow_rays", "reflection_rays",
                                           "hit_shadow_rays", "occluded_rays", "node_visits",
                                           "tri_tests")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        L.or_tea.restype = u32
        L.or_tea.argtypes = [u32, u32]
        L.or_lcg.restype = u32
        L.or_lcg.argtypes = [C.POINTER(u32)]
        L.or_rnd.restype = f32
        L.or_rnd.argtypes = [C.POINTER(u32)]
        L.or_pcg32_seq.argtypes = [C.c_uint64, C.c_uint64, u32, vp]
        L.or_u01_stream.restype = f32
        L.or_u01_stream.argtypes = [u32, u32, u32, u32]
        L.or_morton3.restype = u32
        L.or_morton3.argtypes = [u32, u32, u32]
        L.or_set_treelets.argtypes = [i32]
        L.or_atan2.argtypes = [vp, vp, u32, vp]
        L.or_asin.argtypes = [vp, u32, vp]
        L.or_sincos2pi.argtypes = [vp, u32, vp, vp]
        L.or_distribution_ggx.restype = f32
        L.or_distribution_ggx.argtypes = [vp, vp, f32]
        L.or_geometry_smith.restype = f32
        L.or_geometry_smith.argtypes = [vp, vp, vp, f32]
        L.or_fresnel_schlick.argtypes = [f32, vp, vp]
        L.or_sample_disk.argtypes = [C.POINTER(abi.hr_light), vp, vp, f32, f32, vp]
        L.or_sample_ggx.argtypes = [vp, f32, f32, f32, vp]
        L.or_sample_rect.argtypes = [vp, C.POINTER(abi.hr_light_shape), vp, u32, u32, f32, f32, vp]
        L.or_glass_raygen.argtypes = [vp, vp, f32, vp]
        L.or_glass_hit.argtypes = [vp, vp, f32, vp]
        L.or_set_plant.argtypes = [i32]
        L.or_nested_cell.argtypes = [u32, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
        L.or_adaptive_stop.restype = i32
        L.or_adaptive_stop.argtypes = [u32, u32, u32]
        L.or_adaptive_next.restype = u32
        L.or_adaptive_next.argtypes = [u32, u32]
        L.or_adaptive_shadow_ray.restype = i32
        L.or_adaptive_shadow_ray.argtypes = [vp, u32, u32, vp, vp, vp, vp, u32, vp, vp, u32, u32, u32, u32, vp]
        L.or_intersect_tri.restype = i32
        L.or_intersect_tri.argtypes = [vp, vp, f32, f32, vp, vp, vp, vp, vp, vp]
        L.or_scene_build.restype = i32
        L.or_scene_build.argtypes = [vp, u32, vp, u32, vp, u32, u32, C.POINTER(vp)]
        L.or_scene_refit.restype = i32
        L.or_scene_refit.argtypes = [vp, vp]
        L.or_scene_set_env.restype = i32
        L.or_scene_set_env.argtypes = [vp, C.POINTER(abi.hr_env)]
        L.or_scene_set_textures.restype = i32
        L.or_scene_set_textures.argtypes = [vp, C.POINTER(abi.hr_texture), u32]
        L.or_post_gamma.argtypes = [vp, C.c_uint64, vp, vp]
        L.or_scene_destroy.argtypes = [vp]
        for n in ("or_scene_tri_count", "or_scene_node_count", "or_scene_depth"):
            getattr(L, n).restype = u32
            getattr(L, n).argtypes = [vp]
        L.or_scene_export.restype = i32
        L.or_scene_export.argtypes = [vp, u32, vp, C.c_uint64]
        L.or_trace_rays.argtypes = [vp, vp, u32, i32, i32, vp, C.POINTER(or_counters), u32]
        L.or_gbuffer_primary.argtypes = [vp, C.POINTER(abi.hr_camera), u32, u32, vp, vp, vp, vp, vp,
                                         u32, u32, C.POINTER(or_counters), u32]
        L.or_trace_shadows.argtypes = [vp, u32, u32, vp, vp, vp, vp, u32, C.POINTER(abi.hr_camera),
                                       C.POINTER(abi.hr_trace_params), vp, u32, u32,
                                       C.POINTER(or_counters), u32]
        L.or_shade_hybrid.argtypes = [vp, u32, u32, vp, vp, vp, vp, vp, vp, vp, u32,
                                      C.POINTER(abi.hr_camera), C.POINTER(abi.hr_trace_params), vp, vp,
                                      vp, u32, u32, C.POINTER(or_counters), u32]
        L.or_render_rt.argtypes = [vp, vp, vp, u32, C.POINTER(abi.hr_camera), C.POINTER(abi.hr_trace_params), u32,
                                   u32, vp, u32, u32, C.POINTER(or_counters), u32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class OracleScene:
    """CPU LBVH scene built from a scenes.Scene (same inputs as hr_scene_build)."""

    def __init__(self, scene, leaf_max=2, threads=None):
        self.scene = scene
        self.threads = default_threads() if threads is None else threads
        self.packed = abi.PackedScene(scene)
        h = C.c_void_p()
        L = lib()
        rc = L.or_scene_build(C.cast(self.packed.meshes, C.c_void_p), self.packed.mesh_count,
                              C.cast(self.packed.prims, C.c_void_p), self.packed.prim_count,
                              C.cast(self.packed.materials, C.c_void_p), self.packed.material_count,
                              leaf_max, C.byref(h))
        if rc != 0:
            raise RuntimeError("or_scene_build failed: %d" % rc)
        self.h = h
        L.or_scene_set_env(h, C.byref(self.packed.env))
        if self.packed.texture_count:
            rc = L.or_scene_set_textures(h, self.packed.textures, self.packed.texture_count)
            if rc != 0:
                raise RuntimeError("or_scene_set_textures failed: %d" % rc)

    def close(self):
        if self.h:
            lib().or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tri_count(self):
        return lib().or_scene_tri_count(self.h)

    @property
    def node_count(self):
        return lib().or_scene_node_count(self.h)

    @property
    def depth(self):
        return lib().or_scene_depth(self.h)

    def refit(self, models):
        m = np.ascontiguousarray(models, dtype=np.float32)
        rc = lib().or_scene_refit(self.h, m.ctypes.data)
        assert rc == 0

    def export(self, which):
        n = self.tri_count
        size, dt = {0: (self.node_count, abi.NODE_DTYPE), 1: (n, abi.STRI_DTYPE), 2: (n, np.uint32),
                    3: (n, np.uint32), 4: (max(n - 1, 0) * 4, np.uint32)}[which]
        out = np.zeros(size, dtype=dt)
        rc = lib().or_scene_export(self.h, which, out.ctypes.data, out.nbytes)
        assert rc == 0
        return out

    def trace_rays(self, rays, any_hit=False, brute=False):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=abi.HIT_DTYPE)
        cnt = or_counters()
        lib().or_trace_rays(self.h, rays.ctypes.data, len(rays), int(any_hit), int(brute), hits.ctypes.data,
                            C.byref(cnt), self.threads)
        return hits, cnt.as_dict()

    def gbuffer(self, camera=None, rows=(0, 1)):
        sc = self.scene
        W, H = sc.width, sc.height
        planes = {k: np.zeros((H, W, 4), dtype=np.float32)
                  for k in ("position", "normal", "albedo", "material", "emissive")}
        cam = abi.pack_camera(sc.camera if camera is None else camera)
        cnt = or_counters()
        lib().or_gbuffer_primary(self.h, C.byref(cam), W, H, *[planes[k].ctypes.data for k in
                                 ("position", "normal", "albedo", "material", "emissive")],
                                 rows[0], rows[1], C.byref(cnt), self.threads)
        return planes, cnt.as_dict()

    def _params(self, params):
        sc = self.scene
        p = dict(sc.params)
        if params:
            p.update(params)
        return abi.hr_trace_params(p["shadow_samples"], p["refl_spp"], p["max_depth"], p.get("flags", 0) & ~1)

    def trace_shadows(self, gb, camera=None, params=None, rows=(0, 1), lights=None, shapes="scene"):
        sc = self.scene
        W, H = sc.width, sc.height
        lt = sc.lights if lights is None else lights
        sh = sc.shapes if isinstance(shapes, str) else shapes
        la, sa = abi.pack_lights(lt, sh)
        cam = abi.pack_camera(sc.camera if camera is None else camera)
        prm = self._params(params)
        out = np.zeros((len(lt), H, W), dtype=np.float32)
        cnt = or_counters()
        lib().or_trace_shadows(self.h, W, H, gb["position"].ctypes.data, gb["normal"].ctypes.data,
                               C.cast(la, C.c_void_p), None if sa is None else C.cast(sa, C.c_void_p),
                               len(lt), C.byref(cam), C.byref(prm), out.ctypes.data, rows[0], rows[1],
                               C.byref(cnt), self.threads)
        return out, cnt.as_dict()

    def shade_hybrid(self, gb, shadow, camera=None, params=None, rows=(0, 1), lights=None,
                     shapes="scene"):
        sc = self.scene
        W, H = sc.width, sc.height
        lt = sc.lights if lights is None else lights
        sh = sc.shapes if isinstance(shapes, str) else shapes
        la, sa = abi.pack_lights(lt, sh)
        cam = abi.pack_camera(sc.camera if camera is None else camera)
        prm = self._params(params)
        lit = np.zeros((H, W, 4), dtype=np.float32)
        refl = np.zeros((H, W, 4), dtype=np.float32)
        shadow = np.ascontiguousarray(shadow, dtype=np.float32)
        cnt = or_counters()
        lib().or_shade_hybrid(self.h, W, H, *[gb[k].ctypes.data for k in
                              ("position", "normal", "albedo", "material", "emissive")],
                              C.cast(la, C.c_void_p), None if sa is None else C.cast(sa, C.c_void_p),
                              len(lt), C.byref(cam), C.byref(prm), shadow.ctypes.data, lit.ctypes.data,
                              refl.ctypes.data, rows[0], rows[1], C.byref(cnt), self.threads)
        return lit, refl, cnt.as_dict()

    def adaptive_shadow_rays(self, gb, items, params=None):
        """kinds, rays of adaptive samples items = [(x, y, light, s), ...]:
        kind 0 = no ray, 1 = sampled light, 2 = hard light"""
        sc = self.scene
        la, sa = abi.pack_lights(sc.lights, sc.shapes)
        cam = abi.pack_camera(sc.camera)
        prm = self._params(params)
        rays = np.zeros(len(items), dtype=abi.RAY_DTYPE)
        kinds = np.zeros(len(items), dtype=np.int32)
        f = lib().or_adaptive_shadow_ray
        args = (self.h, sc.width, sc.height, gb["position"].ctypes.data, gb["normal"].ctypes.data,
                C.cast(la, C.c_void_p), None if sa is None else C.cast(sa, C.c_void_p), len(sc.lights),
                C.byref(cam), C.byref(prm))
        base, step = rays.ctypes.data, abi.RAY_DTYPE.itemsize
        for n, (x, y, li, s) in enumerate(items):
            kinds[n] = f(*args, int(x), int(y), int(li), int(s), base + n * step)
        return kinds, rays

    def render_rt(self, accum=None, camera=None, params=None, rows=(0, 1), lights=None, shapes="scene"):
        """pure ray tracing mode (raygen.rgen + closesthit.rchit) into accum
        (float32 H x W x 4, updated in place; a new array when None)"""
        sc = self.scene
        W, H = sc.width, sc.height
        lt = sc.lights if lights is None else lights
        sh = sc.shapes if isinstance(shapes, str) else shapes
        la, sa = abi.pack_lights(lt, sh)
        cam = abi.pack_camera(sc.camera if camera is None else camera)
        prm = self._params(params)
        if accum is None:
            accum = np.zeros((H, W, 4), dtype=np.float32)
        assert accum.dtype == np.float32 and accum.flags.c_contiguous and accum.shape == (H, W, 4)
        cnt = or_counters()
        lib().or_render_rt(self.h, C.cast(la, C.c_void_p), None if sa is None else C.cast(sa, C.c_void_p), len(lt),
                           C.byref(cam), C.byref(prm), W, H, accum.ctypes.data, rows[0], rows[1], C.byref(cnt),
                           self.threads)
        return accum, cnt.as_dict()


# ---- small scalar hooks ----------------------------------------------------
def tea(v0, v1):
    return lib().or_tea(v0, v1)


def lcg_seq(state, n):
    s = C.c_uint32(state)
    return [lib().or_lcg(C.byref(s)) for _ in range(n)]


def rnd_seq(state, n):
    s = C.c_uint32(state)
    return [lib().or_rnd(C.byref(s)) for _ in range(n)]


def pcg32(initstate, initseq, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().or_pcg32_seq(initstate, initseq, n, out.ctypes.data)
    return out


def morton3(x, y, z):
    return lib().or_morton3(x, y, z)


def atan2(y, x):
    y = np.ascontiguousarray(y, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(y)
    lib().or_atan2(y.ctypes.data, x.ctypes.data, y.size, out.ctypes.data)
    return out


def asin(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    lib().or_asin(x.ctypes.data, x.size, out.ctypes.data)
    return out


def sincos2pi(u):
    u = np.ascontiguousarray(u, dtype=np.float32)
    s, c = np.zeros_like(u), np.zeros_like(u)
    lib().or_sincos2pi(u.ctypes.data, u.size, s.ctypes.data, c.ctypes.data)
    return s, c


def _v(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def distribution_ggx(N, H, r):
    N, H = _v(N), _v(H)
    return lib().or_distribution_ggx(N.ctypes.data, H.ctypes.data, r)


def geometry_smith(N, V, L, r):
    N, V, L = _v(N), _v(V), _v(L)
    return lib().or_geometry_smith(N.ctypes.data, V.ctypes.data, L.ctypes.data, r)


def fresnel_schlick(ct, F0):
    F0 = _v(F0)
    out = np.zeros(3, dtype=np.float32)
    lib().or_fresnel_schlick(ct, F0.ctypes.data, out.ctypes.data)
    return out


def sample_disk(light_row, pos, L, r1, r2):
    la, _ = abi.pack_lights(np.asarray(light_row).reshape(1, 12))
    pos, L = _v(pos), _v(L)
    out = np.zeros(3, dtype=np.float32)
    lib().or_sample_disk(la, pos.ctypes.data, L.ctypes.data, r1, r2, out.ctypes.data)
    return out


def sample_ggx(N, roughness, r1, r2):
    N = _v(N)
    out = np.zeros(3, dtype=np.float32)
    lib().or_sample_ggx(N.ctypes.data, roughness, r1, r2, out.ctypes.data)
    return out


def sample_rect(centre, edge_u, edge_v, pos, s, S, r1, r2):
    sh = abi.hr_light_shape()
    sh.type = 1
    sh.edge_u[:] = [float(x) for x in edge_u]
    sh.edge_v[:] = [float(x) for x in edge_v]
    c, p = _v(centre), _v(pos)
    out = np.zeros(3, dtype=np.float32)
    lib().or_sample_rect(c.ctypes.data, C.byref(sh), p.ctypes.data, s, S, r1, r2, out.ctypes.data)
    return out


def glass_raygen(N, V, ior):
    """direction of hybridRaygen.rgen:131-144 (V: unit vector towards the eye)"""
    N, V = _v(N), _v(V)
    out = np.zeros(3, dtype=np.float32)
    lib().or_glass_raygen(N.ctypes.data, V.ctypes.data, ior, out.ctypes.data)
    return out


def glass_hit(N, rd, ior):
    """direction of hybridHit.rchit:158-173 (rd: the incoming ray direction)"""
    N, rd = _v(N), _v(rd)
    out = np.zeros(3, dtype=np.float32)
    lib().or_glass_hit(N.ctypes.data, rd.ctypes.data, ior, out.ctypes.data)
    return out


class treelets:
    """Context manager (tests only): scenes created inside the block build
    with (on=True, the default of the oracle and of the GPU library) or
    without the treelet restructuring of the Karras tree."""

    def __init__(self, on):
        self.on = on

    def __enter__(self):
        lib().or_set_treelets(1 if self.on else 0)
        return self

    def __exit__(self, *exc):
        lib().or_set_treelets(1)
        return False


class planted:
    """Context manager (tests only): the oracle with one deliberate mistake
    (PLANT_* in hr_oracle.cpp) for the duration of the block."""
    GLASS_ETA, GLASS_NORMAL, FRESNEL_NDOTH, GGX_ALPHA, RECT_STRATA, TIR_REFRACT = 1, 2, 3, 4, 5, 6
    ADAPTIVE_STOP, NESTED_ORDER, TREELET_COST = 7, 8, 9

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().or_set_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().or_set_plant(0)
        return False


def nested_cell(s, S):
    v = [C.c_uint32() for _ in range(4)]
    lib().or_nested_cell(s, S, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)  # (cx, nx, cy, ny)


def adaptive_stop(k, N, S):
    return bool(lib().or_adaptive_stop(k, N, S))


def adaptive_next(N, S):
    return lib().or_adaptive_next(N, S)


def intersect_tri(org, d, tmin, tmax, v0, v1, v2):
    org, d, v0, v1, v2 = (_v(a) for a in (org, d, v0, v1, v2))
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    ok = lib().or_intersect_tri(org.ctypes.data, d.ctypes.data, tmin, tmax, v0.ctypes.data, v1.ctypes.data,
                                v2.ctypes.data, C.byref(t), C.byref(u), C.byref(v))
    return (t.value, u.value, v.value) if ok else None


def post_gamma(img):
    """postFragment.frag over a float32 (..., 4) image: (float result, RGBA8)."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    n = a.size // 4
    out = np.empty_like(a)
    q = np.empty(a.shape, dtype=np.uint8)
    lib().or_post_gamma(a.ctypes.data, n, out.ctypes.data, q.ctypes.data)
    return out, q
