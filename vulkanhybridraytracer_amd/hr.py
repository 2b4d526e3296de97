"""Python host binding of libhr.so (the C ABI of include/hr/hr.h).

The classes mirror the reference's host objects for the hybrid pass:
  Context         the device + queue the reference gets from vk-bootstrap
  HybridScene     Scene + its acceleration structure (create_bottom/top_
                  acceleration_structure, renderer.cpp:1643-1688), refit as
                  buildTlas(update=true) (vk_engine.cpp:174-187)
  GBuffer         the six offscreen attachments of renderer.cpp:252-269 (FP32)
  HybridRenderer  Renderer::rasterize_hybrid (renderer.cpp:530-600): G-buffer,
                  shadow trace, hybrid trace, each on the GPU

The library must be built (python -m vulkanhybridraytracer_amd.build); there
is no fallback: every call goes to the HIP kernels.
"""
import atexit
import ctypes as C
import os
import re
import weakref

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhr.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "hr", "hr.h")


class HRError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, status), msg))
        self.status = status


def exported_functions():
    """Names of the functions declared in include/hr/hr.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hr_status|void|uint32_t|uint64_t|const char\*)\s+(hr_\w+)\s*\(", text, re.M)))


_lib = None
_contexts = weakref.WeakSet()


@atexit.register
def _close_contexts():
    """Release every live context (children first) before the interpreter
    tears down modules, so no library object is destroyed in an arbitrary
    order during finalisation."""
    for c in list(_contexts):
        try:
            c.close()
        except Exception:
            pass


def load(path=LIB_PATH):
    """Load libhr.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libhr.so is not built: run python -m vulkanhybridraytracer_amd.build")
    L = C.CDLL(path)
    vp, u32, i32, sz, u64 = C.c_void_p, C.c_uint32, C.c_int, C.c_size_t, C.c_uint64
    st = C.c_int
    sigs = {
        "hr_abi_version": (u32, []),
        "hr_context_create": (st, [i32, C.POINTER(vp)]),
        "hr_context_destroy": (None, [vp]),
        "hr_last_error": (C.c_char_p, [vp]),
        "hr_synchronize": (st, [vp]),
        "hr_device_alloc": (st, [vp, sz, C.POINTER(vp)]),
        "hr_device_free": (st, [vp, vp]),
        "hr_memcpy_h2d": (st, [vp, vp, vp, sz]),
        "hr_memcpy_d2h": (st, [vp, vp, vp, sz]),
        "hr_memset_d": (st, [vp, vp, i32, sz]),
        "hr_scene_build": (st, [vp, vp, u32, vp, u32, vp, u32, C.POINTER(abi.hr_build_params), C.POINTER(vp)]),
        "hr_scene_refit": (st, [vp, vp]),
        "hr_scene_set_materials": (st, [vp, vp, u32]),
        "hr_scene_set_env": (st, [vp, C.POINTER(abi.hr_env)]),
        "hr_scene_set_textures": (st, [vp, C.POINTER(abi.hr_texture), u32]),
        "hr_post_gamma": (st, [vp, u32, u32, vp, vp, vp]),
        "hr_image_load": (st, [C.c_char_p, C.POINTER(abi.hr_image), C.c_char_p, sz]),
        "hr_image_decode": (st, [vp, sz, C.POINTER(abi.hr_image), C.c_char_p, sz]),
        "hr_image_free": (None, [C.POINTER(abi.hr_image)]),
        "hr_asset_load": (st, [C.c_char_p, u32, C.POINTER(vp), C.c_char_p, sz]),
        "hr_asset_view_get": (st, [vp, C.POINTER(abi.hr_asset_view)]),
        "hr_asset_destroy": (None, [vp]),
        "hr_scene_destroy": (None, [vp]),
        "hr_scene_info": (st, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_float)]),
        "hr_scene_memory": (st, [vp, C.POINTER(u32), C.POINTER(u64)]),
        "hr_scene_export": (st, [vp, u32, vp, sz]),
        "hr_gbuffer_primary": (st, [vp, C.POINTER(abi.hr_camera), C.POINTER(abi.hr_gbuffer),
                                    C.POINTER(abi.hr_stats)]),
        "hr_trace_shadows": (st, [vp, C.POINTER(abi.hr_gbuffer), vp, vp, u32, C.POINTER(abi.hr_camera),
                                  C.POINTER(abi.hr_trace_params), vp, C.POINTER(abi.hr_stats)]),
        "hr_trace_shadows_noalloc": (st, [vp, C.POINTER(abi.hr_gbuffer), vp, vp, u32, C.POINTER(abi.hr_camera),
                                          C.POINTER(abi.hr_trace_params), vp, C.POINTER(abi.hr_stats)]),
        "hr_shade_hybrid": (st, [vp, C.POINTER(abi.hr_gbuffer), vp, vp, u32, C.POINTER(abi.hr_camera),
                                 C.POINTER(abi.hr_trace_params), vp, vp, vp, C.POINTER(abi.hr_stats)]),
        "hr_trace_rays": (st, [vp, vp, u32, i32, vp, C.POINTER(abi.hr_stats)]),
        "hr_trace_shadows_batch": (st, [vp, vp, u32, C.POINTER(abi.hr_stats)]),
        "hr_shade_hybrid_batch": (st, [vp, vp, u32, C.POINTER(abi.hr_stats)]),
        "hr_graph_capture": (st, [vp, vp, u32, C.POINTER(vp)]),
        "hr_graph_launch": (st, [vp, C.POINTER(abi.hr_stats)]),
        "hr_graph_destroy": (None, [vp]),
        "hr_context_set_tuning": (st, [vp, C.POINTER(abi.hr_tuning)]),
        "hr_context_get_tuning": (st, [vp, C.POINTER(abi.hr_tuning)]),
        "hr_render_rt": (st, [vp, vp, vp, u32, C.POINTER(abi.hr_camera), C.POINTER(abi.hr_trace_params), u32, u32,
                              vp, C.POINTER(abi.hr_stats)]),
        "hr_gbuffer_ingest": (st, [vp, C.POINTER(abi.hr_raster_gbuffer), C.POINTER(abi.hr_gbuffer), vp]),
        "hr_gbuffer_quantize": (st, [vp, C.POINTER(abi.hr_gbuffer), vp, C.POINTER(abi.hr_raster_gbuffer)]),
        "hr_filter_shadows": (st, [vp, u32, u32, u32, C.c_int32, vp, vp, vp]),
        "hr_motion_vectors": (st, [vp, C.POINTER(abi.hr_gbuffer), vp, vp, vp]),
        "hr_denoiser_param_count": (u64, []),
        "hr_denoiser_layer": (st, [u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
        "hr_denoiser_create": (st, [vp, vp, u64, u64, C.POINTER(vp)]),
        "hr_denoiser_destroy": (None, [vp]),
        "hr_denoiser_get_weights": (st, [vp, vp, u64]),
        "hr_denoiser_set_weights": (st, [vp, vp, u64]),
        "hr_denoiser_get_state": (st, [vp, vp, vp, vp, u64, C.POINTER(u64)]),
        "hr_denoiser_set_state": (st, [vp, vp, vp, vp, u64, u64]),
        "hr_denoiser_set_loss": (st, [vp, u32]),
        "hr_denoise": (st, [vp, C.POINTER(abi.hr_dn_image), C.POINTER(C.c_float)]),
        "hr_denoiser_train_step": (st, [vp, vp, u32, C.c_float, C.POINTER(C.c_float)]),
        "hr_denoiser_gradients": (st, [vp, vp, u32, C.POINTER(C.c_float), vp, u64]),
        "hr_dn_conv3x3_forward": (st, [vp, vp, u32, u32, u32, u32, vp, vp, u32, i32, vp]),
        "hr_dn_conv3x3_backward": (st, [vp, vp, vp, u32, u32, u32, u32, vp, u32, vp, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class Context:
    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.hr_context_create(device, C.byref(h))
        if rc != 0:
            raise HRError(rc, "hr_context_create(%d) failed (no HIP device?)" % device)
        self.h = h
        self._children = weakref.WeakSet()  # scenes, graphs, denoisers: closed before the context
        _contexts.add(self)

    def own(self, obj):
        self._children.add(obj)

    def check(self, rc):
        if rc != 0:
            raise HRError(rc, self.L.hr_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            for c in list(self._children):
                c.close()
            self.L.hr_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.hr_synchronize(self.h))

    def last_error(self):
        """hr_last_error: the message of the context's last failure (bytes)"""
        return self.L.hr_last_error(self.h)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def post_gamma(self, width, height, src, out=None, rgba8=None):
        """the display pass (postFragment.frag) of a float4[H][W] DeviceBuffer
        into out (float4, may be src) and/or rgba8 (uint8[H][W][4])"""
        self.check(self.L.hr_post_gamma(self.h, width, height, src.ptr, None if out is None else out.ptr,
                                        None if rgba8 is None else rgba8.ptr))

    def _tuning(self):
        t = abi.hr_tuning()
        self.check(self.L.hr_context_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, blocks_per_cu=8, tile_w=8, wave_budget_mb=40960):
        """launch shapes (the hierarchy flags are kept)"""
        t = abi.hr_tuning(blocks_per_cu, tile_w, wave_budget_mb, self._tuning().bvh_flags)
        self.check(self.L.hr_context_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = self._tuning()
        return dict(blocks_per_cu=t.blocks_per_cu, tile_w=t.tile_w, wave_budget_mb=t.wave_budget_mb)

    def set_bvh_flags(self, flags):
        """HR_BVH_* flags of the scenes created afterwards"""
        t = self._tuning()
        t.bvh_flags = flags
        self.check(self.L.hr_context_set_tuning(self.h, C.byref(t)))

    def bvh_flags(self):
        return self._tuning().bvh_flags


class DeviceBuffer:
    """A device allocation made through hr_device_alloc."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(ctx.L.hr_device_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx.check(self.ctx.L.hr_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(self.ctx.L.hr_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        self.ctx.check(self.ctx.L.hr_memset_d(self.ctx.h, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            self.ctx.check(self.ctx.L.hr_device_free(self.ctx.h, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.L.hr_device_free(self.ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


class HybridScene:
    """A scene on the device: world-space triangles and the GPU LBVH."""

    def __init__(self, ctx, scene, leaf_max=2):
        self.ctx = ctx
        self.scene = scene
        self.packed = abi.PackedScene(scene)
        bp = abi.hr_build_params(leaf_max)
        h = C.c_void_p()
        ctx.check(ctx.L.hr_scene_build(ctx.h, C.cast(self.packed.meshes, C.c_void_p), self.packed.mesh_count,
                                       C.cast(self.packed.prims, C.c_void_p), self.packed.prim_count,
                                       C.cast(self.packed.materials, C.c_void_p), self.packed.material_count,
                                       C.byref(bp), C.byref(h)))
        self.h = h
        ctx.own(self)
        ctx.check(ctx.L.hr_scene_set_env(h, C.byref(self.packed.env)))
        if self.packed.texture_count:
            ctx.check(ctx.L.hr_scene_set_textures(h, self.packed.textures, self.packed.texture_count))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.hr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        t, n, d, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        self.ctx.check(self.ctx.L.hr_scene_info(self.h, C.byref(t), C.byref(n), C.byref(d), C.byref(ms)))
        return dict(tri_count=t.value, node_count=n.value, depth=d.value, ms_build=ms.value)

    def memory(self):
        """4-wide node count and the scene's device bytes"""
        w, b = C.c_uint32(), C.c_uint64()
        self.ctx.check(self.ctx.L.hr_scene_memory(self.h, C.byref(w), C.byref(b)))
        return dict(wide_count=w.value, device_bytes=b.value)

    def export(self, which):
        info = self.info()
        n = info["tri_count"]
        size, dt = {0: (info["node_count"], abi.NODE_DTYPE), 1: (n, abi.STRI_DTYPE), 2: (n, np.uint32),
                    3: (n, np.uint32), 4: (max(n - 1, 0) * 4, np.uint32),
                    5: (self.memory()["wide_count"] * 32, np.uint32)}[which]
        out = np.zeros(size, dtype=dt)
        self.ctx.check(self.ctx.L.hr_scene_export(self.h, which, out.ctypes.data, out.nbytes))
        return out

    def refit(self, models):
        m = np.ascontiguousarray(models, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_scene_refit(self.h, m.ctypes.data))

    def set_textures(self, textures):
        """material textures: a list of (height, width, 4) uint8 RGBA8 arrays"""
        ts = [np.ascontiguousarray(t, dtype=np.uint8) for t in textures]
        arr = (abi.hr_texture * max(1, len(ts)))()
        for i, t in enumerate(ts):
            arr[i] = abi.hr_texture(t.ctypes.data, t.shape[1], t.shape[0])
        self.ctx.check(self.ctx.L.hr_scene_set_textures(self.h, arr, len(ts)))

    def set_env(self, sky, irradiance):
        """environment maps: (height, width, 4) float32 equirect skybox and
        irradiance (hr_scene_set_env)"""
        sky = np.ascontiguousarray(sky, dtype=np.float32)
        irr = np.ascontiguousarray(irradiance, dtype=np.float32)
        env = abi.hr_env(sky.ctypes.data, sky.shape[1], sky.shape[0], irr.ctypes.data, irr.shape[1], irr.shape[0])
        self.ctx.check(self.ctx.L.hr_scene_set_env(self.h, C.byref(env)))

    def set_materials(self, materials):
        m = np.ascontiguousarray(materials, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_scene_set_materials(self.h, m.ctypes.data, m.shape[0]))

    def trace_rays(self, rays, any_hit=False, stats=False, count=None):
        """Bare ray queries (hr_trace_rays). count (default: stats) walks the
        binary nodes and counts node visits and triangle tests."""
        count = stats if count is None else count
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        n = len(rays)
        rb = self.ctx.alloc(max(rays.nbytes, 32))
        hb = self.ctx.alloc(max(n * 16, 16))
        rb.upload(rays)
        st = abi.hr_stats() if stats else None
        flags = (abi.HR_TRACE_ANY_HIT if any_hit else 0) | (abi.HR_TRACE_COUNT_VISITS if count else 0)
        self.ctx.check(self.ctx.L.hr_trace_rays(self.h, rb.ptr, n, flags, hb.ptr,
                                                C.byref(st) if st is not None else None))
        hits = hb.download(n, abi.HIT_DTYPE)
        return (hits, st.as_dict()) if stats else hits


PLANES = ("position", "normal", "albedo", "material", "emissive")


class GBuffer:
    """FP32 G-buffer planes (float4 per pixel each) in device memory."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        self.bufs = {k: ctx.alloc(width * height * 16) for k in PLANES}
        self.c = abi.hr_gbuffer(width, height, *[self.bufs[k].ptr for k in PLANES])

    def download(self):
        return {k: self.bufs[k].download((self.height, self.width, 4)) for k in PLANES}

    def upload(self, planes):
        for k in PLANES:
            self.bufs[k].upload(np.ascontiguousarray(planes[k], dtype=np.float32))

    def ingest(self, raster, motion_out=None):
        """decode a RasterGBuffer (the reference's attachment formats) into
        these FP32 planes; motion_out: DeviceBuffer of float2[H][W] or None"""
        self.ctx.check(self.ctx.L.hr_gbuffer_ingest(self.ctx.h, C.byref(raster.c), C.byref(self.c),
                                                    None if motion_out is None else motion_out.ptr))

    def quantize(self, raster, motion_in=None):
        """encode these planes into a RasterGBuffer as the raster pass writes them"""
        self.ctx.check(self.ctx.L.hr_gbuffer_quantize(self.ctx.h, C.byref(self.c),
                                                      None if motion_in is None else motion_in.ptr,
                                                      C.byref(raster.c)))


# attachment formats of renderer.cpp:252-261: (numpy dtype, components)
RASTER_PLANES = {"position": (np.float16, 4), "normal": (np.float16, 4), "albedo": (np.uint8, 4),
                 "motion": (np.float16, 2), "material": (np.uint8, 4), "emissive": (np.uint8, 4)}


class RasterGBuffer:
    """The reference's raster G-buffer attachments (RGBA16F position and
    normal, RGBA8 albedo / material / emissive, RG16F motion) in device memory."""

    def __init__(self, ctx, width, height, with_motion=True):
        self.ctx, self.width, self.height = ctx, width, height
        self.bufs = {k: ctx.alloc(width * height * np.dtype(t).itemsize * n) for k, (t, n) in RASTER_PLANES.items()
                     if with_motion or k != "motion"}
        self.c = abi.hr_raster_gbuffer(width, height, *[self.bufs[k].ptr if k in self.bufs else None
                                                        for k in RASTER_PLANES])

    def download(self):
        return {k: self.bufs[k].download((self.height, self.width, RASTER_PLANES[k][1]), RASTER_PLANES[k][0])
                for k in self.bufs}

    def upload(self, planes):
        for k in self.bufs:
            t, n = RASTER_PLANES[k]
            self.bufs[k].upload(np.ascontiguousarray(planes[k], dtype=t))


class HybridRenderer:
    """The hybrid pass of Renderer::rasterize_hybrid (renderer.cpp:530-600):
    G-buffer (here: primary visibility), shadow trace (shadowRaygen), hybrid
    trace (hybridRaygen + hybridHit). Buffers stay in device memory."""

    def __init__(self, ctx, hscene, width=None, height=None):
        self.ctx = ctx
        self.hs = hscene
        sc = hscene.scene
        self.width = width or sc.width
        self.height = height or sc.height
        self.gb = GBuffer(ctx, self.width, self.height)
        nl = len(sc.lights)
        self.shadow = ctx.alloc(nl * self.width * self.height * 4)
        self.lit = ctx.alloc(self.width * self.height * 16)
        self.refl = ctx.alloc(self.width * self.height * 16)
        self.set_frame(sc.camera, sc.lights, sc.shapes, sc.params)

    def set_frame(self, camera, lights, shapes, params):
        self.cam = abi.pack_camera(camera)
        self.lights, self.shapes = abi.pack_lights(lights, shapes)
        self.nl = len(lights)
        self.params = abi.hr_trace_params(params["shadow_samples"], params["refl_spp"], params["max_depth"],
                                          params.get("flags", 0))

    def _shapes_ptr(self):
        return None if self.shapes is None else C.cast(self.shapes, C.c_void_p)

    def gbuffer_primary(self, stats=None):
        L = self.ctx.L
        self.ctx.check(L.hr_gbuffer_primary(self.hs.h, C.byref(self.cam), C.byref(self.gb.c),
                                            C.byref(stats) if stats is not None else None))

    def trace_shadows(self, stats=None):
        L = self.ctx.L
        self.ctx.check(L.hr_trace_shadows(self.hs.h, C.byref(self.gb.c), C.cast(self.lights, C.c_void_p),
                                          self._shapes_ptr(), self.nl, C.byref(self.cam), C.byref(self.params),
                                          self.shadow.ptr, C.byref(stats) if stats is not None else None))

    def shade_hybrid(self, stats=None):
        L = self.ctx.L
        self.ctx.check(L.hr_shade_hybrid(self.hs.h, C.byref(self.gb.c), C.cast(self.lights, C.c_void_p),
                                         self._shapes_ptr(), self.nl, C.byref(self.cam), C.byref(self.params),
                                         self.shadow.ptr, self.lit.ptr, self.refl.ptr,
                                         C.byref(stats) if stats is not None else None))

    def rasterize_hybrid(self, stats=False):
        """G-buffer, shadow trace, hybrid trace; with stats, returns the
        summed ray counts and the per-pass milliseconds."""
        if not stats:
            self.gbuffer_primary()
            self.trace_shadows()
            self.shade_hybrid()
            return None
        parts = [abi.hr_stats() for _ in range(3)]
        self.gbuffer_primary(parts[0])
        self.trace_shadows(parts[1])
        self.shade_hybrid(parts[2])
        out = {k: sum(int(getattr(p, k)) for p in parts) for k in
               ("primary_rays", "shadow_rays", "reflection_rays", "hit_shadow_rays", "occluded_rays",
                "node_visits", "tri_tests")}
        out.update(ms_gbuffer=parts[0].ms_gbuffer, ms_shadow=parts[1].ms_shadow, ms_shade=parts[2].ms_shade)
        return out

    def download_shadow(self):
        return self.shadow.download((self.nl, self.height, self.width))

    def download_lit(self):
        return self.lit.download((self.height, self.width, 4))

    def download_refl(self):
        return self.refl.download((self.height, self.width, 4))

    # pure ray tracing mode (Renderer::raytrace, renderer.cpp:465-528)
    def render_rt(self, accum, stats=None):
        """raygen.rgen + closesthit.rchit into accum (DeviceBuffer of
        float4[H][W]) with this renderer's camera, lights and parameters:
        refl_spp camera samples, shadow_samples per light per hit, max_depth
        rays per chain; camera frame < 0 stores, else accumulates"""
        L = self.ctx.L
        self.ctx.check(L.hr_render_rt(self.hs.h, C.cast(self.lights, C.c_void_p), self._shapes_ptr(), self.nl,
                                      C.byref(self.cam), C.byref(self.params), self.width, self.height, accum.ptr,
                                      C.byref(stats) if stats is not None else None))

    # temporal shadow filter (blur.comp) between the shadow and hybrid passes
    def filter_shadows(self, history, frame, motion=None):
        """history: DeviceBuffer of float[nl][H][W] (previous output, replaced);
        motion: DeviceBuffer of float2[H][W], needed when frame <= 0."""
        self.ctx.check(self.ctx.L.hr_filter_shadows(self.ctx.h, self.width, self.height, self.nl, int(frame),
                                                    self.shadow.ptr, None if motion is None else motion.ptr,
                                                    history.ptr))

    def motion_vectors(self, view_proj, prev_view_proj, motion):
        """motion vectors of the current G-buffer into a float2[H][W] buffer;
        view_proj / prev_view_proj: 16 floats, column-major projection * view"""
        vp = np.ascontiguousarray(view_proj, dtype=np.float32)
        pv = np.ascontiguousarray(prev_view_proj, dtype=np.float32)
        self.ctx.check(self.ctx.L.hr_motion_vectors(self.ctx.h, C.byref(self.gb.c), vp.ctypes.data, pv.ctypes.data,
                                                    motion.ptr))

    def shade_hybrid_with(self, shadow, stats=None):
        """hr_shade_hybrid reading the given shadow buffer (e.g. the filter's
        history, as renderer.cpp:3086-3090 binds the filtered images)"""
        L = self.ctx.L
        self.ctx.check(L.hr_shade_hybrid(self.hs.h, C.byref(self.gb.c), C.cast(self.lights, C.c_void_p),
                                         self._shapes_ptr(), self.nl, C.byref(self.cam), C.byref(self.params),
                                         shadow.ptr, self.lit.ptr, self.refl.ptr,
                                         C.byref(stats) if stats is not None else None))


def _merge_stats(st):
    d = st.as_dict()
    return {k: (int(v) if isinstance(v, int) else float(v)) for k, v in d.items()}


class BatchRenderer:
    """Several frames (HybridRenderers with their own scenes, G-buffers and
    outputs) traced together: one persistent shadow kernel over all frames,
    then the shading pass of each; or the whole batch captured in a hipGraph."""

    def __init__(self, ctx, renderers):
        self.ctx = ctx
        self.renderers = list(renderers)
        self._build_descs()

    def _build_descs(self):
        n = len(self.renderers)
        self.descs = (abi.hr_frame_desc * n)()
        for i, rd in enumerate(self.renderers):
            d = self.descs[i]
            d.scene = rd.hs.h.value
            d.gbuffer = rd.gb.c
            d.lights = C.cast(rd.lights, C.c_void_p)
            d.shapes = None if rd.shapes is None else C.cast(rd.shapes, C.c_void_p)
            d.light_count = rd.nl
            d.camera = rd.cam
            d.params = rd.params
            d.shadow = rd.shadow.ptr
            d.lit = rd.lit.ptr
            d.refl = rd.refl.ptr

    def gbuffers(self):
        for rd in self.renderers:
            rd.gbuffer_primary()

    def trace_shadows(self, stats=False):
        st = abi.hr_stats() if stats else None
        self.ctx.check(self.ctx.L.hr_trace_shadows_batch(self.ctx.h, C.cast(self.descs, C.c_void_p),
                                                         len(self.renderers), C.byref(st) if st else None))
        return _merge_stats(st) if st else None

    def shade(self, stats=False):
        st = abi.hr_stats() if stats else None
        self.ctx.check(self.ctx.L.hr_shade_hybrid_batch(self.ctx.h, C.cast(self.descs, C.c_void_p),
                                                        len(self.renderers), C.byref(st) if st else None))
        return _merge_stats(st) if st else None

    def capture(self):
        return Graph(self)


class Graph:
    """A batch frame (shadows + shading) captured once and replayed."""

    def __init__(self, batch):
        self.batch = batch
        self.ctx = batch.ctx
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.hr_graph_capture(self.ctx.h, C.cast(batch.descs, C.c_void_p),
                                                   len(batch.renderers), C.byref(h)))
        self.h = h
        self.ctx.own(self)

    def launch(self, stats=False):
        st = abi.hr_stats() if stats else None
        self.ctx.check(self.ctx.L.hr_graph_launch(self.h, C.byref(st) if st else None))
        return _merge_stats(st) if st else None

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.hr_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
