/*
 * hr.h - C ABI of the MI355X hybrid pass (G-buffer-seeded shadow and
 * reflection ray tracing plus deferred PBR shading) of VulkanHybridRaytracer.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an hr_status. Errors are returned, never aborted on; the message of the last
 * failure is kept per context (hr_last_error).
 *
 * Pointer conventions
 *   - Scene inputs (meshes, primitives, materials, environment maps) are HOST
 *     pointers; the library copies them into device memory once, as the
 *     reference uploads its vertex/index/material buffers through staging
 *     buffers (vk_mesh.cpp:265-359, renderer.cpp:1477-1530).
 *   - Per-frame uniforms (lights, camera, parameters) are HOST pointers to a
 *     few hundred bytes; they travel as kernel arguments, so no copy is made.
 *   - Images (G-buffer planes, shadow, lit, reflection buffers) and ray/hit
 *     arrays are DEVICE pointers (hr_device_alloc or any HIP allocation on the
 *     context's device). They replace the Vulkan storage images of
 *     renderer.cpp:1532-1624.
 *   - Handles are not thread-safe: use one context per host thread.
 *   - All work of a context is ordered on the context's HIP stream; entry
 *     points that return results to the host synchronise that stream.
 *
 * Reference interfaces replaced (citations into the reference tree):
 *   hr_scene_build    create_bottom_acceleration_structure + buildBlas +
 *                     create_top_acceleration_structure + buildTlas
 *                     (renderer.cpp:1643-1688, 1696-1865), fed by
 *                     Node::node_to_geometry / node_to_instance
 *                     (vk_mesh.cpp:424-492) and the per-primitive id/matrix
 *                     tables (vk_mesh.cpp:506-533, renderer.cpp:3018-3048).
 *   hr_scene_refit    per-frame TLAS update (vk_engine.cpp:174-187,
 *                     renderer.cpp:1757-1865 with update = true).
 *   hr_gbuffer_primary  the raster G-buffer pass (renderer.cpp:1377-1434,
 *                     geometry_shader.frag:51-76, skybox.frag:14-21), replaced
 *                     by primary visibility at pixel centres with the camera
 *                     convention of raygen.rgen:29-40.
 *   hr_trace_shadows  shadowRaygen.rgen + shadow.rmiss + shadowHit.rchit
 *                     dispatched by build_shadow_command_buffer
 *                     (renderer.cpp:2618-2659), descriptor set of
 *                     renderer.cpp:1919-2028.
 *   hr_shade_hybrid   hybridRaygen.rgen + hybridHit.rchit + miss.rmiss +
 *                     shadow.rmiss dispatched by build_hybrid_command_buffers
 *                     (renderer.cpp:3131-3175), descriptor set of
 *                     renderer.cpp:2917-3129.
 *   hr_trace_rays     a bare traceRayEXT (hybridRaygen.rgen:158 closest hit,
 *                     shadowRaygen.rgen:78 terminate-on-first-hit).
 */
#ifndef HR_HR_H
#define HR_HR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HR_ABI_VERSION 1u
#define HR_MAX_LIGHTS 16u

typedef enum hr_status {
    HR_OK = 0,
    HR_E_INVALID = 1,     /* bad argument: null pointer, size, index out of range */
    HR_E_OOM = 2,         /* device allocation failed */
    HR_E_DEVICE = 3,      /* HIP runtime error */
    HR_E_UNSUPPORTED = 4, /* valid request outside what this library implements */
    HR_E_OVERFLOW = 5     /* a fixed-capacity structure (traversal stack, queue) overflowed */
} hr_status;

/* == Vertex (vk_mesh.h:21-27): 44 bytes, position at offset 0 read with stride
 * 44 exactly as the BLAS input does (vk_mesh.cpp:436-442). */
typedef struct hr_vertex {
    float position[3];
    float normal[3];
    float color[3];
    float uv[2];
} hr_vertex;

/* == GPUMaterial (material.h:5-10), filled as Material::materialToShader
 * (material.cpp:19-26): diffuse = (rgb, ior); textures = (diffuse, normal,
 * emissive, metallic-roughness) indices into hr_scene_set_textures as
 * floats, -1 = none (whole numbers only); shading_metal_rough_idx =
 * (shadingModel, metallic, roughness, material index). Shading models 0
 * (PBR), 3 (mirror/glossy) and 4 (glass) exist (hybridRaygen.rgen:105-148).
 * A metallic-roughness texture gives (r, roughness = g, metallic = b), the
 * layout the shaders read (hybridHit.rchit:65-68). */
typedef struct hr_material {
    float diffuse_ior[4];
    float textures[4];
    float shading_metal_rough_idx[4];
} hr_material;

/* == uboLight (vk_engine.h:76-80), packed as vk_engine.cpp:146-164:
 * pos_maxdist = (position, maxDistance) for point lights and
 * (direction towards the light, -1) for directional lights;
 * color_intensity = (rgb, intensity); radius = disk radius for soft shadows. */
typedef struct hr_light {
    float pos_maxdist[4];
    float color_intensity[4];
    float radius;
    float _pad[3];
} hr_light;

/* Optional per-light shape (no counterpart in the reference, whose lights are
 * disks sampled by helpers.glsl:82-93). type HR_LIGHT_DISK keeps that
 * behaviour; HR_LIGHT_RECT is a rectangle centred at pos_maxdist.xyz spanned by
 * edge_u and edge_v (full extents), sampled on a stratified grid. */
#define HR_LIGHT_DISK 0u
#define HR_LIGHT_RECT 1u
typedef struct hr_light_shape {
    uint32_t type;
    float edge_u[3];
    float edge_v[3];
    uint32_t _pad;
} hr_light_shape;

/* == RTCameraData (vk_engine.h:56-62), column-major glm matrices, filled as
 * vk_engine.cpp:730-740: inverse(view), inverse(projection with [1][1] *= -1),
 * frame = (frame index, ...). frame[0] seeds the per-pixel random streams
 * through uint(int(frame)) (hybridRaygen.rgen:29). */
typedef struct hr_camera {
    float inv_view[16];
    float inv_proj[16];
    float frame[4];
} hr_camera;

/* One mesh: its vertex array and uint32 index array (Mesh::_vertices,
 * Mesh::_indices, vk_mesh.h:73-77). */
typedef struct hr_mesh {
    const hr_vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
} hr_mesh;

/* One drawable primitive (Primitive, vk_mesh.h:60-70) placed in the world by
 * model = entity.m_matrix * node.global (vk_mesh.cpp:476, column-major).
 * Its triangles are indices[first_index .. first_index + index_count). */
typedef struct hr_primitive {
    uint32_t mesh;
    uint32_t first_index;
    uint32_t index_count;
    int32_t material;
    float model[16];
} hr_primitive;

/* Environment maps, RGBA float texels, row-major, row 0 at v = 0.
 * skybox = environmentTexture[0] / skybox[0] (miss.rmiss:15-17),
 * irradiance = environmentTexture[1] (hybridRaygen.rgen:59-60,
 * hybridHit.rchit:71-73). Looked up at the nearest texel with wrap in u and v
 * (NEAREST/REPEAT sampler, renderer.cpp:3060, vk_initializers.h:86). */
typedef struct hr_env {
    const float* skybox;
    uint32_t sky_width, sky_height;
    const float* irradiance;
    uint32_t irr_width, irr_height;
} hr_env;

/* FP32 G-buffer, one float4 per pixel per plane, row-major, device pointers.
 *   position = (world xyz, material index)          [gbuffers[0]]
 *   normal   = (world unit normal xyz, 1; 0 for background) [gbuffers[1]]
 *   albedo   = (diffuse rgb before linearisation, 1) [gbuffers[2]]
 *   material = (0, roughness, metallic, 0)           [gbuffers[4]]
 *   emissive = (rgb, 0)                              [gbuffers[5]]
 * Background pixels have position.w == 0 and normal.w == 0 and the sky colour
 * in albedo (skybox.frag:16-21). The normal is stored decoded (the reference
 * stores 0.5 n + 0.5 in RGBA16F, geometry_shader.frag:67). With textures,
 * hr_gbuffer_primary writes what geometry_shader.frag:54-70 writes: albedo =
 * diffuse texel * vertex colour, the emissive texel, the metallic-roughness
 * texel as material (w = 0), and the normal perturbed by the normal-map texel
 * in the screen-space cotangent frame. */
typedef struct hr_gbuffer {
    uint32_t width, height;
    float* position;
    float* normal;
    float* albedo;
    float* material;
    float* emissive;
} hr_gbuffer;

/* Trace parameters.
 * shadow_samples: shadow rays per light per pixel (UBO `samples`,
 *   vk_engine.cpp:167-170, GUI range 1..64, renderer.cpp:656).
 * refl_spp: reflection/refraction paths per shading-model 3/4 pixel
 *   (1 in the reference).
 * max_depth: bounce limit (MAX_RECURSION = 8, helpers.glsl:8). */
typedef struct hr_trace_params {
    uint32_t shadow_samples;  /* with HR_FLAG_ADAPTIVE: the penumbra sample count (<= 64) */
    uint32_t refl_spp;
    uint32_t max_depth;
    uint32_t flags;
} hr_trace_params;

/* A ray for hr_trace_rays: 32 bytes, as traceRayEXT's (origin, tmin,
 * direction, tmax) arguments. */
typedef struct hr_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} hr_ray;

/* A hit: 16 bytes. t = gl_HitTEXT or -1 on a miss; (u, v) = the barycentric
 * hit attributes (weights of the triangle's 2nd and 3rd vertex,
 * hybridHit.rchit:27); tri = global triangle id (primitive order, then the
 * primitive's triangle order; 0xFFFFFFFF on a miss). For any-hit queries only
 * t (>= 0 means occluded) and tri are meaningful. */
typedef struct hr_hit {
    float t;
    float u;
    float v;
    uint32_t tri;
} hr_hit;

/* Ray counts and timings of the last frame call. */
typedef struct hr_stats {
    uint64_t primary_rays;
    uint64_t shadow_rays;      /* G-buffer-surface shadow rays (shadowRaygen) */
    uint64_t reflection_rays;  /* closest-hit bounce rays (hybridRaygen:158) */
    uint64_t hit_shadow_rays;  /* shadow rays cast at bounce hits (hybridHit:123) */
    uint64_t occluded_rays;    /* shadow rays of both kinds that found a hit */
    uint64_t node_visits;      /* only with HR_FLAG_COUNT_VISITS */
    uint64_t tri_tests;        /* only with HR_FLAG_COUNT_VISITS */
    float ms_gbuffer;
    float ms_shadow;
    float ms_shade;
    float ms_build;
} hr_stats;

#define HR_FLAG_COUNT_VISITS 1u  /* count node visits / triangle tests (slower kernel variant) */
#define HR_FLAG_TIMING 2u        /* reserved */
/* Adaptive sampling (no counterpart in the reference, whose sample count is a
 * uniform, renderer.cpp:656, GUI range 1..64):
 *  - shadows: a disk light of radius 0 takes 1 ray. Other lights draw from
 *    one nested stratified sequence of S = shadow_samples (a power of two, at
 *    most 64) cells on the light, in stages of 4, 16, 32, 64 samples (up to S),
 *    each stage adding to the samples before it. After N samples of which k
 *    are unoccluded the count stops when N = S or when the binomial variance
 *    estimate of the visible fraction, k (N - k) / N^3, is at most that of S
 *    samples at one half, 1 / (4 S): a lit or umbral pixel takes 4 rays, a
 *    penumbra 16, 32 or 64 by how mixed its samples are. The factor is k / N;
 *  - reflections: a glossy pixel (shading model 3, roughness r > 0) traces
 *    1 + floor((refl_spp - 1) * r) paths. */
#define HR_FLAG_ADAPTIVE 4u
/* The reference's 8-bit images (no other change to the results):
 *  HR_FLAG_SHADOW_R8: the shadow buffer is the R8_UNORM image of
 *    renderer.cpp:1538 (uint8[nl][H][W]): hr_trace_shadows stores
 *    rint(clamp(factor, 0, 1) * 255) and hr_shade_hybrid reads code / 255, as
 *    imageStore / imageLoad do (shadowRaygen.rgen:93, hybridRaygen.rgen:83);
 *  HR_FLAG_LIT_BGRA8: hr_shade_hybrid stores lit as the B8G8R8A8_UNORM image of
 *    renderer.cpp:1535 (uint8[H][W][4], blue first, alpha 255):
 *    rint(clamp(colour, 0, 1) * 255) before the post gamma (hybridRaygen.rgen:
 *    186). refl stays FP32. */
#define HR_FLAG_SHADOW_R8 8u
#define HR_FLAG_LIT_BGRA8 16u

typedef struct hr_build_params {
    uint32_t leaf_max;   /* largest Karras subtree stored as one leaf (1..8, default 2) */
    uint32_t _reserved[3];
} hr_build_params;

typedef struct hr_context hr_context;
typedef struct hr_scene hr_scene;

/* ---- context ---------------------------------------------------------- */
uint32_t hr_abi_version(void);
hr_status hr_context_create(int device, hr_context** out);
void hr_context_destroy(hr_context* ctx);
const char* hr_last_error(const hr_context* ctx);
hr_status hr_synchronize(hr_context* ctx);

/* Device memory helpers for hosts without their own allocator. */
hr_status hr_device_alloc(hr_context* ctx, size_t bytes, void** out);
hr_status hr_device_free(hr_context* ctx, void* ptr);
hr_status hr_memcpy_h2d(hr_context* ctx, void* dst_dev, const void* src_host, size_t bytes);
hr_status hr_memcpy_d2h(hr_context* ctx, void* dst_host, const void* src_dev, size_t bytes);
hr_status hr_memset_d(hr_context* ctx, void* dst_dev, int value, size_t bytes);

/* ---- scene (replaces the BLAS/TLAS build, renderer.cpp:1643-1916) ------- */
hr_status hr_scene_build(hr_context* ctx,
                         const hr_mesh* meshes, uint32_t mesh_count,
                         const hr_primitive* prims, uint32_t prim_count,
                         const hr_material* materials, uint32_t material_count,
                         const hr_build_params* params, /* may be NULL */
                         hr_scene** out);
/* New model matrices (prim_count x 16 floats, host): re-transform triangles and
 * refit node boxes on the existing topology (TLAS update,
 * vk_engine.cpp:174-187). */
hr_status hr_scene_refit(hr_scene* scene, const float* models);
/* Material re-upload (renderer.cpp:761-774). */
hr_status hr_scene_set_materials(hr_scene* scene, const hr_material* materials, uint32_t count);
hr_status hr_scene_set_env(hr_scene* scene, const hr_env* env);
/* Material textures: the reference's Texture::_textures list (vk_textures.cpp:
 * 128-160), bound as sampler2D[] (renderer.cpp:936-950, 3060-3070) and indexed
 * by hr_material.textures (diffuse, normal, emissive, metallic-roughness; -1 =
 * none). Each texture is RGBA8 as stbi_load(..., STBI_rgb_alpha) returns it
 * (vk_textures.cpp:13-27), row-major, row 0 at v = 0, host pointer; the
 * library copies them. Sampled as VK_FORMAT_R8G8B8A8_UNORM (c / 255) with the
 * NEAREST / REPEAT sampler (vk_initializers.cpp:305-316): texel
 * (floor(u * w) mod w, floor(v * h) mod h). Replaces the previous set. Every
 * texture index the scene's materials use must be below count when a pass
 * that shades runs (HR_E_INVALID otherwise). */
typedef struct hr_texture {
    const uint8_t* rgba8;
    uint32_t width, height;
} hr_texture;
hr_status hr_scene_set_textures(hr_scene* scene, const hr_texture* textures, uint32_t count);
void hr_scene_destroy(hr_scene* scene);

/* Scene layout queries and device-array export (for validation). */
hr_status hr_scene_info(const hr_scene* scene, uint32_t* tri_count, uint32_t* node_count,
                        uint32_t* depth, float* ms_build);
/* The number of 4-wide traversal nodes and the device memory the scene holds
 * (geometry, both node arrays, build scratch kept for refits, textures,
 * environment). No counterpart in the reference, whose acceleration
 * structures live in Vulkan memory the driver sizes (renderer.cpp:1643-1916). */
hr_status hr_scene_memory(const hr_scene* scene, uint32_t* wide_count, uint64_t* device_bytes);
/* which: 0 = traversal nodes (node_count x 64 B), 1 = sorted world-space
 * triangles (tri_count x 48 B), 2 = Morton codes in sorted order (tri_count x
 * 4 B), 3 = sorted triangle ids, in the order of the sorted triangles
 * (tri_count x 4 B), 4 = internal nodes of the binary tree (left, right,
 * first, last) ((tri_count - 1) x 16 B), 5 = 4-wide traversal
 * nodes (wide_count x 128 B of hr_scene_memory, the root first: from the root
 * down, a wide node takes its binary node's two children and twice replaces
 * the internal child of largest box area, (x y + y z) + z x of the extents,
 * first on ties, by that child's two children; the wide nodes are numbered
 * densely in the order of the binary nodes they are rooted at, internal child
 * codes are those numbers, and word 28 of a node holds its binary node).
 * The wide topology is chosen at build time; a refit keeps it and updates
 * the boxes. */
hr_status hr_scene_export(const hr_scene* scene, uint32_t which, void* dst_host, size_t bytes);

/* ---- per-frame passes ---------------------------------------------------- */
hr_status hr_gbuffer_primary(hr_scene* scene, const hr_camera* cam, const hr_gbuffer* gb,
                             hr_stats* stats);

/* shadow: device float[nl][H][W], fraction of unoccluded samples per light.
 * nl may be 0 (lights may then be NULL), as the reference's light buffer may be
 * empty (shadowRaygen.rgen:52): nothing is traced or written. */
hr_status hr_trace_shadows(hr_scene* scene, const hr_gbuffer* gb,
                           const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                           const hr_camera* cam, const hr_trace_params* params,
                           float* shadow, hr_stats* stats);

/* hr_trace_shadows without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc). Same
 * arguments, checks and results (bit for bit) as hr_trace_shadows. Its scratch
 * (the frame's shadow job, the work-band counters and the sample bit arrays)
 * is one static device array of HR_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; a frame needing more returns HR_E_UNSUPPORTED before any
 * device work (1920x1080 with 4 lights fits at every sample count up to 64,
 * adaptive or not). Everything the pass needs zeroed is cleared by kernels on
 * the context's stream. Calls that overlap on one GPU (other streams, contexts
 * or threads) race on the shared array. Under graph capture pass stats = NULL
 * (stats synchronise the stream). */
#define HR_NOALLOC_SCRATCH_BYTES (128u << 20)
hr_status hr_trace_shadows_noalloc(hr_scene* scene, const hr_gbuffer* gb,
                                   const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                                   const hr_camera* cam, const hr_trace_params* params,
                                   float* shadow, hr_stats* stats);

/* lit: device float4[H][W] final colour (imageStore, hybridRaygen.rgen:186,
 * unquantised); refl: device float4[H][W], the multiplicative bounce factor
 * averaged over refl_spp paths (1 where nothing is traced), so that
 * lit = direct * refl + ambient + emissive for non-background pixels.
 * With nl == 0 the direct term is 0 and no bounce is traced (the reference
 * bounces a zero direction, hybridRaygen.rgen:66-67, whose colour multiplies a
 * zero rayColor): lit = ambient + emissive, refl = 1. */
hr_status hr_shade_hybrid(hr_scene* scene, const hr_gbuffer* gb,
                          const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                          const hr_camera* cam, const hr_trace_params* params,
                          const float* shadow, float* lit, float* refl, hr_stats* stats);

/* Bare ray queries over device arrays. flags: HR_TRACE_ANY_HIT stops at the
 * first hit (gl_RayFlagsTerminateOnFirstHitEXT); otherwise the closest hit,
 * ties on t broken by the smaller triangle id. HR_TRACE_COUNT_VISITS walks the
 * binary nodes in the oracle's near-first order and counts node visits and
 * triangle tests into stats (a slower kernel); without it the 4-wide nodes
 * are walked. Both give the same hits. stats (may be NULL) receives the
 * counts and the time of whichever kernel ran. */
#define HR_TRACE_ANY_HIT 1u
#define HR_TRACE_COUNT_VISITS 2u
hr_status hr_trace_rays(hr_scene* scene, const hr_ray* rays, uint32_t n, int flags,
                        hr_hit* hits, hr_stats* stats);

/* ---- batches of frames (no counterpart in the reference, which renders one
 * scene per frame) ----------------------------------------------------------
 * A frame is one scene with its G-buffer, lights, camera, parameters and
 * device output buffers, as hr_trace_shadows / hr_shade_hybrid take them. All
 * scenes must belong to the same context. */
typedef struct hr_frame_desc {
    hr_scene* scene;
    hr_gbuffer gbuffer;
    const hr_light* lights;
    const hr_light_shape* shapes; /* may be NULL */
    uint32_t light_count;
    hr_camera camera;
    hr_trace_params params;
    float* shadow;  /* device float[light_count][H][W] */
    float* lit;     /* device float4[H][W] */
    float* refl;    /* device float4[H][W] */
} hr_frame_desc;

/* The shadow pass of every frame in one persistent kernel; work is taken
 * from one device-side counter shared by all frames. */
hr_status hr_trace_shadows_batch(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_stats* stats);
/* The hybrid shading pass of every frame (after hr_trace_shadows_batch). */
hr_status hr_shade_hybrid_batch(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_stats* stats);

/* A frame of a batch (shadows, then shading) captured once into a hipGraph
 * and replayed by hr_graph_launch. The graph owns its device copies of the
 * frames' parameters and its own workspace; the frames' scenes and buffers
 * must outlive it. The graph holds the scenes' device addresses as they were
 * at capture: after hr_scene_set_materials, hr_scene_set_textures or
 * hr_scene_set_env on one of its scenes, hr_graph_launch returns HR_E_INVALID
 * without launching (capture again). hr_scene_refit keeps the addresses, so a
 * graph replays refitted scenes. */
typedef struct hr_graph hr_graph;
hr_status hr_graph_capture(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_graph** out);
hr_status hr_graph_launch(hr_graph* graph, hr_stats* stats);
void hr_graph_destroy(hr_graph* graph);

/* Launch shapes, set per context (hr_context_set_tuning) and chosen by the
 * autotuner of the Python package:
 *  blocks_per_cu   persistent trace grid, 256-thread workgroups per CU (1..16)
 *  tile_w          shadow work tile width in pixels (1..64, power of two;
 *                  height 64 / tile_w): one wave's 64 groups
 *  wave_budget_mb  reflection-queue workspace per pixel slab, MiB (>= 64;
 *                  default 40960: a 1080p frame at 64 samples per pixel is one
 *                  slab). The workspace is allocated when a frame first needs
 *                  it and kept by the context; when the device cannot hold it,
 *                  the budget is halved (down to 64 MiB) and the frame runs in
 *                  more, smaller slabs, with the same results.
 *  bvh_flags       how scenes created afterwards build their hierarchy: 0 (the
 *                  default) restructures the LBVH by treelets (Karras and Aila
 *                  2013: two passes of 7-leaf treelets, fewer node visits per
 *                  ray for a longer build); HR_BVH_NO_TREELETS keeps the plain
 *                  Karras tree. A restructured tree deeper than the traversal
 *                  stack (64 levels) is replaced by the plain tree. Other bits:
 *                  HR_E_INVALID. Hits do not depend on the hierarchy. */
#define HR_BVH_NO_TREELETS 1u
typedef struct hr_tuning {
    uint32_t blocks_per_cu;
    uint32_t tile_w;
    uint32_t wave_budget_mb;
    uint32_t bvh_flags;
} hr_tuning;
hr_status hr_context_set_tuning(hr_context* ctx, const hr_tuning* tuning);
hr_status hr_context_get_tuning(const hr_context* ctx, hr_tuning* tuning);

/* ---- pure ray tracing mode ---------------------------------------------------
 * The reference's RAYTRACING render mode: raygen.rgen + closesthit.rchit +
 * miss.rmiss + shadow.rmiss, dispatched by build_raytracing_command_buffers
 * (renderer.cpp:2572-2616) and submitted by Renderer::raytrace
 * (renderer.cpp:465-528). Per pixel, params->refl_spp camera samples (the
 * first through the pixel centre, raygen.rgen:27-33; NSAMPLES in the
 * reference), each a closest-hit chain of at most params->max_depth rays
 * (MAX_RECURSION) shaded by closesthit.rchit with params->shadow_samples
 * shadow rays per light at every hit (closesthit.rchit:99-121; a light
 * beyond its range, and a directional light, stays shadowed there).
 * accum: device float4[height][width]; camera->frame[0] < 0 stores the new
 * colour, otherwise accum = mix(accum, colour, 1 / (1 + frame))
 * (raygen.rgen:69-80). stats: primary_rays, reflection_rays (the chain's
 * later rays), hit_shadow_rays, occluded_rays, ms_shade. Without
 * HR_FLAG_COUNT_VISITS and with at most 64 shadow samples the frame runs as
 * wavefront kernels in the context's workspace (about 128 B per pixel plus
 * light_count * 2^ceil(log2 samples) bits per pixel); otherwise one lane per
 * pixel walks the whole chain. Both give the same image. */
hr_status hr_render_rt(hr_scene* scene, const hr_light* lights, const hr_light_shape* shapes, uint32_t light_count,
                       const hr_camera* camera, const hr_trace_params* params, uint32_t width, uint32_t height,
                       float* accum, hr_stats* stats);

/* ---- raster G-buffer ingest ------------------------------------------------
 * The offscreen attachments of the reference's raster G-buffer pass
 * (renderer.cpp:252-261, written by geometry_shader.frag:66-75 and
 * skybox.frag:16-21), device pointers, row-major, tightly packed:
 *   position  R16G16B16A16_SFLOAT  (world xyz, material index)
 *   normal    R16G16B16A16_SFLOAT  (0.5 n + 0.5, 1; 0 on background)
 *   albedo    R8G8B8A8_UNORM       (diffuse rgb before linearisation)
 *   motion    R16G16_SFLOAT        ((ndc - ndc_prev) * 0.5 + 0.5), may be NULL
 *   material  R8G8B8A8_UNORM       (0, roughness, metallic)
 *   emissive  R8G8B8A8_UNORM
 * hr_gbuffer_ingest decodes them into the FP32 planes exactly as the shaders
 * read them (shadowRaygen.rgen:36-41, hybridRaygen.rgen:40-47): half to
 * float, normal = n * 2 - 1 (w kept), UNORM8 = c / 255; motion_out (device
 * float2[H][W], may be NULL) receives the motion vectors for
 * hr_filter_shadows. hr_gbuffer_quantize is the inverse the raster pass
 * performs: half by round to nearest even, UNORM8 = rint(clamp(v, 0, 1) *
 * 255), normal * 0.5 + 0.5 (0 on background). */
typedef struct hr_raster_gbuffer {
    uint32_t width, height;
    void* position;
    void* normal;
    void* albedo;
    void* motion;
    void* material;
    void* emissive;
} hr_raster_gbuffer;
hr_status hr_gbuffer_ingest(hr_context* ctx, const hr_raster_gbuffer* in, const hr_gbuffer* out, float* motion_out);
hr_status hr_gbuffer_quantize(hr_context* ctx, const hr_gbuffer* in, const float* motion_in,
                              const hr_raster_gbuffer* out);

/* ---- temporal shadow filter ----------------------------------------------
 * hr_filter_shadows replaces the compute pass blur.comp (dispatch
 * renderer.cpp:2661-2675, descriptor set renderer.cpp:2030-2082) that runs
 * between the shadow and hybrid passes; the hybrid pass then reads its output
 * (renderer.cpp:3086-3090): pass `history` as the shadow buffer of
 * hr_shade_hybrid. Any width, height and light count (the reference
 * hard-codes 1700x900 and 3 lights, blur.comp:7-10) and every pixel (the
 * reference dispatches W/16 x H/16 whole workgroups).
 *   shadow   device float[light_count][H][W], this frame's traced factors
 *   history  device float[light_count][H][W], the previous output, replaced by
 *            this frame's (the reference's _denoisedImages)
 *   frame    the FrameCount value (vk_engine.cpp:669-686, 722): -1 is set on a
 *            view change and incremented before use, so the first frame after
 *            a change has frame 0
 *   motion   device float2[H][W] = (ndc - ndc_prev) * 0.5 + 0.5 as the raster
 *            G-buffer stores it (geometry_shader.frag:72-75), read at the
 *            pixel's own texel; required when frame <= 0, else may be NULL
 * frame > 0: history = mix(history, shadow, 1 / (frame + 1)).
 * frame <= 0: the previous output at the reprojected pixel (x + r.x * W,
 * y + r.x * H), r = motion * 2 - 1 (the y offset uses r.x as blur.comp:52
 * evaluates it), truncated toward zero, 0 outside the frame, clamped to the
 * 3x3 min/max of this frame's factors (0 outside the frame), then
 * mix(old, shadow, 0.4). Reprojected reads see the previous output (the
 * reference reads and writes one image in place). */
hr_status hr_filter_shadows(hr_context* ctx, uint32_t width, uint32_t height, uint32_t light_count, int32_t frame,
                            const float* shadow, const float* motion, float* history);
/* Motion vectors of a G-buffer for the filter: for each pixel with geometry,
 * (ndc - ndc_prev).xy * 0.5 + 0.5 of its world position under view_proj and
 * prev_view_proj (projection * view, column-major 4x4, as basic.vert:38-46
 * uses them); 0 for background pixels (skybox.frag:19). motion: device
 * float2[H][W]. */
hr_status hr_motion_vectors(hr_context* ctx, const hr_gbuffer* gbuffer, const float* view_proj,
                            const float* prev_view_proj, float* motion);

/* ---- display ---------------------------------------------------------------
 * The post pass (postFragment.frag:9-20, drawn by renderer.cpp's post
 * pipeline over the final image): rgb = pow(max(rgb, 0.001), 1 / 2.2), alpha
 * kept. in: device float4[H][W]. out (device float4[H][W], may be NULL, may
 * equal in) receives the float result; rgba8 (device uint8[H][W][4], may be
 * NULL) receives it stored to an 8-bit UNORM image: rint(clamp(v, 0, 1) * 255). */
hr_status hr_post_gamma(hr_context* ctx, uint32_t width, uint32_t height, const float* in, float* out,
                        uint8_t* rgba8);

/* ---- asset ingest (host only: no context, no device work) -----------------
 * Images: what stbi_load(path, &w, &h, &c, STBI_rgb_alpha) returns to
 * Texture::GET (vk_textures.cpp:11-27) - RGBA8, row 0 first - for PNG (all
 * colour types and depths, Adam7), JPEG (baseline and progressive, grey or
 * YCbCr) and Radiance HDR (converted to 8 bits as stb does:
 * pow(v, 1 / 2.2) * 255 + 0.5, clamped, truncated). The pixels are
 * malloc'ed: release them with hr_image_free. err (may be NULL) receives a
 * message on failure. */
typedef struct hr_image {
    uint8_t* rgba8;
    uint32_t width, height;
} hr_image;
hr_status hr_image_load(const char* path, hr_image* out, char* err, size_t err_len);
hr_status hr_image_decode(const uint8_t* data, size_t size, hr_image* out, char* err, size_t err_len);
void hr_image_free(hr_image* image);

/* Models: Prefab::GET(filename, invertNormals) (vk_mesh.cpp:843-901) for .obj
 * (Mesh::load_from_obj, vk_mesh.cpp:96-150), .gltf and .glb
 * (Prefab::loadNode / loadMaterial / loadTextures, vk_mesh.cpp:646-830).
 * The view holds one mesh (the prefab's shared vertex / index buffer), its
 * primitives in the reference's instance order (Node::node_to_instance,
 * vk_mesh.cpp:494-515) with model = the node's global matrix (multiply by the
 * entity matrix before hr_scene_build), the materials in first-use order
 * (material index = position in this asset's list; the reference numbers
 * them in one program-wide list) and the decoded textures the materials
 * index. Image URIs are resolved against the model file's directory. The view
 * stays valid until hr_asset_destroy. */
typedef struct hr_asset hr_asset;
#define HR_ASSET_INVERT_NORMALS 1u /* negate the root nodes' normals (the reference's invertNormals) */
#define HR_ASSET_TEXTURE_SOURCE 2u /* glTF: image = textures[i].source (the reference uses images[i]) */
typedef struct hr_asset_view {
    hr_mesh mesh;
    const hr_primitive* primitives;
    uint32_t primitive_count;
    const hr_material* materials;
    uint32_t material_count;
    const hr_texture* textures;
    uint32_t texture_count;
} hr_asset_view;
hr_status hr_asset_load(const char* path, uint32_t flags, hr_asset** out, char* err, size_t err_len);
hr_status hr_asset_view_get(const hr_asset* asset, hr_asset_view* out);
void hr_asset_destroy(hr_asset* asset);

/* ---- learned denoiser (no counterpart in the reference, whose only
 * "denoise" is the temporal shadow filter blur.comp) --------------------------
 * A kernel-predicting U-Net: 5 levels of two 3x3 convolutions with ReLU
 * (32, 48, 64, 64, 64 channels), 2x2 average pooling, nearest upsampling with
 * the encoder output concatenated in front, and a 3x3 head predicting 25
 * logits per pixel; the softmax of the logits weights the 5x5 neighbourhood
 * (clamped at the frame border) of the noisy lit colour. Input features per
 * pixel: albedo (3), normal (3), distance to the eye / 100 (1), roughness (1),
 * log(1 + luminance of the noisy colour) (1). BF16 activations and weights,
 * FP32 accumulation; FP32 master weights and Adam state.
 *
 * Weights are exchanged as FP32 in layer order e0a e0b e1a e1b e2a e2b e3a
 * e3b e4a e4b d3a d3b d2a d2b d1a d1b d0a d0b head, each as W[cout][3][3][cin]
 * (tap (dy + 1) * 3 + dx + 1) followed by b[cout]; hr_denoiser_layer gives
 * the shapes. */
typedef struct hr_denoiser hr_denoiser;
#define HR_DN_LAYERS 19
uint64_t hr_denoiser_param_count(void);
/* Shape of layer i: input / output channels and pyramid level (0 = full resolution). */
hr_status hr_denoiser_layer(uint32_t i, uint32_t* cin, uint32_t* cout, uint32_t* level);
/* weights may be NULL: He-normal initialisation from seed. */
hr_status hr_denoiser_create(hr_context* ctx, const float* weights, uint64_t count, uint64_t seed, hr_denoiser** out);
void hr_denoiser_destroy(hr_denoiser* dn);
hr_status hr_denoiser_get_weights(hr_denoiser* dn, float* weights, uint64_t count);
/* Replaces the weights and resets the optimiser state. */
hr_status hr_denoiser_set_weights(hr_denoiser* dn, const float* weights, uint64_t count);
/* Training state for a resumable checkpoint: the FP32 master weights, Adam's
 * first and second moments (each `count` values in the weight exchange
 * layout) and the number of Adam steps taken, which the bias corrections
 * continue from. Training resumed from a state gives the weights that
 * uninterrupted training gives, bit for bit. adam_m / adam_v / step may be
 * NULL in get_state; set_state needs all three arrays. */
hr_status hr_denoiser_get_state(hr_denoiser* dn, float* weights, float* adam_m, float* adam_v, uint64_t count,
                                uint64_t* step);
hr_status hr_denoiser_set_state(hr_denoiser* dn, const float* weights, const float* adam_m, const float* adam_v,
                                uint64_t count, uint64_t step);

/* One image: its G-buffer (position, normal, albedo and material planes are
 * read), the noisy lit colour, and for training the reference colour. All
 * pointers are device float4[height][width]. */
typedef struct hr_dn_image {
    hr_gbuffer gbuffer;
    const float* noisy;
    const float* target; /* training only, else NULL */
    float* out;          /* hr_denoise only */
    float eye[3];
    float _reserved;
} hr_dn_image;

/* Training loss, per pixel and colour channel, averaged:
 *   HR_DN_LOSS_L1_TONEMAPPED  |T(out) - T(target)|, T(v) = max(v, 0) / (1 + max(v, 0))
 *   HR_DN_LOSS_REL_MSE        (out - target)^2 / (target^2 + 0.01)  (the rMSE metric) */
#define HR_DN_LOSS_L1_TONEMAPPED 0u
#define HR_DN_LOSS_REL_MSE 1u
hr_status hr_denoiser_set_loss(hr_denoiser* dn, uint32_t loss);
/* Denoises one frame into image->out; ms (optional) receives the device time. */
hr_status hr_denoise(hr_denoiser* dn, const hr_dn_image* image, float* ms);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on n images of one
 * size for the loss chosen by hr_denoiser_set_loss (default tone-mapped L1).
 * loss (optional) receives the loss before the step. */
hr_status hr_denoiser_train_step(hr_denoiser* dn, const hr_dn_image* images, uint32_t n, float lr, float* loss);
/* Loss and FP32 weight gradients (weight exchange layout, on the host) of n
 * images without updating the weights. */
hr_status hr_denoiser_gradients(hr_denoiser* dn, const hr_dn_image* images, uint32_t n, float* loss, float* grad,
                                uint64_t count);

/* One 3x3 convolution layer (the denoiser's building block) on device BF16
 * NHWC tensors, zero padding. weight: host FP32 W[cout][3][3][cin], bias:
 * host FP32 [cout] or NULL. cin must be a multiple of 16.
 *   forward:  y[n][h][w][cout] = relu?(conv(x, W) + b)
 *   backward: dx = gradient of conv(x, W) w.r.t. x for output gradient dy,
 *             dweight / dbias (host FP32, W's layout) the weight and bias
 *             gradients; cout must be a multiple of 16. */
hr_status hr_dn_conv3x3_forward(hr_context* ctx, const void* x, uint32_t n, uint32_t h, uint32_t w, uint32_t cin,
                                const float* weight, const float* bias, uint32_t cout, int relu, void* y);
hr_status hr_dn_conv3x3_backward(hr_context* ctx, const void* x, const void* dy, uint32_t n, uint32_t h, uint32_t w,
                                 uint32_t cin, const float* weight, uint32_t cout, void* dx, float* dweight,
                                 float* dbias);

#ifdef __cplusplus
}
#endif
#endif /* HR_HR_H */
