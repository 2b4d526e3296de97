// hr_api.hip - the extern "C" entry points of include/hr/hr.h.
//
// Host side: argument validation (errors are returned with a message, never
// aborted on), the upload of the scene inputs in the layouts the kernels read,
// the per-primitive constants (normal matrices), and kernel launches on the
// context's stream. No computation of the hybrid pass happens on the host.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_internal.h"

using namespace hr;

struct hr_scene {
    hr_context* ctx = nullptr;
    uint32_t mesh_count = 0, prim_count = 0, ntri = 0, nmat = 0, leaf_max = 2, depth = 0, node_count = 0;
    float ms_build = 0.0f;
    std::vector<PrimInfo> h_prims;
    std::vector<float> h_models;
    std::vector<uint32_t> h_prim_tri_offset;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // size of each allocs entry
    uint8_t* d_verts = nullptr;
    float4* d_vnormal = nullptr;
    uint32_t* d_indices = nullptr;
    PrimInfo* d_prims = nullptr;
    float* d_models = nullptr;
    uint32_t* d_prim_tri_offset = nullptr;
    hr_material* d_mats = nullptr;
    float4* d_sky = nullptr;
    float4* d_irr = nullptr;
    uint32_t sky_w = 0, sky_h = 0, irr_w = 0, irr_h = 0;
    uint32_t* d_texels = nullptr;  // owned here, replaced by hr_scene_set_textures
    TexInfo* d_tex = nullptr;
    uint32_t ntex = 0;
    int64_t max_tex = -1;          // largest texture index the materials use
    // bumped by every setter that replaces device buffers or what the
    // materials index (materials, textures, environment); a captured graph
    // refuses to replay a scene whose generation moved
    uint64_t generation = 0;
    BuildBuffers bb{};
};

namespace {

template <class T>
hr_status scene_alloc(hr_scene* s, T** p, size_t count) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 64;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "hipMalloc");
    s->allocs.push_back(q);
    s->alloc_bytes.push_back(bytes);
    *p = static_cast<T*>(q);
    return HR_OK;
}
// releases one buffer of s->allocs (device idle: the caller synchronised)
void scene_free(hr_scene* s, void* p) {
    if (!p) return;
    for (size_t i = 0; i < s->allocs.size(); ++i) {
        if (s->allocs[i] == p) {
            (void)hipFree(p);
            s->allocs.erase(s->allocs.begin() + (long)i);
            s->alloc_bytes.erase(s->alloc_bytes.begin() + (long)i);
            return;
        }
    }
}
// transpose(inverse(mat3(model))) as cofactors / determinant, row-major
// (hybridHit.rchit:50); evaluated as written (no contraction)
void normal_matrix(const float* m, float* nm) {
    float A00 = m[0], A01 = m[4], A02 = m[8];
    float A10 = m[1], A11 = m[5], A12 = m[9];
    float A20 = m[2], A21 = m[6], A22 = m[10];
    float C00 = A11 * A22 - A12 * A21;
    float C01 = A12 * A20 - A10 * A22;
    float C02 = A10 * A21 - A11 * A20;
    float C10 = A02 * A21 - A01 * A22;
    float C11 = A00 * A22 - A02 * A20;
    float C12 = A01 * A20 - A00 * A21;
    float C20 = A01 * A12 - A02 * A11;
    float C21 = A02 * A10 - A00 * A12;
    float C22 = A00 * A11 - A01 * A10;
    float det = (A00 * C00 + A01 * C01) + A02 * C02;
    float inv = 1.0f / det;
    nm[0] = C00 * inv; nm[1] = C01 * inv; nm[2] = C02 * inv;
    nm[3] = C10 * inv; nm[4] = C11 * inv; nm[5] = C12 * inv;
    nm[6] = C20 * inv; nm[7] = C21 * inv; nm[8] = C22 * inv;
}

DevScene dev_scene(const hr_scene* s) {
    DevScene d;
    d.nodes = s->bb.nodes;
    d.nodes4 = s->bb.nodes4;
    d.tris = s->bb.stris;
    d.tri_prim = s->bb.tri_prim;
    d.prims = s->d_prims;
    d.indices = s->d_indices;
    d.vnormal = s->d_vnormal;
    d.verts = s->d_verts;
    d.wtri = s->bb.wtri;
    d.texels = s->d_texels;
    d.tex = s->d_tex;
    d.ntex = s->ntex;
    d.mats = s->d_mats;
    d.sky = s->d_sky;
    d.irr = s->d_irr;
    d.node_count = s->node_count;
    d.wide_count = s->bb.wide_count;
    d.tri_count = s->ntri;
    d.nmat = s->nmat;
    d.prim_count = s->prim_count;
    d.sky_w = s->sky_w;
    d.sky_h = s->sky_h;
    d.irr_w = s->irr_w;
    d.irr_h = s->irr_h;
    return d;
}

hr_status check_material(hr_context* ctx, const hr_material& m, uint32_t i, uint32_t nmat) {
    int mode = (int)m.shading_metal_rough_idx[0];
    if ((float)mode != m.shading_metal_rough_idx[0] || (mode != 0 && mode != 3 && mode != 4))
        return fail(ctx, HR_E_INVALID, "material %u: shading model %g is not 0, 3 or 4", i,
                    m.shading_metal_rough_idx[0]);
    // a texture slot is used when its value is above -1 (hybridHit.rchit:63-65)
    for (int k = 0; k < 4; ++k) {
        float t = m.textures[k];
        if (t > -1.0f && !(t < 16777216.0f && (float)(int)t == t))
            return fail(ctx, HR_E_INVALID, "material %u: texture index %g is not a whole number", i, t);
    }
    float w = m.shading_metal_rough_idx[3];
    if (!(w >= 0.0f) || w >= (float)nmat || (float)(int)w != w)
        return fail(ctx, HR_E_INVALID, "material %u: material index %g out of range", i, w);
    for (int k = 0; k < 12; ++k) {
        const float* f = &m.diffuse_ior[0];
        if (!std::isfinite(f[k])) return fail(ctx, HR_E_INVALID, "material %u: non-finite value", i);
    }
    return HR_OK;
}

int64_t max_texture(const hr_material* m, uint32_t n) {
    int64_t mx = -1;
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k)
            if (m[i].textures[k] > -1.0f && (int64_t)m[i].textures[k] > mx) mx = (int64_t)m[i].textures[k];
    return mx;
}
// every texture a material names must be set before a pass that shades
hr_status check_textures(hr_context* ctx, const hr_scene* s) {
    if (s->max_tex >= (int64_t)s->ntex)
        return fail(ctx, HR_E_INVALID, "materials use texture %lld but %u textures are set",
                    (long long)s->max_tex, s->ntex);
    return HR_OK;
}

hr_status fill_frame(hr_context* ctx, FrameArgs& f, const hr_camera* cam, uint32_t W, uint32_t H) {
    memset(&f, 0, sizeof(f));
    if (!cam) return fail(ctx, HR_E_INVALID, "camera is NULL");
    memcpy(f.inv_view, cam->inv_view, sizeof(f.inv_view));
    memcpy(f.inv_proj, cam->inv_proj, sizeof(f.inv_proj));
    f.frame = (uint32_t)(int)cam->frame[0];  // uint(int(cam.frame.x)), hybridRaygen.rgen:29
    f.width = W;
    f.height = H;
    return HR_OK;
}
hr_status fill_lights(hr_context* ctx, FrameArgs& f, const hr_light* lights, const hr_light_shape* shapes,
                      uint32_t nl, const hr_trace_params* prm) {
    if (!lights && nl != 0) return fail(ctx, HR_E_INVALID, "lights is NULL");
    if (nl > HR_MAX_LIGHTS) return fail(ctx, HR_E_UNSUPPORTED, "%u lights exceed HR_MAX_LIGHTS", nl);
    if (!prm) return fail(ctx, HR_E_INVALID, "trace params are NULL");
    if (prm->shadow_samples == 0 || prm->shadow_samples > 4096)
        return fail(ctx, HR_E_INVALID, "shadow_samples must be in 1..4096");
    if ((prm->flags & HR_FLAG_ADAPTIVE) && (prm->shadow_samples > 64 || (prm->shadow_samples & (prm->shadow_samples - 1u))))
        return fail(ctx, HR_E_INVALID, "adaptive sampling takes a power of two of at most 64 shadow samples");
    if (prm->refl_spp > 4096 || prm->max_depth > 64)
        return fail(ctx, HR_E_INVALID, "refl_spp must be <= 4096 and max_depth <= 64");
    for (uint32_t i = 0; i < nl; ++i) {
        f.lights[i] = lights[i];
        const float* v = lights[i].pos_maxdist;
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(v[k])) return fail(ctx, HR_E_INVALID, "light %u: non-finite value", i);
        if (shapes) {
            if (shapes[i].type > HR_LIGHT_RECT) return fail(ctx, HR_E_INVALID, "light %u: unknown shape", i);
            f.shapes[i] = shapes[i];
        }
    }
    f.nl = nl;
    f.has_shapes = shapes ? 1u : 0u;
    f.samples = prm->shadow_samples;
    f.spp = prm->refl_spp;
    f.max_depth = prm->max_depth;
    f.adaptive = (prm->flags & HR_FLAG_ADAPTIVE) ? 1u : 0u;
    const uint32_t known = HR_FLAG_COUNT_VISITS | HR_FLAG_TIMING | HR_FLAG_ADAPTIVE | HR_FLAG_SHADOW_R8 |
                           HR_FLAG_LIT_BGRA8;
    if (prm->flags & ~known) return fail(ctx, HR_E_INVALID, "unknown trace flags 0x%x", prm->flags & ~known);
    f.unorm8 = ((prm->flags & HR_FLAG_SHADOW_R8) ? UNORM8_SHADOW : 0u) | ((prm->flags & HR_FLAG_LIT_BGRA8) ? UNORM8_LIT : 0u);
    return HR_OK;
}
hr_status check_gbuffer(hr_context* ctx, const hr_gbuffer* gb) {
    if (!gb || !gb->position || !gb->normal || !gb->albedo || !gb->material || !gb->emissive)
        return fail(ctx, HR_E_INVALID, "G-buffer planes must be non-NULL device pointers");
    if (gb->width == 0 || gb->height == 0 || gb->width > 16384 || gb->height > 16384)
        return fail(ctx, HR_E_INVALID, "G-buffer size %ux%u out of range", gb->width, gb->height);
    return HR_OK;
}
GBufView gview(const hr_gbuffer* gb) {
    return GBufView{reinterpret_cast<float4*>(gb->position), reinterpret_cast<float4*>(gb->normal),
                    reinterpret_cast<float4*>(gb->albedo), reinterpret_cast<float4*>(gb->material),
                    reinterpret_cast<float4*>(gb->emissive)};
}

// Begin / end of a frame call: zero the counters and time with events; with
// stats the call synchronises and reports counts and milliseconds.
hr_status frame_begin(hr_context* ctx, bool timing) {
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, launch_zero_u64(ctx->d_counters, C_COUNT * C_STRIDE, ctx->stream));
    if (timing) HR_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    return HR_OK;
}
hr_status frame_end(hr_context* ctx, hr_stats* stats, float* ms_slot) {
    if (!stats) return HR_OK;
    HR_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    HR_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
    unsigned long long cs[C_COUNT * C_STRIDE], c[C_COUNT];
    HR_HIP(ctx, hipMemcpy(cs, ctx->d_counters, sizeof(cs), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < C_COUNT; ++k) c[k] = cs[k * C_STRIDE];
    float ms = 0.0f;
    HR_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    float keep[4] = {stats->ms_gbuffer, stats->ms_shadow, stats->ms_shade, stats->ms_build};
    memset(stats, 0, sizeof(*stats));
    stats->ms_gbuffer = keep[0];
    stats->ms_shadow = keep[1];
    stats->ms_shade = keep[2];
    stats->ms_build = keep[3];
    stats->primary_rays = c[C_PRIMARY];
    stats->shadow_rays = c[C_SHADOW];
    stats->reflection_rays = c[C_REFL];
    stats->hit_shadow_rays = c[C_HITSHADOW];
    stats->occluded_rays = c[C_OCCLUDED];
    stats->node_visits = c[C_NODES];
    stats->tri_tests = c[C_TRIS];
    if (ms_slot) *ms_slot = ms;
    if (c[C_OVERFLOW]) return fail(ctx, HR_E_OVERFLOW, "traversal stack overflow");
    return HR_OK;
}

// Workspace budget of one slab of the reflection pipeline; larger frames are
// processed in slabs of pixels.

// Grows the context's workspace to `bytes`. With `oom` non-NULL an
// out-of-memory failure is reported there (and cleared) instead of failing.
hr_status ensure_work(hr_context* ctx, size_t bytes, bool* oom = nullptr) {
    if (oom) *oom = false;
    if (bytes <= ctx->work_bytes) return HR_OK;
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->work) {
        HR_HIP(ctx, hipFree(ctx->work));
        ctx->work = nullptr;
        ctx->work_bytes = 0;
    }
    hipError_t e = hipMalloc(&ctx->work, bytes);
    if (e == hipErrorOutOfMemory && oom) {
        (void)hipGetLastError();  // not sticky: clear it for the next call
        ctx->work = nullptr;
        *oom = true;
        return HR_OK;
    }
    if (e != hipSuccess) {
        ctx->work = nullptr;
        return hip_fail(ctx, e, "hipMalloc workspace");
    }
    ctx->work_bytes = bytes;
    return HR_OK;
}

// Layout of the reflection workspace for frame f: returns its size and, with
// a base address, fills w. Frames larger than the budget run in pixel slabs.
size_t wave_layout(const hr_context* ctx, const FrameArgs& f, WaveArgs& w, char* b, uint32_t budget_mb) {
    uint32_t spp_max = f.spp > 1 ? f.spp : 1;
    size_t per_slot = wave_bytes_per_slot(f.nl);
    size_t npix = (size_t)f.width * f.height;
    size_t budget = (size_t)budget_mb << 20;
    size_t slab = budget / (per_slot * spp_max + 32);
    if (slab < 256) slab = 256;
    slab = slab < npix ? slab : npix;
    slab = (slab + 255) / 256 * 256;
    size_t slots = slab * spp_max;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    size_t o_p0 = carve(slots * 48), o_p1 = carve(slots * 48), o_hits = carve(slots * 16);
    size_t o_sr = carve(slots * f.nl * 32), o_vis = carve(slots * 4), o_pend = carve(slots * 4);
    size_t o_col = carve(slots * 16), o_pix = carve(slab * 32), o_cnt = carve(4 * BAND_STRIDE * 4);
    size_t o_take = carve(2 * BAND_COUNTER_BYTES);
    if (b) {
        w.paths0 = reinterpret_cast<float4*>(b + o_p0);
        w.paths1 = reinterpret_cast<float4*>(b + o_p1);
        w.hits = reinterpret_cast<float4*>(b + o_hits);
        w.srays = reinterpret_cast<float4*>(b + o_sr);
        w.vis = reinterpret_cast<uint32_t*>(b + o_vis);
        w.pending = reinterpret_cast<uint32_t*>(b + o_pend);
        w.pathcol = reinterpret_cast<float4*>(b + o_col);
        w.pix = reinterpret_cast<float4*>(b + o_pix);
        w.counts = reinterpret_cast<uint32_t*>(b + o_cnt);
        w.takes = reinterpret_cast<uint32_t*>(b + o_take);
        w.p0 = 0;
        w.p1 = 0;
        w.spp_max = spp_max;
        w.slab_pixels = (uint32_t)slab;
        // the trace stages' persistent grid, at most one lane per queue slot:
        // every wave of it takes (and at the end fails) at least once
        w.blocks = (uint32_t)std::min<size_t>(ctx->resident_blocks, (slots + 255) / 256);
    }
    return slots >= (size_t(1) << 28) ? 0 : off;
}

// The reflection workspace budget a frame runs with: the tuned budget, halved
// while the device cannot hold the workspace it needs (64 MiB at least).
// Smaller budgets only cut the frame into more pixel slabs. `need` gives the
// workspace bytes at a budget (0: the frame does not fit any slab).
template <class Need>
hr_status fit_budget(hr_context* ctx, Need need, uint32_t& budget_mb) {
    budget_mb = ctx->tuning.wave_budget_mb;
    for (;;) {
        const size_t bytes = need(budget_mb);
        if (bytes == 0) return fail(ctx, HR_E_UNSUPPORTED, "reflection slab too large");
        bool oom = false;
        HR_TRY(ensure_work(ctx, bytes, budget_mb > 64u ? &oom : nullptr));
        if (!oom) return HR_OK;
        budget_mb = budget_mb / 2u < 64u ? 64u : budget_mb / 2u;
    }
}

hr_status wave_args(hr_context* ctx, const FrameArgs& f, WaveArgs& w) {
    uint32_t mb = 0;
    HR_TRY(fit_budget(ctx, [&](uint32_t b) { return wave_layout(ctx, f, w, nullptr, b); }, mb));
    wave_layout(ctx, f, w, static_cast<char*>(ctx->work), mb);
    return HR_OK;
}

hr_status ensure_jobs(hr_context* ctx, uint32_t n) {
    if (n <= ctx->jobs_cap) return HR_OK;
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_jobs) {
        HR_HIP(ctx, hipFree(ctx->d_jobs));
        ctx->d_jobs = nullptr;
        ctx->jobs_cap = 0;
    }
    HR_HIP(ctx, hipMalloc(&ctx->d_jobs, n * sizeof(ShadowJob)));
    ctx->jobs_cap = n;
    return HR_OK;
}

// One frame of a batch, validated and converted to kernel arguments.
struct FramePlan {
    hr_scene* scene;
    FrameArgs f;
    GBufView g;
    float* shadow;
    float* lit;
    float* refl;
    bool count;
};

hr_status plan_frame(hr_context* ctx, const hr_frame_desc& d, FramePlan& fp, bool need_shading) {
    if (!d.scene) return fail(ctx, HR_E_INVALID, "frame without a scene");
    if (d.scene->ctx != ctx) return fail(ctx, HR_E_INVALID, "scene belongs to another context");
    HR_TRY(check_gbuffer(ctx, &d.gbuffer));
    if (!d.shadow) return fail(ctx, HR_E_INVALID, "shadow output is NULL");
    if (need_shading && (!d.lit || !d.refl)) return fail(ctx, HR_E_INVALID, "lit/refl output is NULL");
    HR_TRY(fill_frame(ctx, fp.f, &d.camera, d.gbuffer.width, d.gbuffer.height));
    HR_TRY(fill_lights(ctx, fp.f, d.lights, d.shapes, d.light_count, &d.params));
    if (need_shading && d.params.refl_spp > 64) return fail(ctx, HR_E_UNSUPPORTED, "refl_spp above 64");
    if (need_shading) HR_TRY(check_textures(ctx, d.scene));
    fp.scene = d.scene;
    fp.g = gview(&d.gbuffer);
    fp.shadow = d.shadow;
    fp.lit = d.lit;
    fp.refl = d.refl;
    fp.count = (d.params.flags & HR_FLAG_COUNT_VISITS) != 0;
    return HR_OK;
}

// The shadow job of one frame whose groups are numbered from `base`;
// returns the frame's group count.
uint32_t fill_job(const hr_context* ctx, const FramePlan& fp, uint32_t base, ShadowJob& j) {
    const uint32_t tw = ctx->tuning.tile_w;
    memset(&j, 0, sizeof(j));
    j.S = dev_scene(fp.scene);
    j.F = fp.f;
    j.G = fp.g;
    j.shadow = fp.shadow;
    j.group_base = base;
    j.tile_w = tw;
    j.tiles_x = (fp.f.width + tw - 1u) / tw;
    return shadow_groups(fp.f, tw);
}
// Shadow jobs of a batch (host side); returns the total group count.
uint32_t make_jobs(const hr_context* ctx, const std::vector<FramePlan>& plans, std::vector<ShadowJob>& jobs) {
    uint32_t base = 0;
    jobs.resize(plans.size());
    for (size_t i = 0; i < plans.size(); ++i) base += fill_job(ctx, plans[i], base, jobs[i]);
    return base;
}

// The sample-parallel shadow kernels take timed frames whose sample counts are
// powers of two in 4..64, all adaptive or none; items are spaced by the
// largest count. Other batches use the group-per-lane kernel.
struct SplPlan {
    bool use = false;
    bool adaptive = false;
    uint32_t lg2s = 0;
    size_t scratch = 0;
};

SplPlan plan_spl(const FramePlan* plans, size_t nplans, uint32_t ngroups) {
    SplPlan sp;
    uint32_t smax = 0;
    for (size_t i = 0; i < nplans; ++i) {
        const FrameArgs& f = plans[i].f;
        const uint32_t S = f.samples;
        if (plans[i].count || S < 4u || S > 64u || (S & (S - 1u)) != 0u) return sp;
        if (i > 0 && (f.adaptive != 0u) != sp.adaptive) return sp;
        sp.adaptive = f.adaptive != 0u;
        smax = S > smax ? S : smax;
    }
    while ((1u << sp.lg2s) < smax) ++sp.lg2s;
    if (((uint64_t)ngroups << sp.lg2s) >= (1ull << 32)) return sp;
    sp.scratch = shadows_spl_scratch(ngroups, sp.lg2s, sp.adaptive);
    sp.use = true;
    return sp;
}

// Kernels of a shadow pass (device jobs already uploaded); scratch holds
// sp.scratch bytes when sp.use.
// (wide_count: the 4-wide nodes of the first frame's scene; used for one frame)
hr_status enqueue_shadows(hr_context* ctx, const ShadowJob* d_jobs, uint32_t n, uint32_t ngroups, const SplPlan& sp,
                          bool count, uint32_t* next_group, uint32_t* scratch, uint32_t wide_count) {
    if (sp.use)
        HR_HIP(ctx, launch_shadows_spl(d_jobs, n, ngroups, sp.lg2s, sp.adaptive, ctx->d_counters, ctx->stream,
                                       next_group, scratch, ctx->resident_blocks, wide_count));
    else
        HR_HIP(ctx, launch_shadows(d_jobs, n, ngroups, ctx->d_counters, count, ctx->stream, next_group,
                                   ctx->resident_blocks));
    return HR_OK;
}

// ---- scratch of the allocation-free shadow pass ----------------------------------
// One array per GPU, held until the process ends: [shadow job | band counters
// | sample bits (| probe bits, refine list, count)]. In host code `scratch`
// names a host-side stand-in, never the device array: kernels and copies use
// the device address that hipGetSymbolAddress gives on each device.
constexpr size_t NOALLOC_BYTES = HR_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];
constexpr size_t NOALLOC_JOB_BYTES = (sizeof(ShadowJob) + 255u) / 256u * 256u;
constexpr size_t NOALLOC_BITS_OFFSET = NOALLOC_JOB_BYTES + BAND_COUNTER_BYTES;
constexpr int NOALLOC_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_DEVICES] = {};

// Device address of `scratch` on `device`, looked up once per device with the
// device current (the caller has made it current).
hipError_t noalloc_scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

uint32_t hr_abi_version(void) { return HR_ABI_VERSION; }

hr_status hr_context_create(int device, hr_context** out) {
    if (!out) return HR_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return HR_E_DEVICE;
    if (device < 0 || device >= n) return HR_E_INVALID;
    hr_context* ctx = new hr_context();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev[1]);
    for (int k = 0; k < SHADE_STREAMS && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_ev[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&ctx->d_counters, C_COUNT * C_STRIDE * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&ctx->d_next_group, BAND_COUNTER_BYTES);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            ctx->cu_count = (uint32_t)prop.multiProcessorCount;
        ctx->resident_blocks = ctx->cu_count * ctx->tuning.blocks_per_cu;
        // only hr_trace_shadows_noalloc needs it: a failure is kept, not fatal
        ctx->noalloc_err = noalloc_scratch_address(device, &ctx->noalloc_scratch);
        if (ctx->noalloc_err != hipSuccess) (void)hipGetLastError();
    }
    if (e != hipSuccess) {
        hr_context_destroy(ctx);
        return e == hipErrorOutOfMemory ? HR_E_OOM : HR_E_DEVICE;
    }
    *out = ctx;
    return HR_OK;
}

void hr_context_destroy(hr_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_next_group) (void)hipFree(ctx->d_next_group);
    if (ctx->d_jobs) (void)hipFree(ctx->d_jobs);
    if (ctx->work) (void)hipFree(ctx->work);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < SHADE_STREAMS; ++k) {
        if (ctx->side[k]) {
            (void)hipStreamSynchronize(ctx->side[k]);
            (void)hipStreamDestroy(ctx->side[k]);
        }
        if (ctx->join_ev[k]) (void)hipEventDestroy(ctx->join_ev[k]);
    }
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* hr_last_error(const hr_context* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

hr_status hr_synchronize(hr_context* ctx) {
    if (!ctx) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HR_OK;
}

hr_status hr_device_alloc(hr_context* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    HR_HIP(ctx, hipMalloc(out, bytes ? bytes : 1));
    return HR_OK;
}
hr_status hr_device_free(hr_context* ctx, void* p) {
    if (!ctx) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (p) HR_HIP(ctx, hipFree(p));
    return HR_OK;
}
hr_status hr_memcpy_h2d(hr_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (bytes) HR_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return HR_OK;
}
hr_status hr_memcpy_d2h(hr_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (bytes) HR_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return HR_OK;
}
hr_status hr_memset_d(hr_context* ctx, void* dst, int value, size_t bytes) {
    if (!ctx || (bytes && !dst)) return HR_E_INVALID;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    if (bytes) HR_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HR_OK;
}

void hr_scene_destroy(hr_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->d_texels) (void)hipFree(s->d_texels);
    if (s->d_tex) (void)hipFree(s->d_tex);
    delete s;
}

hr_status hr_scene_build(hr_context* ctx, const hr_mesh* meshes, uint32_t mesh_count, const hr_primitive* prims,
                         uint32_t prim_count, const hr_material* materials, uint32_t material_count,
                         const hr_build_params* params, hr_scene** out) {
    if (!ctx) return HR_E_INVALID;
    if (!out) return fail(ctx, HR_E_INVALID, "out is NULL");
    *out = nullptr;
    if ((mesh_count && !meshes) || (prim_count && !prims))
        return fail(ctx, HR_E_INVALID, "meshes/primitives pointer is NULL");
    if (!materials || material_count == 0) return fail(ctx, HR_E_INVALID, "at least one material is required");
    uint32_t leaf_max = params && params->leaf_max ? params->leaf_max : 2u;
    if (leaf_max > LEAF_MAX_LIMIT) return fail(ctx, HR_E_INVALID, "leaf_max %u exceeds %u", leaf_max, LEAF_MAX_LIMIT);
    for (uint32_t i = 0; i < material_count; ++i) HR_TRY(check_material(ctx, materials[i], i, material_count));

    // concatenated vertices / indices with per-mesh bases
    std::vector<uint64_t> vbase(mesh_count + 1, 0), ibase(mesh_count + 1, 0);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        if ((meshes[m].vertex_count && !meshes[m].vertices) || (meshes[m].index_count && !meshes[m].indices))
            return fail(ctx, HR_E_INVALID, "mesh %u: NULL arrays", m);
        vbase[m + 1] = vbase[m] + meshes[m].vertex_count;
        ibase[m + 1] = ibase[m] + meshes[m].index_count;
        for (uint32_t v = 0; v < meshes[m].vertex_count; ++v) {
            const float* p = meshes[m].vertices[v].position;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
                return fail(ctx, HR_E_INVALID, "mesh %u vertex %u: non-finite position", m, v);
        }
    }
    if (vbase[mesh_count] >= 0xFFFFFFFFull || ibase[mesh_count] >= 0xFFFFFFFFull)
        return fail(ctx, HR_E_UNSUPPORTED, "too many vertices or indices");

    hr_scene* s = new hr_scene();
    s->ctx = ctx;
    s->mesh_count = mesh_count;
    s->prim_count = prim_count;
    s->nmat = material_count;
    s->max_tex = max_texture(materials, material_count);
    s->leaf_max = leaf_max;
    uint64_t tri = 0;
    s->h_prims.resize(prim_count);
    s->h_models.resize((size_t)prim_count * 16);
    s->h_prim_tri_offset.resize(prim_count + 1);
    for (uint32_t i = 0; i < prim_count; ++i) {
        const hr_primitive& hp = prims[i];
        if (hp.mesh >= mesh_count || hp.index_count % 3 != 0 ||
            (uint64_t)hp.first_index + hp.index_count > meshes[hp.mesh].index_count || hp.material < 0 ||
            (uint32_t)hp.material >= material_count) {
            hr_scene_destroy(s);
            return fail(ctx, HR_E_INVALID, "primitive %u: mesh, index range or material out of range", i);
        }
        const hr_mesh& mh = meshes[hp.mesh];
        for (uint32_t k = 0; k < hp.index_count; ++k) {
            if (mh.indices[hp.first_index + k] >= mh.vertex_count) {
                hr_scene_destroy(s);
                return fail(ctx, HR_E_INVALID, "primitive %u: vertex index out of range", i);
            }
        }
        for (int k = 0; k < 16; ++k) {
            if (!std::isfinite(hp.model[k])) {
                hr_scene_destroy(s);
                return fail(ctx, HR_E_INVALID, "primitive %u: non-finite model matrix", i);
            }
        }
        PrimInfo& p = s->h_prims[i];
        memset(&p, 0, sizeof(p));
        p.tri_offset = (uint32_t)tri;
        p.index_base = (uint32_t)(ibase[hp.mesh] + hp.first_index);
        p.vertex_base = (uint32_t)vbase[hp.mesh];
        p.material = hp.material;
        normal_matrix(hp.model, p.nm);
        memcpy(&s->h_models[(size_t)i * 16], hp.model, 16 * sizeof(float));
        s->h_prim_tri_offset[i] = (uint32_t)tri;
        tri += hp.index_count / 3;
    }
    if (tri >= MAX_TRIS) {
        hr_scene_destroy(s);
        return fail(ctx, HR_E_UNSUPPORTED, "%llu triangles exceed the 2^27 limit", (unsigned long long)tri);
    }
    s->h_prim_tri_offset[prim_count] = (uint32_t)tri;
    s->ntri = (uint32_t)tri;
    uint32_t n = s->ntri;
    s->node_count = n >= 2 ? n - 1 : (n == 1 ? 1 : 0);

    hr_status st = HR_OK;
    auto check = [&](hr_status x) {
        if (st == HR_OK) st = x;
    };
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) {
        hr_scene_destroy(s);
        return hip_fail(ctx, he, "hipSetDevice");
    }
    uint64_t nv = vbase[mesh_count], ni = ibase[mesh_count];
    check(scene_alloc(s, &s->d_verts, nv * sizeof(hr_vertex)));
    check(scene_alloc(s, &s->d_vnormal, nv));
    check(scene_alloc(s, &s->d_indices, ni));
    check(scene_alloc(s, &s->d_prims, prim_count));
    check(scene_alloc(s, &s->d_models, (size_t)prim_count * 16));
    check(scene_alloc(s, &s->d_prim_tri_offset, prim_count + 1));
    check(scene_alloc(s, &s->d_mats, material_count));
    BuildBuffers& b = s->bb;
    size_t nn = n >= 2 ? n - 1 : 1;
    check(scene_alloc(s, &b.wtri, (size_t)n * 3));
    check(scene_alloc(s, &b.tbox, (size_t)n * 2));
    check(scene_alloc(s, &b.keys, n));
    check(scene_alloc(s, &b.ids, n));
    check(scene_alloc(s, &b.keys_tmp, n));
    check(scene_alloc(s, &b.ids_tmp, n));
    check(scene_alloc(s, &b.hist, radix_hist_words(n)));
    check(scene_alloc(s, &b.bounds, 8));
    check(scene_alloc(s, &b.kl, nn));
    check(scene_alloc(s, &b.kr, nn));
    check(scene_alloc(s, &b.kfirst, nn));
    check(scene_alloc(s, &b.klast, nn));
    check(scene_alloc(s, &b.kparent, nn));
    check(scene_alloc(s, &b.lparent, n));
    check(scene_alloc(s, &b.flags, nn));
    check(scene_alloc(s, &b.lvl, nn + 2 * (size_t)REFIT_LEVELS + 1));
    check(scene_alloc(s, &b.kbox, nn * 2));
    check(scene_alloc(s, &b.nodes, nn));
    check(scene_alloc(s, &b.stris, n));
    check(scene_alloc(s, &b.tri_prim, n));
    check(scene_alloc(s, &b.depth, 1));
    if (st != HR_OK) {
        hr_scene_destroy(s);
        return st;
    }
    // uploads: vertices as-is (44-byte stride), normals as float4 (the
    // rtVertexAttribute repack of renderer.cpp:3018-3048), indices as-is
    {
        std::vector<float4> nrm(nv);
        std::vector<uint32_t> idx(ni);
        for (uint32_t m = 0; m < mesh_count; ++m) {
            if (meshes[m].vertex_count) {
                he = hipMemcpy(s->d_verts + vbase[m] * sizeof(hr_vertex), meshes[m].vertices,
                               (size_t)meshes[m].vertex_count * sizeof(hr_vertex), hipMemcpyHostToDevice);
                if (he != hipSuccess) break;
            }
            for (uint32_t v = 0; v < meshes[m].vertex_count; ++v) {
                const float* q = meshes[m].vertices[v].normal;
                nrm[vbase[m] + v] = make_float4(q[0], q[1], q[2], 0.0f);
            }
            if (meshes[m].index_count)
                memcpy(&idx[ibase[m]], meshes[m].indices, (size_t)meshes[m].index_count * sizeof(uint32_t));
        }
        if (he == hipSuccess && nv)
            he = hipMemcpy(s->d_vnormal, nrm.data(), nv * sizeof(float4), hipMemcpyHostToDevice);
        if (he == hipSuccess && ni)
            he = hipMemcpy(s->d_indices, idx.data(), ni * sizeof(uint32_t), hipMemcpyHostToDevice);
    }
    if (he == hipSuccess && prim_count)
        he = hipMemcpy(s->d_prims, s->h_prims.data(), prim_count * sizeof(PrimInfo), hipMemcpyHostToDevice);
    if (he == hipSuccess && prim_count)
        he = hipMemcpy(s->d_models, s->h_models.data(), s->h_models.size() * sizeof(float), hipMemcpyHostToDevice);
    if (he == hipSuccess)
        he = hipMemcpy(s->d_prim_tri_offset, s->h_prim_tri_offset.data(), (prim_count + 1) * sizeof(uint32_t),
                       hipMemcpyHostToDevice);
    if (he == hipSuccess)
        he = hipMemcpy(s->d_mats, materials, material_count * sizeof(hr_material), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        hr_scene_destroy(s);
        return hip_fail(ctx, he, "scene upload");
    }
    b.verts = s->d_verts;
    b.indices = s->d_indices;
    b.prims = s->d_prims;
    b.models = s->d_models;
    b.prim_tri_offset = s->d_prim_tri_offset;
    b.prim_count = prim_count;
    b.tri_count = n;
    b.leaf_max = leaf_max;

    he = hipEventRecord(ctx->ev[0], ctx->stream);
    if (he == hipSuccess) he = lbvh_build(b, ctx->stream, !(ctx->tuning.bvh_flags & HR_BVH_NO_TREELETS));
    if (he == hipSuccess && n > 0) {  // the 4-wide nodes, dense: as many as the build chose
        st = HR_OK;
        check(scene_alloc(s, &b.nodes4, b.wide_count));
        check(scene_alloc(s, &b.wsrc, b.wide_count));
        if (st != HR_OK) {
            hr_scene_destroy(s);
            return st;
        }
        he = lbvh_build_wide(b, ctx->stream);
    }
    if (he == hipSuccess) he = hipEventRecord(ctx->ev[1], ctx->stream);
    if (he == hipSuccess) he = hipEventSynchronize(ctx->ev[1]);
    if (he == hipSuccess) he = hipEventElapsedTime(&s->ms_build, ctx->ev[0], ctx->ev[1]);
    uint32_t depth = 0;
    if (he == hipSuccess && n > 0) he = hipMemcpy(&depth, b.depth, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (he != hipSuccess) {
        hr_scene_destroy(s);
        return hip_fail(ctx, he, "LBVH build");
    }
    s->depth = depth;
    if (depth > MAX_DEPTH) {
        hr_scene_destroy(s);
        return fail(ctx, HR_E_OVERFLOW, "BVH depth %u exceeds the traversal stack (%u)", depth, MAX_DEPTH);
    }
    *out = s;
    return HR_OK;
}

hr_status hr_scene_refit(hr_scene* s, const float* models) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (!models) return fail(ctx, HR_E_INVALID, "models is NULL");
    for (uint32_t i = 0; i < s->prim_count; ++i) {
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(models[16 * i + k]))
                return fail(ctx, HR_E_INVALID, "primitive %u: non-finite model matrix", i);
    }
    for (uint32_t i = 0; i < s->prim_count; ++i) {
        memcpy(&s->h_models[(size_t)i * 16], models + 16 * (size_t)i, 16 * sizeof(float));
        normal_matrix(models + 16 * (size_t)i, s->h_prims[i].nm);
    }
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->prim_count) {
        HR_HIP(ctx, hipMemcpy(s->d_models, s->h_models.data(), s->h_models.size() * sizeof(float),
                              hipMemcpyHostToDevice));
        HR_HIP(ctx, hipMemcpy(s->d_prims, s->h_prims.data(), s->prim_count * sizeof(PrimInfo),
                              hipMemcpyHostToDevice));
    }
    HR_HIP(ctx, lbvh_refit(s->bb, ctx->stream));
    uint32_t depth = 0;
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->ntri) HR_HIP(ctx, hipMemcpy(&depth, s->bb.depth, sizeof(uint32_t), hipMemcpyDeviceToHost));
    s->depth = depth;
    return HR_OK;
}

hr_status hr_scene_set_materials(hr_scene* s, const hr_material* materials, uint32_t count) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (!materials || count != s->nmat)
        return fail(ctx, HR_E_INVALID, "expected %u materials", s->nmat);
    for (uint32_t i = 0; i < count; ++i) HR_TRY(check_material(ctx, materials[i], i, count));
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HR_HIP(ctx, hipMemcpy(s->d_mats, materials, count * sizeof(hr_material), hipMemcpyHostToDevice));
    s->max_tex = max_texture(materials, count);
    ++s->generation;
    return HR_OK;
}

hr_status hr_scene_set_textures(hr_scene* s, const hr_texture* textures, uint32_t count) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (count && !textures) return fail(ctx, HR_E_INVALID, "textures is NULL");
    std::vector<TexInfo> info(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const hr_texture& t = textures[i];
        if (!t.rgba8 || t.width == 0 || t.height == 0 || t.width > 65536 || t.height > 65536)
            return fail(ctx, HR_E_INVALID, "texture %u: empty, NULL or larger than 65536 texels a side", i);
        info[i].offset = total;
        info[i].w = t.width;
        info[i].h = t.height;
        total += (uint64_t)t.width * t.height;
    }
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t* texels = nullptr;
    TexInfo* tex = nullptr;
    hipError_t e = hipMalloc(&texels, (total ? total : 1) * sizeof(uint32_t));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc textures");
    e = hipMalloc(&tex, (count ? count : 1) * sizeof(TexInfo));
    for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
        e = hipMemcpy(texels + info[i].offset, textures[i].rgba8, (size_t)info[i].w * info[i].h * 4,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && count) e = hipMemcpy(tex, info.data(), count * sizeof(TexInfo), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(texels);
        if (tex) (void)hipFree(tex);
        return hip_fail(ctx, e, "texture upload");
    }
    if (s->d_texels) (void)hipFree(s->d_texels);
    if (s->d_tex) (void)hipFree(s->d_tex);
    s->d_texels = texels;
    s->d_tex = tex;
    s->ntex = count;
    ++s->generation;
    return HR_OK;
}

hr_status hr_scene_set_env(hr_scene* s, const hr_env* env) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (!env || !env->skybox || !env->irradiance || !env->sky_width || !env->sky_height || !env->irr_width ||
        !env->irr_height)
        return fail(ctx, HR_E_INVALID, "environment maps must be non-empty");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t ns = (size_t)env->sky_width * env->sky_height, ni = (size_t)env->irr_width * env->irr_height;
    // a map of another size gets a new buffer; the old one is released once
    // the new contents are in place
    float4* sky = s->d_sky;
    float4* irr = s->d_irr;
    if (ns != (size_t)s->sky_w * s->sky_h) HR_TRY(scene_alloc(s, &sky, ns));
    if (ni != (size_t)s->irr_w * s->irr_h) {
        hr_status st = scene_alloc(s, &irr, ni);
        if (st != HR_OK) {
            if (sky != s->d_sky) scene_free(s, sky);
            return st;
        }
    }
    hipError_t e = hipMemcpy(sky, env->skybox, ns * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(irr, env->irradiance, ni * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (sky != s->d_sky) scene_free(s, sky);
        if (irr != s->d_irr) scene_free(s, irr);
        return hip_fail(ctx, e, "environment upload");
    }
    if (sky != s->d_sky) scene_free(s, s->d_sky);
    if (irr != s->d_irr) scene_free(s, s->d_irr);
    s->d_sky = sky;
    s->d_irr = irr;
    s->sky_w = env->sky_width;
    s->sky_h = env->sky_height;
    s->irr_w = env->irr_width;
    s->irr_h = env->irr_height;
    ++s->generation;
    return HR_OK;
}

hr_status hr_scene_info(const hr_scene* s, uint32_t* tri_count, uint32_t* node_count, uint32_t* depth,
                        float* ms_build) {
    if (!s) return HR_E_INVALID;
    if (tri_count) *tri_count = s->ntri;
    if (node_count) *node_count = s->node_count;
    if (depth) *depth = s->depth;
    if (ms_build) *ms_build = s->ms_build;
    return HR_OK;
}

hr_status hr_scene_memory(const hr_scene* s, uint32_t* wide_count, uint64_t* device_bytes) {
    if (!s) return HR_E_INVALID;
    if (wide_count) *wide_count = s->bb.wide_count;
    if (device_bytes) {
        uint64_t t = 0;
        for (size_t b : s->alloc_bytes) t += b;
        *device_bytes = t;
    }
    return HR_OK;
}

hr_status hr_scene_export(const hr_scene* s, uint32_t which, void* dst, size_t bytes) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    const void* src = nullptr;
    size_t n = 0;
    uint32_t nt = s->ntri;
    std::vector<uint32_t> tmp;
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    switch (which) {
        case 0: src = s->bb.nodes; n = (size_t)s->node_count * sizeof(TNode); break;
        case 1: src = s->bb.stris; n = (size_t)nt * sizeof(STri); break;
        case 2: src = s->bb.keys; n = (size_t)nt * 4; break;
        case 3: src = s->bb.ids; n = (size_t)nt * 4; break;
        case 5: src = s->bb.nodes4; n = (size_t)s->bb.wide_count * sizeof(TNode4); break;
        case 4: {
            if (nt >= 2) {
                size_t m = nt - 1;
                std::vector<uint32_t> a(m), b(m), c(m), d(m);
                HR_HIP(ctx, hipMemcpy(a.data(), s->bb.kl, m * 4, hipMemcpyDeviceToHost));
                HR_HIP(ctx, hipMemcpy(b.data(), s->bb.kr, m * 4, hipMemcpyDeviceToHost));
                HR_HIP(ctx, hipMemcpy(c.data(), s->bb.kfirst, m * 4, hipMemcpyDeviceToHost));
                HR_HIP(ctx, hipMemcpy(d.data(), s->bb.klast, m * 4, hipMemcpyDeviceToHost));
                tmp.resize(m * 4);
                for (size_t i = 0; i < m; ++i) {
                    tmp[4 * i] = a[i]; tmp[4 * i + 1] = b[i]; tmp[4 * i + 2] = c[i]; tmp[4 * i + 3] = d[i];
                }
            }
            if (bytes < tmp.size() * 4) return fail(ctx, HR_E_INVALID, "export buffer too small");
            if (!tmp.empty()) memcpy(dst, tmp.data(), tmp.size() * 4);
            return HR_OK;
        }
        default: return fail(ctx, HR_E_INVALID, "unknown export %u", which);
    }
    if (bytes < n) return fail(ctx, HR_E_INVALID, "export buffer too small");
    if (n) HR_HIP(ctx, hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
    return HR_OK;
}

hr_status hr_gbuffer_primary(hr_scene* s, const hr_camera* cam, const hr_gbuffer* gb, hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    HR_TRY(check_gbuffer(ctx, gb));
    FrameArgs f;
    HR_TRY(fill_frame(ctx, f, cam, gb->width, gb->height));
    HR_TRY(check_textures(ctx, s));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, launch_gbuffer(dev_scene(s), f, gview(gb), ctx->d_counters, false, ctx->stream));
    return frame_end(ctx, stats, stats ? &stats->ms_gbuffer : nullptr);
}

hr_status hr_render_rt(hr_scene* s, const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                       const hr_camera* cam, const hr_trace_params* params, uint32_t width, uint32_t height,
                       float* accum, hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (!accum) return fail(ctx, HR_E_INVALID, "accum must be a device pointer");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(ctx, HR_E_INVALID, "frame size %ux%u out of range", width, height);
    FrameArgs f;
    HR_TRY(fill_frame(ctx, f, cam, width, height));
    HR_TRY(fill_lights(ctx, f, lights, shapes, nl, params));
    if (f.max_depth == 0) return fail(ctx, HR_E_INVALID, "max_depth must be at least 1");
    if (f.adaptive) return fail(ctx, HR_E_UNSUPPORTED, "adaptive sampling is a hybrid-pass option");
    HR_TRY(check_textures(ctx, s));
    const bool count = (params->flags & HR_FLAG_COUNT_VISITS) != 0;
    const size_t wbytes = count ? 0 : rt_work_bytes(f);
    if (wbytes) HR_TRY(ensure_work(ctx, wbytes));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    if (wbytes)
        HR_HIP(ctx, launch_rt_wavefront(dev_scene(s), f, reinterpret_cast<float4*>(accum),
                                        (int32_t)(int)cam->frame[0], ctx->d_counters, ctx->stream, ctx->work,
                                        ctx->d_next_group, ctx->resident_blocks));
    else
        HR_HIP(ctx, launch_rt(dev_scene(s), f, reinterpret_cast<float4*>(accum), (int32_t)(int)cam->frame[0],
                              ctx->d_counters, count, ctx->stream));
    return frame_end(ctx, stats, stats ? &stats->ms_shade : nullptr);
}

hr_status hr_trace_shadows(hr_scene* s, const hr_gbuffer* gb, const hr_light* lights, const hr_light_shape* shapes,
                           uint32_t nl, const hr_camera* cam, const hr_trace_params* params, float* shadow,
                           hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    HR_TRY(check_gbuffer(ctx, gb));
    if (!shadow) return fail(ctx, HR_E_INVALID, "shadow output is NULL");
    FrameArgs f;
    HR_TRY(fill_frame(ctx, f, cam, gb->width, gb->height));
    HR_TRY(fill_lights(ctx, f, lights, shapes, nl, params));
    bool count = (params->flags & HR_FLAG_COUNT_VISITS) != 0;
    std::vector<FramePlan> plans(1);
    plans[0] = FramePlan{s, f, gview(gb), shadow, nullptr, nullptr, count};
    std::vector<ShadowJob> jobs;
    uint32_t ngroups = make_jobs(ctx, plans, jobs);
    HR_TRY(ensure_jobs(ctx, 1));
    const SplPlan sp = plan_spl(plans.data(), plans.size(), ngroups);
    if (sp.use) HR_TRY(ensure_work(ctx, sp.scratch));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, hipMemcpyAsync(ctx->d_jobs, jobs.data(), sizeof(ShadowJob), hipMemcpyHostToDevice, ctx->stream));
    HR_TRY(enqueue_shadows(ctx, ctx->d_jobs, 1, ngroups, sp, count, ctx->d_next_group,
                           reinterpret_cast<uint32_t*>(ctx->work), s->bb.wide_count));
    return frame_end(ctx, stats, stats ? &stats->ms_shadow : nullptr);
}

hr_status hr_trace_shadows_noalloc(hr_scene* s, const hr_gbuffer* gb, const hr_light* lights,
                                   const hr_light_shape* shapes, uint32_t nl, const hr_camera* cam,
                                   const hr_trace_params* params, float* shadow, hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (!ctx->noalloc_scratch)
        return fail(ctx, HR_E_DEVICE, "the allocation-free scratch is unavailable on device %d: %s", ctx->device,
                    hipGetErrorString(ctx->noalloc_err));
    HR_TRY(check_gbuffer(ctx, gb));
    if (!shadow) return fail(ctx, HR_E_INVALID, "shadow output is NULL");
    // as hr_trace_shadows, with the job on the stack instead of in vectors
    FramePlan fp;
    fp.scene = s;
    HR_TRY(fill_frame(ctx, fp.f, cam, gb->width, gb->height));
    HR_TRY(fill_lights(ctx, fp.f, lights, shapes, nl, params));
    fp.g = gview(gb);
    fp.shadow = shadow;
    fp.lit = nullptr;
    fp.refl = nullptr;
    fp.count = (params->flags & HR_FLAG_COUNT_VISITS) != 0;
    ShadowJob job;
    const uint32_t ngroups = fill_job(ctx, fp, 0u, job);
    const SplPlan sp = plan_spl(&fp, 1, ngroups);
    const size_t need = NOALLOC_BITS_OFFSET + (sp.use ? sp.scratch : 0);
    if (need > NOALLOC_BYTES)
        return fail(ctx, HR_E_UNSUPPORTED, "the frame needs %zu bytes of scratch, above HR_NOALLOC_SCRATCH_BYTES (%zu)",
                    need, NOALLOC_BYTES);
    char* base = static_cast<char*>(ctx->noalloc_scratch);
    ShadowJob* d_job = reinterpret_cast<ShadowJob*>(base);
    uint32_t* d_bands = reinterpret_cast<uint32_t*>(base + NOALLOC_JOB_BYTES);
    uint32_t* d_bits = reinterpret_cast<uint32_t*>(base + NOALLOC_BITS_OFFSET);
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, launch_store_job(job, d_job, ctx->stream));
    HR_TRY(enqueue_shadows(ctx, d_job, 1, ngroups, sp, fp.count, d_bands, d_bits, s->bb.wide_count));
    return frame_end(ctx, stats, stats ? &stats->ms_shadow : nullptr);
}

hr_status hr_shade_hybrid(hr_scene* s, const hr_gbuffer* gb, const hr_light* lights, const hr_light_shape* shapes,
                          uint32_t nl, const hr_camera* cam, const hr_trace_params* params, const float* shadow,
                          float* lit, float* refl, hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    HR_TRY(check_gbuffer(ctx, gb));
    if (!shadow || !lit || !refl) return fail(ctx, HR_E_INVALID, "shadow/lit/refl pointer is NULL");
    FrameArgs f;
    HR_TRY(fill_frame(ctx, f, cam, gb->width, gb->height));
    HR_TRY(fill_lights(ctx, f, lights, shapes, nl, params));
    if (params->refl_spp > 64) return fail(ctx, HR_E_UNSUPPORTED, "refl_spp above 64");
    bool count = (params->flags & HR_FLAG_COUNT_VISITS) != 0;
    HR_TRY(check_textures(ctx, s));
    WaveArgs w;
    HR_TRY(wave_args(ctx, f, w));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, launch_shade(dev_scene(s), f, gview(gb), shadow, reinterpret_cast<float4*>(lit),
                             reinterpret_cast<float4*>(refl), ctx->d_counters, count, ctx->stream, w));
    return frame_end(ctx, stats, stats ? &stats->ms_shade : nullptr);
}

hr_status hr_trace_rays(hr_scene* s, const hr_ray* rays, uint32_t n, int flags, hr_hit* hits, hr_stats* stats) {
    if (!s) return HR_E_INVALID;
    hr_context* ctx = s->ctx;
    if (n && (!rays || !hits)) return fail(ctx, HR_E_INVALID, "rays/hits pointer is NULL");
    if ((uint32_t)flags & ~(HR_TRACE_ANY_HIT | HR_TRACE_COUNT_VISITS))
        return fail(ctx, HR_E_INVALID, "unknown trace flags 0x%x", (unsigned)flags);
    const int any_hit = ((uint32_t)flags & HR_TRACE_ANY_HIT) ? 1 : 0;
    const bool count = ((uint32_t)flags & HR_TRACE_COUNT_VISITS) != 0;
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, launch_trace_rays(dev_scene(s), rays, n, any_hit, hits, ctx->d_counters, count, ctx->stream));
    return frame_end(ctx, stats, stats ? &stats->ms_shadow : nullptr);
}

hr_status hr_context_set_tuning(hr_context* ctx, const hr_tuning* t) {
    if (!ctx) return HR_E_INVALID;
    if (!t) return fail(ctx, HR_E_INVALID, "tuning is NULL");
    bool pow2 = t->tile_w && (t->tile_w & (t->tile_w - 1)) == 0;
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 16 || !pow2 || t->tile_w > 64 || t->wave_budget_mb < 64)
        return fail(ctx, HR_E_INVALID, "tuning out of range");
    if (t->bvh_flags & ~HR_BVH_NO_TREELETS) return fail(ctx, HR_E_INVALID, "unknown bvh_flags 0x%x", t->bvh_flags);
    ctx->tuning = *t;
    ctx->resident_blocks = ctx->cu_count * t->blocks_per_cu;
    return HR_OK;
}

hr_status hr_context_get_tuning(const hr_context* ctx, hr_tuning* t) {
    if (!ctx || !t) return HR_E_INVALID;
    *t = ctx->tuning;
    return HR_OK;
}

}  // extern "C"

// ---- batches and graphs --------------------------------------------------------
namespace {

hr_status plan_batch(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, bool shading,
                     std::vector<FramePlan>& plans) {
    if (!frames || n == 0) return fail(ctx, HR_E_INVALID, "empty batch");
    plans.resize(n);
    for (uint32_t i = 0; i < n; ++i) HR_TRY(plan_frame(ctx, frames[i], plans[i], shading));
    return HR_OK;
}

// Kernels of a batch's shading pass, frames one after another on a shared
// workspace at base (laid out per frame).
// How the frames of a batch share the shading streams: frames by decreasing
// work (pixels x paths per pixel; the lower index first on ties), each to the
// stream with the least work so far; each stream's workspace holds its
// largest frame at budget_mb / streams, the streams' workspaces one after
// another (256-byte aligned). total = 0: a frame does not fit.
struct ShadeSplit {
    uint32_t streams = 1;
    std::vector<uint32_t> stream_of;
    size_t offset[SHADE_STREAMS] = {};
    size_t total = 0;
};
ShadeSplit split_shading(hr_context* ctx, const std::vector<FramePlan>& plans, uint32_t budget_mb) {
    ShadeSplit sp;
    const uint32_t n = (uint32_t)plans.size();
    sp.streams = n < (uint32_t)SHADE_STREAMS ? (n > 0 ? n : 1u) : (uint32_t)SHADE_STREAMS;
    sp.stream_of.assign(n, 0u);
    std::vector<uint64_t> cost(n);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) {
        cost[i] = (uint64_t)plans[i].f.width * plans[i].f.height * (plans[i].f.spp ? plans[i].f.spp : 1u);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    uint64_t load[SHADE_STREAMS] = {};
    size_t need[SHADE_STREAMS] = {};
    const uint32_t mb = budget_mb / sp.streams > 64u ? budget_mb / sp.streams : 64u;
    for (uint32_t i : order) {
        uint32_t s = 0;
        for (uint32_t k = 1; k < sp.streams; ++k)
            if (load[k] < load[s]) s = k;
        sp.stream_of[i] = s;
        load[s] += cost[i];
        WaveArgs w;
        const size_t b = wave_layout(ctx, plans[i].f, w, nullptr, mb);
        if (b == 0) return ShadeSplit{};
        need[s] = b > need[s] ? b : need[s];
    }
    size_t off = 0;
    for (uint32_t k = 0; k < sp.streams; ++k) {
        sp.offset[k] = off;
        off += (need[k] + 255u) & ~(size_t)255u;
    }
    sp.total = off;
    return sp;
}

// Shading of a batch's frames: on the context's stream for one frame, else
// forked over the split's streams and joined back (in a captured graph the
// branches become parallel paths).
hr_status enqueue_shading(hr_context* ctx, const std::vector<FramePlan>& plans, char* base, uint32_t budget_mb) {
    const ShadeSplit sp = split_shading(ctx, plans, budget_mb);
    if (sp.total == 0) return fail(ctx, HR_E_UNSUPPORTED, "reflection slab too large");
    const uint32_t mb = budget_mb / sp.streams > 64u ? budget_mb / sp.streams : 64u;
    const bool fork = sp.streams > 1u;
    if (fork) {
        HR_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
        for (uint32_t k = 0; k < sp.streams; ++k) HR_HIP(ctx, hipStreamWaitEvent(ctx->side[k], ctx->fork_ev, 0));
    }
    for (size_t i = 0; i < plans.size(); ++i) {
        const FramePlan& fp = plans[i];
        const uint32_t s = sp.stream_of[i];
        WaveArgs w;
        wave_layout(ctx, fp.f, w, base + sp.offset[s], mb);
        HR_HIP(ctx, launch_shade(dev_scene(fp.scene), fp.f, fp.g, fp.shadow, reinterpret_cast<float4*>(fp.lit),
                                 reinterpret_cast<float4*>(fp.refl), ctx->d_counters, fp.count,
                                 fork ? ctx->side[s] : ctx->stream, w));
    }
    if (fork) {
        for (uint32_t k = 0; k < sp.streams; ++k) {
            HR_HIP(ctx, hipEventRecord(ctx->join_ev[k], ctx->side[k]));
            HR_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->join_ev[k], 0));
        }
    }
    return HR_OK;
}

size_t batch_work_bytes(hr_context* ctx, const std::vector<FramePlan>& plans, uint32_t budget_mb) {
    return split_shading(ctx, plans, budget_mb).total;
}

}  // namespace

struct hr_graph {
    hr_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ShadowJob* d_jobs = nullptr;
    uint32_t* d_next = nullptr;
    void* work = nullptr;
    // the captured scenes and their generations at capture
    std::vector<std::pair<const hr_scene*, uint64_t>> scenes;
};

extern "C" {

hr_status hr_trace_shadows_batch(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_stats* stats) {
    if (!ctx) return HR_E_INVALID;
    std::vector<FramePlan> plans;
    HR_TRY(plan_batch(ctx, frames, n, false, plans));
    std::vector<ShadowJob> jobs;
    uint32_t ngroups = make_jobs(ctx, plans, jobs);
    bool count = false;
    for (auto& fp : plans) count = count || fp.count;
    HR_TRY(ensure_jobs(ctx, n));
    const SplPlan sp = plan_spl(plans.data(), plans.size(), ngroups);
    if (sp.use) HR_TRY(ensure_work(ctx, sp.scratch));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_HIP(ctx, hipMemcpyAsync(ctx->d_jobs, jobs.data(), n * sizeof(ShadowJob), hipMemcpyHostToDevice, ctx->stream));
    HR_TRY(enqueue_shadows(ctx, ctx->d_jobs, n, ngroups, sp, count, ctx->d_next_group,
                           reinterpret_cast<uint32_t*>(ctx->work), plans[0].scene->bb.wide_count));
    return frame_end(ctx, stats, stats ? &stats->ms_shadow : nullptr);
}

hr_status hr_shade_hybrid_batch(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_stats* stats) {
    if (!ctx) return HR_E_INVALID;
    std::vector<FramePlan> plans;
    HR_TRY(plan_batch(ctx, frames, n, true, plans));
    uint32_t mb = 0;
    HR_TRY(fit_budget(ctx, [&](uint32_t b) { return batch_work_bytes(ctx, plans, b); }, mb));
    HR_TRY(frame_begin(ctx, stats != nullptr));
    HR_TRY(enqueue_shading(ctx, plans, static_cast<char*>(ctx->work), mb));
    return frame_end(ctx, stats, stats ? &stats->ms_shade : nullptr);
}

void hr_graph_destroy(hr_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_jobs) (void)hipFree(g->d_jobs);
    if (g->d_next) (void)hipFree(g->d_next);
    if (g->work) (void)hipFree(g->work);
    delete g;
}

hr_status hr_graph_capture(hr_context* ctx, const hr_frame_desc* frames, uint32_t n, hr_graph** out) {
    if (!ctx) return HR_E_INVALID;
    if (!out) return fail(ctx, HR_E_INVALID, "out is NULL");
    *out = nullptr;
    std::vector<FramePlan> plans;
    HR_TRY(plan_batch(ctx, frames, n, true, plans));
    std::vector<ShadowJob> jobs;
    uint32_t ngroups = make_jobs(ctx, plans, jobs);
    bool count = false;
    for (auto& fp : plans) count = count || fp.count;
    // the shadow pass's scratch and the shading workspace share g->work: the
    // passes run one after the other on the stream
    const SplPlan sp = plan_spl(plans.data(), plans.size(), ngroups);
    auto need = [&](uint32_t b) {
        size_t bytes = batch_work_bytes(ctx, plans, b);
        return (bytes != 0 && sp.use && sp.scratch > bytes) ? sp.scratch : bytes;
    };
    uint32_t mb = ctx->tuning.wave_budget_mb;
    if (need(mb) == 0) return fail(ctx, HR_E_UNSUPPORTED, "reflection slab too large");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hr_graph* g = new hr_graph();
    g->ctx = ctx;
    for (const FramePlan& fp : plans) g->scenes.emplace_back(fp.scene, fp.scene->generation);
    hipError_t e = hipMalloc(&g->d_jobs, n * sizeof(ShadowJob));
    if (e == hipSuccess) e = hipMalloc(&g->d_next, BAND_COUNTER_BYTES);
    // the workspace at the tuned budget, halved while the device cannot hold it
    while (e == hipSuccess) {
        const size_t bytes = need(mb);
        if (bytes == 0) {
            hr_graph_destroy(g);
            return fail(ctx, HR_E_UNSUPPORTED, "reflection slab too large");
        }
        e = hipMalloc(&g->work, bytes);
        if (e != hipErrorOutOfMemory || mb <= 64u) break;
        (void)hipGetLastError();
        g->work = nullptr;
        e = hipSuccess;
        mb = mb / 2u < 64u ? 64u : mb / 2u;
    }
    if (e == hipSuccess) e = hipMemcpy(g->d_jobs, jobs.data(), n * sizeof(ShadowJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hr_graph_destroy(g);
        return hip_fail(ctx, e, "graph buffers");
    }
    // capture: counters cleared by a kernel (no memset nodes), shadows, shading
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hr_graph_destroy(g);
        return hip_fail(ctx, e, "hipStreamBeginCapture");
    }
    hr_status st = HR_OK;
    hipError_t le = launch_zero_u64(ctx->d_counters, C_COUNT * C_STRIDE, ctx->stream);
    if (le != hipSuccess) st = hip_fail(ctx, le, "launch");
    if (st == HR_OK)
        st = enqueue_shadows(ctx, g->d_jobs, n, ngroups, sp, count, g->d_next, static_cast<uint32_t*>(g->work),
                             plans[0].scene->bb.wide_count);
    if (st == HR_OK) st = enqueue_shading(ctx, plans, static_cast<char*>(g->work), mb);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(ctx->stream, &graph);
    g->graph = graph;
    if (st != HR_OK || e != hipSuccess) {
        hr_graph_destroy(g);
        return st != HR_OK ? st : hip_fail(ctx, e, "hipStreamEndCapture");
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hr_graph_destroy(g);
        return hip_fail(ctx, e, "hipGraphInstantiate");
    }
    *out = g;
    return HR_OK;
}

hr_status hr_graph_launch(hr_graph* g, hr_stats* stats) {
    if (!g) return HR_E_INVALID;
    hr_context* ctx = g->ctx;
    for (size_t i = 0; i < g->scenes.size(); ++i)
        if (g->scenes[i].first->generation != g->scenes[i].second)
            return fail(ctx, HR_E_INVALID, "frame %zu: its scene's materials, textures or environment changed "
                        "since capture; capture the graph again", i);
    HR_HIP(ctx, hipSetDevice(ctx->device));
    if (stats) HR_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    HR_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return frame_end(ctx, stats, stats ? &stats->ms_shade : nullptr);
}

hr_status hr_filter_shadows(hr_context* ctx, uint32_t width, uint32_t height, uint32_t nl, int32_t frame,
                            const float* shadow, const float* motion, float* history) {
    if (!ctx) return HR_E_INVALID;
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(ctx, HR_E_INVALID, "frame size %ux%u out of range", width, height);
    if (nl > HR_MAX_LIGHTS) return fail(ctx, HR_E_INVALID, "light_count %u out of range", nl);
    if (nl == 0) return HR_OK;  // nothing to filter
    if (!shadow || !history) return fail(ctx, HR_E_INVALID, "shadow and history must be device pointers");
    if (frame <= 0 && !motion) return fail(ctx, HR_E_INVALID, "motion vectors are needed when frame <= 0");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    const float* prev = history;
    if (frame <= 0) {  // reprojected reads must see the previous output
        const size_t bytes = (size_t)nl * width * height * sizeof(float);
        HR_TRY(ensure_work(ctx, bytes));
        HR_HIP(ctx, hipMemcpyAsync(ctx->work, history, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        prev = static_cast<const float*>(ctx->work);
    }
    HR_HIP(ctx, launch_filter_shadows(width, height, nl, frame, shadow, motion, prev, history, ctx->stream));
    return HR_OK;
}

hr_status hr_gbuffer_ingest(hr_context* ctx, const hr_raster_gbuffer* in, const hr_gbuffer* out, float* motion_out) {
    if (!ctx) return HR_E_INVALID;
    if (!in || !in->position || !in->normal || !in->albedo || !in->material || !in->emissive)
        return fail(ctx, HR_E_INVALID, "raster G-buffer planes must be non-NULL device pointers");
    HR_TRY(check_gbuffer(ctx, out));
    if (in->width != out->width || in->height != out->height)
        return fail(ctx, HR_E_INVALID, "raster G-buffer %ux%u does not match %ux%u", in->width, in->height,
                    out->width, out->height);
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, launch_gbuffer_ingest(in->width * in->height, in->position, in->normal, in->albedo, in->motion,
                                      in->material, in->emissive, gview(out), motion_out, ctx->stream));
    return HR_OK;
}

hr_status hr_gbuffer_quantize(hr_context* ctx, const hr_gbuffer* in, const float* motion_in,
                              const hr_raster_gbuffer* out) {
    if (!ctx) return HR_E_INVALID;
    HR_TRY(check_gbuffer(ctx, in));
    if (!out || !out->position || !out->normal || !out->albedo || !out->material || !out->emissive)
        return fail(ctx, HR_E_INVALID, "raster G-buffer planes must be non-NULL device pointers");
    if (in->width != out->width || in->height != out->height)
        return fail(ctx, HR_E_INVALID, "raster G-buffer %ux%u does not match %ux%u", out->width, out->height,
                    in->width, in->height);
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, launch_gbuffer_quantize(in->width * in->height, gview(in), motion_in, out->position, out->normal,
                                        out->albedo, out->motion, out->material, out->emissive, ctx->stream));
    return HR_OK;
}

hr_status hr_motion_vectors(hr_context* ctx, const hr_gbuffer* gb, const float* view_proj,
                            const float* prev_view_proj, float* motion) {
    if (!ctx) return HR_E_INVALID;
    HR_TRY(check_gbuffer(ctx, gb));
    if (!view_proj || !prev_view_proj || !motion)
        return fail(ctx, HR_E_INVALID, "view_proj, prev_view_proj and motion are required");
    MotionArgs a;
    for (int i = 0; i < 16; ++i) {
        a.vp[i] = view_proj[i];
        a.vp_prev[i] = prev_view_proj[i];
    }
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, launch_motion(gb->width, gb->height, reinterpret_cast<const float4*>(gb->position),
                              reinterpret_cast<const float4*>(gb->normal), a, motion, ctx->stream));
    return HR_OK;
}

hr_status hr_post_gamma(hr_context* ctx, uint32_t width, uint32_t height, const float* in, float* out,
                        uint8_t* rgba8) {
    if (!ctx) return HR_E_INVALID;
    if (!in || (!out && !rgba8)) return fail(ctx, HR_E_INVALID, "in and at least one output are required");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(ctx, HR_E_INVALID, "image size %ux%u out of range", width, height);
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, launch_post_gamma(width * height, reinterpret_cast<const float4*>(in), reinterpret_cast<float4*>(out),
                                  reinterpret_cast<uint32_t*>(rgba8), ctx->stream));
    return HR_OK;
}

}  // extern "C"
