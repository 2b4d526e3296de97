// trace.hip - traversal and shading kernels of the hybrid pass (gfx950).
//
//   k_gbuffer   primary visibility at pixel centres (raygen.rgen:29-40) into
//               the FP32 G-buffer (geometry_shader.frag:51-76, skybox.frag:16-21)
//   k_shadows   shadowRaygen.rgen:28-93 (+ shadow.rmiss, shadowHit.rchit)
//   k_trace     bare traceRayEXT queries
//
// k_gbuffer and k_trace: one lane per pixel (ray); a 256-thread workgroup
// covers 16x16 pixels and each 64-lane wave an 8x8 tile, so a wave's rays start
// close together and walk similar nodes. Workgroups are renumbered so that
// consecutive screen tiles land on the same XCD (dispatch assigns workgroup i
// to XCD i mod 8), which keeps a screen region's working set in one L2.
// k_shadows is persistent and fetches its work dynamically (see below).
//
// Traversal: 64-byte nodes holding both child boxes (four 16-byte loads),
// near child first, far child pushed on a per-lane stack whose first 16
// entries live in LDS (one column per lane, conflict free) and the rest in
// private memory. Leaves hold up to leaf_max sorted triangles tested with the
// watertight test; any-hit queries stop at the first hit.
#include "trace_common.h"

namespace hr {
using namespace hrd;
using namespace tc;

namespace {
// Pixel (x, y) of pixel q of shadow tile t (tiles of tile_w x 64 / tile_w
// pixels). Tiles run row by row down vertical bands of TILE_BAND tile
// columns, one band after another, so the tiles in flight at once cover a
// compact patch of the screen rather than a thin strip. (x, y) may lie
// beyond the frame's right or bottom edge.
constexpr uint32_t TILE_BAND = 8u;
__device__ __forceinline__ void tile_pixel(const ShadowJob& J, uint32_t t, uint32_t q, uint32_t& x, uint32_t& y) {
    const uint32_t tw = J.tile_w, th = 64u / J.tile_w;
    const uint32_t tiles_y = (J.F.height + th - 1u) / th;
    const uint32_t per_band = TILE_BAND * tiles_y;
    const uint32_t band = t / per_band, k = t - band * per_band;
    const uint32_t bw = min(TILE_BAND, J.tiles_x - band * TILE_BAND);
    const uint32_t ty = k / bw, tx = band * TILE_BAND + (k - ty * bw);
    x = tx * tw + (q % tw);
    y = ty * th + (q / tw);
}

// World-space direction of the primary ray through the centre of pixel (x, y)
// (raygen.rgen:29-40); x and y may lie outside the image.
__device__ __forceinline__ v3 primary_dir(const FrameArgs& F, uint32_t x, uint32_t y) {
    float ux = ((float)x + 0.5f) / (float)F.width;
    float uy = ((float)y + 0.5f) / (float)F.height;
    float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    float tgt[4], dir[4];
    mul4(F.inv_proj, dx, dy, 1.0f, 1.0f, tgt);
    v3 tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
    mul4(F.inv_view, tn.x, tn.y, tn.z, 0.0f, dir);
    return mk(dir[0], dir[1], dir[2]);
}
// Point where the ray from o through pixel (x, y) meets the plane of the
// triangle (v0, ng): the world position a rasteriser interpolates for that
// pixel, helper pixels included.
__device__ __forceinline__ v3 plane_point(const FrameArgs& F, v3 o, v3 v0, v3 ng, uint32_t x, uint32_t y) {
    v3 d = primary_dir(F, x, y);
    float t = dot(sub(v0, o), ng) / dot(d, ng);
    return add(o, scl(d, t));
}
// perturbNormal / cotangent_frame of geometry_shader.frag:28-49 for pixel
// (x, y): the screen-space derivatives dFdx(p), dFdy(p) are the fine
// differences across the pixel's 2x2 quad of the world positions on the hit
// triangle's plane; the frame uses dFdx(p).xy and dFdy(p).xy as the uv
// derivatives, as the shader does. Nin is the interpolated, transformed,
// unnormalised vertex normal; tx the normal-map texel. A degenerate frame
// (zero or non-finite) keeps normalize(Nin).
__device__ v3 perturb_normal(const FrameArgs& F, v3 o, const float4* wtri, uint32_t tri, v3 Nin, v3 tx,
                             uint32_t x, uint32_t y) {
    v3 np = mk(tx.x * (255.0f / 127.0f) - 128.0f / 127.0f, tx.y * (255.0f / 127.0f) - 128.0f / 127.0f,
               tx.z * (255.0f / 127.0f) - 128.0f / 127.0f);
    v3 v0 = ld3(wtri[3 * (size_t)tri]), v1 = ld3(wtri[3 * (size_t)tri + 1]), v2 = ld3(wtri[3 * (size_t)tri + 2]);
    v3 ng = cross(sub(v1, v0), sub(v2, v0));
    uint32_t x0 = x & ~1u, x1 = x | 1u, y0 = y & ~1u, y1 = y | 1u;
    v3 dp1 = sub(plane_point(F, o, v0, ng, x1, y), plane_point(F, o, v0, ng, x0, y));
    v3 dp2 = sub(plane_point(F, o, v0, ng, x, y1), plane_point(F, o, v0, ng, x, y0));
    v3 dp2perp = cross(dp2, Nin);
    v3 dp1perp = cross(Nin, dp1);
    v3 T = add(scl(dp2perp, dp1.x), scl(dp1perp, dp2.x));
    v3 B = add(scl(dp2perp, dp1.y), scl(dp1perp, dp2.y));
    float tt = dot(T, T), bb = dot(B, B);
    float m = (bb > tt) ? bb : tt;
    if (!(isfinite(tt) && isfinite(bb) && m > 0.0f)) return normalize(Nin);
    float invmax = 1.0f / sqrtf(m);
    T = scl(T, invmax);
    B = scl(B, invmax);
    return normalize(add(add(scl(T, np.x), scl(B, np.y)), scl(Nin, np.z)));
}

// ---- kernels ------------------------------------------------------------------
template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_gbuffer(DevScene S, FrameArgs F, GBufView G,
                                                   unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t x, y;
    pixel_of(F.width, x, y);
    if (x < F.width && y < F.height) {
        size_t p = (size_t)y * F.width + x;
        float org[4];
        mul4(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, org);
        v3 o = mk(org[0], org[1], org[2]), d = primary_dir(F, x, y);
        v3 ro = add(o, scl(d, 1e-2f));
        RayP r = prep_ray(ro, d);
        HitD h;
        ++c.primary;
        bool hit = traverse<false, COUNT>(S, r, 0.001f, 10000.0f, lds, h, c);
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f), Nn = P, A = P, Mt = P, E = P;
        if (!hit) {
            v3 sky = env_lookup(S.sky, S.sky_w, S.sky_h, normalize(d), PI_M);
            A = make_float4(sky.x, sky.y, sky.z, 1.0f);
        } else {
            const PrimInfo& pr = S.prims[S.tri_prim[h.tri]];
            uint32_t local = h.tri - pr.tri_offset;
            uint32_t ib = pr.index_base + 3 * local;
            uint32_t vi[3] = {pr.vertex_base + S.indices[ib + 0], pr.vertex_base + S.indices[ib + 1],
                              pr.vertex_base + S.indices[ib + 2]};
            v3 n0 = ld3(S.vnormal[vi[0]]);
            v3 n1 = ld3(S.vnormal[vi[1]]);
            v3 n2 = ld3(S.vnormal[vi[2]]);
            float b1 = h.u, b2 = h.v, b0 = 1.0f - b1 - b2;
            v3 nrm = add(add(scl(n0, b0), scl(n1, b1)), scl(n2, b2));
            const float* M = pr.nm;
            v3 Nin = mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                        (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z);
            v3 N = normalize(Nin);
            v3 wp = add(ro, scl(d, h.t));
            const hr_material& mt = S.mats[pr.material];
            v3 alb = ld3p(mt.diffuse_ior);
            v3 emi = mk(0.0f, 0.0f, 0.0f);
            v3 mtl = mk(0.0f, mt.shading_metal_rough_idx[2], mt.shading_metal_rough_idx[1]);
            if (textured(mt)) {  // geometry_shader.frag:54-64
                float u, v;
                interp_uv(S, vi, b0, b1, b2, u, v);
                if (mt.textures[0] > -1.0f) {
                    const float* a = vtx(S, vi[0]);
                    const float* b = vtx(S, vi[1]);
                    const float* e = vtx(S, vi[2]);
                    v3 col = add(add(scl(mk(a[6], a[7], a[8]), b0), scl(mk(b[6], b[7], b[8]), b1)),
                                 scl(mk(e[6], e[7], e[8]), b2));
                    alb = mul(tex_rgba8(S, mt.textures[0], u, v), col);
                }
                if (mt.textures[2] > -1.0f) emi = tex_rgba8(S, mt.textures[2], u, v);
                if (mt.textures[3] > -1.0f) mtl = tex_rgba8(S, mt.textures[3], u, v);
                if (mt.textures[1] > -1.0f)
                    N = perturb_normal(F, o, S.wtri, h.tri, Nin, tex_rgba8(S, mt.textures[1], u, v), x, y);
            }
            P = make_float4(wp.x, wp.y, wp.z, mt.shading_metal_rough_idx[3]);
            Nn = make_float4(N.x, N.y, N.z, 1.0f);
            A = make_float4(alb.x, alb.y, alb.z, 1.0f);
            Mt = make_float4(mtl.x, mtl.y, mtl.z, 0.0f);
            E = make_float4(emi.x, emi.y, emi.z, 0.0f);
        }
        G.pos[p] = P;
        G.nrm[p] = Nn;
        G.alb[p] = A;
        G.mat[p] = Mt;
        G.emi[p] = E;
    }
    flush_counters(c, counters, COUNT);
}

// Shadow pass, persistent waves with dynamic fetch, over one frame or a batch
// of frames (each with its own scene, G-buffer and lights). The work unit is a
// group = (frame, pixel, light): its rays are traced by one lane in sample
// order and its shadow factor is written once by that lane. Groups are
// numbered frame by frame, then tile by tile (tile_w x 64/tile_w pixels), then
// light by light, so the 64 groups a wave takes together are one light's rays
// from one tile. The group list is cut into SHADOW_QUEUES contiguous bands
// (screen bands), one per XCD, each with its own counter. A lane whose group
// is finished takes the next group from its wave's current band (one atomic
// per wave for all lanes that need work): first the band of its XCD, so
// neighbouring tiles share that XCD's L2, then the other bands in turn, so
// lanes stay busy until every band is empty. With BATCH = false the single
// frame's job is read with uniform (scalar) loads.
// Traversal: near-first stack walk, internal nodes in an inner loop until the
// lane reaches a leaf, then the leaf's triangles.
template <bool COUNT, bool BATCH>
// 6 waves per SIMD: the register budget of 80 VGPRs costs two spills and
// measured faster than the 5 waves the unconstrained 86 VGPRs allow
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_shadows(const ShadowJob* __restrict__ jobs, uint32_t njobs,
                                                   uint32_t ngroups, uint32_t* next_group,
                                                   unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    // Workgroups are dispatched round-robin over the XCDs, so workgroup b runs
    // on XCD b mod 8; its home band is the one of that XCD.
    const uint32_t home = blockIdx.x & (SHADOW_QUEUES - 1u);
    uint32_t emptied = 0;  // bands this lane has seen empty, in visiting order
    // current group
    uint32_t job = 0;
    const TNode* nodes = jobs[0].S.nodes;
    const STri* tris = jobs[0].S.tris;
    uint32_t node_count = jobs[0].S.node_count;
    float* out_base = nullptr;  // the group's factor: out_base[out_idx], FP32 or R8 by out_fmt
    size_t out_idx = 0;
    uint32_t out_fmt = 0;
    uint32_t li = 0, seed = 0, s = 0, unocc = 0, SS = 1;
    uint32_t target = 1;    // rays of the current stage (adaptive) or all rays
    bool hard = false;      // adaptive: a light of one sample
    v3 P = mk(0.0f, 0.0f, 0.0f), L = P;
    float dist = 0.0f;
    bool isDir = false, in_group = false;
    // current ray
    uint32_t cur = DONE;
    bool hit = false;
    RayP r;
    float tmax = 0.0f;
    const float tmin = 0.001f;
    int sp = 0;
    // The timed kernel walks the 4-wide nodes (up to 3 pushes per level), the
    // counting kernel the binary ones.
    constexpr uint32_t STACK_CAP = COUNT ? MAX_DEPTH : MAX_STACK4;
    uint32_t spill[STACK_CAP - STACK_LDS];
    const TNode4* nodes4 = jobs[0].S.nodes4;
    // Speculative traversal (Aila & Laine 2009) in the timed kernel: the
    // shadow result does not depend on the order leaves are tested in. The
    // counting kernel keeps the canonical near-first order, so its node and
    // triangle counts are the oracle's.
    constexpr bool SPEC = !COUNT;
    uint32_t leaf = 0u;  // postponed leaf code (0: none; leaf codes have LEAF_BIT)
    for (;;) {
        if (cur == DONE && leaf == 0u) {
            if (in_group) {
                if (s > 0) {  // a ray of this group just finished
                    if (hit) ++c.occluded; else ++unocc;
                }
                if (s == target) {
                    // adaptive: the next stage of the nested sequence while the
                    // estimate is not yet as tight as S samples at one half
                    if (jobs[job].F.adaptive != 0u && !hard && !adaptive_stop(unocc, target, SS)) {
                        target = adaptive_next(target, SS);
                    } else {
                        put_factor(out_base, out_idx, (float)unocc / (float)target, out_fmt);
                        in_group = false;
                    }
                }
            }
            if (!in_group) {
                // take groups until one has rays to trace: from the home queue
                // first, then from the others in turn
                bool exhausted = false;
                for (;;) {
                    // Bands below a lane's count are known to be empty, so the
                    // wave uses the largest count among its active lanes: one
                    // band for every lane of this fetch, the one the leader's
                    // atomic counts in.
                    uint32_t k = 0;
                    for (uint32_t b = 1; b <= SHADOW_QUEUES; ++b)
                        if (__ballot(emptied >= b)) k = b;
                    emptied = k;
                    if (k >= SHADOW_QUEUES) {
                        exhausted = true;
                        break;
                    }
                    const uint32_t band = (home + k) & (SHADOW_QUEUES - 1u);
                    const uint32_t band_end = (uint32_t)(((uint64_t)ngroups * (band + 1u)) / SHADOW_QUEUES);
                    unsigned long long need = __ballot(1);
                    uint32_t rank = (uint32_t)__popcll(need & lt_mask);
                    uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(next_group + band * BAND_STRIDE, (uint32_t)__popcll(need));
                    base = (uint32_t)__shfl((int)base, (int)leader, 64);
                    uint32_t g = (uint32_t)(((uint64_t)ngroups * band) / SHADOW_QUEUES) + base + rank;
                    if (g >= band_end) {  // the band is empty: move to the next one
                        emptied = k + 1u;
                        continue;
                    }
                    if (BATCH) {  // the frame holding group g
                        uint32_t lo = 0, hi = njobs;
                        while (hi - lo > 1) {
                            uint32_t mid = (lo + hi) >> 1;
                            if (jobs[mid].group_base <= g) lo = mid; else hi = mid;
                        }
                        job = lo;
                    }
                    const ShadowJob& J = jobs[job];
                    const FrameArgs& F = J.F;
                    uint32_t lg = g - J.group_base;
                    uint32_t per_tile = F.nl * 64u;
                    uint32_t t = lg / per_tile, rem = lg - t * per_tile;
                    li = rem >> 6;
                    uint32_t q = rem & 63u;
                    uint32_t x, y;
                    tile_pixel(J, t, q, x, y);
                    if (x >= F.width || y >= F.height) continue;
                    size_t p = (size_t)y * F.width + x;
                    out_base = J.shadow;
                    out_idx = li * ((size_t)F.width * F.height) + p;
                    out_fmt = F.unorm8;
                    float4 P4 = J.G.pos[p], N4 = J.G.nrm[p];
                    bool background = P4.w == 0.0f && N4.w == 0.0f;
                    int mode = (int)J.S.mats[mat_index(J.S, P4.w)].shading_metal_rough_idx[0];
                    if (background || mode == 4) {
                        put_factor(out_base, out_idx, background ? 0.0f : 1.0f, out_fmt);
                        continue;
                    }
                    P = ld3(P4);
                    v3 N = normalize(ld3(N4));
                    const hr_light& lt = F.lights[li];
                    isDir = lt.pos_maxdist[3] < 0.0f;
                    L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), P);
                    dist = length(L);
                    L = normalize(L);
                    float NdotL = clamp01(dot(N, L));
                    if (!(NdotL > 0.0f)) {
                        put_factor(out_base, out_idx, 0.0f, out_fmt);
                        continue;
                    }
                    if (!(isDir || dist < lt.pos_maxdist[3])) {
                        put_factor(out_base, out_idx, 1.0f, out_fmt);
                        continue;
                    }
                    seed = tea(y * F.width + x, F.frame);
                    SS = F.samples;
                    s = 0;
                    unocc = 0;
                    hard = F.adaptive && hard_light(F, li);
                    target = F.adaptive ? (hard ? 1u : adaptive_first(SS)) : SS;
                    if (BATCH) {
                        nodes = J.S.nodes;
                        nodes4 = J.S.nodes4;
                        tris = J.S.tris;
                        node_count = J.S.node_count;
                    }
                    in_group = true;
                    break;
                }
                if (exhausted) break;
            }
            const FrameArgs& F = jobs[job].F;
            v3 dir;
            if (F.adaptive) {
                Pcg g = ray_stream(seed, PURPOSE_ADAPTIVE_SHADOW, li * 128u + s);
                dir = light_sample_dir_nested(F, li, P, L, g, s, hard ? 1u : SS);
            } else {
                Pcg g = ray_stream(seed, PURPOSE_SURFACE_SHADOW, li * SS + s);
                dir = light_sample_dir(F, li, P, L, g, s, SS);
            }
            tmax = isDir ? 1000.0f : dist + 1.0f;
            r = prep_ray(add(P, scl(dir, 1e-1f)), dir);
            ++c.shadow;
            ++s;
            hit = false;
            sp = 0;
            cur = node_count ? 0u : DONE;
            continue;
        }
        while (!COUNT && cur < LEAF_BIT) {
            cur = wide_step<STACK_CAP>(nodes4, cur, r, tmin, tmax, lds, spill, sp, c);
            // postpone the first leaf found and keep walking until every lane
            // still in this loop holds one
            if (leaf == 0u && cur >= LEAF_BIT && cur != DONE) {
                leaf = cur;
                if (sp == 0) {
                    cur = DONE;
                } else {
                    --sp;
                    cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
                }
            }
            if (__ballot(leaf == 0u) == 0ull) break;
        }
        while (COUNT && cur < LEAF_BIT) {
            const float4* np = reinterpret_cast<const float4*>(nodes + cur);
            float4 a = np[0], b = np[1], cc = np[2], d = np[3];
            if (COUNT) ++c.nodes;
            float tn0, tn1;
            bool h0 = box_hit(a.x, a.y, a.z, b.x, b.y, b.z, r, tmin, tmax, tn0);
            bool h1 = box_hit(cc.x, cc.y, cc.z, d.x, d.y, d.z, r, tmin, tmax, tn1);
            uint32_t c0 = __float_as_uint(a.w), c1 = __float_as_uint(b.w);
            if (h0 && h1) {
                uint32_t nearc = c0, farc = c1;
                if (tn1 < tn0) {
                    nearc = c1;
                    farc = c0;
                }
                if (sp < (int)STACK_LDS) lds[sp * BLOCK] = farc;
                else if (sp < (int)STACK_CAP) spill[sp - STACK_LDS] = farc;
                else c.overflow = 1;
                sp = min(sp + 1, (int)STACK_CAP);
                cur = nearc;
            } else if (h0) {
                cur = c0;
            } else if (h1) {
                cur = c1;
            } else if (sp == 0) {
                cur = DONE;
            } else {
                --sp;
                cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
            }
            if (SPEC) {
                // postpone the first leaf found and keep walking until every
                // lane still in this loop holds one
                if (leaf == 0u && cur >= LEAF_BIT && cur != DONE) {
                    leaf = cur;
                    if (sp == 0) {
                        cur = DONE;
                    } else {
                        --sp;
                        cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
                    }
                }
                if (__ballot(leaf == 0u) == 0ull) break;
            }
        }
        // leaf phase: the postponed leaf first, else a leaf reached in order
        uint32_t lf = DONE;
        bool in_order = false;
        if (SPEC && leaf != 0u) {
            lf = leaf;
            leaf = 0u;
        } else if (cur >= LEAF_BIT && cur != DONE) {
            lf = cur;
            in_order = true;
        }
        if (lf != DONE) {
            uint32_t first = lf & 0x07FFFFFFu;
            uint32_t cnt = ((lf >> 27) & 0xFu) + 1u;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const gf4 tp = gptr(tris + k);
                float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                if (COUNT) ++c.tris;
                float t, b1, b2;
                if (intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tmax, t, b1, b2)) {
                    hit = true;
                    break;
                }
            }
            if (hit) {  // any hit ends the ray
                cur = DONE;
                leaf = 0u;
            } else if (in_order) {
                if (sp == 0) {
                    cur = DONE;
                } else {
                    --sp;
                    cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
                }
            }
        }
    }
    flush_counters(c, counters, COUNT);
}

template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace(DevScene S, const hr_ray* rays, uint32_t n, int any_hit,
                                                 hr_hit* hits, unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) {
        hr_ray ry = rays[i];
        RayP r = prep_ray(ld3p(ry.org), ld3p(ry.dir));
        HitD h;
        bool hit = any_hit ? traverse<true, COUNT>(S, r, ry.tmin, ry.tmax, lds, h, c)
                           : traverse<false, COUNT>(S, r, ry.tmin, ry.tmax, lds, h, c);
        hr_hit o;
        if (hit) {
            o.t = h.t;
            o.u = h.u;
            o.v = h.v;
            o.tri = h.tri;
        } else {
            o.t = -1.0f;
            o.u = 0.0f;
            o.v = 0.0f;
            o.tri = 0xFFFFFFFFu;
        }
        hits[i] = o;
    }
    flush_counters(c, counters, COUNT);
}

// zero the band counters (word b * BAND_STRIDE for band b)
__global__ void k_reset_u32(uint32_t* p) {
    if (threadIdx.x < SHADOW_QUEUES) p[threadIdx.x * BAND_STRIDE] = 0u;
}

__global__ void k_fill_bits(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// copies a shadow job passed by value (a kernel argument, so a captured
// graph keeps its own copy) into device memory, one 32-bit word per lane
static_assert(sizeof(ShadowJob) % 4 == 0, "ShadowJob is copied in 32-bit words");
constexpr uint32_t JOB_WORDS = (uint32_t)(sizeof(ShadowJob) / 4);
__global__ __launch_bounds__(256) void k_store_job(ShadowJob job, uint32_t* dst) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&job);
    for (uint32_t i = threadIdx.x; i < JOB_WORDS; i += blockDim.x) dst[i] = src[i];
}

inline uint32_t pixel_blocks(const FrameArgs& f) { return ((f.width + 15u) / 16u) * ((f.height + 15u) / 16u); }


// Sample-parallel shadow pass for a power-of-two sample count S in [4, 64]:
// the work items are (group, sample) pairs, numbered so a group's S samples
// are consecutive, and a lane traces one item at a time. Lanes that take work
// together take consecutive items, i.e. the nearly parallel rays of one pixel
// towards one light, which keeps a wave's rays coherent. An unoccluded sample
// sets bit (group << lg2s) | sample of a bit array (atomicOr: the result does
// not depend on the order); k_shadow_bits turns each group's bits into the
// factor count / S. Items are fetched from the same per-XCD bands as in
// k_shadows.
// Without adaptive sampling, k_group_kinds / k_tile_totals / k_scan_counts /
// k_tile_offsets / k_group_scatter first settle the groups that need no ray
// and list the others in order, and MODE 3 traces the S samples of each
// listed group (item = list slot * 2^lg2s + sample; the item count is read
// from *list_count on the device). With adaptive sampling
// (the nested sequence of adaptive_stop / adaptive_next) each group owns a
// window of 2^lg2s bits, bit s for sample s, that every stage adds to: MODE 1
// traces the probes (items numbered 4 per group; 1 probe for a hard light),
// k_adaptive_decide<true> settles the groups whose estimate is tight enough and
// lists the others, and MODE 2 traces a stage's samples [n0, n0 + size) for
// each listed group (item = list slot * size + sample - n0; the item count is
// read from *list_count on the device), after which k_adaptive_decide<false>
// settles or lists the groups again, up to 2^lg2s samples.
// Several frames (njobs > 1): groups are numbered frame after frame (the
// frames' group_base), every frame uses the stride 2^lg2s (the largest sample
// count) and skips samples at or beyond its own count.
__device__ __forceinline__ uint32_t job_of(const ShadowJob* __restrict__ jobs, uint32_t njobs, uint32_t g) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].group_base <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// The (pixel, light) part of a shadow item, the same for all its samples:
// group g of job J (tile t, light li, pixel q of the tile). ITEM_SKIP (factor
// 0: outside the frame, background, the light behind the surface), ITEM_LIT
// (factor 1: glass, beyond the light's reach) or ITEM_TRACE with the surface
// point, the unit vector to the light and its distance.
struct GroupRay {
    v3 P, L;
    float dist;
    bool isDir;
    uint32_t x, y;
};
// the unit vector from P to light lt and its distance (a directional light:
// its direction)
__device__ __forceinline__ void light_vector(const hr_light& lt, v3 P, GroupRay& o) {
    o.isDir = lt.pos_maxdist[3] < 0.0f;
    v3 L = o.isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), P);
    o.dist = length(L);
    o.L = normalize(L);
    o.P = P;
}
__device__ __forceinline__ int group_kind(const ShadowJob& J, uint32_t t, uint32_t li, uint32_t q, GroupRay& o) {
    const FrameArgs& F = J.F;
    uint32_t x, y;
    tile_pixel(J, t, q, x, y);
    if (x >= F.width || y >= F.height) return ITEM_SKIP;
    const size_t p = (size_t)y * F.width + x;
    const float4 P4 = J.G.pos[p], N4 = J.G.nrm[p];
    const bool background = P4.w == 0.0f && N4.w == 0.0f;
    if (background) return ITEM_SKIP;  // factor 0: no bit
    const int mode = (int)J.S.mats[mat_index(J.S, P4.w)].shading_metal_rough_idx[0];
    if (mode == 4) return ITEM_LIT;  // glass: factor 1
    const v3 N = normalize(ld3(N4));
    const hr_light& lt = F.lights[li];
    light_vector(lt, ld3(P4), o);
    const float NdotL = clamp01(dot(N, o.L));
    if (!(NdotL > 0.0f)) return ITEM_SKIP;  // factor 0
    if (!(o.isDir || o.dist < lt.pos_maxdist[3])) return ITEM_LIT;  // beyond the light's reach: factor 1
    o.x = x;
    o.y = y;
    return ITEM_TRACE;
}
// group_kind's ray data of a group it found ITEM_TRACE (the checks skipped)
__device__ __forceinline__ void group_ray(const ShadowJob& J, uint32_t t, uint32_t li, uint32_t q, GroupRay& o) {
    const FrameArgs& F = J.F;
    tile_pixel(J, t, q, o.x, o.y);
    light_vector(F.lights[li], ld3(J.G.pos[(size_t)o.y * F.width + o.x]), o);
}

// Ray source of the hybrid shadow items (MODE and MULTI as above).
template <int MODE, bool MULTI>
struct HybridItems {
    const ShadowJob* __restrict__ jobs;
    uint32_t njobs, lg2s;
    const uint32_t* __restrict__ list;
    const uint32_t* __restrict__ seeds;  // per group: tea of its pixel and frame
    uint32_t n0, size;  // MODE 2: the stage's first sample and sample count
    const TNode4* nodes4;
    const STri* tris;
    uint32_t node_count;

    __device__ __forceinline__ int setup(uint32_t item, uint32_t& bit, RayP& r, float& tmax, const TNode4*& n4,
                                         const STri*& tr, uint32_t& nc, Cnt& c) const {
        static_assert(MODE >= 1 && MODE <= 3, "MODE 1 and 2: adaptive stages; MODE 3: listed groups");
        uint32_t G, s;  // global group, sample
        if (MODE == 1) {
            G = list[item >> 2];
            s = item & 3u;
        } else if (MODE == 3) {
            G = list[item >> lg2s];
            s = item & ((1u << lg2s) - 1u);
        } else {
            const uint32_t slot = item / size;
            G = list[slot];
            s = n0 + (item - slot * size);
        }
        bit = (G << lg2s) | s;
        const ShadowJob& J = jobs[MULTI ? job_of(jobs, njobs, G) : 0u];
        const FrameArgs& F = J.F;
        const uint32_t SS = F.samples;
        if (s >= SS) return ITEM_SKIP;  // beyond this frame's sample count
        const uint32_t g = G - J.group_base;
        const uint32_t per_tile = F.nl * 64u;
        const uint32_t t = g / per_tile, rem = g - t * per_tile;
        const uint32_t li = rem >> 6, q = rem & 63u;
        const bool hard = (MODE == 1 || MODE == 2) && hard_light(F, li);
        if (hard && s > 0u) return ITEM_SKIP;  // a hard light takes one probe
        // a listed group needs rays (k_group_kinds / k_adaptive_decide): its
        // ray data without the checks, its pixel's seed as k_group_kinds stored it
        GroupRay gr;
        group_ray(J, t, li, q, gr);
        const v3 P = gr.P, L = gr.L;
        const float dist = gr.dist;
        const bool isDir = gr.isDir;
        const uint32_t seed = seeds[G];
        v3 dir;
        if (MODE == 3) {
            Pcg gs = ray_stream(seed, PURPOSE_SURFACE_SHADOW, li * SS + s);
            dir = light_sample_dir(F, li, P, L, gs, s, SS);
        } else {  // adaptive: sample s of the light's nested sequence
            Pcg gs = ray_stream(seed, PURPOSE_ADAPTIVE_SHADOW, li * 128u + s);
            dir = light_sample_dir_nested(F, li, P, L, gs, s, hard ? 1u : SS);
        }
        tmax = isDir ? 1000.0f : dist + 1.0f;
        r = prep_ray(add(P, scl(dir, 1e-1f)), dir);
        ++c.shadow;
        if (MULTI) {  // one frame: the scene stays uniform across the wave
            n4 = J.S.nodes4;
            tr = J.S.tris;
            nc = J.S.node_count;
        }
        return ITEM_TRACE;
    }
};

// OFF32 (one frame whose 4-wide node array is below 4 GiB): nodes are
// addressed by 32-bit offsets from the wave-uniform array base.
template <int MODE, bool MULTI, bool OFF32>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_shadows_spl(
    const ShadowJob* __restrict__ jobs, uint32_t njobs, uint32_t nitems, uint32_t lg2s, uint32_t* next_item,
    uint32_t* bits, const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count,
    const uint32_t* __restrict__ seeds, uint32_t n0, uint32_t size, unsigned long long* counters) {
    static_assert(!(MULTI && OFF32), "frames of a batch have their own node arrays");
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    Cnt c = {};
    if (MODE == 1) nitems = *list_count << 2;
    if (MODE == 2) nitems = *list_count * size;
    if (MODE == 3) nitems = *list_count << lg2s;
    HybridItems<MODE, MULTI> src{jobs, njobs, lg2s, list, seeds, n0, size, jobs[0].S.nodes4, jobs[0].S.tris,
                                 jobs[0].S.node_count};
    item_loop<OFF32>(src, nitems, next_item, bits, stk + threadIdx.x, c);
    flush_counters(c, counters, false);
}

// launches k_shadows_spl<MODE, ...> for njobs frames
template <int MODE>
void launch_spl_mode(bool multi, bool off32, uint32_t nb, hipStream_t st, const ShadowJob* jobs, uint32_t njobs,
                     uint32_t nitems, uint32_t lg2s, uint32_t* next_item, uint32_t* bits, const uint32_t* list,
                     const uint32_t* list_count, const uint32_t* seeds, uint32_t n0, uint32_t size,
                     unsigned long long* counters) {
    if (multi)
        k_shadows_spl<MODE, true, false><<<nb, BLOCK, 0, st>>>(jobs, njobs, nitems, lg2s, next_item, bits, list,
                                                               list_count, seeds, n0, size, counters);
    else if (off32)
        k_shadows_spl<MODE, false, true><<<nb, BLOCK, 0, st>>>(jobs, njobs, nitems, lg2s, next_item, bits, list,
                                                               list_count, seeds, n0, size, counters);
    else
        k_shadows_spl<MODE, false, false><<<nb, BLOCK, 0, st>>>(jobs, njobs, nitems, lg2s, next_item, bits, list,
                                                                list_count, seeds, n0, size, counters);
}

// shadow factor of every group from the bits of its samples
__global__ __launch_bounds__(256) void k_shadow_bits(const ShadowJob* __restrict__ jobs, uint32_t njobs,
                                                     uint32_t ngroups, uint32_t lg2s,
                                                     const uint32_t* __restrict__ bits) {
    const uint32_t G = blockIdx.x * blockDim.x + threadIdx.x;
    if (G >= ngroups) return;
    const ShadowJob& J = jobs[njobs > 1u ? job_of(jobs, njobs, G) : 0u];
    const FrameArgs& F = J.F;
    const uint32_t g = G - J.group_base;
    const uint32_t per_tile = F.nl * 64u;
    const uint32_t t = g / per_tile, rem = g - t * per_tile;
    const uint32_t li = rem >> 6, q = rem & 63u;
    uint32_t x, y;
    tile_pixel(J, t, q, x, y);
    if (x >= F.width || y >= F.height) return;
    const uint32_t n = window_bits(bits, (uint64_t)G << lg2s, lg2s);
    put_factor(J.shadow, li * ((size_t)F.width * F.height) + (size_t)y * F.width + x, (float)n / (float)F.samples,
               F.unorm8);
}

// Non-adaptive passes: every (pixel, light) group decided once before the
// samples are traced. A group lit without rays gets the first F.samples bits
// of its window (the bits its samples would set one by one), a group without
// rays and factor 0 gets nothing, and the groups to trace are listed in group
// order (so the trace's work bands keep neighbouring tiles together): each
// chunk of 64 groups records the mask of its groups to trace, the chunks'
// counts are scanned, and the groups are scattered to their places. The
// trace then takes only the listed groups' samples (MODE 3).
// probes: 0 without adaptive sampling (a lit group's F.samples bits), else
// the probe count of the adaptive first stage (a lit group's probe bits: one
// for a hard light, else min(F.samples, probes)).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_group_kinds(const ShadowJob* __restrict__ jobs, uint32_t njobs,
                                                     uint32_t ngroups, uint32_t lg2s, uint32_t probes,
                                                     uint32_t* bits, uint32_t* seeds, unsigned long long* masks,
                                                     uint32_t* counts) {
    const uint32_t G = blockIdx.x * blockDim.x + threadIdx.x;
    int kind = ITEM_SKIP;
    if (G < ngroups) {
        const ShadowJob& J = jobs[MULTI ? job_of(jobs, njobs, G) : 0u];
        const FrameArgs& F = J.F;
        const uint32_t g = G - J.group_base;
        const uint32_t per_tile = F.nl * 64u;
        const uint32_t t = g / per_tile, rem = g - t * per_tile;
        const uint32_t li = rem >> 6;
        GroupRay gr;
        kind = group_kind(J, t, li, rem & 63u, gr);
        if (kind == ITEM_TRACE) seeds[G] = tea(gr.y * F.width + gr.x, F.frame);
        if (kind == ITEM_LIT) {
            uint64_t b = (uint64_t)G << lg2s;
            const uint32_t lit = probes == 0u ? F.samples : (hard_light(F, li) ? 1u : min(F.samples, probes));
            for (uint32_t n = lit; n > 0u;) {
                const uint32_t off = (uint32_t)(b & 31u), take = min(n, 32u - off);
                const uint32_t m = (take == 32u ? 0xFFFFFFFFu : ((1u << take) - 1u)) << off;
                atomicOr(bits + (b >> 5), m);
                b += take;
                n -= take;
            }
        }
    }
    const unsigned long long want = __ballot(kind == ITEM_TRACE);
    if ((threadIdx.x & 63u) == 0u && G < ngroups) {
        masks[G >> 6] = want;
        counts[G >> 6] = (uint32_t)__popcll(want);
    }
}

// unoccluded samples among the first N (<= 2^lg2s <= 64) of the window at bit b0
__device__ __forceinline__ uint32_t prefix_bits(const uint32_t* __restrict__ bits, uint64_t b0, uint32_t lg2s,
                                                uint32_t N) {
    auto mask = [](uint32_t n) { return n >= 32u ? 0xFFFFFFFFu : (1u << n) - 1u; };
    if (lg2s >= 5u) {
        uint32_t k = (uint32_t)__popc(bits[b0 >> 5] & mask(N));
        if (N > 32u) k += (uint32_t)__popc(bits[(b0 >> 5) + 1u] & mask(N - 32u));
        return k;
    }
    return (uint32_t)__popc((bits[b0 >> 5] >> (uint32_t)(b0 & 31u)) & mask(N));
}

// Adaptive sampling between stages. FIRST: every group after its probes (a
// hard light's factor is its one probe); else the groups of list_in after the
// stage that ended at sample n_end. A group whose estimate k / N is tight
// enough (adaptive_stop, N = min(n_end, its frame's S)) gets its factor, the
// others go to list_out for the next stage.
template <bool FIRST>
__global__ __launch_bounds__(256) void k_adaptive_decide(const ShadowJob* __restrict__ jobs, uint32_t njobs,
                                                         uint32_t ngroups, uint32_t lg2s,
                                                         const uint32_t* __restrict__ bits,
                                                         const uint32_t* __restrict__ list_in,
                                                         const uint32_t* __restrict__ count_in, uint32_t n_end,
                                                         uint32_t* list_out, uint32_t* count_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (FIRST ? ngroups : *count_in)) return;
    const uint32_t G = FIRST ? i : list_in[i];
    const ShadowJob& J = jobs[njobs > 1u ? job_of(jobs, njobs, G) : 0u];
    const FrameArgs& F = J.F;
    const uint32_t SS = F.samples;
    const uint32_t g = G - J.group_base;
    const uint32_t per_tile = F.nl * 64u;
    const uint32_t t = g / per_tile, rem = g - t * per_tile;
    const uint32_t li = rem >> 6, q = rem & 63u;
    uint32_t x, y;
    tile_pixel(J, t, q, x, y);
    if (x >= F.width || y >= F.height) return;
    const size_t idx = li * ((size_t)F.width * F.height) + (size_t)y * F.width + x;
    const uint64_t b0 = (uint64_t)G << lg2s;
    if (FIRST && hard_light(F, li)) {
        put_factor(J.shadow, idx, (float)prefix_bits(bits, b0, lg2s, 1u) / 1.0f, F.unorm8);
        return;
    }
    const uint32_t N = FIRST ? adaptive_first(SS) : min(n_end, SS);
    const uint32_t k = prefix_bits(bits, b0, lg2s, N);
    if (adaptive_stop(k, N, SS)) {
        put_factor(J.shadow, idx, (float)k / (float)N, F.unorm8);
        return;
    }
    list_out[atomicAdd(count_out, 1u)] = G;
}

}  // namespace

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st) {
    k_zero_u64<<<(n + 63) / 64, 64, 0, st>>>(p, n);
    return hipGetLastError();
}
hipError_t launch_store_job(const ShadowJob& job, ShadowJob* dst, hipStream_t st) {
    k_store_job<<<1, 256, 0, st>>>(job, reinterpret_cast<uint32_t*>(dst));
    return hipGetLastError();
}
hipError_t launch_gbuffer(const DevScene& s, const FrameArgs& f, GBufView gb, unsigned long long* counters, bool count,
                          hipStream_t st) {
    if (count) k_gbuffer<true><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, counters);
    else k_gbuffer<false><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, counters);
    return hipGetLastError();
}
uint32_t shadow_groups(const FrameArgs& f, uint32_t tile_w) {
    uint32_t tw = tile_w, th = 64u / tile_w;
    uint32_t tiles = ((f.width + tw - 1u) / tw) * ((f.height + th - 1u) / th);
    return tiles * f.nl * 64u;
}

// the stage after one that ended at sample N of at most S (adaptive_next)
static uint32_t host_adaptive_next(uint32_t N, uint32_t S) {
    const uint32_t n = N == 4u ? 16u : 2u * N;
    return n < S ? n : S;
}

size_t shadows_spl_scratch(uint32_t ngroups, uint32_t lg2s, bool adaptive) {
    const size_t w = ((((size_t)ngroups << lg2s) + 31u) >> 5) * 4u;
    // bits, the seeds, the list of groups to trace and its count, the tail of
    // list_groups
    if (!adaptive) return w + 2u * (size_t)ngroups * 4u + 4u + group_list_tail_bytes(ngroups);
    // bits, the seeds, two lists of groups, their two counts on lines of their
    // own, the tail of list_groups
    return w + 3u * (size_t)ngroups * 4u + 2u * 256u + group_list_tail_bytes(ngroups);
}

// The groups that need rays, listed in group order into list / *count (the
// others settled in bits, see k_group_kinds); tail: scratch for the chunk
// masks, counts and tile totals (list_tail_bytes).
static void list_groups(bool multi, const ShadowJob* jobs_dev, uint32_t njobs, uint32_t ngroups, uint32_t lg2s,
                        uint32_t probes, uint32_t* bits, uint32_t* seeds, uint32_t* list, uint32_t* count,
                        void* tail, hipStream_t st) {
    const GroupListTail t = group_list_tail(tail, ngroups);
    const uint32_t gblocks = (ngroups + 255u) / 256u;
    if (multi)
        k_group_kinds<true><<<gblocks, 256, 0, st>>>(jobs_dev, njobs, ngroups, lg2s, probes, bits, seeds,
                                                     t.masks, t.counts);
    else
        k_group_kinds<false><<<gblocks, 256, 0, st>>>(jobs_dev, njobs, ngroups, lg2s, probes, bits, seeds,
                                                      t.masks, t.counts);
    scan_and_scatter(ngroups, t, count, list, st);
}

hipError_t launch_shadows_spl(const ShadowJob* jobs_dev, uint32_t njobs, uint32_t ngroups, uint32_t lg2s,
                              bool adaptive, unsigned long long* counters, hipStream_t st, uint32_t* next_item,
                              uint32_t* scratch, uint32_t resident_blocks, uint32_t wide_count) {
    const bool multi = njobs > 1u;
    const bool off32 = !multi && (uint64_t)wide_count * sizeof(TNode4) <= 0xFFFFFFFFull;
    const uint64_t items = (uint64_t)ngroups << lg2s;
    if (items >= (1ull << 32)) return hipErrorInvalidValue;
    const uint32_t words = (uint32_t)((items + 31u) >> 5);
    uint32_t nb = min(resident_blocks, (uint32_t)((items + BLOCK - 1) / BLOCK));
    if (nb == 0) return hipSuccess;
    uint32_t* bits = scratch;
    k_reset_u32<<<1, 64, 0, st>>>(next_item);
    k_fill_bits<<<(words + 255u) / 256u, 256, 0, st>>>(bits, words);
    uint32_t* seeds = bits + words;  // per group, written for the groups that need rays
    if (!adaptive) {
        // scratch: bits, seeds, the list of groups to trace, its count, the tail
        uint32_t* list = seeds + ngroups;
        uint32_t* count = list + ngroups;
        list_groups(multi, jobs_dev, njobs, ngroups, lg2s, 0u, bits, seeds, list, count, count + 1, st);
        launch_spl_mode<3>(multi, off32, nb, st, jobs_dev, njobs, 0u, lg2s, next_item, bits, list, count, seeds, 0u,
                           0u, counters);
        k_shadow_bits<<<(ngroups + 255u) / 256u, 256, 0, st>>>(jobs_dev, njobs, ngroups, lg2s, bits);
        return hipGetLastError();
    }
    // scratch: bits (2^lg2s per group, every stage adds its own), seeds, lists
    // A and B, their counts 64 words apart, the tail of list_groups. The probe
    // stage traces the groups listed in B (free until the first refinement
    // stage lists into it).
    uint32_t* lists[2] = {seeds + ngroups, seeds + 2u * (size_t)ngroups};
    uint32_t* counts[2] = {seeds + 3u * (size_t)ngroups, seeds + 3u * (size_t)ngroups + 64u};
    const uint32_t S = 1u << lg2s;
    const uint32_t gb = (ngroups + 255u) / 256u;
    k_fill_bits<<<1, 64, 0, st>>>(counts[0], 1u);
    list_groups(multi, jobs_dev, njobs, ngroups, lg2s, 4u, bits, seeds, lists[1], counts[1], counts[1] + 64u, st);
    const uint32_t pitems = ngroups << 2;
    launch_spl_mode<1>(multi, off32, min(resident_blocks, (pitems + BLOCK - 1) / BLOCK), st, jobs_dev, njobs, 0u,
                       lg2s, next_item, bits, lists[1], counts[1], seeds, 0u, 4u, counters);
    k_adaptive_decide<true><<<gb, 256, 0, st>>>(jobs_dev, njobs, ngroups, lg2s, bits, nullptr, nullptr, 4u,
                                                lists[0], counts[0]);
    int cur = 0;
    for (uint32_t N = 4u; N < S;) {
        const uint32_t Nn = host_adaptive_next(N, S);
        const uint32_t size = Nn - N;
        const uint64_t most = (uint64_t)ngroups * size;
        k_reset_u32<<<1, 64, 0, st>>>(next_item);
        k_fill_bits<<<1, 64, 0, st>>>(counts[cur ^ 1], 1u);
        const uint64_t want = (most + BLOCK - 1) / BLOCK;
        launch_spl_mode<2>(multi, off32, want < resident_blocks ? (uint32_t)want : resident_blocks, st,
                           jobs_dev, njobs, 0u, lg2s, next_item, bits, lists[cur], counts[cur], seeds, N, size,
                           counters);
        k_adaptive_decide<false><<<gb, 256, 0, st>>>(jobs_dev, njobs, 0u, lg2s, bits, lists[cur], counts[cur], Nn,
                                                     lists[cur ^ 1], counts[cur ^ 1]);
        cur ^= 1;
        N = Nn;
    }
    return hipGetLastError();
}

hipError_t launch_shadows(const ShadowJob* jobs_dev, uint32_t njobs, uint32_t ngroups, unsigned long long* counters,
                          bool count, hipStream_t st, uint32_t* next_group, uint32_t resident_blocks) {
    uint32_t nb = min(resident_blocks, (ngroups + BLOCK - 1) / BLOCK);
    if (nb == 0) return hipSuccess;
    k_reset_u32<<<1, 64, 0, st>>>(next_group);
    bool batch = njobs > 1;
    if (count && batch) k_shadows<true, true><<<nb, BLOCK, 0, st>>>(jobs_dev, njobs, ngroups, next_group, counters);
    else if (count) k_shadows<true, false><<<nb, BLOCK, 0, st>>>(jobs_dev, njobs, ngroups, next_group, counters);
    else if (batch) k_shadows<false, true><<<nb, BLOCK, 0, st>>>(jobs_dev, njobs, ngroups, next_group, counters);
    else k_shadows<false, false><<<nb, BLOCK, 0, st>>>(jobs_dev, njobs, ngroups, next_group, counters);
    return hipGetLastError();
}
hipError_t launch_trace_rays(const DevScene& s, const hr_ray* rays, uint32_t n, int any_hit, hr_hit* hits,
                             unsigned long long* counters, bool count, hipStream_t st) {
    uint32_t nb = (n + BLOCK - 1) / BLOCK;
    if (nb == 0) return hipSuccess;
    if (count) k_trace<true><<<nb, BLOCK, 0, st>>>(s, rays, n, any_hit, hits, counters);
    else k_trace<false><<<nb, BLOCK, 0, st>>>(s, rays, n, any_hit, hits, counters);
    return hipGetLastError();
}

}  // namespace hr
