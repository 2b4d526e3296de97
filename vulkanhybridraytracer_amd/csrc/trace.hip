// trace.hip - traversal and shading kernels of the hybrid pass (gfx950).
//
//   k_gbuffer   primary visibility at pixel centres (raygen.rgen:29-40) into
//               the FP32 G-buffer (geometry_shader.frag:51-76, skybox.frag:16-21)
//   k_shadows   shadowRaygen.rgen:28-93 (+ shadow.rmiss, shadowHit.rchit)
//   k_trace     bare traceRayEXT queries
//
// One lane per pixel; a 256-thread workgroup covers 16x16 pixels and each
// 64-lane wave an 8x8 tile, so a wave's rays start close together and walk
// similar nodes. Workgroups are renumbered so that consecutive screen tiles
// land on the same XCD (dispatch assigns workgroup i to XCD i mod 8), which
// keeps a screen region's working set in one L2.
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
        float ux = ((float)x + 0.5f) / (float)F.width;
        float uy = ((float)y + 0.5f) / (float)F.height;
        float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
        float org[4], tgt[4], dir[4];
        mul4(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, org);
        mul4(F.inv_proj, dx, dy, 1.0f, 1.0f, tgt);
        v3 tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
        mul4(F.inv_view, tn.x, tn.y, tn.z, 0.0f, dir);
        v3 o = mk(org[0], org[1], org[2]), d = mk(dir[0], dir[1], dir[2]);
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
            v3 n0 = ld3(S.vnormal[pr.vertex_base + S.indices[ib + 0]]);
            v3 n1 = ld3(S.vnormal[pr.vertex_base + S.indices[ib + 1]]);
            v3 n2 = ld3(S.vnormal[pr.vertex_base + S.indices[ib + 2]]);
            float b1 = h.u, b2 = h.v, b0 = 1.0f - b1 - b2;
            v3 nrm = add(add(scl(n0, b0), scl(n1, b1)), scl(n2, b2));
            const float* M = pr.nm;
            v3 N = normalize(mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z,
                                (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                                (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z));
            v3 wp = add(ro, scl(d, h.t));
            const hr_material& mt = S.mats[pr.material];
            P = make_float4(wp.x, wp.y, wp.z, mt.shading_metal_rough_idx[3]);
            Nn = make_float4(N.x, N.y, N.z, 1.0f);
            A = make_float4(mt.diffuse_ior[0], mt.diffuse_ior[1], mt.diffuse_ior[2], 1.0f);
            Mt = make_float4(0.0f, mt.shading_metal_rough_idx[2], mt.shading_metal_rough_idx[1], 0.0f);
        }
        G.pos[p] = P;
        G.nrm[p] = Nn;
        G.alb[p] = A;
        G.mat[p] = Mt;
        G.emi[p] = E;
    }
    flush_counters(c, counters, COUNT);
}

template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_shadows(DevScene S, FrameArgs F, GBufView G, float* shadow,
                                                   unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t x, y;
    pixel_of(F.width, x, y);
    if (x < F.width && y < F.height) {
        size_t plane = (size_t)F.width * F.height;
        size_t p = (size_t)y * F.width + x;
        uint32_t seed = tea(y * F.width + x, F.frame);
        float4 P4 = G.pos[p], N4 = G.nrm[p];
        bool background = P4.w == 0.0f && N4.w == 0.0f;
        int mode = (int)S.mats[mat_index(S, P4.w)].shading_metal_rough_idx[0];
        if (background || mode == 4) {
            float v = background ? 0.0f : 1.0f;
            for (uint32_t i = 0; i < F.nl; ++i) shadow[i * plane + p] = v;
        } else {
            v3 P = ld3(P4);
            v3 N = normalize(ld3(N4));
            uint32_t SS = F.samples;
            for (uint32_t i = 0; i < F.nl; ++i) {
                const hr_light& lt = F.lights[i];
                bool isDir = lt.pos_maxdist[3] < 0.0f;
                v3 L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), P);
                float maxd = lt.pos_maxdist[3];
                float dist = length(L);
                L = normalize(L);
                float NdotL = clamp01(dot(N, L));
                float factor = 0.0f;
                if (NdotL > 0.0f) {
                    for (uint32_t s = 0; s < SS; ++s) {
                        bool shadowed;
                        if (isDir || dist < maxd) {
                            Pcg g = ray_stream(seed, PURPOSE_SURFACE_SHADOW, i * SS + s);
                            v3 dir = light_sample_dir(F, i, P, L, g, s, SS);
                            float tmax = isDir ? 1000.0f : dist + 1.0f;
                            ++c.shadow;
                            shadowed = occluded<COUNT>(S, add(P, scl(dir, 1e-1f)), dir, 0.001f, tmax, lds, c);
                            if (shadowed) ++c.occluded;
                        } else {
                            shadowed = false;
                        }
                        if (!shadowed) factor += 1.0f;
                    }
                    factor /= (float)SS;
                }
                shadow[i * plane + p] = factor;
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

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

inline uint32_t pixel_blocks(const FrameArgs& f) { return ((f.width + 15u) / 16u) * ((f.height + 15u) / 16u); }

}  // namespace

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st) {
    k_zero_u64<<<(n + 63) / 64, 64, 0, st>>>(p, n);
    return hipGetLastError();
}
hipError_t launch_gbuffer(const DevScene& s, const FrameArgs& f, GBufView gb, unsigned long long* counters, bool count,
                          hipStream_t st) {
    if (count) k_gbuffer<true><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, counters);
    else k_gbuffer<false><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, counters);
    return hipGetLastError();
}
hipError_t launch_shadows(const DevScene& s, const FrameArgs& f, GBufView gb, float* shadow,
                          unsigned long long* counters, bool count, hipStream_t st) {
    if (count) k_shadows<true><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, shadow, counters);
    else k_shadows<false><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, shadow, counters);
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
