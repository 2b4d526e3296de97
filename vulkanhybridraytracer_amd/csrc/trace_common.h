// trace_common.h - traversal, texture, BRDF and sampling device functions
// shared by the trace kernels (trace.hip) and the wavefront reflection
// pipeline (wavefront.hip).
#pragma once

#include "device_math.h"
#include "internal.h"

namespace hr {
namespace tc {
using namespace hrd;


constexpr int BLOCK = 256;
constexpr float PI_H = 3.14159265359f;  // helpers.glsl:5
constexpr float PI_M = 3.141592f;       // miss.rmiss:11

struct Cnt {
    uint32_t shadow, refl, hitshadow, occluded, nodes, tris, primary, overflow;
};

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    uint32_t xcd = b & 7u, slot = b >> 3;
    uint32_t per = n >> 3, rem = n & 7u;
    return xcd * per + min(xcd, rem) + slot;
}
__device__ __forceinline__ void pixel_of(uint32_t W, uint32_t& x, uint32_t& y) {
    uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
    uint32_t tx = (W + 15u) / 16u;
    uint32_t bx = bid % tx, by = bid / tx;
    uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    x = bx * 16u + (w & 1u) * 8u + (lane & 7u);
    y = by * 16u + (w >> 1) * 8u + (lane >> 3);
}

struct HitD {
    float t, u, v;
    uint32_t tri;
};

// compare-exchange of (key, code) pairs, ascending by key (selects, no branch)
__device__ __forceinline__ void sort_pair(float& ka, uint32_t& ca, float& kb, uint32_t& cb) {
    const bool sw = kb < ka;
    const float k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const uint32_t c0 = sw ? cb : ca, c1 = sw ? ca : cb;
    ka = k0;
    kb = k1;
    ca = c0;
    cb = c1;
}

constexpr uint32_t DONE = 0xFFFFFFFFu;  // no node: no leaf code reaches it (fewer than 2^27 triangles)
// The inner traversal loops also end when fewer lanes than this would keep
// walking, so lanes whose ray ended or that hold two leaves get back to work
// (leaf tests, new rays) instead of idling; results do not depend on it.
// consume_rays (queues of closest-hit and hit-shadow rays) and item_loop
// (sample-parallel shadow items) have their own. The item loop refills early:
// lanes that start their rays together walk the same nodes, which a
// wave-uniform node step reads through the scalar cache.
constexpr int REFILL_MIN_LANES = 16;
constexpr int REFILL_MIN_ITEMS = 56;

// Node and triangle arrays live in global memory; reading them through
// global-address-space pointers gives global (not flat) loads, whose waits do
// not also wait for the LDS stack traffic.
// (The host compilation pass of this file sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const float4 __attribute__((address_space(1)))* gf4;
typedef const uint4 __attribute__((address_space(1)))* gu4;
// read-only for the kernel's lifetime: a wave-uniform address loads through
// the scalar cache
typedef const float4 __attribute__((address_space(4)))* cf4;
typedef const uint4 __attribute__((address_space(4)))* cu4;
#else
typedef const float4* gf4;
typedef const uint4* gu4;
typedef const float4* cf4;
typedef const uint4* cu4;
#endif
__device__ __forceinline__ gf4 gptr(const void* p) { return (gf4)(p); }

// One step over a 4-wide node: its children's boxes are tested against
// [tmin, tmax], the hits sorted near to far; all but the nearest are pushed
// (far first, so the nearer ones pop first) and the nearest is returned. With
// no hit the next entry is popped (DONE when the stack is empty). The stack
// holds STACK_LDS entries per lane in LDS and the rest in `spill`, CAP in
// all; a push beyond CAP sets the overflow flag.
template <uint32_t CAP>
__device__ __forceinline__ uint32_t wide_step(const TNode4* nodes4, uint32_t cur, const RayP& r, float tmin,
                                              float tmax, uint32_t* lds, uint32_t* spill, int& sp, Cnt& c) {
    const gf4 np = gptr(nodes4 + cur);
    const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
    const uint4 cd = ((gu4)np)[6];
    const float inf = __uint_as_float(0x7f800000u);
    float t0, t1, t2, t3;
    const bool b0 = box_hit(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, r, tmin, tmax, t0) && cd.x != EMPTY4;
    const bool b1 = box_hit(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, r, tmin, tmax, t1) && cd.y != EMPTY4;
    const bool b2 = box_hit(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, r, tmin, tmax, t2) && cd.z != EMPTY4;
    const bool b3 = box_hit(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, r, tmin, tmax, t3) && cd.w != EMPTY4;
    float k0 = b0 ? t0 : inf, k1 = b1 ? t1 : inf, k2 = b2 ? t2 : inf, k3 = b3 ? t3 : inf;
    uint32_t c0 = cd.x, c1 = cd.y, c2 = cd.z, c3 = cd.w;
    sort_pair(k0, c0, k1, c1);
    sort_pair(k2, c2, k3, c3);
    sort_pair(k0, c0, k2, c2);
    sort_pair(k1, c1, k3, c3);
    sort_pair(k1, c1, k2, c2);
    const uint32_t nh = (uint32_t)b0 + (uint32_t)b1 + (uint32_t)b2 + (uint32_t)b3;
    if (nh == 0u) {
        if (sp == 0) return DONE;
        --sp;
        return sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
    }
    auto push = [&](uint32_t code) {
        if (sp < (int)STACK_LDS) lds[sp * BLOCK] = code;
        else if (sp < (int)CAP) spill[sp - STACK_LDS] = code;
        else c.overflow = 1;
        sp = min(sp + 1, (int)CAP);
    };
    if (nh > 3u) push(c3);
    if (nh > 2u) push(c2);
    if (nh > 1u) push(c1);
    return c0;
}

// Entry sp of a lane's stack (16 in LDS, the rest in `spill`): the LDS word is
// read whatever sp is (an LDS read, not a generic load that waits on memory
// too), and replaced from `spill` in the rare deep case.
__device__ __forceinline__ uint32_t stack_at(const uint32_t* lds, const uint32_t* spill, int sp) {
    uint32_t v = lds[min(sp, (int)STACK_LDS - 1) * BLOCK];
    if (sp >= (int)STACK_LDS) v = spill[sp - STACK_LDS];
    return v;
}

// Byte offsets in a 4-wide node of a ray's near x, y and z planes: the lo
// planes where the direction component is positive, the hi planes where it is
// negative (the inverse direction keeps the sign even for a zero component);
// the far plane of each axis is at offset ^ 16.
struct NearPlanes {
    uint32_t x, y, z;
};
__device__ __forceinline__ NearPlanes near_planes(const RayP& r) {
    return NearPlanes{r.idir.x < 0.0f ? 16u : 0u, r.idir.y < 0.0f ? 48u : 32u, r.idir.z < 0.0f ? 80u : 64u};
}

// One any-hit step over a 4-wide node, as wide_step but with each axis's
// near and far planes loaded by the ray's direction signs. The slab distances
// are those of box_hit (fma is monotone in the plane for a fixed inverse
// direction, so the near plane's distance is the min of the two and the far
// plane's the max): the same boxes are entered. An empty slot's inverted box
// (lo = +inf, hi = -inf) then misses by itself: its near distance is +inf or
// its far distance -inf (at most one of the two can be NaN), so no code check
// is needed. OFF32: the node array is addressed by 32-bit byte offsets from
// its (wave-uniform) base, which holds for arrays below 4 GiB.
template <bool OFF32>
__device__ __forceinline__ uint32_t wide_step_any(const TNode4* nodes4, uint32_t cur, const RayP& r,
                                                  const NearPlanes& pl, float tmin, float tmax, uint32_t* lds,
                                                  uint32_t* spill, int& sp, Cnt& c) {
    float4 nx, fx, ny, fy, nz, fz;
    uint4 cd;
    const uint32_t u = __builtin_amdgcn_readfirstlane(cur);
    if (OFF32 && __ballot(cur != u) == 0ull) {
        // every walking lane is at node u: it is read once, through the
        // scalar cache, and each lane picks its near and far planes
        const cf4 q = (cf4)(nodes4 + u);
        const float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
        cd = ((cu4)q)[6];
        const bool sx = pl.x != 0u, sy = pl.y != 32u, sz = pl.z != 64u;
        nx = sx ? hx : lx;
        fx = sx ? lx : hx;
        ny = sy ? hy : ly;
        fy = sy ? ly : hy;
        nz = sz ? hz : lz;
        fz = sz ? lz : hz;
    } else if (OFF32) {
        const char* b = reinterpret_cast<const char*>(nodes4);
        const uint32_t o = cur << 7;
        nx = *gptr(b + (o + pl.x));
        fx = *gptr(b + (o + (pl.x ^ 16u)));
        ny = *gptr(b + (o + pl.y));
        fy = *gptr(b + (o + (pl.y ^ 16u)));
        nz = *gptr(b + (o + pl.z));
        fz = *gptr(b + (o + (pl.z ^ 16u)));
        cd = *(gu4)gptr(b + (o + 96u));
    } else {
        const char* b = reinterpret_cast<const char*>(nodes4 + cur);
        nx = *gptr(b + pl.x);
        fx = *gptr(b + (pl.x ^ 16u));
        ny = *gptr(b + pl.y);
        fy = *gptr(b + (pl.y ^ 16u));
        nz = *gptr(b + pl.z);
        fz = *gptr(b + (pl.z ^ 16u));
        cd = *(gu4)gptr(b + 96u);
    }
    auto slab = [&](float ax, float bx, float ay, float by, float az, float bz, float& tn) {
        const float tnx = fmaf(ax, r.idir.x, -r.ood.x), tfx = fmaf(bx, r.idir.x, -r.ood.x);
        const float tny = fmaf(ay, r.idir.y, -r.ood.y), tfy = fmaf(by, r.idir.y, -r.ood.y);
        const float tnz = fmaf(az, r.idir.z, -r.ood.z), tfz = fmaf(bz, r.idir.z, -r.ood.z);
        const float n = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        const float f = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        tn = n;
        return n <= f * 1.0000002384185791f;
    };
    const float inf = __uint_as_float(0x7f800000u);
    float t0, t1, t2, t3;
    const bool b0 = slab(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, t0);
    const bool b1 = slab(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, t1);
    const bool b2 = slab(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, t2);
    const bool b3 = slab(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, t3);
    float k0 = b0 ? t0 : inf, k1 = b1 ? t1 : inf, k2 = b2 ? t2 : inf, k3 = b3 ? t3 : inf;
    uint32_t c0 = cd.x, c1 = cd.y, c2 = cd.z, c3 = cd.w;
    sort_pair(k0, c0, k1, c1);
    sort_pair(k2, c2, k3, c3);
    sort_pair(k0, c0, k2, c2);
    sort_pair(k1, c1, k3, c3);
    sort_pair(k1, c1, k2, c2);
    const uint32_t nh = (uint32_t)b0 + (uint32_t)b1 + (uint32_t)b2 + (uint32_t)b3;
    if (nh == 0u) {
        if (sp == 0) return DONE;
        --sp;
        return stack_at(lds, spill, sp);
    }
    auto push = [&](uint32_t code) {
        if (sp < (int)STACK_LDS) lds[sp * BLOCK] = code;
        else if (sp < (int)MAX_STACK4) spill[sp - STACK_LDS] = code;
        else c.overflow = 1;
        sp = min(sp + 1, (int)MAX_STACK4);
    };
    if (nh > 3u) push(c3);
    if (nh > 2u) push(c2);
    if (nh > 1u) push(c1);
    return c0;
}

// Closest-hit (ANY = false) or any-hit traversal of one ray. The counting
// variant walks the binary nodes in near-first order (the oracle's order and
// counts); the timed variant walks the 4-wide nodes. Results are the same:
// any hit is any hit, and the closest hit is the smallest (t, triangle id).
template <bool ANY, bool COUNT>
__device__ bool traverse(const DevScene& S, const RayP& r, float tmin, float tmax, uint32_t* lds, HitD& h,
                         Cnt& c) {
    if (!COUNT) {
        if (S.node_count == 0) return false;
        uint32_t spill[MAX_STACK4 - STACK_LDS];
        int sp = 0;
        uint32_t cur = 0;
        float tbest = tmax;
        uint32_t best = 0xFFFFFFFFu;
        float bu = 0.0f, bv = 0.0f;
        for (;;) {
            if (cur < LEAF_BIT) {
                cur = wide_step<MAX_STACK4>(S.nodes4, cur, r, tmin, tbest, lds, spill, sp, c);
                continue;
            }
            if (cur == DONE) break;
            uint32_t first = cur & 0x07FFFFFFu;
            uint32_t cnt = ((cur >> 27) & 0xFu) + 1u;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const gf4 tp = gptr(S.tris + k);
                float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                float t, b1, b2;
                if (intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tbest, t, b1, b2)) {
                    uint32_t id = __float_as_uint(v0.w);
                    if (ANY) {
                        h.t = t;
                        h.u = b1;
                        h.v = b2;
                        h.tri = id;
                        return true;
                    }
                    if (t < tbest || id < best) {
                        tbest = t;
                        best = id;
                        bu = b1;
                        bv = b2;
                    }
                }
            }
            if (sp == 0) break;
            --sp;
            cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
        }
        if (best == 0xFFFFFFFFu) return false;
        h.t = tbest;
        h.u = bu;
        h.v = bv;
        h.tri = best;
        return true;
    }
    if (S.node_count == 0) return false;
    uint32_t spill[STACK_SPILL];
    int sp = 0;
    uint32_t cur = 0;
    float tbest = tmax;
    uint32_t best = 0xFFFFFFFFu;
    float bu = 0.0f, bv = 0.0f;
    for (;;) {
        if (!(cur & LEAF_BIT)) {
            const float4* np = reinterpret_cast<const float4*>(S.nodes + cur);
            float4 a = np[0], b = np[1], cc = np[2], d = np[3];
            if (COUNT) ++c.nodes;
            float tn0, tn1;
            bool h0 = box_hit(a.x, a.y, a.z, b.x, b.y, b.z, r, tmin, tbest, tn0);
            bool h1 = box_hit(cc.x, cc.y, cc.z, d.x, d.y, d.z, r, tmin, tbest, tn1);
            uint32_t c0 = __float_as_uint(a.w), c1 = __float_as_uint(b.w);
            if (h0 && h1) {
                uint32_t nearc = c0, farc = c1;
                if (tn1 < tn0) {
                    nearc = c1;
                    farc = c0;
                }
                if (sp < (int)STACK_LDS) {
                    lds[sp * BLOCK] = farc;
                    ++sp;
                } else if (sp < (int)MAX_DEPTH) {
                    spill[sp - STACK_LDS] = farc;
                    ++sp;
                } else {
                    c.overflow = 1;
                }
                cur = nearc;
                continue;
            }
            if (h0) {
                cur = c0;
                continue;
            }
            if (h1) {
                cur = c1;
                continue;
            }
        } else {
            uint32_t first = cur & 0x07FFFFFFu;
            uint32_t cnt = ((cur >> 27) & 0xFu) + 1u;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const gf4 tp = gptr(S.tris + k);
                float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                if (COUNT) ++c.tris;
                float t, b1, b2;
                if (intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tbest, t, b1, b2)) {
                    uint32_t id = __float_as_uint(v0.w);
                    if (ANY) {
                        h.t = t;
                        h.u = b1;
                        h.v = b2;
                        h.tri = id;
                        return true;
                    }
                    if (t < tbest || id < best) {
                        tbest = t;
                        best = id;
                        bu = b1;
                        bv = b2;
                    }
                }
            }
        }
        if (sp == 0) break;
        --sp;
        cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
    }
    if (best == 0xFFFFFFFFu) return false;
    h.t = tbest;
    h.u = bu;
    h.v = bv;
    h.tri = best;
    return true;
}

template <bool COUNT>
__device__ __forceinline__ bool occluded(const DevScene& S, v3 o, v3 d, float tmin, float tmax, uint32_t* lds,
                                         Cnt& c) {
    RayP r = prep_ray(o, d);
    HitD h;
    return traverse<true, COUNT>(S, r, tmin, tmax, lds, h, c);
}

// ---- textures ---------------------------------------------------------------
__device__ __forceinline__ int wrap_index(float f, uint32_t n) {
    int i = (f == f && fabsf(f) < 1.0e9f) ? (int)f : 0;
    i %= (int)n;
    if (i < 0) i += (int)n;
    return i;
}
__device__ __forceinline__ v3 tex_nearest(const float4* t, uint32_t w, uint32_t h, float u, float v) {
    if (w == 0 || h == 0) return mk(0.0f, 0.0f, 0.0f);
    int x = wrap_index(floorf(u * (float)w), w);
    int y = wrap_index(floorf(v * (float)h), h);
    return ld3(t[(size_t)y * w + (size_t)x]);
}
__device__ __forceinline__ v3 env_lookup(const float4* t, uint32_t w, uint32_t h, v3 d, float pi) {
    float u = 0.5f + atan2_hr(d.x, d.z) / (2.0f * pi);
    float v = 0.5f - asin_hr(d.y) / pi;
    return tex_nearest(t, w, h, u, v);
}

// ---- helpers.glsl -----------------------------------------------------------
__device__ __forceinline__ float DistributionGGX(v3 N, v3 H, float roughness) {
    float a = roughness * roughness;
    float a2 = a * a;
    float NdotH = fmaxf(dot(N, H), 0.0f);
    float NdotH2 = NdotH * NdotH;
    float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
    denom = PI_H * denom * denom;
    return a2 / denom;
}
__device__ __forceinline__ float GeometrySchlickGGX(float NdotV, float roughness) {
    float r = roughness + 1.0f;
    float k = (r * r) / 8.0f;
    float denom = NdotV * (1.0f - k) + k;
    return NdotV / denom;
}
__device__ __forceinline__ float GeometrySmith(v3 N, v3 V, v3 L, float roughness) {
    float NdotV = fmaxf(dot(N, V), 0.0f);
    float NdotL = fmaxf(dot(N, L), 0.0f);
    return GeometrySchlickGGX(NdotV, roughness) * GeometrySchlickGGX(NdotL, roughness);
}
__device__ __forceinline__ v3 FresnelSchlick(float cosTheta, v3 F0) {
    float p = powf(1.0f - cosTheta, 5.0f);
    return mk(F0.x + (1.0f - F0.x) * p, F0.y + (1.0f - F0.y) * p, F0.z + (1.0f - F0.z) * p);
}
__device__ __forceinline__ v3 sampleDisk(const hr_light& light, v3 position, v3 L, float r1, float r2) {
    float radius = light.radius * sqrtf(r1);
    float sn, cs;
    sincos2pi_hr(r2, &sn, &cs);
    float px = radius * cs, py = radius * sn;
    v3 tangent = normalize(cross(L, mk(0.0f, 1.0f, 0.0f)));
    if (L.x == 0.0f && L.y == 1.0f && L.z == 0.0f) tangent = mk(0.0f, 0.0f, 1.0f);
    v3 bitangent = normalize(cross(tangent, L));
    v3 target = add(add(add(position, L), scl(tangent, px)), scl(bitangent, py));
    return normalize(sub(target, position));
}
__device__ __forceinline__ v3 sampleRect(v3 centre, const hr_light_shape& sh, v3 position, uint32_t s, uint32_t S,
                                         float r1, float r2) {
    uint32_t n = 1;
    while ((n + 1) * (n + 1) <= S) ++n;
    uint32_t m = (S + n - 1) / n;
    float fu = ((float)(s % n) + r1) / (float)n - 0.5f;
    float fv = ((float)(s / n) + r2) / (float)m - 0.5f;
    v3 eu = ld3p(sh.edge_u), ev = ld3p(sh.edge_v);
    v3 target = add(add(centre, scl(eu, fu)), scl(ev, fv));
    return normalize(sub(target, position));
}
__device__ __forceinline__ void createCoordSystem(v3 N, v3& T, v3& B) {  // random.glsl:46-53
    if (fabsf(N.x) > fabsf(N.y))
        T = divs(mk(N.z, 0.0f, -N.x), sqrtf(N.x * N.x + N.z * N.z));
    else
        T = divs(mk(0.0f, -N.z, N.y), sqrtf(N.y * N.y + N.z * N.z));
    B = cross(N, T);
}
__device__ __forceinline__ v3 sampleGGX(v3 N, float roughness, float r1, float r2) {
    float a = roughness * roughness;
    float cos2 = (1.0f - r1) / (1.0f + (a * a - 1.0f) * r1);
    float ct = sqrtf(cos2);
    float st = sqrtf(fmaxf(1.0f - cos2, 0.0f));
    float sn, cs;
    sincos2pi_hr(r2, &sn, &cs);
    v3 T, B;
    createCoordSystem(N, T, B);
    return add(add(scl(T, st * cs), scl(B, st * sn)), scl(N, ct));
}
__device__ __forceinline__ v3 light_sample_dir(const FrameArgs& F, uint32_t i, v3 P, v3 L, Pcg& g, uint32_t s,
                                               uint32_t S) {
    const hr_light& lt = F.lights[i];
    float r1 = u01(g);
    float r2 = u01(g);
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    if (!isDir && F.has_shapes && F.shapes[i].type == HR_LIGHT_RECT)
        return sampleRect(ld3p(lt.pos_maxdist), F.shapes[i], P, s, S, r1, r2);
    return sampleDisk(lt, P, L, r1, r2);
}
// ---- adaptive shadow sampling --------------------------------------------------
// Samples of a (pixel, light) come from one nested stratified sequence of S (a
// power of two <= 64) cells: sample s lies in column cx of nx = 2^ceil(lg S / 2)
// and row cy of ny = S / nx, where the bits of s at even positions, reversed,
// give cx and those at odd positions, reversed, give cy. Every prefix of 4^j
// samples then holds one sample per cell of the 2^j x 2^j coarsening (of the
// 2^(j+1) x 2^j one for 2 * 4^j), so each stage extends a stratified set.
__device__ __forceinline__ void nested_cell(uint32_t s, uint32_t S, uint32_t& cx, uint32_t& nx, uint32_t& cy,
                                            uint32_t& ny) {
    const uint32_t lg = 31u - (uint32_t)__clz(S);
    const uint32_t a = (lg + 1u) >> 1, b = lg >> 1;
    cx = 0u;
    cy = 0u;
    for (uint32_t k = 0; k < a; ++k) cx |= ((s >> (2u * k)) & 1u) << (a - 1u - k);
    for (uint32_t k = 0; k < b; ++k) cy |= ((s >> (2u * k + 1u)) & 1u) << (b - 1u - k);
    nx = 1u << a;
    ny = 1u << b;
}
// Sample s of the sequence: the cell's point jittered by (r1, r2), on the
// rectangle (as sampleRect) or on the (radius^2, angle) square of sampleDisk.
__device__ __forceinline__ v3 light_sample_dir_nested(const FrameArgs& F, uint32_t i, v3 P, v3 L, Pcg& g,
                                                      uint32_t s, uint32_t S) {
    const hr_light& lt = F.lights[i];
    const float r1 = u01(g);
    const float r2 = u01(g);
    uint32_t cx, nx, cy, ny;
    nested_cell(s, S, cx, nx, cy, ny);
    const float a = ((float)cx + r1) / (float)nx;
    const float b = ((float)cy + r2) / (float)ny;
    const bool isDir = lt.pos_maxdist[3] < 0.0f;
    if (!isDir && F.has_shapes && F.shapes[i].type == HR_LIGHT_RECT) {
        const v3 eu = ld3p(F.shapes[i].edge_u), ev = ld3p(F.shapes[i].edge_v);
        const v3 target = add(add(ld3p(lt.pos_maxdist), scl(eu, a - 0.5f)), scl(ev, b - 0.5f));
        return normalize(sub(target, P));
    }
    return sampleDisk(lt, P, L, a, b);
}
// Stages: 4 probes (all S when S < 4), then 16, then doubling up to S.
__device__ __forceinline__ uint32_t adaptive_first(uint32_t S) { return S < 4u ? S : 4u; }
__device__ __forceinline__ uint32_t adaptive_next(uint32_t N, uint32_t S) {
    const uint32_t n = N == 4u ? 16u : 2u * N;
    return n < S ? n : S;
}
// After N samples with k unoccluded: stop when N reaches S or when the
// binomial variance estimate of the visible fraction, k (N - k) / N^3, is at
// most that of S samples at one half, 1 / (4 S): exact in integers.
__device__ __forceinline__ bool adaptive_stop(uint32_t k, uint32_t N, uint32_t S) {
    return N >= S || 4ull * S * k * (N - k) <= (uint64_t)N * N * N;
}
// A light whose adaptive sample count is 1: a disk (or direction) of radius 0.
__device__ __forceinline__ bool hard_light(const FrameArgs& F, uint32_t i) {
    const hr_light& lt = F.lights[i];
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    bool rect = !isDir && F.has_shapes && F.shapes[i].type == HR_LIGHT_RECT;
    return !rect && lt.radius == 0.0f;
}

// ---- the reference's 8-bit images --------------------------------------------
// imageStore to a UNORM8 channel: rint(clamp(v, 0, 1) * 255) (NaN stores 0)
__device__ __forceinline__ uint32_t unorm8(float v) { return (uint32_t)rintf(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f); }
// shadow factor idx of an FP32 or R8_UNORM shadow buffer
__device__ __forceinline__ void put_factor(float* shadow, size_t idx, float v, uint32_t fmt) {
    if (fmt & UNORM8_SHADOW) reinterpret_cast<uint8_t*>(shadow)[idx] = (uint8_t)unorm8(v);
    else shadow[idx] = v;
}
__device__ __forceinline__ float get_factor(const float* shadow, size_t idx, uint32_t fmt) {
    if (fmt & UNORM8_SHADOW) return (float)reinterpret_cast<const uint8_t*>(shadow)[idx] / 255.0f;
    return shadow[idx];
}
// lit colour of pixel p (alpha 1) to an FP32 RGBA or a B8G8R8A8_UNORM image
__device__ __forceinline__ void put_lit(float4* lit, size_t p, v3 c, uint32_t fmt) {
    if (fmt & UNORM8_LIT)
        reinterpret_cast<uint32_t*>(lit)[p] = unorm8(c.z) | (unorm8(c.y) << 8) | (unorm8(c.x) << 16) | (255u << 24);
    else lit[p] = make_float4(c.x, c.y, c.z, 1.0f);
}

__device__ __forceinline__ int mat_index(const DevScene& S, float w) {
    int i = (int)w;
    return (i < 0 || i >= (int)S.nmat) ? 0 : i;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    return v;
}
__device__ __forceinline__ void flush_counters(const Cnt& c, unsigned long long* counters, bool count) {
    uint32_t v[8] = {c.primary, c.shadow, c.refl, c.hitshadow, c.occluded, c.nodes, c.tris, c.overflow};
    const int slot[8] = {C_PRIMARY, C_SHADOW, C_REFL, C_HITSHADOW, C_OCCLUDED, C_NODES, C_TRIS, C_OVERFLOW};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!count && (k == 5 || k == 6)) continue;
        uint32_t s = wave_sum(v[k]);
        if ((threadIdx.x & 63u) == 0 && s) atomicAdd(&counters[slot[k] * C_STRIDE], (unsigned long long)s);
    }
}

// Wave-aggregated take of one index of [0, n) per calling lane. The range is
// cut into SHADOW_QUEUES bands with their counters BAND_STRIDE words apart in
// `next` (one atomic per band for all lanes of the wave that reach this call
// together; counters on separate lines do not serialise one another). A wave
// starts at its home band and moves on when a band runs dry; `emptied` (per
// lane, reduced over the wave by ballot so all lanes agree) counts the bands
// found empty. Returns n once every band is empty.
template <uint32_t NB = SHADOW_QUEUES>
__device__ __forceinline__ uint32_t band_take(uint32_t* next, uint32_t n, uint32_t home, uint32_t& emptied) {
    const uint32_t lane = threadIdx.x & 63u;
    for (;;) {
        uint32_t k = 0;
        for (uint32_t b = 1; b <= NB; ++b)
            if (__ballot(emptied >= b)) k = b;
        emptied = k;
        if (k >= NB) return n;
        const uint32_t band = (home + k) & (NB - 1u);
        const uint32_t band_end = (uint32_t)(((uint64_t)n * (band + 1u)) / NB);
        const unsigned long long need = __ballot(1);
        const uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
        const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(next + band * BAND_STRIDE, (uint32_t)__popcll(need));
        base = (uint32_t)__shfl((int)base, (int)leader, 64);
        const uint32_t idx = (uint32_t)(((uint64_t)n * band) / NB) + base + rank;
        if (idx < band_end) return idx;
        emptied = k + 1u;
    }
}

// Persistent consumer of a ray queue of length n. Each lane traces ray after
// ray; when its ray finishes it takes the next index (band_take over the
// SHADOW_QUEUES band counters at `next`), so lanes do not wait for the
// slowest ray of the wave. Per ray the walk is the shared
// near-first stack traversal (same order and counts as traverse()).
//   get(i, o, d, tmin, tmax) -> false for an empty queue entry (skipped)
//   put(i, hit, HitD)         result of entry i
template <bool ANY, bool COUNT, class GetRay, class PutHit>
__device__ void consume_rays(const DevScene& S, uint32_t n, uint32_t* next, uint32_t* lds, Cnt& c, GetRay get,
                             PutHit put) {
    uint32_t cur = DONE, idx = 0;
    bool have = false;
    RayP r;
    float tmin = 0.0f, tbest = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t best = 0xFFFFFFFFu;
    int sp = 0;
    // the timed variant walks the 4-wide nodes, the counting one the binary nodes
    constexpr uint32_t STACK_CAP = COUNT ? MAX_DEPTH : MAX_STACK4;
    uint32_t spill[STACK_CAP - STACK_LDS];
    // speculative traversal (postponed leaves) except when counting: the
    // closest hit (smallest t, then smallest triangle id) and any-hit results
    // do not depend on the order leaves are tested in
    constexpr bool SPEC = !COUNT;
    uint32_t leaf = 0u;  // postponed leaf code (0: none)
    const uint32_t home = blockIdx.x & (SHADOW_QUEUES - 1u);
    uint32_t emptied = 0;
    for (;;) {
        if (cur == DONE && leaf == 0u) {
            if (have) {
                HitD h;
                h.t = tbest;
                h.u = bu;
                h.v = bv;
                h.tri = best;
                put(idx, best != 0xFFFFFFFFu, h);
                have = false;
            }
            bool got = false;
            for (;;) {
                idx = band_take(next, n, home, emptied);
                if (idx >= n) break;
                v3 o, d;
                float tmax;
                if (get(idx, o, d, tmin, tmax)) {
                    r = prep_ray(o, d);
                    tbest = tmax;
                    got = true;
                    break;
                }
            }
            if (!got) break;
            have = true;
            best = 0xFFFFFFFFu;
            sp = 0;
            cur = S.node_count ? 0u : DONE;
            continue;
        }
        while (!COUNT && cur < LEAF_BIT) {
            cur = wide_step<STACK_CAP>(S.nodes4, cur, r, tmin, tbest, lds, spill, sp, c);
            // any hit: postpone the first leaf found and keep walking until
            // every lane still in this loop holds one. Closest hit: leaves in
            // walk order, so each one can shrink tbest before the next boxes
            if (ANY && leaf == 0u && cur >= LEAF_BIT && cur != DONE) {
                leaf = cur;
                if (sp == 0) {
                    cur = DONE;
                } else {
                    --sp;
                    cur = sp < (int)STACK_LDS ? lds[sp * BLOCK] : spill[sp - STACK_LDS];
                }
            }
            if (__ballot(leaf == 0u) == 0ull || __popcll(__ballot(cur < LEAF_BIT)) < REFILL_MIN_LANES) break;
        }
        while (COUNT && cur < LEAF_BIT) {
            const float4* np = reinterpret_cast<const float4*>(S.nodes + cur);
            float4 a = np[0], b = np[1], cc = np[2], d = np[3];
            if (COUNT) ++c.nodes;
            float tn0, tn1;
            bool h0 = box_hit(a.x, a.y, a.z, b.x, b.y, b.z, r, tmin, tbest, tn0);
            bool h1 = box_hit(cc.x, cc.y, cc.z, d.x, d.y, d.z, r, tmin, tbest, tn1);
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
            bool stop = false;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const gf4 tp = gptr(S.tris + k);
                float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                if (COUNT) ++c.tris;
                float t, b1, b2;
                if (intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tbest, t, b1, b2)) {
                    uint32_t id = __float_as_uint(v0.w);
                    if (ANY) {
                        tbest = t;
                        best = id;
                        bu = b1;
                        bv = b2;
                        stop = true;
                        break;
                    }
                    if (t < tbest || id < best) {
                        tbest = t;
                        best = id;
                        bu = b1;
                        bv = b2;
                    }
                }
            }
            if (stop) {
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
}

// ---- material textures ----------------------------------------------------------
// Texel of material texture `id` (a whole number >= 0) at (u, v): NEAREST /
// REPEAT, R8G8B8A8_UNORM decoded as c / 255.
__device__ __forceinline__ v3 tex_rgba8(const DevScene& S, float id, float u, float v) {
    const TexInfo t = S.tex[(int)id];
    int x = wrap_index(floorf(u * (float)t.w), t.w);
    int y = wrap_index(floorf(v * (float)t.h), t.h);
    uint32_t c = __ldg(S.texels + t.offset + (size_t)y * t.w + (size_t)x);
    return mk((float)(c & 0xFFu) / 255.0f, (float)((c >> 8) & 0xFFu) / 255.0f, (float)((c >> 16) & 0xFFu) / 255.0f);
}
__device__ __forceinline__ const float* vtx(const DevScene& S, uint32_t vi) {
    return reinterpret_cast<const float*>(S.verts + (size_t)vi * 44u);
}
// barycentric interpolation (b0 v0 + b1 v1) + b2 v2 of the uv (offset 36) of
// three vertices
__device__ __forceinline__ void interp_uv(const DevScene& S, const uint32_t* vi, float b0, float b1, float b2,
                                          float& u, float& v) {
    const float* a = vtx(S, vi[0]);
    const float* b = vtx(S, vi[1]);
    const float* c = vtx(S, vi[2]);
    u = (a[9] * b0 + b[9] * b1) + c[9] * b2;
    v = (a[10] * b0 + b[10] * b1) + c[10] * b2;
}
__device__ __forceinline__ bool textured(const hr_material& m) {
    return m.textures[0] > -1.0f || m.textures[1] > -1.0f || m.textures[2] > -1.0f || m.textures[3] > -1.0f;
}

// Geometry and surface of a closest hit (hybridHit.rchit:26-68,
// closesthit.rchit:26-70): albedo is the diffuse texel or the material colour
// (before linearisation), emissive the emissive texel or 0, roughness and
// metallic the g and b of the metallic-roughness texel or the material's.
struct HitGeo {
    v3 N, V, worldPos;
    float NdotV;
    const hr_material* mat;
    v3 albedo, emissive;
    float roughness, metallic;
};
__device__ __forceinline__ HitGeo hit_geometry(const DevScene& S, v3 ro, v3 rd, const HitD& hit) {
    HitGeo g;
    float b1 = hit.u, b2 = hit.v;
    float b0 = 1.0f - b1 - b2;
    const PrimInfo& p = S.prims[S.tri_prim[hit.tri]];
    uint32_t local = hit.tri - p.tri_offset;
    uint32_t ib = p.index_base + 3 * local;
    uint32_t vi[3] = {p.vertex_base + S.indices[ib + 0], p.vertex_base + S.indices[ib + 1],
                      p.vertex_base + S.indices[ib + 2]};
    v3 n0 = ld3(S.vnormal[vi[0]]);
    v3 n1 = ld3(S.vnormal[vi[1]]);
    v3 n2 = ld3(S.vnormal[vi[2]]);
    v3 nrm = add(add(scl(n0, b0), scl(n1, b1)), scl(n2, b2));
    const float* M = p.nm;
    g.N = normalize(mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                       (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z));
    g.V = normalize(neg(rd));
    g.NdotV = clamp01(dot(g.N, g.V));
    g.worldPos = add(ro, scl(rd, hit.t));
    const hr_material& m = S.mats[p.material];
    g.mat = &m;
    g.albedo = ld3p(m.diffuse_ior);
    g.emissive = mk(0.0f, 0.0f, 0.0f);
    g.roughness = m.shading_metal_rough_idx[2];
    g.metallic = m.shading_metal_rough_idx[1];
    if (textured(m)) {
        float u, v;
        interp_uv(S, vi, b0, b1, b2, u, v);
        if (m.textures[0] > -1.0f) g.albedo = tex_rgba8(S, m.textures[0], u, v);
        if (m.textures[2] > -1.0f) g.emissive = tex_rgba8(S, m.textures[2], u, v);
        if (m.textures[3] > -1.0f) {
            v3 rm = tex_rgba8(S, m.textures[3], u, v);
            g.roughness = rm.y;
            g.metallic = rm.z;
        }
    }
    return g;
}

// ---- persistent any-hit tracing over numbered work items ----------------------
// Items 0..nitems-1 are handed out from SHADOW_QUEUES bands (one atomic counter
// each in next_item; workgroup b starts at band b mod SHADOW_QUEUES and moves
// to the next band when its own runs dry). Lanes that take work together take
// consecutive items. A lane whose ray ended takes its next item at once, while
// the rest of the wave keeps traversing. src.setup(item, bit, ...) decides an
// item and the bit of `bits` it owns: ITEM_SKIP (no bit), ITEM_LIT (bit set
// without a ray) or ITEM_TRACE (ray, tmax and scene set); an unoccluded traced
// item sets its bit (atomicOr, so the result does not depend on the order).
enum : int { ITEM_SKIP = 0, ITEM_LIT = 1, ITEM_TRACE = 2 };

// Bit `item` of `bits` for every calling lane with `set`, one global atomicOr
// per distinct word: lanes that finish together mostly hold consecutive items.
// The word's bits are gathered with LDS ORs in the wave's `slot` (a shuffle
// reduction would read the stale registers of inactive lanes).
__device__ __forceinline__ void or_item_bits(uint32_t* bits, uint32_t item, bool set, uint32_t* slot) {
    unsigned long long pend = __ballot(set);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = item >> 5;
    while (pend) {
        const uint32_t leader = (uint32_t)__ffsll((long long)pend) - 1u;
        const uint32_t word = (uint32_t)__shfl((int)w, (int)leader, 64);
        const bool mine = set && w == word;
        if (lane == leader) *slot = 0u;
        if (mine) atomicOr(slot, 1u << (item & 31u));
        if (lane == leader) atomicOr(bits + word, *slot);
        pend &= ~__ballot(mine);
    }
}

template <bool OFF32, class Src>
__device__ __forceinline__ void item_loop(Src& src, uint32_t nitems, uint32_t* next_item, uint32_t* bits,
                                          uint32_t* lds, Cnt& c) {
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const uint32_t home = blockIdx.x & (SHADOW_QUEUES - 1u);
    uint32_t emptied = 0;
    const TNode4* nodes4 = src.nodes4;
    const STri* tris = src.tris;
    uint32_t node_count = src.node_count;
    uint32_t item = 0, bit = 0;
    bool tracing = false, hit = false;
    uint32_t cur = DONE, leaf = 0u;
    RayP r;
    NearPlanes pl{0u, 32u, 64u};
    float tmax = 0.0f;
    const float tmin = 0.001f;
    int sp = 0;
    uint32_t spill[MAX_STACK4 - STACK_LDS];
    __shared__ uint32_t or_slot[BLOCK / 64];
    uint32_t* slot = or_slot + (threadIdx.x >> 6);
    for (;;) {
        if (cur == DONE && leaf == 0u) {
            // the ray of `item` ended
            if (tracing && hit) ++c.occluded;
            or_item_bits(bits, bit, tracing && !hit, slot);
            tracing = false;
            bool exhausted = false;
            for (;;) {
                uint32_t k = 0;
                for (uint32_t b = 1; b <= SHADOW_QUEUES; ++b)
                    if (__ballot(emptied >= b)) k = b;
                emptied = k;
                if (k >= SHADOW_QUEUES) {
                    exhausted = true;
                    break;
                }
                const uint32_t band = (home + k) & (SHADOW_QUEUES - 1u);
                const uint32_t band_end = (uint32_t)(((uint64_t)nitems * (band + 1u)) / SHADOW_QUEUES);
                unsigned long long need = __ballot(1);
                uint32_t rank = (uint32_t)__popcll(need & lt_mask);
                uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(next_item + band * BAND_STRIDE, (uint32_t)__popcll(need));
                base = (uint32_t)__shfl((int)base, (int)leader, 64);
                item = (uint32_t)(((uint64_t)nitems * band) / SHADOW_QUEUES) + base + rank;
                if (item >= band_end) {
                    emptied = k + 1u;
                    continue;
                }
                const int kind = src.setup(item, bit, r, tmax, nodes4, tris, node_count, c);
                or_item_bits(bits, bit, kind == ITEM_LIT, slot);
                if (kind != ITEM_TRACE) continue;
                if (OFF32) pl = near_planes(r);
                tracing = true;
                hit = false;
                sp = 0;
                cur = node_count ? 0u : DONE;
                break;
            }
            if (exhausted) break;
            continue;
        }
        while (cur < LEAF_BIT) {
            // per-lane 64-bit plane addresses cost more registers than the
            // min/max they save: the sign-selected step only with 32-bit offsets
            if (OFF32) cur = wide_step_any<true>(nodes4, cur, r, pl, tmin, tmax, lds, spill, sp, c);
            else cur = wide_step<MAX_STACK4>(nodes4, cur, r, tmin, tmax, lds, spill, sp, c);
            if (leaf == 0u && cur >= LEAF_BIT && cur != DONE) {
                leaf = cur;
                if (sp == 0) {
                    cur = DONE;
                } else {
                    --sp;
                    cur = stack_at(lds, spill, sp);
                }
            }
            if (__ballot(leaf == 0u) == 0ull || __popcll(__ballot(cur < LEAF_BIT)) < REFILL_MIN_ITEMS) break;
        }
        uint32_t lf = DONE;
        bool in_order = false;
        if (leaf != 0u) {
            lf = leaf;
            leaf = 0u;
        } else if (cur >= LEAF_BIT && cur != DONE) {
            lf = cur;
            in_order = true;
        }
        if (lf != DONE) {
            const uint32_t first = lf & 0x07FFFFFFu;
            const uint32_t end = first + ((lf >> 27) & 0xFu) + 1u;
            const uint32_t ul = __builtin_amdgcn_readfirstlane(lf);
            if (OFF32 && __ballot(lf != ul) == 0ull) {
                // every lane at one leaf: its triangles through the scalar cache
                const uint32_t ufirst = ul & 0x07FFFFFFu;
                const uint32_t uend = ufirst + ((ul >> 27) & 0xFu) + 1u;
                for (uint32_t k = ufirst; k < uend; ++k) {
                    const cf4 tp = (cf4)(tris + k);
                    const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                    float t, b1, b2;
                    if (intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tmax, t, b1, b2)) hit = true;
                    if (__ballot(!hit) == 0ull) break;
                }
            } else {
                // Two triangles per step, both loaded before either test: a leaf
                // costs half as many memory round trips (the second of an odd
                // leaf's last pair repeats the first).
                for (uint32_t k = first; k < end; k += 2u) {
                    const gf4 tp = gptr(tris + k), tq = gptr(tris + (k + 1u < end ? k + 1u : k));
                    const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                    const float4 w0 = tq[0], w1 = tq[1], w2 = tq[2];
                    float t, b1, b2;
                    const bool h0 = intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tmax, t, b1, b2);
                    const bool h1 = intersect_tri(r, ld3(w0), ld3(w1), ld3(w2), tmin, tmax, t, b1, b2);
                    if (h0 || h1) {
                        hit = true;
                        break;
                    }
                }
            }
            if (hit) {
                cur = DONE;
                leaf = 0u;
            } else if (in_order) {
                if (sp == 0) {
                    cur = DONE;
                } else {
                    --sp;
                    cur = stack_at(lds, spill, sp);
                }
            }
        }
    }
}

// the number of set bits of a group's item window (stride 2^lg2s, lg2s <= 6):
// whole words for 32 and 64, a masked part of one word below (a shift by 32
// would not give a full mask)
__device__ __forceinline__ uint32_t window_bits(const uint32_t* __restrict__ bits, uint64_t b0, uint32_t lg2s) {
    if (lg2s == 6u) return (uint32_t)__popc(bits[b0 >> 5]) + (uint32_t)__popc(bits[(b0 >> 5) + 1u]);
    if (lg2s == 5u) return (uint32_t)__popc(bits[b0 >> 5]);
    const uint32_t w = bits[b0 >> 5] >> (uint32_t)(b0 & 31u);
    return (uint32_t)__popc(w & ((1u << (1u << lg2s)) - 1u));
}

// ---- ordered lists of the groups that need rays -------------------------------
// A pre-pass writes, per chunk of 64 groups, the ballot of its groups that need
// rays (masks) and its count (counts); scan_and_scatter turns the counts into
// offsets and lists the groups in group order (so a trace's work bands keep
// neighbouring groups together), with the list length at *count.
namespace {

// Exclusive prefix sums of the chunk counts in place, in three launches:
// k_tile_totals sums each tile of SCAN_TILE counts, k_scan_counts scans the
// tile totals (one workgroup; *total gets the sum of all), k_tile_offsets
// scans each tile through LDS and adds its tile's offset.
constexpr uint32_t SCAN_TILE = 1024, SCAN_TILE_THREADS = 256;
__global__ __launch_bounds__(SCAN_TILE_THREADS) void k_tile_totals(const uint32_t* counts, uint32_t n,
                                                                    uint32_t* totals) {
    __shared__ uint32_t red[SCAN_TILE_THREADS];
    const uint32_t base = blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_TILE / SCAN_TILE_THREADS; ++k) {
        const uint32_t i = base + k * SCAN_TILE_THREADS + threadIdx.x;
        if (i < n) s += counts[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t d = SCAN_TILE_THREADS / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(SCAN_TILE_THREADS) void k_tile_offsets(uint32_t* counts, uint32_t n,
                                                                     const uint32_t* tile_offsets) {
    constexpr uint32_t PER = SCAN_TILE / SCAN_TILE_THREADS;
    __shared__ uint32_t v[SCAN_TILE];
    __shared__ uint32_t part[SCAN_TILE_THREADS];
    const uint32_t base = blockIdx.x * SCAN_TILE;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t j = k * SCAN_TILE_THREADS + threadIdx.x, i = base + j;
        v[j] = i < n ? counts[i] : 0u;
    }
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) s += v[threadIdx.x * PER + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t d = 1; d < SCAN_TILE_THREADS; d <<= 1) {
        const uint32_t x = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = tile_offsets[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t j = threadIdx.x * PER + k;
        const uint32_t c = v[j];
        v[j] = run;
        run += c;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t j = k * SCAN_TILE_THREADS + threadIdx.x, i = base + j;
        if (i < n) counts[i] = v[j];
    }
}

// exclusive prefix sums of counts[0, n) in place, the total to *total: one
// workgroup, each thread a contiguous run (for the tile totals)
constexpr uint32_t SCAN_BLOCK = 1024;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_counts(uint32_t* counts, uint32_t n, uint32_t* total) {
    __shared__ uint32_t part[SCAN_BLOCK];
    const uint32_t per = (n + SCAN_BLOCK - 1u) / SCAN_BLOCK;
    const uint32_t b = threadIdx.x * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += counts[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t d = 1; d < SCAN_BLOCK; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;  // exclusive
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (threadIdx.x == SCAN_BLOCK - 1u) *total = part[SCAN_BLOCK - 1u];
}

__global__ __launch_bounds__(256) void k_group_scatter(uint32_t ngroups, const unsigned long long* masks,
                                                       const uint32_t* offsets, uint32_t* list) {
    const uint32_t G = blockIdx.x * blockDim.x + threadIdx.x;
    if (G >= ngroups) return;
    const unsigned long long m = masks[G >> 6];
    const uint32_t lane = G & 63u;
    if ((m >> lane) & 1ull) list[offsets[G >> 6] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = G;
}

// scratch after a list and its count: up to 7 alignment bytes, per chunk an
// 8-byte mask and a count, the chunk counts' tile totals
inline size_t group_list_tail_bytes(uint32_t ngroups) {
    const size_t nchunks = (ngroups + 63u) / 64u;
    return 7u + nchunks * 12u + ((nchunks + SCAN_TILE - 1u) / SCAN_TILE) * 4u;
}
struct GroupListTail {
    unsigned long long* masks;
    uint32_t *counts, *tile_totals;
    uint32_t nchunks;
};
inline GroupListTail group_list_tail(void* tail, uint32_t ngroups) {
    GroupListTail t;
    t.nchunks = (ngroups + 63u) / 64u;
    t.masks = reinterpret_cast<unsigned long long*>(reinterpret_cast<uintptr_t>(tail) + 7u & ~(uintptr_t)7u);
    t.counts = reinterpret_cast<uint32_t*>(t.masks + t.nchunks);
    t.tile_totals = t.counts + t.nchunks;
    return t;
}
inline void scan_and_scatter(uint32_t ngroups, const GroupListTail& t, uint32_t* count, uint32_t* list,
                             hipStream_t st) {
    const uint32_t ntiles = (t.nchunks + SCAN_TILE - 1u) / SCAN_TILE;
    k_tile_totals<<<ntiles, SCAN_TILE_THREADS, 0, st>>>(t.counts, t.nchunks, t.tile_totals);
    k_scan_counts<<<1, SCAN_BLOCK, 0, st>>>(t.tile_totals, ntiles, count);
    k_tile_offsets<<<ntiles, SCAN_TILE_THREADS, 0, st>>>(t.counts, t.nchunks, t.tile_totals);
    k_group_scatter<<<(ngroups + 255u) / 256u, 256, 0, st>>>(ngroups, t.masks, t.counts, list);
}

}  // namespace

}  // namespace tc
}  // namespace hr
