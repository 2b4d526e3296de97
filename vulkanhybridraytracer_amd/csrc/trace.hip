// trace.hip - traversal and shading kernels of the hybrid pass (gfx950).
//
//   k_gbuffer   primary visibility at pixel centres (raygen.rgen:29-40) into
//               the FP32 G-buffer (geometry_shader.frag:51-76, skybox.frag:16-21)
//   k_shadows   shadowRaygen.rgen:28-93 (+ shadow.rmiss, shadowHit.rchit)
//   k_shade     hybridRaygen.rgen:27-186 with the closest-hit shading of
//               hybridHit.rchit:24-183 and miss.rmiss:13-18 inlined
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
#include "device_math.h"
#include "internal.h"

namespace hr {
using namespace hrd;

namespace {

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

template <bool ANY, bool COUNT>
__device__ bool traverse(const DevScene& S, const RayP& r, float tmin, float tmax, uint32_t* lds, HitD& h,
                         Cnt& c) {
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
                const float4* tp = reinterpret_cast<const float4*>(S.tris + k);
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
__device__ __forceinline__ int mat_index(const DevScene& S, float w) {
    int i = (int)w;
    return (i < 0 || i >= (int)S.nmat) ? 0 : i;
}

// closest-hit shading, hybridHit.rchit:24-183
struct HitOut {
    v3 color, dir, origin;
    bool scatter;
};
template <bool COUNT>
__device__ HitOut shade_hit(const DevScene& S, const FrameArgs& F, v3 ro, v3 rd, const HitD& hit, uint32_t seed,
                            uint32_t sample, uint32_t depth, uint32_t* lds, Cnt& c) {
    HitOut out;
    float b1 = hit.u, b2 = hit.v;
    float b0 = 1.0f - b1 - b2;
    uint32_t pidx = S.tri_prim[hit.tri];
    const PrimInfo& p = S.prims[pidx];
    uint32_t local = hit.tri - p.tri_offset;
    uint32_t ib = p.index_base + 3 * local;
    v3 n0 = ld3(S.vnormal[p.vertex_base + S.indices[ib + 0]]);
    v3 n1 = ld3(S.vnormal[p.vertex_base + S.indices[ib + 1]]);
    v3 n2 = ld3(S.vnormal[p.vertex_base + S.indices[ib + 2]]);
    v3 nrm = add(add(scl(n0, b0), scl(n1, b1)), scl(n2, b2));
    const float* M = p.nm;
    v3 N = normalize(mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                        (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z));
    v3 V = normalize(neg(rd));
    float NdotV = clamp01(dot(N, V));
    v3 worldPos = add(ro, scl(rd, hit.t));

    v3 Lo = mk(0.0f, 0.0f, 0.0f);
    const hr_material& mat = S.mats[p.material];
    int mode = (int)mat.shading_metal_rough_idx[0];
    v3 albedo = pow3(ld3p(mat.diffuse_ior), 2.2f);
    float roughness = mat.shading_metal_rough_idx[2];
    float metallic = mat.shading_metal_rough_idx[1];
    v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
    v3 irradiance = env_lookup(S.irr, S.irr_w, S.irr_h, N, PI_H);

    v3 direction = mk(1.0f, 1.0f, 1.0f);
    bool scatter = false;
    v3 rdn = normalize(rd);
    v3 reflected = mk(0.0f, 0.0f, 0.0f);
    if (mode == 3) {
        v3 Nr = N;
        if (roughness > 0.0f) {
            Pcg g = ray_stream(seed, PURPOSE_HIT_REFLECT, sample * F.max_depth + depth);
            float r1 = u01(g), r2 = u01(g);
            Nr = sampleGGX(N, roughness, r1, r2);
        }
        reflected = reflect(rdn, Nr);
    }
    for (uint32_t i = 0; i < F.nl; ++i) {
        const hr_light& lt = F.lights[i];
        bool isDir = lt.pos_maxdist[3] < 0.0f;
        v3 L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), worldPos);
        float maxd = lt.pos_maxdist[3];
        float dist = length(L);
        L = normalize(L);
        float intensity = isDir ? 1.0f : (lt.color_intensity[3] / (dist * dist));
        v3 H = normalize(add(V, L));
        float NdotL = clamp01(dot(N, L));
        float NdotH = clamp01(dot(N, H));
        float shadowFactor = 1.0f;
        float attenuation;
        if (intensity == 0.0f) {
            attenuation = 0.0f;
        } else {
            attenuation = maxd - dist;
            attenuation /= maxd;
            attenuation = fmaxf(attenuation, 0.0f);
            attenuation = isDir ? 0.3f : attenuation * attenuation;
        }
        v3 lc = ld3p(lt.color_intensity);
        if (mode == 0) {
            if (NdotL > 0.0f) {
                bool shadowed = true;
                if (isDir || dist < maxd) {
                    Pcg g = ray_stream(seed, PURPOSE_HIT_SHADOW, (sample * F.max_depth + depth) * F.nl + i);
                    v3 dir = light_sample_dir(F, i, worldPos, L, g, 0, 1);
                    float tmax = isDir ? 1000.0f : dist + 1.0f;
                    ++c.hitshadow;
                    shadowed = occluded<COUNT>(S, add(worldPos, scl(dir, 1e-2f)), dir, 0.001f, tmax, lds, c);
                    if (shadowed) ++c.occluded;
                }
                if (!shadowed) shadowFactor += 1.0f;
            }
            v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
            v3 Fr = FresnelSchlick(NdotH, F0);
            float D = DistributionGGX(N, H, roughness);
            float G = GeometrySmith(N, V, L, roughness);
            v3 numerator = scl(Fr, D * G);
            float denominator = fmaxf(4.0f * clamp01(dot(N, V)) * NdotL, 0.000001f);
            v3 specular = divs(numerator, denominator);
            v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fr), 1.0f - metallic);
            v3 term = add(divs(mul(kD, albedo), PI_H), specular);
            Lo = add(Lo, scl(mul(term, radiance), NdotL));
            direction = mk(1.0f, 1.0f, 1.0f);
            scatter = false;
        } else if (mode == 3) {
            bool sc = dot(reflected, N) > 0.0f;
            v3 add_t = (NdotL > 0.0f && intensity > 0.0f)
                           ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                           : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, add_t);
            direction = reflected;
            scatter = sc;
        } else if (mode == 4) {
            float ior = mat.diffuse_ior[3];
            float ndv = dot(N, rdn);
            v3 refrNormal = ndv > 0.0f ? neg(N) : N;
            float refrEta = ndv > 0.0f ? 1.0f / ior : ior;
            v3 add_t = (intensity > 0.0f) ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                                          : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, add_t);
            float radicand = 1.0f + refrEta * refrEta * (ndv * ndv - 1.0f);
            direction = radicand < 0.0f ? reflect(rd, N) : refract(rdn, refrNormal, refrEta);
            scatter = true;
        }
    }
    v3 Fa = FresnelSchlick(NdotV, F0);
    v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fa), 1.0f - metallic);
    v3 ambient = mul(mul(kD, albedo), irradiance);
    out.color = add(Lo, ambient);
    out.dir = direction;
    out.scatter = scatter;
    out.origin = worldPos;
    return out;
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
        if ((threadIdx.x & 63u) == 0 && s) atomicAdd(&counters[slot[k]], (unsigned long long)s);
    }
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
__global__ __launch_bounds__(BLOCK) void k_shade(DevScene S, FrameArgs F, GBufView G, const float* shadow,
                                                 float4* lit, float4* refl, unsigned long long* counters) {
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
        const hr_material& mat = S.mats[mat_index(S, P4.w)];
        int mode = (int)mat.shading_metal_rough_idx[0];
        v3 albedo = pow3(ld3(G.alb[p]), 2.2f);
        v3 mtl = ld3(G.mat[p]);
        v3 emis = ld3(G.emi[p]);
        bool background = P4.w == 0.0f && N4.w == 0.0f;
        if (background) {
            lit[p] = make_float4(albedo.x, albedo.y, albedo.z, 1.0f);
            refl[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        } else {
            float cpos[4];
            mul4(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, cpos);
            v3 camPos = mk(cpos[0], cpos[1], cpos[2]);
            v3 position3 = ld3(P4);
            float roughness = mtl.y, metallic = mtl.z;
            v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
            v3 V = normalize(sub(camPos, position3));
            v3 N = normalize(ld3(N4));
            float NdotV = clamp01(dot(N, V));
            v3 irradiance = env_lookup(S.irr, S.irr_w, S.irr_h, N, PI_H);
            float attenuation = 1.0f;
            v3 rayColor = mk(0.0f, 0.0f, 0.0f);
            v3 origin = position3;
            v3 direction = mk(0.0f, 0.0f, 0.0f);
            if (mode == 3) direction = reflect(normalize(neg(V)), N);
            if (mode == 4) {
                float ior = mat.diffuse_ior[3];
                float cosAlpha = dot(N, V);
                v3 I = neg(V);
                float NdotI = dot(N, I);
                v3 refrNormal = NdotI > 0.0f ? neg(N) : N;
                float refrEta = NdotI > 0.0f ? 1.0f / ior : ior;
                float radicand = 1.0f + refrEta * refrEta * (cosAlpha * cosAlpha - 1.0f);
                direction = radicand < 0.0f ? reflect(I, N) : refract(I, refrNormal, refrEta);
            }
            for (uint32_t i = 0; i < F.nl; ++i) {
                const hr_light& lt = F.lights[i];
                bool isDir = lt.pos_maxdist[3] < 0.0f;
                v3 L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), position3);
                float maxd = lt.pos_maxdist[3];
                float dist = length(L);
                float intensity = isDir ? 1.0f : (lt.color_intensity[3] / (dist * dist));
                L = normalize(L);
                float NdotL = clamp01(dot(N, L));
                float shadowFactor = shadow[i * plane + p];
                if (NdotL > 0.0f) {
                    if (intensity == 0.0f) {
                        attenuation = 0.0f;
                    } else {
                        attenuation = maxd - dist;
                        attenuation /= maxd;
                        attenuation = fmaxf(attenuation, 0.0f);
                        attenuation = isDir ? 0.3f : attenuation * attenuation;
                    }
                }
                v3 lc = ld3p(lt.color_intensity);
                if (mode == 0) {
                    v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
                    v3 Hh = normalize(add(V, L));
                    float D = DistributionGGX(N, Hh, roughness);
                    float Gs = GeometrySmith(N, V, L, roughness);
                    v3 Fr = FresnelSchlick(fmaxf(dot(Hh, V), 0.0f), F0);
                    v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fr), 1.0f - metallic);
                    v3 numerator = scl(Fr, D * Gs);
                    float denominator = 4.0f * NdotV * NdotL;
                    v3 specular = divs(numerator, fmaxf(denominator, 0.001f));
                    v3 term = add(divs(mul(kD, albedo), PI_H), specular);
                    rayColor = add(rayColor, scl(mul(term, radiance), NdotL));
                } else if (mode == 3) {
                    v3 t = (NdotL > 0.0f && intensity > 0.0f)
                               ? scl(mul(scl(scl(scl(lc, intensity), attenuation), shadowFactor), albedo), metallic)
                               : scl(mul(irradiance, albedo), metallic);
                    rayColor = add(rayColor, t);
                } else if (mode == 4) {
                    v3 t = (intensity > 0.0f) ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                                              : scl(mul(irradiance, albedo), metallic);
                    rayColor = add(rayColor, t);
                }
            }
            v3 rfac = mk(1.0f, 1.0f, 1.0f);
            if ((mode == 3 || mode == 4) && F.max_depth > 0) {
                uint32_t spp = (mode == 3 && roughness > 0.0f) ? max(F.spp, 1u) : 1u;
                v3 acc = mk(0.0f, 0.0f, 0.0f);
                for (uint32_t s = 0; s < spp; ++s) {
                    v3 d = direction;
                    v3 prod = mk(1.0f, 1.0f, 1.0f);
                    bool alive = true;
                    if (mode == 3 && roughness > 0.0f) {
                        Pcg g = ray_stream(seed, PURPOSE_REFLECT, s);
                        float r1 = u01(g), r2 = u01(g);
                        v3 Hm = sampleGGX(N, roughness, r1, r2);
                        d = reflect(normalize(neg(V)), Hm);
                        if (!(dot(d, N) > 0.0f)) {
                            prod = mk(0.0f, 0.0f, 0.0f);
                            alive = false;
                        }
                    }
                    v3 o = origin;
                    for (uint32_t depth = 0; alive && depth < F.max_depth; ++depth) {
                        v3 ro = add(o, scl(d, 1e-2f));
                        RayP r = prep_ray(ro, d);
                        HitD h;
                        ++c.refl;
                        if (!traverse<false, COUNT>(S, r, 0.001f, 1000.0f, lds, h, c)) {
                            v3 sky = pow3(env_lookup(S.sky, S.sky_w, S.sky_h, normalize(d), PI_M), 2.2f);
                            prod = mul(prod, sky);
                            break;
                        }
                        HitOut ho = shade_hit<COUNT>(S, F, ro, d, h, seed, s, depth, lds, c);
                        prod = mul(prod, ho.color);
                        if (!ho.scatter) break;
                        o = ho.origin;
                        d = ho.dir;
                    }
                    acc = add(acc, prod);
                }
                rfac = spp == 1 ? acc : divs(acc, (float)spp);
                rayColor = mul(rayColor, rfac);
            }
            v3 Fa = FresnelSchlick(NdotV, F0);
            v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fa), 1.0f - metallic);
            v3 ambient = mul(mul(kD, albedo), irradiance);
            v3 fc = add(add(rayColor, ambient), emis);
            lit[p] = make_float4(fc.x, fc.y, fc.z, 1.0f);
            refl[p] = make_float4(rfac.x, rfac.y, rfac.z, 1.0f);
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
hipError_t launch_shade(const DevScene& s, const FrameArgs& f, GBufView gb, const float* shadow, float4* lit,
                        float4* refl, unsigned long long* counters, bool count, hipStream_t st) {
    if (count) k_shade<true><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, shadow, lit, refl, counters);
    else k_shade<false><<<pixel_blocks(f), BLOCK, 0, st>>>(s, f, gb, shadow, lit, refl, counters);
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
