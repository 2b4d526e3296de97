// rt.hip - the reference's pure ray tracing mode (raygen.rgen + closesthit.rchit
// + miss.rmiss + shadow.rmiss): primary rays from the camera, a closest-hit
// bounce chain, `samples` shadow rays per light at every hit, and progressive
// accumulation over frames (raygen.rgen:69-80).
//
// Two implementations with identical results:
// - the wavefront path (timed frames with 1..64 shadow samples): per camera
//   sample and depth, one closest-hit kernel over the live chains (a list),
//   the hit shadow rays as sample-parallel work items (item_loop), then one
//   shading kernel per hit that extends the live list;
// - the megakernel (visit counting, or more than 64 samples): one lane per
//   pixel walks its samples and bounces with the shadow rays inline.
// Per ray, the random numbers come from the pixel's PCG32 stream for
// (purpose, index), as in the hybrid pass; the hit shadow samples of one
// light share a stream, two numbers per sample, so the wavefront path jumps
// the stream ahead by 2 * sample (pcg_advance).
#include "trace_common.h"

namespace hr {
using namespace hrd;
using namespace tc;

namespace {

struct RtOut {
    v3 color, dir;
    bool scatter;
};

// Light i seen from a hit (closesthit.rchit:88-97).
struct RtLight {
    v3 L;
    float dist, maxd, NdotL;
    bool isDir;
};
__device__ __forceinline__ RtLight rt_light(const FrameArgs& F, uint32_t i, v3 worldPos, v3 N) {
    const hr_light& lt = F.lights[i];
    RtLight o;
    o.isDir = lt.pos_maxdist[3] < 0.0f;
    v3 L = o.isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), worldPos);
    o.maxd = lt.pos_maxdist[3];
    o.dist = length(L);
    o.L = normalize(L);
    o.NdotL = clamp01(dot(N, o.L));
    return o;
}

// stream index of light i's hit shadow samples at (camera sample, depth)
__device__ __forceinline__ uint32_t rt_shadow_stream_index(const FrameArgs& F, uint32_t sample, uint32_t depth,
                                                           uint32_t i) {
    return (sample * F.max_depth + depth) * F.nl + i;
}

// closesthit.rchit:24-186 for one closest hit, with the shadow factor of
// light i given by sfac(i, light).
template <class SF>
__device__ RtOut shade_rt(const DevScene& S, const FrameArgs& F, v3 rd, const HitGeo& g, uint32_t seed,
                          uint32_t sample, uint32_t depth, const SF& sfac) {
    RtOut out;
    const hr_material& mat = *g.mat;
    const v3 N = g.N, V = g.V;
    const float NdotV = g.NdotV;
    v3 Lo = mk(0.0f, 0.0f, 0.0f);
    const int mode = (int)mat.shading_metal_rough_idx[0];
    const v3 albedo = pow3(g.albedo, 2.2f);
    const float roughness = g.roughness;
    const float metallic = g.metallic;
    const v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
    const v3 irradiance = env_lookup(S.irr, S.irr_w, S.irr_h, N, PI_H);
    v3 direction = mk(1.0f, 1.0f, 1.0f);
    bool scatter = false;
    v3 reflected = mk(0.0f, 0.0f, 0.0f);
    if (mode == 3) {
        v3 Nr = N;
        if (roughness > 0.0f) {  // glossy extension: GGX microfacet normal
            Pcg gg = ray_stream(seed, PURPOSE_RT_REFLECT, sample * F.max_depth + depth);
            float r1 = u01(gg), r2 = u01(gg);
            Nr = sampleGGX(N, roughness, r1, r2);
        }
        reflected = reflect(normalize(rd), Nr);
    }
    for (uint32_t i = 0; i < F.nl; ++i) {
        const hr_light& lt = F.lights[i];
        const RtLight lg = rt_light(F, i, g.worldPos, N);
        const bool isDir = lg.isDir;
        const v3 L = lg.L;
        const float maxd = lg.maxd;
        const float dist = lg.dist;
        const float intensity = isDir ? 1.0f : (lt.color_intensity[3] / (dist * dist));
        const v3 H = normalize(add(V, L));
        const float NdotL = lg.NdotL;
        const float NdotH = clamp01(dot(N, H));
        const float shadowFactor = sfac(i, lg);
        float attenuation;
        if (intensity == 0.0f) {
            attenuation = 0.0f;
        } else {
            attenuation = maxd - dist;
            attenuation /= maxd;
            attenuation = fmaxf(attenuation, 0.0f);
            attenuation = attenuation * attenuation;
        }
        const v3 lc = ld3p(lt.color_intensity);
        if (mode == 0) {
            const v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
            const v3 Fr = FresnelSchlick(NdotH, F0);
            const float D = DistributionGGX(N, H, roughness);
            const float G = GeometrySmith(N, V, L, roughness);
            const v3 numerator = scl(Fr, D * G);
            const float denominator = fmaxf(4.0f * clamp01(dot(N, V)) * NdotL, 0.000001f);
            const v3 specular = divs(numerator, denominator);
            const v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fr), 1.0f - metallic);
            const v3 term = add(divs(mul(kD, albedo), PI_H), specular);
            Lo = add(Lo, scl(mul(term, radiance), NdotL));
            direction = mk(1.0f, 1.0f, 1.0f);
            scatter = false;
        } else if (mode == 3) {
            const bool sc = dot(reflected, N) > 0.0f;
            const v3 t = (NdotL > 0.0f && intensity > 0.0f)
                             ? mul(scl(scl(scl(scl(lc, intensity), attenuation), shadowFactor), metallic), albedo)
                             : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, t);
            direction = reflected;
            scatter = sc;
        } else if (mode == 4) {
            const float ior = mat.diffuse_ior[3];
            const float ndv = dot(N, normalize(rd));
            const v3 refrNormal = ndv > 0.0f ? neg(N) : N;
            const float refrEta = ndv > 0.0f ? 1.0f / ior : ior;
            const v3 t = (intensity > 0.0f) ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                                            : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, t);
            const float radicand = 1.0f + refrEta * refrEta * (ndv * ndv - 1.0f);
            direction = radicand < 0.0f ? reflect(rd, N) : refract(rd, refrNormal, refrEta);
            scatter = true;
        }
    }
    const v3 Fa = FresnelSchlick(NdotV, F0);
    const v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fa), 1.0f - metallic);
    const v3 ambient = mul(mul(kD, albedo), irradiance);
    out.color = add(add(Lo, ambient), g.emissive);
    out.dir = direction;
    out.scatter = scatter;
    return out;
}

// Shadow samples of light i at a hit: traced only with NdotL > 0 and within
// the light's range; a light beyond its range counts as shadowed
// (closesthit.rchit:103-121). The factor is unshadowed / samples.
// The megakernel's shadow factor: samples traced inline, in order.
template <bool COUNT>
struct InlineShadows {
    const DevScene& S;
    const FrameArgs& F;
    const HitGeo& g;
    uint32_t seed, sample, depth;
    uint32_t* lds;
    Cnt& c;
    __device__ __forceinline__ float operator()(uint32_t i, const RtLight& lg) const {
        float shadowFactor = 0.0f;
        if (lg.NdotL > 0.0f) {
            const uint32_t SS = F.samples;
            Pcg gs = ray_stream(seed, PURPOSE_RT_SHADOW, rt_shadow_stream_index(F, sample, depth, i));
            for (uint32_t a = 0; a < SS; ++a) {
                bool shadowed = true;  // lights beyond their range stay shadowed
                if (lg.dist < lg.maxd) {
                    const v3 dir = light_sample_dir(F, i, g.worldPos, lg.L, gs, a, SS);
                    ++c.hitshadow;
                    shadowed = occluded<COUNT>(S, add(g.worldPos, scl(dir, 1e-2f)), dir, 0.001f, lg.dist + 1.0f,
                                               lds, c);
                    if (shadowed) ++c.occluded;
                }
                if (!shadowed) shadowFactor += 1.0f;
            }
            shadowFactor /= (float)SS;
        }
        return shadowFactor;
    }
};

// primary ray of camera sample nS through pixel (x, y) (raygen.rgen:29-40)
__device__ __forceinline__ void rt_primary(const FrameArgs& F, uint32_t x, uint32_t y, uint32_t seed, uint32_t nS,
                                           v3& o, v3& d) {
    float ox = 0.5f, oy = 0.5f;
    if (nS > 0) {
        Pcg gp = ray_stream(seed, PURPOSE_RT_PIXEL, nS);
        ox = u01(gp);
        oy = u01(gp);
    }
    const float ux = ((float)x + ox) / (float)F.width;
    const float uy = ((float)y + oy) / (float)F.height;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    float org[4], tgt[4], dir[4];
    mul4(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, org);
    mul4(F.inv_proj, dx, dy, 1.0f, 1.0f, tgt);
    const v3 tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
    mul4(F.inv_view, tn.x, tn.y, tn.z, 0.0f, dir);
    o = mk(org[0], org[1], org[2]);
    d = mk(dir[0], dir[1], dir[2]);
}

// progressive accumulation (raygen.rgen:69-80)
__device__ __forceinline__ void rt_store(float4* accum, size_t p, v3 finalColor, int32_t frame_signed) {
    if (frame_signed < 0) {
        accum[p] = make_float4(finalColor.x, finalColor.y, finalColor.z, 1.0f);
    } else {
        const float a = 1.0f / (1.0f + (float)frame_signed);
        const float4 last = accum[p];
        const v3 m = add(scl(mk(last.x, last.y, last.z), 1.0f - a), scl(finalColor, a));
        accum[p] = make_float4(m.x, m.y, m.z, 1.0f);
    }
}

// raygen.rgen:20-81
template <bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_rt(DevScene S, FrameArgs F, float4* accum, int32_t frame_signed,
                                              unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t x, y;
    pixel_of(F.width, x, y);
    if (x < F.width && y < F.height) {
        const size_t p = (size_t)y * F.width + x;
        const uint32_t seed = tea(y * F.width + x, F.frame);
        v3 finalColor = mk(0.0f, 0.0f, 0.0f);
        const uint32_t spp = max(F.spp, 1u);
        for (uint32_t nS = 0; nS < spp; ++nS) {
            v3 o, d;
            rt_primary(F, x, y, seed, nS, o, d);
            v3 rayColor = mk(1.0f, 1.0f, 1.0f);
            for (uint32_t depth = 0; depth < F.max_depth; ++depth) {
                const v3 ro = add(o, scl(d, 1e-2f));
                const RayP r = prep_ray(ro, d);
                HitD h;
                if (depth == 0) ++c.primary; else ++c.refl;
                const bool hit = traverse<false, COUNT>(S, r, 0.001f, 10000.0f, lds, h, c);
                if (!hit) {  // miss.rmiss:13-18
                    rayColor = mul(rayColor, pow3(env_lookup(S.sky, S.sky_w, S.sky_h, normalize(d), PI_M), 2.2f));
                    break;
                }
                const HitGeo g = hit_geometry(S, ro, d, h);
                const InlineShadows<COUNT> sfac{S, F, g, seed, nS, depth, lds, c};
                const RtOut hc = shade_rt(S, F, d, g, seed, nS, depth, sfac);
                rayColor = mul(rayColor, hc.color);
                if (!hc.scatter) break;
                o = g.worldPos;
                d = hc.dir;
            }
            finalColor = add(finalColor, rayColor);
        }
        finalColor = divs(finalColor, (float)spp);
        rt_store(accum, p, finalColor, frame_signed);
    }
    flush_counters(c, counters, COUNT);
}

// ---- wavefront path --------------------------------------------------------------
// Per pixel: chain origin and direction, ray colour (product of the hit
// colours so far), sum over camera samples. Per hit slot (one per closest hit
// of the current depth): t/u/v/triangle, world position, shading normal and
// pixel. Shadow sample bits: item ((light * hits + slot) << lg2s) + sample, so
// a wave's items are the samples of neighbouring hits towards one light.
struct RtWork {
    float4 *o, *d, *color, *fin;
    uint32_t* list[2];  // live chains by depth parity (pixel ids, ~0u = none); depth 0 in 8x8 tile order
    uint4* hit;
    float4 *hpos, *hnrm;
    uint32_t* hpix;
    uint32_t* bits;
    uint32_t* cnt;  // [0], [1]: live list lengths, [2]: hits of the current depth, [3]: listed groups
    uint32_t* glist;  // the (light, hit) groups of the current depth that need shadow rays, in group order
    void* gtail;      // scratch of their listing
};

__device__ __forceinline__ uint32_t rt_tiles(const FrameArgs& F) {
    return ((F.width + 7u) / 8u) * ((F.height + 7u) / 8u);
}

// camera sample nS: the primary ray of every pixel, the live list in tile order
__global__ __launch_bounds__(BLOCK) void k_rt_begin(FrameArgs F, RtWork w, uint32_t nS) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t T = rt_tiles(F) * 64u;
    if (i == 0) {
        w.cnt[0] = T;
        w.cnt[1] = 0u;
        w.cnt[2] = 0u;
        w.cnt[3] = 0u;
    }
    if (i >= T) return;
    const uint32_t t = i >> 6, q = i & 63u, tx = (F.width + 7u) / 8u;
    const uint32_t x = (t % tx) * 8u + (q & 7u), y = (t / tx) * 8u + (q >> 3);
    if (x >= F.width || y >= F.height) {
        w.list[0][i] = ~0u;
        return;
    }
    const uint32_t pix = y * F.width + x;
    w.list[0][i] = pix;
    v3 o, d;
    rt_primary(F, x, y, tea(pix, F.frame), nS, o, d);
    w.o[pix] = make_float4(o.x, o.y, o.z, 0.0f);
    w.d[pix] = make_float4(d.x, d.y, d.z, 0.0f);
    w.color[pix] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    if (nS == 0u) w.fin[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// before depth `depth`: no hits or listed groups yet, the next live list
// empty, work bands reset
__global__ void k_rt_reset(RtWork w, uint32_t depth, uint32_t* next_item) {
    if (threadIdx.x == 0) {
        w.cnt[2] = 0u;
        w.cnt[3] = 0u;
        w.cnt[(depth + 1u) & 1u] = 0u;
    }
    if (threadIdx.x < SHADOW_QUEUES) next_item[threadIdx.x * BAND_STRIDE] = 0u;
}

// wave-aggregated append: slot of this lane among the wave's takers
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool take) {
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return 0u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// closest hit of every live chain (raygen.rgen:52-67; miss.rmiss:13-18)
__global__ __launch_bounds__(BLOCK) void k_rt_closest(DevScene S, FrameArgs F, RtWork w, uint32_t depth,
                                                      unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    Cnt c = {};
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t n = w.cnt[depth & 1u];
    if (blockIdx.x * BLOCK >= n) return;
    const uint32_t pix = i < n ? w.list[depth & 1u][i] : ~0u;
    bool is_hit = false;
    HitD h;
    HitGeo g;
    if (pix != ~0u) {
        const float4 o4 = w.o[pix], d4 = w.d[pix];
        const v3 d = ld3(d4);
        const v3 ro = add(ld3(o4), scl(d, 1e-2f));
        const RayP r = prep_ray(ro, d);
        if (depth == 0) ++c.primary; else ++c.refl;
        is_hit = traverse<false, false>(S, r, 0.001f, 10000.0f, stk + threadIdx.x, h, c);
        if (!is_hit) {
            const float4 cl = w.color[pix];
            const v3 col = mul(ld3(cl), pow3(env_lookup(S.sky, S.sky_w, S.sky_h, normalize(d), PI_M), 2.2f));
            w.color[pix] = make_float4(col.x, col.y, col.z, 0.0f);
        } else {
            g = hit_geometry(S, ro, d, h);
        }
    }
    const uint32_t slot = wave_append(w.cnt + 2, is_hit);
    if (is_hit) {
        w.hit[slot] = make_uint4(__float_as_uint(h.t), __float_as_uint(h.u), __float_as_uint(h.v), h.tri);
        w.hpos[slot] = make_float4(g.worldPos.x, g.worldPos.y, g.worldPos.z, 0.0f);
        w.hnrm[slot] = make_float4(g.N.x, g.N.y, g.N.z, 0.0f);
        w.hpix[slot] = pix;
    }
    flush_counters(c, counters, false);
}

// The (light, hit) groups G = light * hits + slot of the current depth whose
// shadow rays are traced (the light in front of the surface and within reach;
// the others' samples set no bit): per chunk of 64 groups the ballot of them
// and its count, for every chunk of the frame's largest group count.
__global__ __launch_bounds__(256) void k_rt_group_kinds(FrameArgs F, RtWork w, uint32_t ngroups_max,
                                                        unsigned long long* masks, uint32_t* counts) {
    const uint32_t G = blockIdx.x * 256u + threadIdx.x;
    const uint32_t nh = w.cnt[2];
    bool trace = false;
    if (G < F.nl * nh) {
        const uint32_t li = G / nh, slot = G - li * nh;
        const RtLight lg = rt_light(F, li, ld3(w.hpos[slot]), ld3(w.hnrm[slot]));
        trace = lg.NdotL > 0.0f && lg.dist < lg.maxd;
    }
    const unsigned long long m = __ballot(trace);
    if ((threadIdx.x & 63u) == 0u && G < ngroups_max) {
        masks[G >> 6] = m;
        counts[G >> 6] = (uint32_t)__popcll(m);
    }
}

// clear the shadow bits of the current depth's items
__global__ __launch_bounds__(256) void k_rt_clear_bits(RtWork w, uint32_t nl, uint32_t lg2s) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t words = (uint32_t)((((uint64_t)nl * w.cnt[2] << lg2s) + 31u) >> 5);
    if (i < words) w.bits[i] = 0u;
}

// Pcg state of sample a of a stream: two numbers per light sample
__device__ __forceinline__ Pcg stream_at_sample(uint32_t seed, uint32_t index, uint32_t a) {
    Pcg g = ray_stream(seed, PURPOSE_RT_SHADOW, index);
    pcg_advance(g, 2u * a);
    return g;
}

// ray source of the hit shadow items: the samples of the listed groups (item
// = list slot * 2^lg2s + sample; its bit (group << lg2s) | sample)
struct RtItems {
    const FrameArgs& F;
    const float4* __restrict__ hpos;
    const float4* __restrict__ hnrm;
    const uint32_t* __restrict__ hpix;
    const uint32_t* __restrict__ glist;
    uint32_t nh, lg2s, nS, depth;
    const TNode4* nodes4;
    const STri* tris;
    uint32_t node_count;

    __device__ __forceinline__ int setup(uint32_t item, uint32_t& bit, RayP& r, float& tmax, const TNode4*&,
                                         const STri*&, uint32_t&, Cnt& c) const {
        const uint32_t G = glist[item >> lg2s], a = item & ((1u << lg2s) - 1u);
        bit = (G << lg2s) | a;
        const uint32_t SS = F.samples;
        if (a >= SS) return ITEM_SKIP;
        const uint32_t li = G / nh, slot = G - li * nh;
        const v3 P = ld3(hpos[slot]);
        const RtLight lg = rt_light(F, li, P, ld3(hnrm[slot]));
        if (!(lg.NdotL > 0.0f) || !(lg.dist < lg.maxd)) return ITEM_SKIP;
        Pcg gs = stream_at_sample(tea(hpix[slot], F.frame), rt_shadow_stream_index(F, nS, depth, li), a);
        const v3 dir = light_sample_dir(F, li, P, lg.L, gs, a, SS);
        ++c.hitshadow;
        r = prep_ray(add(P, scl(dir, 1e-2f)), dir);
        tmax = lg.dist + 1.0f;
        return ITEM_TRACE;
    }
};

template <bool OFF32>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_rt_shadows(
    DevScene S, FrameArgs F, RtWork w, uint32_t lg2s, uint32_t nS, uint32_t depth, uint32_t* next_item,
    unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * BLOCK];
    Cnt c = {};
    const uint32_t nh = w.cnt[2];
    RtItems src{F, w.hpos, w.hnrm, w.hpix, w.glist, nh, lg2s, nS, depth, S.nodes4, S.tris, S.node_count};
    item_loop<OFF32>(src, w.cnt[3] << lg2s, next_item, w.bits, stk + threadIdx.x, c);
    flush_counters(c, counters, false);
}

// the shadow factor from the sample bits
struct BitShadows {
    const uint32_t* __restrict__ bits;
    uint32_t nh, slot, lg2s, samples;
    __device__ __forceinline__ float operator()(uint32_t i, const RtLight&) const {
        return (float)window_bits(bits, (uint64_t)(i * nh + slot) << lg2s, lg2s) / (float)samples;
    }
};

// shading of every hit (closesthit.rchit:24-186), the chain's next ray
__global__ __launch_bounds__(BLOCK) void k_rt_shade(DevScene S, FrameArgs F, RtWork w, uint32_t lg2s, uint32_t nS,
                                                    uint32_t depth) {
    const uint32_t slot = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t nh = w.cnt[2];
    if (blockIdx.x * BLOCK >= nh) return;
    bool next = false;
    uint32_t pix = 0u;
    if (slot < nh) {
        pix = w.hpix[slot];
        const uint4 hr = w.hit[slot];
        const HitD h{__uint_as_float(hr.x), __uint_as_float(hr.y), __uint_as_float(hr.z), hr.w};
        const v3 d = ld3(w.d[pix]);
        const v3 ro = add(ld3(w.o[pix]), scl(d, 1e-2f));
        const HitGeo g = hit_geometry(S, ro, d, h);
        const BitShadows sfac{w.bits, nh, slot, lg2s, F.samples};
        const RtOut hc = shade_rt(S, F, d, g, tea(pix, F.frame), nS, depth, sfac);
        const v3 col = mul(ld3(w.color[pix]), hc.color);
        w.color[pix] = make_float4(col.x, col.y, col.z, 0.0f);
        if (hc.scatter && depth + 1u < F.max_depth) {
            w.o[pix] = make_float4(g.worldPos.x, g.worldPos.y, g.worldPos.z, 0.0f);
            w.d[pix] = make_float4(hc.dir.x, hc.dir.y, hc.dir.z, 0.0f);
            next = true;
        }
    }
    const uint32_t k = wave_append(w.cnt + ((depth + 1u) & 1u), next);
    if (next) w.list[(depth + 1u) & 1u][k] = pix;
}

// after camera sample nS: the sum over samples; after the last, the average
// and the progressive accumulation
__global__ __launch_bounds__(BLOCK) void k_rt_end(FrameArgs F, RtWork w, uint32_t nS, float4* accum,
                                                  int32_t frame_signed) {
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= F.width * F.height) return;
    const v3 sum = add(ld3(w.fin[p]), ld3(w.color[p]));
    const uint32_t spp = max(F.spp, 1u);
    if (nS + 1u < spp) {
        w.fin[p] = make_float4(sum.x, sum.y, sum.z, 0.0f);
        return;
    }
    rt_store(accum, p, divs(sum, (float)spp), frame_signed);
}

inline uint32_t rt_lg2s(uint32_t S) {
    uint32_t l = 0;
    while ((1u << l) < S) ++l;
    return l;
}

size_t rt_layout(const FrameArgs& f, char* base, RtWork& w) {
    const size_t P = (size_t)f.width * f.height;
    const size_t T = (size_t)((f.width + 7u) / 8u) * ((f.height + 7u) / 8u) * 64u;
    const uint32_t lg2s = rt_lg2s(f.samples);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255u) & ~(size_t)255u;
        return p;
    };
    w.o = reinterpret_cast<float4*>(take(P * 16));
    w.d = reinterpret_cast<float4*>(take(P * 16));
    w.color = reinterpret_cast<float4*>(take(P * 16));
    w.fin = reinterpret_cast<float4*>(take(P * 16));
    w.hit = reinterpret_cast<uint4*>(take(P * 16));
    w.hpos = reinterpret_cast<float4*>(take(P * 16));
    w.hnrm = reinterpret_cast<float4*>(take(P * 16));
    w.hpix = reinterpret_cast<uint32_t*>(take(P * 4));
    w.list[0] = reinterpret_cast<uint32_t*>(take(T * 4));
    w.list[1] = reinterpret_cast<uint32_t*>(take(T * 4));
    w.bits = reinterpret_cast<uint32_t*>(take((((uint64_t)f.nl * P << lg2s) + 31u) / 32u * 4u));
    w.cnt = reinterpret_cast<uint32_t*>(take(64));
    const uint32_t groups = f.nl * (uint32_t)P;
    w.glist = reinterpret_cast<uint32_t*>(take((size_t)groups * 4u));
    w.gtail = take(group_list_tail_bytes(groups));
    return off;
}

}  // namespace

size_t rt_work_bytes(const FrameArgs& f) {
    if (f.samples == 0u || f.samples > 64u) return 0;
    const uint64_t items = ((uint64_t)f.nl * f.width * f.height) << rt_lg2s(f.samples);
    if (items >= (1ull << 32)) return 0;
    RtWork w;
    return rt_layout(f, nullptr, w);
}

hipError_t launch_rt_wavefront(const DevScene& s, const FrameArgs& f, float4* accum, int32_t frame_signed,
                               unsigned long long* counters, hipStream_t st, void* work, uint32_t* next_item,
                               uint32_t resident_blocks) {
    RtWork w;
    rt_layout(f, static_cast<char*>(work), w);
    const uint32_t lg2s = rt_lg2s(f.samples);
    const uint32_t P = f.width * f.height;
    const uint32_t T = ((f.width + 7u) / 8u) * ((f.height + 7u) / 8u) * 64u;
    const uint32_t nbT = (T + BLOCK - 1u) / BLOCK, nbP = (P + BLOCK - 1u) / BLOCK;
    const uint64_t max_words = ((((uint64_t)f.nl * P) << lg2s) + 31u) >> 5;
    const uint32_t nbW = (uint32_t)((max_words + 255u) / 256u);
    const uint32_t spp = max(f.spp, 1u);
    const uint32_t gmax = f.nl * P;  // (light, hit) groups of a depth at most
    const GroupListTail gt = group_list_tail(w.gtail, gmax);
    for (uint32_t nS = 0; nS < spp; ++nS) {
        k_rt_begin<<<nbT, BLOCK, 0, st>>>(f, w, nS);
        for (uint32_t depth = 0; depth < f.max_depth; ++depth) {
            k_rt_reset<<<1, 64, 0, st>>>(w, depth, next_item);
            k_rt_closest<<<nbT, BLOCK, 0, st>>>(s, f, w, depth, counters);
            if (nbW) k_rt_clear_bits<<<nbW, 256, 0, st>>>(w, f.nl, lg2s);
            if (gmax) {
                k_rt_group_kinds<<<(gmax + 255u) / 256u, 256, 0, st>>>(f, w, gmax, gt.masks, gt.counts);
                scan_and_scatter(gmax, gt, w.cnt + 3, w.glist, st);
            }
            if ((uint64_t)s.wide_count * sizeof(TNode4) <= 0xFFFFFFFFull)
                k_rt_shadows<true><<<resident_blocks, BLOCK, 0, st>>>(s, f, w, lg2s, nS, depth, next_item, counters);
            else
                k_rt_shadows<false><<<resident_blocks, BLOCK, 0, st>>>(s, f, w, lg2s, nS, depth, next_item, counters);
            k_rt_shade<<<nbT, BLOCK, 0, st>>>(s, f, w, lg2s, nS, depth);
        }
        k_rt_end<<<nbP, BLOCK, 0, st>>>(f, w, nS, accum, frame_signed);
    }
    return hipGetLastError();
}

hipError_t launch_rt(const DevScene& s, const FrameArgs& f, float4* accum, int32_t frame_signed,
                     unsigned long long* counters, bool count, hipStream_t st) {
    const uint32_t nb = ((f.width + 15u) / 16u) * ((f.height + 15u) / 16u);
    if (count) k_rt<true><<<nb, BLOCK, 0, st>>>(s, f, accum, frame_signed, counters);
    else k_rt<false><<<nb, BLOCK, 0, st>>>(s, f, accum, frame_signed, counters);
    return hipGetLastError();
}

}  // namespace hr
