// wavefront.hip - the hybrid shading pass (hybridRaygen.rgen:27-186 with the
// closest-hit shading of hybridHit.rchit:24-183 and miss.rmiss:13-18) as a
// wavefront pipeline over device-side ray queues.
//
//   k_wf_direct   per pixel: G-buffer decode, direct light terms, IBL ambient;
//                 pixels that reflect or refract (shading models 3 / 4) push
//                 their paths (refl_spp GGX samples or the mirror / glass
//                 direction) onto queue 0
//   per bounce depth d < max_depth:
//     k_wf_trace  closest hit of every queued path
//     k_wf_shade  miss -> sky; model 3/4 hit -> shaded now, pushed onto the
//                 next queue while it scatters; model 0 hit -> its hit shadow
//                 rays (one per light, hybridHit.rchit:119-129) pushed onto the
//                 shadow queue and the path onto the pending list
//     k_wf_shadow any hit of every shadow ray; unoccluded rays set their light's
//                 bit in the path's visibility mask
//     k_wf_finish pending paths shaded with the visibility mask
//   k_wf_resolve  per pixel: the samples' bounce factors summed in sample
//                 order, lit = direct * refl + ambient + emissive
//
// Pushes are wave-aggregated (one atomic per wave, ranks from a wave prefix
// sum). Queue lengths stay in device memory and every stage runs a fixed
// grid that strides over the current length, so a frame needs no host
// synchronisation. Each path owns a result slot (pixel, sample), so the
// output does not depend on queue order.
#include "trace_common.h"

namespace hr {
using namespace hrd;
using namespace tc;

namespace {

constexpr int WF_BLOCK = 256;

// path record: (origin, pixel), (direction, sample | depth << 16), (throughput, 0)
struct PathRec {
    v3 o, d, thr;
    uint32_t pixel, sample, depth;
};
__device__ __forceinline__ PathRec load_path(const float4* q, uint32_t i) {
    float4 a = q[3 * (size_t)i], b = q[3 * (size_t)i + 1], c = q[3 * (size_t)i + 2];
    PathRec r;
    r.o = ld3(a);
    r.d = ld3(b);
    r.thr = ld3(c);
    r.pixel = __float_as_uint(a.w);
    uint32_t sd = __float_as_uint(b.w);
    r.sample = sd & 0xFFFFu;
    r.depth = sd >> 16;
    return r;
}
__device__ __forceinline__ void store_path(float4* q, uint32_t i, v3 o, v3 d, v3 thr, uint32_t pixel,
                                           uint32_t sample, uint32_t depth) {
    q[3 * (size_t)i] = make_float4(o.x, o.y, o.z, __uint_as_float(pixel));
    q[3 * (size_t)i + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(sample | (depth << 16)));
    q[3 * (size_t)i + 2] = make_float4(thr.x, thr.y, thr.z, 0.0f);
}

// Wave-aggregated reservation of n slots per lane on counter *cnt; returns
// the lane's first slot. Must be reached by the whole wave.
__device__ __forceinline__ uint32_t wave_push(uint32_t* cnt, uint32_t n) {
    uint32_t lane = threadIdx.x & 63u;
    uint32_t inc = n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= (uint32_t)off) inc += t;
    }
    uint32_t total = (uint32_t)__shfl((int)inc, 63, 64);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(cnt, total);
    base = (uint32_t)__shfl((int)base, 0, 64);
    return base + inc - n;
}

// Does light i need a hit-point shadow ray (shading model 0, NdotL > 0 and
// in range, hybridHit.rchit:309-324)?
__device__ __forceinline__ bool hit_shadow_needed(const FrameArgs& F, uint32_t i, const HitGeo& g, v3& L, float& dist) {
    const hr_light& lt = F.lights[i];
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), g.worldPos);
    dist = length(L);
    L = normalize(L);
    float NdotL = clamp01(dot(g.N, L));
    return NdotL > 0.0f && (isDir || dist < lt.pos_maxdist[3]);
}

// hybridHit.rchit:24-183 with the hit shadow rays' outcomes given as a mask
// (bit i set = the ray towards light i was unoccluded).
struct HitOut {
    v3 color, dir;
    bool scatter;
};
__device__ HitOut shade_hit(const DevScene& S, const FrameArgs& F, v3 rd, const HitGeo& g, uint32_t seed,
                            uint32_t sample, uint32_t depth, uint32_t vis) {
    HitOut out;
    const hr_material& mat = *g.mat;
    v3 N = g.N, V = g.V;
    float NdotV = g.NdotV;
    v3 Lo = mk(0.0f, 0.0f, 0.0f);
    int mode = (int)mat.shading_metal_rough_idx[0];
    v3 albedo = pow3(g.albedo, 2.2f);
    float roughness = g.roughness;
    float metallic = g.metallic;
    v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
    v3 irradiance = env_lookup(S.irr, S.irr_w, S.irr_h, N, PI_H);
    v3 direction = mk(1.0f, 1.0f, 1.0f);
    bool scatter = false;
    v3 rdn = normalize(rd);
    v3 reflected = mk(0.0f, 0.0f, 0.0f);
    if (mode == 3) {
        v3 Nr = N;
        if (roughness > 0.0f) {
            Pcg gg = ray_stream(seed, PURPOSE_HIT_REFLECT, sample * F.max_depth + depth);
            float r1 = u01(gg), r2 = u01(gg);
            Nr = sampleGGX(N, roughness, r1, r2);
        }
        reflected = reflect(rdn, Nr);
    }
    for (uint32_t i = 0; i < F.nl; ++i) {
        const hr_light& lt = F.lights[i];
        bool isDir = lt.pos_maxdist[3] < 0.0f;
        v3 L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), g.worldPos);
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
                if (isDir || dist < maxd) shadowed = ((vis >> i) & 1u) == 0u;
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
    out.color = add(add(Lo, ambient), g.emissive);
    out.dir = direction;
    out.scatter = scatter;
    return out;
}

__device__ __forceinline__ size_t slot_of(const WaveArgs& W, uint32_t pixel, uint32_t sample) {
    return (size_t)(pixel - W.p0) * W.spp_max + sample;
}
// paths of a reflecting pixel: refl_spp for glossy surfaces (1 for mirrors
// and glass); adaptive: 1 + floor((refl_spp - 1) * roughness)
__device__ __forceinline__ uint32_t pixel_spp(const FrameArgs& F, int mode, float roughness) {
    if (!(mode == 3 && roughness > 0.0f)) return 1u;
    uint32_t m = max(F.spp, 1u);
    if (!F.adaptive) return m;
    uint32_t k = 1u + (uint32_t)floorf((float)(m - 1u) * roughness);
    return min(k, m);
}

// ---- stage 1: direct light and path emission (hybridRaygen.rgen:27-149) ------
__global__ __launch_bounds__(WF_BLOCK) void k_wf_direct(DevScene S, FrameArgs F, GBufView G, const float* shadow,
                                                        float4* lit, float4* refl, WaveArgs W,
                                                        unsigned long long* counters) {
    uint32_t p = W.p0 + blockIdx.x * WF_BLOCK + threadIdx.x;
    bool active = p < W.p1;
    uint32_t npush = 0;
    v3 position3 = mk(0.f, 0.f, 0.f), N = position3, V = position3, direction = position3;
    float roughness = 0.0f;
    int mode = 0;
    uint32_t seed = 0;
    uint32_t nspp = 0;
    if (active) {
        size_t plane = (size_t)F.width * F.height;
        uint32_t x = p % F.width, y = p / F.width;
        seed = tea(y * F.width + x, F.frame);
        float4 P4 = G.pos[p], N4 = G.nrm[p];
        const hr_material& mat = S.mats[mat_index(S, P4.w)];
        mode = (int)mat.shading_metal_rough_idx[0];
        v3 albedo = pow3(ld3(G.alb[p]), 2.2f);
        v3 mtl = ld3(G.mat[p]);
        v3 emis = ld3(G.emi[p]);
        bool background = P4.w == 0.0f && N4.w == 0.0f;
        if (background) {
            put_lit(lit, p, albedo, F.unorm8);
            refl[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            W.pix[2 * (size_t)(p - W.p0) + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            float cpos[4];
            mul4(F.inv_view, 0.0f, 0.0f, 0.0f, 1.0f, cpos);
            v3 camPos = mk(cpos[0], cpos[1], cpos[2]);
            position3 = ld3(P4);
            roughness = mtl.y;
            float metallic = mtl.z;
            v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
            V = normalize(sub(camPos, position3));
            N = normalize(ld3(N4));
            float NdotV = clamp01(dot(N, V));
            v3 irradiance = env_lookup(S.irr, S.irr_w, S.irr_h, N, PI_H);
            float attenuation = 1.0f;
            v3 rayColor = mk(0.0f, 0.0f, 0.0f);
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
                float shadowFactor = get_factor(shadow, i * plane + p, F.unorm8);
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
            v3 Fa = FresnelSchlick(NdotV, F0);
            v3 kD = scl(sub(mk(1.0f, 1.0f, 1.0f), Fa), 1.0f - metallic);
            v3 ambient = mul(mul(kD, albedo), irradiance);
            // no lights: the reference bounces a zero direction (set only in
            // its light loop) into a zero rayColor, so nothing is traced
            bool traced = (mode == 3 || mode == 4) && F.max_depth > 0 && F.nl > 0;
            if (!traced) {
                v3 fc = add(add(rayColor, ambient), emis);
                put_lit(lit, p, fc, F.unorm8);
                refl[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
                W.pix[2 * (size_t)(p - W.p0) + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                nspp = pixel_spp(F, mode, roughness);
                W.pix[2 * (size_t)(p - W.p0)] = make_float4(rayColor.x, rayColor.y, rayColor.z, 0.0f);
                W.pix[2 * (size_t)(p - W.p0) + 1] = make_float4(ambient.x, ambient.y, ambient.z, (float)nspp);
                npush = nspp;
            }
        }
    }
    // reserve queue slots for this wave's paths, then write them
    uint32_t base = wave_push(&W.counts[0 * BAND_STRIDE], npush);
    if (npush) {
        bool glossy = mode == 3 && roughness > 0.0f;
        uint32_t k = 0;
        for (uint32_t s = 0; s < nspp; ++s) {
            v3 d = direction;
            if (glossy) {
                Pcg g = ray_stream(seed, PURPOSE_REFLECT, s);
                float r1 = u01(g), r2 = u01(g);
                v3 Hm = sampleGGX(N, roughness, r1, r2);
                d = reflect(normalize(neg(V)), Hm);
            }
            if (glossy && !(dot(d, N) > 0.0f)) {
                W.pathcol[slot_of(W, p, s)] = make_float4(0.f, 0.f, 0.f, 0.f);  // below the surface
            } else {
                store_path(W.paths0, base + k, position3, d, mk(1.0f, 1.0f, 1.0f), p, s, 0);
                ++k;
            }
        }
        for (; k < npush; ++k)  // unused reserved slots become empty paths
            store_path(W.paths0, base + k, mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), 0xFFFFFFFFu, 0, 0);
    }
    (void)counters;
}

// ---- bounce stages -------------------------------------------------------------
// closest hit of every queued path (persistent, dynamic fetch)
template <bool COUNT>
__global__ __launch_bounds__(WF_BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_wf_trace(DevScene S, const float4* paths, const uint32_t* count,
                                                       float4* hits, uint32_t* next, unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * WF_BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t n = *count;
    auto get = [&](uint32_t i, v3& o, v3& d, float& tmin, float& tmax) {
        float4 a = paths[3 * (size_t)i], b = paths[3 * (size_t)i + 1];
        if (__float_as_uint(a.w) == 0xFFFFFFFFu) {  // reserved but empty slot
            hits[i] = make_float4(-1.0f, 0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu));
            return false;
        }
        d = ld3(b);
        o = add(ld3(a), scl(d, 1e-2f));
        tmin = 0.001f;
        tmax = 1000.0f;
        ++c.refl;
        return true;
    };
    auto put = [&](uint32_t i, bool hit, const HitD& h) {
        hits[i] = hit ? make_float4(h.t, h.u, h.v, __uint_as_float(h.tri))
                      : make_float4(-1.0f, 0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu));
    };
    consume_rays<false, COUNT>(S, n, next, lds, c, get, put);
    flush_counters(c, counters, COUNT);
}

__global__ __launch_bounds__(WF_BLOCK) void k_wf_shade(DevScene S, FrameArgs F, WaveArgs W, const float4* paths,
                                                       float4* paths_next, const uint32_t* count,
                                                       uint32_t* count_next, uint32_t depth) {
    uint32_t n = *count;
    uint32_t stride = gridDim.x * WF_BLOCK;
    for (uint32_t base = blockIdx.x * WF_BLOCK; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        PathRec r;
        HitD h;
        HitGeo g;
        uint32_t seed = 0;
        bool is_hit = false, pending = false, next = false;
        uint32_t nshadow = 0;
        HitOut ho;
        if (valid) {
            r = load_path(paths, i);
            valid = r.pixel != 0xFFFFFFFFu;
        }
        if (valid) {
            float4 h4 = W.hits[i];
            h.t = h4.x;
            h.u = h4.y;
            h.v = h4.z;
            h.tri = __float_as_uint(h4.w);
            is_hit = h.tri != 0xFFFFFFFFu;
            if (!is_hit) {
                v3 sky = pow3(env_lookup(S.sky, S.sky_w, S.sky_h, normalize(r.d), PI_M), 2.2f);
                v3 prod = mul(r.thr, sky);
                W.pathcol[slot_of(W, r.pixel, r.sample)] = make_float4(prod.x, prod.y, prod.z, 0.0f);
            } else {
                uint32_t x = r.pixel % F.width, y = r.pixel / F.width;
                seed = tea(y * F.width + x, F.frame);
                v3 ro = add(r.o, scl(r.d, 1e-2f));
                g = hit_geometry(S, ro, r.d, h);
                int mode = (int)g.mat->shading_metal_rough_idx[0];
                if (mode == 0) {
                    pending = true;
                    for (uint32_t l = 0; l < F.nl; ++l) {
                        v3 L;
                        float dist;
                        if (hit_shadow_needed(F, l, g, L, dist)) ++nshadow;
                    }
                } else {
                    ho = shade_hit(S, F, r.d, g, seed, r.sample, depth, 0u);
                    v3 prod = mul(r.thr, ho.color);
                    if (ho.scatter && depth + 1 < F.max_depth) {
                        next = true;
                        r.thr = prod;
                    } else {
                        W.pathcol[slot_of(W, r.pixel, r.sample)] = make_float4(prod.x, prod.y, prod.z, 0.0f);
                    }
                }
            }
        }
        // pushes (whole wave participates)
        uint32_t pbase = wave_push(&W.counts[3 * BAND_STRIDE], pending ? 1u : 0u);
        if (pending) {
            W.pending[pbase] = i;
            W.vis[i] = 0u;
        }
        uint32_t sbase = wave_push(&W.counts[2 * BAND_STRIDE], nshadow);
        if (nshadow) {
            uint32_t k = 0;
            for (uint32_t l = 0; l < F.nl; ++l) {
                v3 L;
                float dist;
                if (!hit_shadow_needed(F, l, g, L, dist)) continue;
                Pcg gg = ray_stream(seed, PURPOSE_HIT_SHADOW, (r.sample * F.max_depth + depth) * F.nl + l);
                v3 dir = light_sample_dir(F, l, g.worldPos, L, gg, 0, 1);
                bool isDir = F.lights[l].pos_maxdist[3] < 0.0f;
                float tmax = isDir ? 1000.0f : dist + 1.0f;
                v3 o = add(g.worldPos, scl(dir, 1e-2f));
                size_t q = (size_t)(sbase + k);
                W.srays[2 * q] = make_float4(o.x, o.y, o.z, tmax);
                W.srays[2 * q + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float((i << 4) | l));
                ++k;
            }
        }
        uint32_t nbase = wave_push(count_next, next ? 1u : 0u);
        if (next) store_path(paths_next, nbase, g.worldPos, ho.dir, r.thr, r.pixel, r.sample, depth + 1);
    }
}

// any hit of every hit shadow ray; unoccluded rays set their light's bit
template <bool COUNT>
__global__ __launch_bounds__(WF_BLOCK) __attribute__((amdgpu_waves_per_eu(7))) void k_wf_shadow(DevScene S, WaveArgs W, uint32_t* next,
                                                        unsigned long long* counters) {
    __shared__ uint32_t stk[STACK_LDS * WF_BLOCK];
    uint32_t* lds = stk + threadIdx.x;
    Cnt c = {};
    uint32_t n = W.counts[2 * BAND_STRIDE];
    auto get = [&](uint32_t i, v3& o, v3& d, float& tmin, float& tmax) {
        float4 a = W.srays[2 * (size_t)i], b = W.srays[2 * (size_t)i + 1];
        o = ld3(a);
        d = ld3(b);
        tmin = 0.001f;
        tmax = a.w;
        ++c.hitshadow;
        return true;
    };
    auto put = [&](uint32_t i, bool hit, const HitD&) {
        if (hit) {
            ++c.occluded;
        } else {
            uint32_t id = __float_as_uint(W.srays[2 * (size_t)i + 1].w);
            atomicOr(&W.vis[id >> 4], 1u << (id & 15u));
        }
    };
    consume_rays<true, COUNT>(S, n, next, lds, c, get, put);
    flush_counters(c, counters, COUNT);
}

__global__ __launch_bounds__(WF_BLOCK) void k_wf_finish(DevScene S, FrameArgs F, WaveArgs W, const float4* paths,
                                                        uint32_t depth) {
    uint32_t n = W.counts[3 * BAND_STRIDE];
    uint32_t stride = gridDim.x * WF_BLOCK;
    for (uint32_t k = blockIdx.x * WF_BLOCK + threadIdx.x; k < n; k += stride) {
        uint32_t i = W.pending[k];
        PathRec r = load_path(paths, i);
        float4 h4 = W.hits[i];
        HitD h;
        h.t = h4.x;
        h.u = h4.y;
        h.v = h4.z;
        h.tri = __float_as_uint(h4.w);
        uint32_t x = r.pixel % F.width, y = r.pixel / F.width;
        uint32_t seed = tea(y * F.width + x, F.frame);
        v3 ro = add(r.o, scl(r.d, 1e-2f));
        HitGeo g = hit_geometry(S, ro, r.d, h);
        HitOut ho = shade_hit(S, F, r.d, g, seed, r.sample, depth, W.vis[i]);
        v3 prod = mul(r.thr, ho.color);
        W.pathcol[slot_of(W, r.pixel, r.sample)] = make_float4(prod.x, prod.y, prod.z, 0.0f);
    }
}

__global__ __launch_bounds__(WF_BLOCK) void k_wf_resolve(FrameArgs F, GBufView G, WaveArgs W, float4* lit,
                                                         float4* refl) {
    uint32_t p = W.p0 + blockIdx.x * WF_BLOCK + threadIdx.x;
    if (p >= W.p1) return;
    float4 b = W.pix[2 * (size_t)(p - W.p0) + 1];
    uint32_t nspp = (uint32_t)b.w;
    if (nspp == 0) return;  // written by k_wf_direct
    float4 a = W.pix[2 * (size_t)(p - W.p0)];
    v3 acc = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t s = 0; s < nspp; ++s) acc = add(acc, ld3(W.pathcol[slot_of(W, p, s)]));
    v3 rfac = nspp == 1 ? acc : divs(acc, (float)nspp);
    v3 rayColor = mul(ld3(a), rfac);
    v3 fc = add(add(rayColor, ld3(b)), ld3(G.emi[p]));
    put_lit(lit, p, fc, F.unorm8);
    refl[p] = make_float4(rfac.x, rfac.y, rfac.z, 1.0f);
}

// queue counter i (word i * BAND_STRIDE) for the bits i of mask; with
// `takes`, also the band take counters of both trace stages
__global__ void k_wf_reset(uint32_t* counts, uint32_t mask, uint32_t* takes) {
    uint32_t i = threadIdx.x;
    if (i < 4 && ((mask >> i) & 1u)) counts[i * BAND_STRIDE] = 0u;
    if (takes && i < 2 * SHADOW_QUEUES) takes[i * BAND_STRIDE] = 0u;
}

constexpr uint32_t STAGE_BLOCKS = 2048;  // 8 workgroups per CU, each striding over the queue

}  // namespace

size_t wave_bytes_per_slot(uint32_t nl) {
    // 2 x path (48 B) + hit (16) + vis + pending (8) + shadow rays nl x 32 + result (16)
    return 96 + 16 + 8 + 32 * (size_t)nl + 16;
}

hipError_t launch_shade(const DevScene& s, const FrameArgs& f, GBufView gb, const float* shadow, float4* lit,
                        float4* refl, unsigned long long* counters, bool count, hipStream_t st, WaveArgs w) {
    uint32_t npix = f.width * f.height;
    uint32_t slab = w.slab_pixels;
    for (uint32_t p0 = 0; p0 < npix; p0 += slab) {
        w.p0 = p0;
        w.p1 = min(npix, p0 + slab);
        k_wf_reset<<<1, 64, 0, st>>>(w.counts, 0xFu, nullptr);
        uint32_t nb = (w.p1 - w.p0 + WF_BLOCK - 1) / WF_BLOCK;
        k_wf_direct<<<nb, WF_BLOCK, 0, st>>>(s, f, gb, shadow, lit, refl, w, counters);
        for (uint32_t d = 0; d < f.max_depth; ++d) {
            float4* cur = (d & 1) ? w.paths1 : w.paths0;
            float4* nxt = (d & 1) ? w.paths0 : w.paths1;
            uint32_t* ccur = &w.counts[(d & 1) * BAND_STRIDE];
            uint32_t* cnxt = &w.counts[((d + 1) & 1) * BAND_STRIDE];
            uint32_t* take_trace = w.takes;
            uint32_t* take_shadow = w.takes + SHADOW_QUEUES * BAND_STRIDE;
            k_wf_reset<<<1, 64, 0, st>>>(w.counts, (1u << ((d + 1) & 1)) | 0xCu, w.takes);
            if (count) k_wf_trace<true><<<w.blocks, WF_BLOCK, 0, st>>>(s, cur, ccur, w.hits, take_trace, counters);
            else k_wf_trace<false><<<w.blocks, WF_BLOCK, 0, st>>>(s, cur, ccur, w.hits, take_trace, counters);
            k_wf_shade<<<STAGE_BLOCKS, WF_BLOCK, 0, st>>>(s, f, w, cur, nxt, ccur, cnxt, d);
            if (count) k_wf_shadow<true><<<w.blocks, WF_BLOCK, 0, st>>>(s, w, take_shadow, counters);
            else k_wf_shadow<false><<<w.blocks, WF_BLOCK, 0, st>>>(s, w, take_shadow, counters);
            k_wf_finish<<<STAGE_BLOCKS, WF_BLOCK, 0, st>>>(s, f, w, cur, d);
        }
        k_wf_resolve<<<nb, WF_BLOCK, 0, st>>>(f, gb, w, lit, refl);
    }
    return hipGetLastError();
}

}  // namespace hr
