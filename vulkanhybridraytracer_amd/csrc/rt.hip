// rt.hip - the reference's pure ray tracing mode (raygen.rgen + closesthit.rchit
// + miss.rmiss + shadow.rmiss): primary rays from the camera, a closest-hit
// bounce chain, `samples` shadow rays per light at every hit, and progressive
// accumulation over frames (raygen.rgen:69-80).
//
// One lane per pixel (16x16 pixels per workgroup, 8x8 per wave, workgroups
// renumbered per XCD as in k_gbuffer); each lane walks its samples and
// bounces with the shared traversal (closest hit for the chain, any hit for
// shadow rays). Per ray, the random numbers come from the pixel's PCG32
// stream for (purpose, index), as in the hybrid pass.
#include "trace_common.h"

namespace hr {
using namespace hrd;
using namespace tc;

namespace {

struct RtOut {
    v3 color, dir;
    bool scatter;
};

// closesthit.rchit:24-186 for one closest hit; the shadow rays of mode-0
// surfaces are traced here (closesthit.rchit:103-121).
template <bool COUNT>
__device__ RtOut shade_rt(const DevScene& S, const FrameArgs& F, v3 rd, const HitGeo& g, uint32_t seed,
                          uint32_t sample, uint32_t depth, uint32_t* lds, Cnt& c) {
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
    const uint32_t SS = F.samples;
    for (uint32_t i = 0; i < F.nl; ++i) {
        const hr_light& lt = F.lights[i];
        const bool isDir = lt.pos_maxdist[3] < 0.0f;
        v3 L = isDir ? ld3p(lt.pos_maxdist) : sub(ld3p(lt.pos_maxdist), g.worldPos);
        const float maxd = lt.pos_maxdist[3];
        const float dist = length(L);
        L = normalize(L);
        const float intensity = isDir ? 1.0f : (lt.color_intensity[3] / (dist * dist));
        const v3 H = normalize(add(V, L));
        const float NdotL = clamp01(dot(N, L));
        const float NdotH = clamp01(dot(N, H));
        float shadowFactor = 0.0f;
        if (NdotL > 0.0f) {
            Pcg gs = ray_stream(seed, PURPOSE_RT_SHADOW, (sample * F.max_depth + depth) * F.nl + i);
            for (uint32_t a = 0; a < SS; ++a) {
                bool shadowed = true;  // lights beyond their range stay shadowed
                if (dist < maxd) {
                    const v3 dir = light_sample_dir(F, i, g.worldPos, L, gs, a, SS);
                    ++c.hitshadow;
                    shadowed = occluded<COUNT>(S, add(g.worldPos, scl(dir, 1e-2f)), dir, 0.001f, dist + 1.0f, lds, c);
                    if (shadowed) ++c.occluded;
                }
                if (!shadowed) shadowFactor += 1.0f;
            }
            shadowFactor /= (float)SS;
        }
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
            v3 o = mk(org[0], org[1], org[2]), d = mk(dir[0], dir[1], dir[2]);
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
                const RtOut hc = shade_rt<COUNT>(S, F, d, g, seed, nS, depth, lds, c);
                rayColor = mul(rayColor, hc.color);
                if (!hc.scatter) break;
                o = g.worldPos;
                d = hc.dir;
            }
            finalColor = add(finalColor, rayColor);
        }
        finalColor = divs(finalColor, (float)spp);
        if (frame_signed < 0) {
            accum[p] = make_float4(finalColor.x, finalColor.y, finalColor.z, 1.0f);
        } else {
            const float a = 1.0f / (1.0f + (float)frame_signed);
            const float4 last = accum[p];
            const v3 m = add(scl(mk(last.x, last.y, last.z), 1.0f - a), scl(finalColor, a));
            accum[p] = make_float4(m.x, m.y, m.z, 1.0f);
        }
    }
    flush_counters(c, counters, COUNT);
}

}  // namespace

hipError_t launch_rt(const DevScene& s, const FrameArgs& f, float4* accum, int32_t frame_signed,
                     unsigned long long* counters, bool count, hipStream_t st) {
    const uint32_t nb = ((f.width + 15u) / 16u) * ((f.height + 15u) / 16u);
    if (count) k_rt<true><<<nb, BLOCK, 0, st>>>(s, f, accum, frame_signed, counters);
    else k_rt<false><<<nb, BLOCK, 0, st>>>(s, f, accum, frame_signed, counters);
    return hipGetLastError();
}

}  // namespace hr
