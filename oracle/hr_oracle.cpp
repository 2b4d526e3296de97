/*
 * hr_oracle.cpp - TEST INFRASTRUCTURE ONLY: the scalar CPU oracle.
 *
 * A restatement, in plain C++17, of VulkanHybridRaytracer's hybrid pass:
 *   random.glsl:3-33        tea, lcg, rnd
 *   helpers.glsl:33-93      GGX / Smith / Fresnel / sampleDisk
 *   raygen.rgen:29-40       primary-ray convention (G-buffer from visibility)
 *   shadowRaygen.rgen:28-93 shadow pass, shadow.rmiss / shadowHit.rchit
 *   hybridRaygen.rgen:27-186 deferred shading and bounce driver
 *   hybridHit.rchit:24-183  closest-hit shading with hit shadow rays
 *   miss.rmiss:13-18        environment miss
 * plus the acceleration structure that replaces the driver's BLAS/TLAS
 * (renderer.cpp:1643-1916): a Karras (HPG 2012) LBVH over world-space
 * triangles with 30-bit Morton codes, stable key order, bottom-up boxes, one
 * two passes of treelet restructuring (Karras and Aila, HPG 2013) and a
 * two-child-box node layout, and a stack traversal with the watertight
 * ray/triangle test of Woop, Benthin and Wald (JCGT 2013).
 *
 * It is the checker of the GPU library and the CPU baseline of the benchmark:
 * the library never links or calls it. It must be compiled with
 * -ffp-contract=off and without fast-math: every float expression below is
 * evaluated in the order written, which is the order the GPU kernels use, so
 * integer results, hit/miss, triangle ids, t and barycentrics agree bit for
 * bit. Transcendental functions the shaders call (pow) come from libm and
 * differ from the GPU's in the last bits; atan2/asin (texel selection) and
 * sin/cos of 2*pi*u (sample directions) are evaluated by the Taylor forms of
 * hr_fmath below so that texel choice and sample directions are exact too.
 *
 * Decisions where the reference is ambiguous or degenerate (documented per
 * function, shared with the GPU library):
 *  - random numbers: PCG32 (O'Neill 2014) streams seeded by the reference's
 *    tea(y*W + x, frame); stream id = (purpose << 32) | index, so each ray's
 *    samples are a pure function of (pixel, purpose, index).
 *  - directional lights (pos.w < 0) are shadow traced with tmax = 1000 (the
 *    reference never traces them, shadowRaygen.rgen:70-83).
 *  - background pixels (position.w == 0 && normal.w == 0) trace nothing:
 *    shadow = 0 and lit = albedo^2.2 (hybridRaygen.rgen:182-183).
 *  - the bounce factor is accumulated separately and applied once:
 *    lit = direct * refl + ambient + emissive.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

/* ------------------------------------------------------------------------ */
/* vector algebra, evaluated left to right                                   */
/* ------------------------------------------------------------------------ */
struct v3 {
    float x, y, z;
};
static inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize(v3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return scl(a, inv);
}
static inline float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
/* ordered selection, first operand on ties: box unions do not depend on a
 * min instruction's sign-of-zero convention */
static inline float minsel(float a, float b) { return (b < a) ? b : a; }
static inline float maxsel(float a, float b) { return (b > a) ? b : a; }
static inline v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
/* GLSL reflect(I, N) = I - 2 dot(N, I) N */
static inline v3 reflect(v3 I, v3 N) {
    float d2 = 2.0f * dot(N, I);
    return sub(I, scl(N, d2));
}
/* GLSL refract(I, N, eta) */
static inline v3 refract(v3 I, v3 N, float eta) {
    float ndi = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float a = eta * ndi + sqrtf(k);
    return sub(scl(I, eta), scl(N, a));
}
/* GLSL mix(x, y, a) = x (1 - a) + y a */
static inline v3 mixs(v3 x, v3 y, float a) {
    float oma = 1.0f - a;
    return add(scl(x, oma), scl(y, a));
}
static inline v3 pow3(v3 a, float e) { return mk(powf(a.x, e), powf(a.y, e), powf(a.z, e)); }

/* column-major 4x4 times (x, y, z, w), sums left to right */
static inline void mul4(const float* m, float x, float y, float z, float w, float* o) {
    for (int r = 0; r < 4; ++r) o[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

/* ------------------------------------------------------------------------ */
/* hr_fmath: Taylor-series atan2 / asin / sincos(2 pi u)                     */
/* ------------------------------------------------------------------------ */
static inline float atan_small(float s) { /* |s| <= tan(pi/8); Taylor series to s^19 */
    float s2 = s * s;
    float p = -1.0f / 19.0f;
    p = p * s2 + 1.0f / 17.0f;
    p = p * s2 + -1.0f / 15.0f;
    p = p * s2 + 1.0f / 13.0f;
    p = p * s2 + -1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + -1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + -1.0f / 3.0f;
    p = p * s2 + 1.0f;
    return s * p;
}
static inline float atan01(float t) { /* t in [0, 1]: atan t = pi/4 + atan((t-1)/(t+1)) above tan(pi/8) */
    if (t > 0.41421356237309503f) {
        float s = (t - 1.0f) / (t + 1.0f);
        return 0.78539816339744831f + atan_small(s);
    }
    return atan_small(t);
}
static inline float hr_atan2(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return 0.0f;
    bool swp = ay > ax;
    float t = swp ? ax / ay : ay / ax;
    float a = atan01(t);
    if (swp) a = 1.5707963267948966f - a;
    if (x < 0.0f) a = 3.1415926535897932f - a;
    if (y < 0.0f) a = -a;
    return a;
}
static inline float hr_asin(float x) {
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return hr_atan2(x, sqrtf((1.0f - x) * (1.0f + x)));
}
/* sin and cos of 2 pi u: quadrant from floor(4u), Taylor series of
 * sin(pi/2 f) (to f^11) and cos(pi/2 f) (to f^12) on the remainder f. */
static inline void hr_sincos2pi(float u, float* so, float* co) {
    float q4 = u * 4.0f;
    float fq = floorf(q4);
    float f = q4 - fq;
    int q = ((int)fq) & 3;
    float f2 = f * f;
    float ps = -3.5988432352120853e-06f;
    ps = ps * f2 + 1.6044118478735982e-04f;
    ps = ps * f2 + -4.6817541353186881e-03f;
    ps = ps * f2 + 7.9692626246167046e-02f;
    ps = ps * f2 + -6.4596409750624625e-01f;
    ps = ps * f2 + 1.5707963267948966f;
    float s = f * ps;
    float pc = 4.7108747788181715e-07f;
    pc = pc * f2 + -2.5202042373060605e-05f;
    pc = pc * f2 + 9.1926027483942659e-04f;
    pc = pc * f2 + -2.0863480763352961e-02f;
    pc = pc * f2 + 2.5366950790104802e-01f;
    pc = pc * f2 + -1.2337005501361697f;
    pc = pc * f2 + 1.0f;
    float c = pc;
    switch (q) {
        case 0: *so = s; *co = c; break;
        case 1: *so = c; *co = -s; break;
        case 2: *so = -s; *co = -c; break;
        default: *so = -c; *co = s; break;
    }
}

/* ------------------------------------------------------------------------ */
/* random.glsl and PCG32                                                     */
/* ------------------------------------------------------------------------ */
/* random.glsl:3-17: 16 rounds of TEA, returns v0 */
static inline uint32_t tea(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (uint32_t n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
/* random.glsl:21-27 */
static inline uint32_t lcg(uint32_t& prev) {
    prev = 1664525u * prev + 1013904223u;
    return prev & 0x00FFFFFFu;
}
/* random.glsl:30-33 */
static inline float rnd(uint32_t& prev) { return (float)lcg(prev) / (float)0x01000000; }

/* PCG32 XSH-RR 64/32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r) */
struct Pcg {
    uint64_t state, inc;
};
static inline uint32_t pcg_next(Pcg& g) {
    uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
static inline Pcg pcg_init(uint64_t initstate, uint64_t initseq) {
    Pcg g;
    g.state = 0u;
    g.inc = (initseq << 1u) | 1u;
    pcg_next(g);
    g.state += initstate;
    pcg_next(g);
    return g;
}
static inline float u01(Pcg& g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }

enum : uint32_t {
    PURPOSE_SURFACE_SHADOW = 1, /* index = light * samples + sample */
    PURPOSE_REFLECT = 2,        /* index = spp sample (GGX lobe at the G-buffer surface) */
    PURPOSE_HIT_SHADOW = 3,     /* index = (sample * max_depth + depth) * nl + light */
    PURPOSE_HIT_REFLECT = 4,    /* index = sample * max_depth + depth */
    PURPOSE_ADAPTIVE_SHADOW = 5, /* index = light * 128 + s (s: sample of the nested sequence, < 64) */
    PURPOSE_RT_PIXEL = 6,        /* pure ray tracing: index = pixel sample (sub-pixel offset) */
    PURPOSE_RT_SHADOW = 7,       /* pure ray tracing: index = (sample * max_depth + depth) * nl + light */
    PURPOSE_RT_REFLECT = 8       /* pure ray tracing: index = sample * max_depth + depth */
};
static inline Pcg ray_stream(uint32_t pixel_seed, uint32_t purpose, uint32_t index) {
    return pcg_init((uint64_t)pixel_seed, ((uint64_t)purpose << 32) | (uint64_t)index);
}

/* ------------------------------------------------------------------------ */
/* helpers.glsl                                                              */
/* ------------------------------------------------------------------------ */
static const float PI_H = 3.14159265359f; /* helpers.glsl:5 */
static const float PI_M = 3.141592f;      /* miss.rmiss:11 */

/* helpers.glsl:33-44 */
static inline float DistributionGGX(v3 N, v3 H, float roughness) {
    float a = roughness * roughness;
    float a2 = a * a;
    float NdotH = fmaxf(dot(N, H), 0.0f);
    float NdotH2 = NdotH * NdotH;
    float denom = NdotH2 * (a2 - 1.0f) + 1.0f;
    denom = PI_H * denom * denom;
    return a2 / denom;
}
/* helpers.glsl:47-55 */
static inline float GeometrySchlickGGX(float NdotV, float roughness) {
    float r = roughness + 1.0f;
    float k = (r * r) / 8.0f;
    float nom = NdotV;
    float denom = NdotV * (1.0f - k) + k;
    return nom / denom;
}
/* helpers.glsl:57-65 */
static inline float GeometrySmith(v3 N, v3 V, v3 L, float roughness) {
    float NdotV = fmaxf(dot(N, V), 0.0f);
    float NdotL = fmaxf(dot(N, L), 0.0f);
    return GeometrySchlickGGX(NdotV, roughness) * GeometrySchlickGGX(NdotL, roughness);
}
/* helpers.glsl:68-71 */
static inline v3 FresnelSchlick(float cosTheta, v3 F0) {
    float p = powf(1.0f - cosTheta, 5.0f);
    return mk(F0.x + (1.0f - F0.x) * p, F0.y + (1.0f - F0.y) * p, F0.z + (1.0f - F0.z) * p);
}
/* helpers.glsl:82-93 with the two rnd() draws supplied as (r1, r2) */
static inline v3 sampleDisk(const hr_light& light, v3 position, v3 L, float r1, float r2) {
    float radius = light.radius * sqrtf(r1);
    float sn, cs;
    hr_sincos2pi(r2, &sn, &cs);
    float px = radius * cs, py = radius * sn;
    v3 tangent = normalize(cross(L, mk(0.0f, 1.0f, 0.0f)));
    if (L.x == 0.0f && L.y == 1.0f && L.z == 0.0f) tangent = mk(0.0f, 0.0f, 1.0f);
    v3 bitangent = normalize(cross(tangent, L));
    v3 target = add(add(add(position, L), scl(tangent, px)), scl(bitangent, py));
    return normalize(sub(target, position));
}
/* Planted mistakes (tests only): or_set_plant selects one deliberate error so
 * that a test can show its independent check fails on it. 0 = none. */
enum : int {
    PLANT_NONE = 0,
    PLANT_GLASS_ETA = 1,      /* refrEta = ior <-> 1 / ior swapped */
    PLANT_GLASS_NORMAL = 2,   /* refrNormal not flipped for rays leaving the surface */
    PLANT_FRESNEL_NDOTH = 3,  /* hybridRaygen.rgen:111 Fresnel of NdotH instead of HdotV */
    PLANT_GGX_ALPHA = 4,      /* sampleGGX with a = roughness instead of roughness^2 */
    PLANT_RECT_STRATA = 5,    /* sampleRect rows divided by n instead of m */
    PLANT_TIR_REFRACT = 6,    /* total internal reflection answered by refract() */
    PLANT_ADAPTIVE_STOP = 7,  /* stopping rule against N^2 instead of N^3 */
    PLANT_NESTED_ORDER = 8,   /* nested cells without the bit reversal */
    PLANT_TREELET_COST = 9    /* treelet split cost without its second part */
};
static int g_plant = PLANT_NONE;

/* stratified point on a rectangle light: cell s of an n x m grid */
static inline v3 sampleRect(v3 centre, const hr_light_shape& sh, v3 position, uint32_t s,
                            uint32_t S, float r1, float r2) {
    uint32_t n = 1;
    while ((n + 1) * (n + 1) <= S) ++n;
    uint32_t m = (S + n - 1) / n;
    float fu = ((float)(s % n) + r1) / (float)n - 0.5f;
    float fv = ((float)(s / n) + r2) / (float)(g_plant == PLANT_RECT_STRATA ? n : m) - 0.5f;
    v3 eu = mk(sh.edge_u[0], sh.edge_u[1], sh.edge_u[2]);
    v3 ev = mk(sh.edge_v[0], sh.edge_v[1], sh.edge_v[2]);
    v3 target = add(add(centre, scl(eu, fu)), scl(ev, fv));
    return normalize(sub(target, position));
}
/* random.glsl:46-53 */
static inline void createCoordSystem(v3 N, v3& T, v3& B) {
    if (fabsf(N.x) > fabsf(N.y))
        T = divs(mk(N.z, 0.0f, -N.x), sqrtf(N.x * N.x + N.z * N.z));
    else
        T = divs(mk(0.0f, -N.z, N.y), sqrtf(N.y * N.y + N.z * N.z));
    B = cross(N, T);
}
/* GGX microfacet normal (Walter et al. 2007) around N with a = roughness^2 */
static inline v3 sampleGGX(v3 N, float roughness, float r1, float r2) {
    float a = g_plant == PLANT_GGX_ALPHA ? roughness : roughness * roughness;
    float cos2 = (1.0f - r1) / (1.0f + (a * a - 1.0f) * r1);
    float ct = sqrtf(cos2);
    float st = sqrtf(fmaxf(1.0f - cos2, 0.0f));
    float sn, cs;
    hr_sincos2pi(r2, &sn, &cs);
    v3 T, B;
    createCoordSystem(N, T, B);
    return add(add(scl(T, st * cs), scl(B, st * sn)), scl(N, ct));
}

/* Glass (shading model 4) directions. Both shaders flip the normal and use
 * 1 / ior as the ratio for rays leaving the surface (N.I > 0), the normal and
 * ior for rays entering it, and reflect about N when the radicand
 * 1 + eta^2 (cos^2 - 1) is negative. */
static inline void glass_frame(v3 N, float NdotI, float ior, v3& refrNormal, float& refrEta) {
    const bool out = NdotI > 0.0f;
    refrNormal = (out != (g_plant == PLANT_GLASS_NORMAL)) ? neg(N) : N;
    refrEta = (out != (g_plant == PLANT_GLASS_ETA)) ? 1.0f / ior : ior;
}
/* hybridRaygen.rgen:131-144: from the surface seen along V (towards the eye) */
static inline v3 glass_raygen_dir(v3 N, v3 V, float ior) {
    float cosAlpha = dot(N, V);
    v3 I = neg(V);
    float NdotI = dot(N, I);
    v3 refrNormal;
    float refrEta;
    glass_frame(N, NdotI, ior, refrNormal, refrEta);
    float radicand = 1.0f + refrEta * refrEta * (cosAlpha * cosAlpha - 1.0f);
    if (radicand < 0.0f && g_plant != PLANT_TIR_REFRACT) return reflect(I, N);
    return refract(I, refrNormal, refrEta);
}
/* hybridHit.rchit:158-173: at a hit reached along rd (gl_WorldRayDirectionEXT);
 * the reflection takes rd as it is, the refraction its normalisation */
static inline v3 glass_hit_dir(v3 N, v3 rd, float ior) {
    v3 rdn = normalize(rd);
    float ndv = dot(N, rdn);
    v3 refrNormal;
    float refrEta;
    glass_frame(N, ndv, ior, refrNormal, refrEta);
    float radicand = 1.0f + refrEta * refrEta * (ndv * ndv - 1.0f);
    if (radicand < 0.0f && g_plant != PLANT_TIR_REFRACT) return reflect(rd, N);
    return refract(rdn, refrNormal, refrEta);
}

/* ------------------------------------------------------------------------ */
/* textures                                                                  */
/* ------------------------------------------------------------------------ */
struct Tex {
    std::vector<float> px;
    uint32_t w = 0, h = 0;
};
static inline int wrap_index(float f, uint32_t n) {
    int i = (f == f && fabsf(f) < 1.0e9f) ? (int)f : 0;
    i %= (int)n;
    if (i < 0) i += (int)n;
    return i;
}
/* NEAREST + REPEAT lookup */
static inline v3 tex_nearest(const Tex& t, float u, float v) {
    if (t.w == 0 || t.h == 0) return mk(0.0f, 0.0f, 0.0f);
    int x = wrap_index(floorf(u * (float)t.w), t.w);
    int y = wrap_index(floorf(v * (float)t.h), t.h);
    const float* p = &t.px[((size_t)y * t.w + (size_t)x) * 4];
    return mk(p[0], p[1], p[2]);
}
/* equirect mapping of miss.rmiss:16 / hybridRaygen.rgen:59 with that site's pi */
static inline v3 env_lookup(const Tex& t, v3 d, float pi) {
    float u = 0.5f + hr_atan2(d.x, d.z) / (2.0f * pi);
    float v = 0.5f - hr_asin(d.y) / pi;
    return tex_nearest(t, u, v);
}

/* ------------------------------------------------------------------------ */
/* watertight ray / triangle (Woop, Benthin, Wald, JCGT 2013)                */
/* ------------------------------------------------------------------------ */
struct RayP {
    v3 o, d;
    float tmin, tmax;
    v3 idir, ood;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
static inline float safe_inv(float d) { return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d); }
static inline RayP prep_ray(v3 o, v3 d, float tmin, float tmax) {
    RayP r;
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    r.tmax = tmax;
    r.idir = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    r.ood = mul(o, r.idir);
    int kz = 0;
    if (fabsf(d.y) > fabsf(comp(d, kz))) kz = 1;
    if (fabsf(d.z) > fabsf(comp(d, kz))) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) std::swap(kx, ky);
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dz = comp(d, kz);
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}
static inline bool intersect_tri(const RayP& r, v3 v0, v3 v1, v3 v2, float tmin, float tmax,
                                 float& t_out, float& b1_out, float& b2_out) {
    v3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.Sx * Akz;
    float Ay = comp(A, r.ky) - r.Sy * Akz;
    float Bx = comp(B, r.kx) - r.Sx * Bkz;
    float By = comp(B, r.ky) - r.Sy * Bkz;
    float Cx = comp(C, r.kx) - r.Sx * Ckz;
    float Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = (U + V) + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = (U * Az + V * Bz) + W * Cz;
    float inv = 1.0f / det;
    float t = T * inv;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    b1_out = V * inv;
    b2_out = W * inv;
    return true;
}

/* ------------------------------------------------------------------------ */
/* scene + LBVH                                                              */
/* ------------------------------------------------------------------------ */
struct TNode { /* 64 bytes: child boxes and codes */
    float lo0[3];
    uint32_t c0;
    float hi0[3];
    uint32_t c1;
    float lo1[3];
    uint32_t pad0;
    float hi1[3];
    uint32_t pad1;
};
struct STri { /* 48 bytes: world-space vertices in sorted order + triangle id */
    float v0[3];
    uint32_t id;
    float v1[3];
    uint32_t pad0;
    float v2[3];
    uint32_t pad1;
};
static_assert(sizeof(TNode) == 64, "node");
static_assert(sizeof(STri) == 48, "tri");

static const uint32_t LEAF_BIT = 0x80000000u;
static inline uint32_t leaf_code(uint32_t first, uint32_t count) {
    return LEAF_BIT | ((count - 1u) << 27) | first;
}

struct Prim {
    uint32_t mesh, first_index, index_count;
    int32_t material;
    float model[16];
    float nm[9]; /* row-major transpose(inverse(mat3(model))) */
    uint32_t tri_offset, tri_count;
};

} // namespace

struct or_scene {
    std::vector<hr_vertex> verts;
    std::vector<uint32_t> vbase;
    std::vector<uint32_t> idx;
    std::vector<uint32_t> ibase;
    std::vector<Prim> prims;
    std::vector<hr_material> mats;
    uint32_t ntri = 0;
    std::vector<uint32_t> tri_prim;
    std::vector<float> wtri;    /* ntri x 9, primitive order */
    std::vector<float> tbox;    /* ntri x 6 (lo, hi), primitive order */
    std::vector<uint32_t> keys; /* sorted Morton codes */
    std::vector<uint32_t> ids;  /* sorted triangle ids */
    std::vector<uint32_t> kl, kr, kfirst, klast, kparent; /* Karras internal nodes */
    std::vector<float> kbox;    /* (ntri-1) x 6 */
    std::vector<TNode> nodes;
    std::vector<STri> stris;
    uint32_t leaf_max = 4;
    uint32_t depth = 0;
    Tex sky, irr;
    std::vector<uint32_t> texels;      /* material textures, RGBA8, red in the low byte */
    std::vector<uint64_t> tex_offset;
    std::vector<uint32_t> tex_w, tex_h;
};

namespace {

/* transpose(inverse(A)) of the upper 3x3 of a column-major model matrix, as
 * cofactors / determinant, row-major out (hybridHit.rchit:50) */
static void normal_matrix(const float* m, float* nm) {
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

static inline uint32_t expand_bits(uint32_t v) { /* 10 bits -> every third bit */
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
static inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand_bits(x) << 2) | (expand_bits(y) << 1) | expand_bits(z);
}
static inline uint32_t quant10(float c, float lo, float scale) {
    float q = floorf((c - lo) * scale);
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}

/* world-space triangles and their boxes (vk_mesh.cpp:436-442: position at
 * offset 0 of the 44-byte vertex, uint32 indices from firstIndex) */
static void transform_tris(or_scene& s) {
    for (const Prim& p : s.prims) {
        for (uint32_t t = 0; t < p.tri_count; ++t) {
            uint32_t g = p.tri_offset + t;
            float* w = &s.wtri[(size_t)g * 9];
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = s.idx[s.ibase[p.mesh] + p.first_index + 3 * t + k];
                const hr_vertex& v = s.verts[s.vbase[p.mesh] + vi];
                float o[4];
                mul4(p.model, v.position[0], v.position[1], v.position[2], 1.0f, o);
                w[3 * k + 0] = o[0];
                w[3 * k + 1] = o[1];
                w[3 * k + 2] = o[2];
            }
            float* b = &s.tbox[(size_t)g * 6];
            for (int a = 0; a < 3; ++a) {
                b[a] = minsel(minsel(w[a], w[3 + a]), w[6 + a]);
                b[3 + a] = maxsel(maxsel(w[a], w[3 + a]), w[6 + a]);
            }
        }
    }
}

/* Karras 2012, Fig. 4: delta = common prefix of (key, index) pairs */
static inline int kdelta(const std::vector<uint32_t>& k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
    return __builtin_clz(a ^ b);
}
static void karras(or_scene& s) {
    int n = (int)s.ntri;
    const std::vector<uint32_t>& k = s.keys;
    s.kl.assign(n - 1, 0);
    s.kr.assign(n - 1, 0);
    s.kfirst.assign(n - 1, 0);
    s.klast.assign(n - 1, 0);
    s.kparent.assign(n - 1, 0xFFFFFFFFu);
    std::vector<uint32_t> lparent(n, 0xFFFFFFFFu);
    for (int i = 0; i < n - 1; ++i) {
        int d = (kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = kdelta(k, n, i, i - d);
        int lmax = 2;
        while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = kdelta(k, n, i, j);
        int sp = 0;
        int div = 2;
        for (;;) {
            int t = (l + div - 1) / div;
            if (kdelta(k, n, i, i + (sp + t) * d) > dnode) sp += t;
            if (t <= 1) break;
            div *= 2;
        }
        int gamma = i + sp * d + std::min(d, 0);
        int lo = std::min(i, j), hi = std::max(i, j);
        uint32_t left = (lo == gamma) ? (LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
        uint32_t right = (hi == gamma + 1) ? (LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        s.kl[i] = left;
        s.kr[i] = right;
        s.kfirst[i] = (uint32_t)lo;
        s.klast[i] = (uint32_t)hi;
        if (left & LEAF_BIT) lparent[left & ~LEAF_BIT] = i; else s.kparent[left] = i;
        if (right & LEAF_BIT) lparent[right & ~LEAF_BIT] = i; else s.kparent[right] = i;
    }
}

static inline void sorted_leaf_box(const or_scene& s, uint32_t k, float* b) {
    const float* tb = &s.tbox[(size_t)s.ids[k] * 6];
    for (int a = 0; a < 6; ++a) b[a] = tb[a];
}
static void child_box(const or_scene& s, uint32_t c, float* b) {
    if (c & LEAF_BIT) sorted_leaf_box(s, c & ~LEAF_BIT, b);
    else for (int a = 0; a < 6; ++a) b[a] = s.kbox[(size_t)c * 6 + a];
}
/* bottom-up boxes: union of the children (exact, order independent) */
static void refit_boxes(or_scene& s) {
    int n = (int)s.ntri;
    s.kbox.assign((size_t)(n - 1) * 6, 0.0f);
    /* post-order over the Karras tree */
    std::vector<std::pair<uint32_t, int>> st;
    st.push_back({0u, 0});
    while (!st.empty()) {
        auto& top = st.back();
        uint32_t i = top.first;
        if (top.second == 0) {
            top.second = 1;
            if (!(s.kl[i] & LEAF_BIT)) st.push_back({s.kl[i], 0});
            if (!(s.kr[i] & LEAF_BIT)) st.push_back({s.kr[i], 0});
            continue;
        }
        st.pop_back();
        float a[6], b[6];
        child_box(s, s.kl[i], a);
        child_box(s, s.kr[i], b);
        float* o = &s.kbox[(size_t)i * 6];
        for (int q = 0; q < 3; ++q) {
            o[q] = minsel(a[q], b[q]);
            o[3 + q] = maxsel(a[3 + q], b[3 + q]);
        }
    }
}
/* ---- treelet restructuring (Karras and Aila, HPG 2013) ---------------------
 * TREELET_PASSES bottom-up passes, the first over the Karras tree: every
 * internal node N, after the nodes below it, roots a treelet grown from N's two children by replacing,
 * while it has fewer than TREELET_LEAVES leaves, its internal leaf of largest
 * half box area (the first in list order on ties) by that node's two
 * children (the first in its place, the second appended). Over the subsets S
 * of the treelet leaves (bit j = leaf j), by increasing S:
 *   cost(S) = SAH_TRI * count(S) * area(S)          if count(S) <= leaf_max
 *           = SAH_NODE * area(S) + min_P cost(P) + cost(S \ P)   otherwise
 * with P over the subsets of S that hold S's lowest leaf (submasks in
 * decreasing order, the first minimum, ties to the split whose parts differ
 * least in triangle count); a single leaf costs its subtree's cost (a triangle or
 * a subtree of at most leaf_max triangles: SAH_TRI * count * area). The
 * treelet is rebuilt in its minimum-cost topology: N keeps the root, the other
 * internal nodes are the removed ones in removal order, assigned in pre-order
 * (node, first part, second part). area = dx*dy + dy*dz + dz*dx of the box,
 * boxes are unions in increasing leaf order. Then leaves are renumbered in
 * depth-first order (first child first), so every subtree again covers a
 * contiguous range of the sorted triangles; keys keep the Morton order. */
constexpr int TREELET_LEAVES = 7, TREELET_PASSES = 2;
constexpr float SAH_NODE = 1.2f, SAH_TRI = 1.0f;
static int g_treelets = 1;

static inline float half_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

static void restructure_treelet(or_scene& s, uint32_t N, std::vector<float>& cost, std::vector<uint32_t>& cnt) {
    const uint32_t leaf_max = s.leaf_max;
    uint32_t tl[TREELET_LEAVES], freed[TREELET_LEAVES - 2];
    int m = 2, nf = 0;
    tl[0] = s.kl[N];
    tl[1] = s.kr[N];
    while (m < TREELET_LEAVES) {
        int bi = -1;
        float ba = -1.0f;
        for (int i = 0; i < m; ++i) {
            if (tl[i] & LEAF_BIT) continue;
            float a = half_area(&s.kbox[(size_t)tl[i] * 6]);
            if (a > ba) { ba = a; bi = i; }
        }
        if (bi < 0) break;
        uint32_t x = tl[bi];
        freed[nf++] = x;
        tl[bi] = s.kl[x];
        tl[m++] = s.kr[x];
    }
    const int ns = 1 << m;
    float lb[TREELET_LEAVES][6], C[1 << TREELET_LEAVES];
    uint32_t K[1 << TREELET_LEAVES];
    int part[1 << TREELET_LEAVES];
    for (int j = 0; j < m; ++j) {
        child_box(s, tl[j], lb[j]);
        uint32_t k = (tl[j] & LEAF_BIT) ? 1u : cnt[tl[j]];
        K[1 << j] = k;
        C[1 << j] = ((tl[j] & LEAF_BIT) || k <= leaf_max) ? SAH_TRI * (float)k * half_area(lb[j]) : cost[tl[j]];
        part[1 << j] = 0;
    }
    auto set_box = [&](int S, float* b) {
        int j = __builtin_ctz((unsigned)S);
        for (int q = 0; q < 6; ++q) b[q] = lb[j][q];
        for (++j; j < m; ++j) {
            if (!((S >> j) & 1)) continue;
            for (int q = 0; q < 3; ++q) {
                b[q] = minsel(b[q], lb[j][q]);
                b[3 + q] = maxsel(b[3 + q], lb[j][3 + q]);
            }
        }
    };
    for (int S = 1; S < ns; ++S) {
        if ((S & (S - 1)) == 0) continue;
        const int low = S & -S;
        K[S] = K[low] + K[S ^ low];
        float best = INFINITY;
        int bp = 0;
        /* ties: the split whose parts differ least in triangle count */
        auto imbalance = [&](int P) -> uint32_t {
            if (P == 0) return 0xFFFFFFFFu;
            const uint32_t a = K[P], b = K[S ^ P];
            return a > b ? a - b : b - a;
        };
        for (int P = (S - 1) & S; P > 0; P = (P - 1) & S) {
            if (!(P & low)) continue;
            const float c = g_plant == PLANT_TREELET_COST ? C[P] : C[P] + C[S ^ P];
            if (c < best || (c == best && imbalance(P) < imbalance(bp))) { best = c; bp = P; }
        }
        float b[6];
        set_box(S, b);
        const float A = half_area(b);
        C[S] = K[S] <= leaf_max ? SAH_TRI * (float)K[S] * A : SAH_NODE * A + best;
        part[S] = bp;
    }
    /* pre-order list of the internal subsets, then their nodes */
    int order[TREELET_LEAVES - 1], no = 0, stk[TREELET_LEAVES], sp = 0;
    stk[sp++] = ns - 1;
    while (sp) {
        const int S = stk[--sp];
        order[no++] = S;
        const int P = part[S] ? part[S] : (S & -S), Q = S ^ P;
        if (Q & (Q - 1)) stk[sp++] = Q;
        if (P & (P - 1)) stk[sp++] = P;
    }
    auto code_of = [&](int S) -> uint32_t {
        if ((S & (S - 1)) == 0) return tl[__builtin_ctz((unsigned)S)];
        for (int t = 0; t < no; ++t)
            if (order[t] == S) return t == 0 ? N : freed[t - 1];
        return 0xFFFFFFFFu;
    };
    for (int t = 0; t < no; ++t) {
        const int S = order[t];
        const int P = part[S] ? part[S] : (S & -S), Q = S ^ P;
        const uint32_t id = t == 0 ? N : freed[t - 1];
        const uint32_t l = code_of(P), r = code_of(Q);
        s.kl[id] = l;
        s.kr[id] = r;
        set_box(S, &s.kbox[(size_t)id * 6]);
        if (!(l & LEAF_BIT)) s.kparent[l] = id;
        if (!(r & LEAF_BIT)) s.kparent[r] = id;
        cost[id] = C[S];
        cnt[id] = K[S];
    }
}

static void restructure(or_scene& s) {
    const uint32_t n = s.ntri;
    if (n < 3) return;
    std::vector<float> cost(n - 1, 0.0f);
    std::vector<uint32_t> cnt(n - 1, 0u);
    for (int pass = 0; pass < TREELET_PASSES; ++pass) {
        /* post-order: every node after the nodes below it */
        std::vector<std::pair<uint32_t, int>> st;
        st.push_back({0u, 0});
        while (!st.empty()) {
            auto& top = st.back();
            const uint32_t N = top.first;
            if (top.second == 0) {
                top.second = 1;
                if (!(s.kl[N] & LEAF_BIT)) st.push_back({s.kl[N], 0});
                if (!(s.kr[N] & LEAF_BIT)) st.push_back({s.kr[N], 0});
                continue;
            }
            st.pop_back();
            restructure_treelet(s, N, cost, cnt);
        }
    }
    /* depth-first leaf order: new leaf positions and node ranges */
    std::vector<uint32_t> ids(n), newpos(n);
    uint32_t next = 0;
    std::vector<std::pair<uint32_t, int>> dfs;
    dfs.push_back({0u, 0});
    while (!dfs.empty()) {
        auto& top = dfs.back();
        const uint32_t c = top.first;
        if (top.second == 0) {
            top.second = 1;
            s.kfirst[c] = next;
            const uint32_t ch[2] = {s.kl[c], s.kr[c]};
            /* the second child is visited after the first's whole subtree */
            dfs.push_back({ch[1], 2});
            dfs.push_back({ch[0], 2});
            continue;
        }
        if (top.second == 2 && (c & LEAF_BIT)) {
            dfs.pop_back();
            newpos[c & ~LEAF_BIT] = next;
            ids[next++] = s.ids[c & ~LEAF_BIT];
            continue;
        }
        if (top.second == 2) {
            top.second = 0;
            continue;
        }
        dfs.pop_back();
        s.klast[c] = next - 1;
    }
    for (uint32_t i = 0; i + 1 < n; ++i) {
        if (s.kl[i] & LEAF_BIT) s.kl[i] = LEAF_BIT | newpos[s.kl[i] & ~LEAF_BIT];
        if (s.kr[i] & LEAF_BIT) s.kr[i] = LEAF_BIT | newpos[s.kr[i] & ~LEAF_BIT];
    }
    s.ids.swap(ids);
}

static void layout_nodes(or_scene& s) {
    uint32_t n = s.ntri;
    s.stris.assign(n, STri{});
    for (uint32_t k = 0; k < n; ++k) {
        const float* w = &s.wtri[(size_t)s.ids[k] * 9];
        STri& t = s.stris[k];
        memset(&t, 0, sizeof(t));
        for (int a = 0; a < 3; ++a) {
            t.v0[a] = w[a];
            t.v1[a] = w[3 + a];
            t.v2[a] = w[6 + a];
        }
        t.id = s.ids[k];
    }
    if (n == 0) {
        s.nodes.clear();
        return;
    }
    if (n == 1) {
        TNode nd;
        memset(&nd, 0, sizeof(nd));
        float b[6];
        sorted_leaf_box(s, 0, b);
        for (int a = 0; a < 3; ++a) {
            nd.lo0[a] = b[a]; nd.hi0[a] = b[3 + a];
            nd.lo1[a] = b[a]; nd.hi1[a] = b[3 + a];
        }
        nd.c0 = leaf_code(0, 1);
        nd.c1 = leaf_code(0, 1);
        s.nodes.assign(1, nd);
        return;
    }
    s.nodes.assign(n - 1, TNode{});
    for (uint32_t i = 0; i < n - 1; ++i) {
        TNode nd;
        memset(&nd, 0, sizeof(nd));
        uint32_t ch[2] = {s.kl[i], s.kr[i]};
        uint32_t code[2];
        float bx[2][6];
        for (int c = 0; c < 2; ++c) {
            child_box(s, ch[c], bx[c]);
            if (ch[c] & LEAF_BIT) {
                code[c] = leaf_code(ch[c] & ~LEAF_BIT, 1);
            } else {
                uint32_t sz = s.klast[ch[c]] - s.kfirst[ch[c]] + 1;
                code[c] = sz <= s.leaf_max ? leaf_code(s.kfirst[ch[c]], sz) : ch[c];
            }
        }
        for (int a = 0; a < 3; ++a) {
            nd.lo0[a] = bx[0][a]; nd.hi0[a] = bx[0][3 + a];
            nd.lo1[a] = bx[1][a]; nd.hi1[a] = bx[1][3 + a];
        }
        nd.c0 = code[0];
        nd.c1 = code[1];
        s.nodes[i] = nd;
    }
}
/* deepest path counted in traversal nodes: ancestors that are the root or
 * hold more than leaf_max triangles */
static void compute_depth(or_scene& s) {
    uint32_t n = s.ntri;
    if (n <= 1) {
        s.depth = n;
        return;
    }
    std::vector<uint32_t> lparent(n);
    for (uint32_t i = 0; i < n - 1; ++i) {
        if (s.kl[i] & LEAF_BIT) lparent[s.kl[i] & ~LEAF_BIT] = i;
        if (s.kr[i] & LEAF_BIT) lparent[s.kr[i] & ~LEAF_BIT] = i;
    }
    uint32_t best = 0;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t d = 0;
        uint32_t a = lparent[k];
        for (;;) {
            uint32_t sz = s.klast[a] - s.kfirst[a] + 1;
            if (a == 0 || sz > s.leaf_max) ++d;
            if (a == 0) break;
            a = s.kparent[a];
        }
        best = std::max(best, d);
    }
    s.depth = best;
}

/* the GPU library's traversal stack bound (internal.h MAX_DEPTH) */
constexpr uint32_t TRAVERSAL_DEPTH_MAX = 64;

/* treelets: restructure the Karras tree; a restructured tree deeper than
 * TRAVERSAL_DEPTH_MAX is rebuilt without (the treelet cost ignores depth) */
static void build_bvh(or_scene& s, bool treelets) {
    uint32_t n = s.ntri;
    s.keys.assign(n, 0);
    s.ids.resize(n);
    if (n > 0) {
        v3 cmin = mk(INFINITY, INFINITY, INFINITY), cmax = mk(-INFINITY, -INFINITY, -INFINITY);
        std::vector<v3> cen(n);
        for (uint32_t g = 0; g < n; ++g) {
            const float* b = &s.tbox[(size_t)g * 6];
            v3 c = mk((b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f);
            cen[g] = c;
            cmin = vmin(cmin, c);
            cmax = vmax(cmax, c);
        }
        // cubic cells, 1024 along the longest extent: a flat scene spends its
        // leading key bits on its long axes
        v3 ext = sub(cmax, cmin);
        const float em = std::max(ext.x, std::max(ext.y, ext.z));
        float sx = ext.x > 0.0f ? 1024.0f / em : 0.0f;
        float sy = ext.y > 0.0f ? 1024.0f / em : 0.0f;
        float sz = ext.z > 0.0f ? 1024.0f / em : 0.0f;
        std::vector<uint32_t> codes(n);
        for (uint32_t g = 0; g < n; ++g)
            codes[g] = morton3(quant10(cen[g].x, cmin.x, sx), quant10(cen[g].y, cmin.y, sy),
                               quant10(cen[g].z, cmin.z, sz));
        for (uint32_t g = 0; g < n; ++g) s.ids[g] = g;
        std::stable_sort(s.ids.begin(), s.ids.end(),
                         [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
        for (uint32_t k = 0; k < n; ++k) s.keys[k] = codes[s.ids[k]];
    }
    if (n >= 2) {
        karras(s);
        refit_boxes(s);
        if (treelets) restructure(s);
    }
    layout_nodes(s);
    compute_depth(s);
    if (treelets && n >= 3 && s.depth > TRAVERSAL_DEPTH_MAX) build_bvh(s, false);
}

/* ------------------------------------------------------------------------ */
/* traversal                                                                 */
/* ------------------------------------------------------------------------ */
struct Cnt {
    uint64_t primary = 0, shadow = 0, refl = 0, hitshadow = 0, occluded = 0, nodes = 0, tris = 0;
    void add(const Cnt& o) {
        primary += o.primary; shadow += o.shadow; refl += o.refl; hitshadow += o.hitshadow;
        occluded += o.occluded; nodes += o.nodes; tris += o.tris;
    }
};
struct Hit {
    float t, u, v;
    uint32_t tri;
};

static inline bool box_hit(const float* lo, const float* hi, const RayP& r, float tbest, float& tn) {
    float tx0 = fmaf(lo[0], r.idir.x, -r.ood.x), tx1 = fmaf(hi[0], r.idir.x, -r.ood.x);
    float ty0 = fmaf(lo[1], r.idir.y, -r.ood.y), ty1 = fmaf(hi[1], r.idir.y, -r.ood.y);
    float tz0 = fmaf(lo[2], r.idir.z, -r.ood.z), tz1 = fmaf(hi[2], r.idir.z, -r.ood.z);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tmin));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tbest));
    tn = n;
    return n <= f * 1.0000002384185791f;
}

/* Returns true on a hit. any_hit stops at the first hit in traversal order;
 * closest hit keeps the smallest (t, triangle id). Children are visited near
 * first (ties: first child), the far child is pushed. */
static bool traverse(const or_scene& s, const RayP& r, bool any_hit, Hit& h, Cnt& c) {
    h.t = -1.0f;
    h.u = h.v = 0.0f;
    h.tri = 0xFFFFFFFFu;
    if (s.nodes.empty()) return false;
    uint32_t stack[128];
    int sp = 0;
    uint32_t cur = 0;
    float tbest = r.tmax;
    uint32_t best = 0xFFFFFFFFu;
    float bu = 0.0f, bv = 0.0f;
    for (;;) {
        if (!(cur & LEAF_BIT)) {
            const TNode& nd = s.nodes[cur];
            ++c.nodes;
            float tn0, tn1;
            bool h0 = box_hit(nd.lo0, nd.hi0, r, tbest, tn0);
            bool h1 = box_hit(nd.lo1, nd.hi1, r, tbest, tn1);
            if (h0 && h1) {
                uint32_t nearc = nd.c0, farc = nd.c1;
                if (tn1 < tn0) std::swap(nearc, farc);
                stack[sp++] = farc;
                cur = nearc;
                continue;
            }
            if (h0) { cur = nd.c0; continue; }
            if (h1) { cur = nd.c1; continue; }
        } else {
            uint32_t first = cur & 0x07FFFFFFu;
            uint32_t cnt = ((cur >> 27) & 0xFu) + 1u;
            for (uint32_t k = first; k < first + cnt; ++k) {
                const STri& tr = s.stris[k];
                ++c.tris;
                float t, b1, b2;
                if (intersect_tri(r, mk(tr.v0[0], tr.v0[1], tr.v0[2]), mk(tr.v1[0], tr.v1[1], tr.v1[2]),
                                  mk(tr.v2[0], tr.v2[1], tr.v2[2]), r.tmin, tbest, t, b1, b2)) {
                    if (any_hit) {
                        h.t = t; h.u = b1; h.v = b2; h.tri = tr.id;
                        return true;
                    }
                    if (t < tbest || tr.id < best) {
                        tbest = t; best = tr.id; bu = b1; bv = b2;
                    }
                }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    if (best == 0xFFFFFFFFu) return false;
    h.t = tbest; h.u = bu; h.v = bv; h.tri = best;
    return true;
}
/* all triangles, same acceptance rules (tests only) */
static bool brute(const or_scene& s, const RayP& r, bool any_hit, Hit& h) {
    h.t = -1.0f; h.u = h.v = 0.0f; h.tri = 0xFFFFFFFFu;
    float tbest = r.tmax;
    uint32_t best = 0xFFFFFFFFu;
    float bu = 0, bv = 0;
    for (uint32_t g = 0; g < s.ntri; ++g) {
        const float* w = &s.wtri[(size_t)g * 9];
        float t, b1, b2;
        if (intersect_tri(r, mk(w[0], w[1], w[2]), mk(w[3], w[4], w[5]), mk(w[6], w[7], w[8]), r.tmin,
                          tbest, t, b1, b2)) {
            if (any_hit) { h.t = t; h.u = b1; h.v = b2; h.tri = g; return true; }
            if (t < tbest || g < best) { tbest = t; best = g; bu = b1; bv = b2; }
        }
    }
    if (best == 0xFFFFFFFFu) return false;
    h.t = tbest; h.u = bu; h.v = bv; h.tri = best;
    return true;
}
static inline bool occluded(const or_scene& s, v3 o, v3 d, float tmin, float tmax, Cnt& c) {
    RayP r = prep_ray(o, d, tmin, tmax);
    Hit h;
    return traverse(s, r, true, h, c);
}

/* ------------------------------------------------------------------------ */
/* passes                                                                    */
/* ------------------------------------------------------------------------ */
template <class F>
static void parallel_rows(uint32_t H, uint32_t row_start, uint32_t row_step, uint32_t threads, F fn,
                          Cnt& total) {
    std::vector<uint32_t> rows;
    for (uint32_t y = row_start; y < H; y += (row_step ? row_step : 1)) rows.push_back(y);
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    threads = std::min<uint32_t>(threads, std::max<size_t>(rows.size(), 1));
    std::atomic<size_t> next{0};
    std::vector<Cnt> cnts(threads);
    auto work = [&](uint32_t tid) {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= rows.size()) break;
            fn(rows[i], cnts[tid]);
        }
    };
    if (threads == 1) {
        work(0);
    } else {
        std::vector<std::thread> ts;
        for (uint32_t t = 0; t < threads; ++t) ts.emplace_back(work, t);
        for (auto& t : ts) t.join();
    }
    for (auto& c : cnts) total.add(c);
}
static void to_counters(const Cnt& c, or_counters* o) {
    if (!o) return;
    o->primary_rays += c.primary;
    o->shadow_rays += c.shadow;
    o->reflection_rays += c.refl;
    o->hit_shadow_rays += c.hitshadow;
    o->occluded_rays += c.occluded;
    o->node_visits += c.nodes;
    o->tri_tests += c.tris;
}

struct Frame {
    const hr_light* lights;
    const hr_light_shape* shapes;
    uint32_t nl;
    uint32_t samples, spp, max_depth;
    uint32_t W, H;
    uint32_t frame;
    bool adaptive;
};
/* Adaptive sampling: sample s of a light's nested stratified sequence of S (a
 * power of two <= 64) cells. Column cx of nx = 2^ceil(lg S / 2) from the bits
 * of s at even positions, reversed; row cy of ny = S / nx from the bits at odd
 * positions, reversed: every prefix of 4^j samples holds one sample per cell
 * of the 2^j x 2^j coarsening. */
static inline void nested_cell(uint32_t s, uint32_t S, uint32_t& cx, uint32_t& nx, uint32_t& cy, uint32_t& ny) {
    uint32_t lg = 0;
    while ((2u << lg) <= S) ++lg;
    const uint32_t a = (lg + 1u) / 2u, b = lg / 2u;
    cx = 0;
    cy = 0;
    const bool rev = g_plant != PLANT_NESTED_ORDER;
    for (uint32_t k = 0; k < a; ++k) cx |= ((s >> (2u * k)) & 1u) << (rev ? a - 1u - k : k);
    for (uint32_t k = 0; k < b; ++k) cy |= ((s >> (2u * k + 1u)) & 1u) << (rev ? b - 1u - k : k);
    nx = 1u << a;
    ny = 1u << b;
}
static inline v3 light_sample_dir_nested(const Frame& f, uint32_t i, v3 P, v3 L, Pcg& g, uint32_t s, uint32_t S) {
    const hr_light& lt = f.lights[i];
    float r1 = u01(g);
    float r2 = u01(g);
    uint32_t cx, nx, cy, ny;
    nested_cell(s, S, cx, nx, cy, ny);
    float a = ((float)cx + r1) / (float)nx;
    float b = ((float)cy + r2) / (float)ny;
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    if (!isDir && f.shapes && f.shapes[i].type == HR_LIGHT_RECT) {
        const hr_light_shape& sh = f.shapes[i];
        v3 eu = mk(sh.edge_u[0], sh.edge_u[1], sh.edge_u[2]);
        v3 ev = mk(sh.edge_v[0], sh.edge_v[1], sh.edge_v[2]);
        v3 target = add(add(mk(lt.pos_maxdist[0], lt.pos_maxdist[1], lt.pos_maxdist[2]), scl(eu, a - 0.5f)),
                        scl(ev, b - 0.5f));
        return normalize(sub(target, P));
    }
    return sampleDisk(lt, P, L, a, b);
}
/* Stages: 4 probes (all S when S < 4), then 16, then doubling up to S; a
 * stage is taken while the estimate k / N of the visible fraction has a
 * binomial variance estimate k (N - k) / N^3 above that of S samples at one
 * half, 1 / (4 S) (in integers). */
static inline uint32_t adaptive_first(uint32_t S) { return S < 4u ? S : 4u; }
static inline uint32_t adaptive_next(uint32_t N, uint32_t S) { return std::min(N == 4u ? 16u : 2u * N, S); }
static inline bool adaptive_stop(uint32_t k, uint32_t N, uint32_t S) {
    const uint64_t rhs = g_plant == PLANT_ADAPTIVE_STOP ? (uint64_t)N * N : (uint64_t)N * N * N;
    return N >= S || 4ull * S * k * (N - k) <= rhs;
}
static inline bool hard_light(const Frame& f, uint32_t i) {
    const hr_light& lt = f.lights[i];
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    bool rect = !isDir && f.shapes && f.shapes[i].type == HR_LIGHT_RECT;
    return !rect && lt.radius == 0.0f;
}
/* glossy paths per pixel (adaptive: 1 + floor((refl_spp - 1) * roughness)) */
static inline uint32_t pixel_spp(const Frame& f, int mode, float roughness) {
    if (!(mode == 3 && roughness > 0.0f)) return 1u;
    uint32_t m = std::max(f.spp, 1u);
    if (!f.adaptive) return m;
    uint32_t k = 1u + (uint32_t)floorf((float)(m - 1u) * roughness);
    return std::min(k, m);
}
static inline uint32_t frame_index(const hr_camera* cam) { return (uint32_t)(int)cam->frame[0]; }
static inline int mat_index(const or_scene& s, float w) {
    int i = (int)w;
    return (i < 0 || i >= (int)s.mats.size()) ? 0 : i;
}
static inline v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

/* Material texture `id` at (u, v): NEAREST / REPEAT sampler
 * (vk_initializers.cpp:305-316 with renderer.cpp:936, 3060), texel
 * (floor(u w) mod w, floor(v h) mod h); R8G8B8A8_UNORM read as c / 255
 * (vk_textures.cpp:24). */
static inline v3 tex_rgba8(const or_scene& s, float id, float u, float v) {
    int i = (int)id;
    uint32_t w = s.tex_w[i], h = s.tex_h[i];
    int x = wrap_index(floorf(u * (float)w), w);
    int y = wrap_index(floorf(v * (float)h), h);
    uint32_t c = s.texels[s.tex_offset[i] + (size_t)y * w + (size_t)x];
    return mk((float)(c & 0xFFu) / 255.0f, (float)((c >> 8) & 0xFFu) / 255.0f, (float)((c >> 16) & 0xFFu) / 255.0f);
}
static inline bool textured(const hr_material& m) {
    return m.textures[0] > -1.0f || m.textures[1] > -1.0f || m.textures[2] > -1.0f || m.textures[3] > -1.0f;
}
/* Surface at a closest hit (hybridHit.rchit:47-68, closesthit.rchit:51-70):
 * uv = (b0 uv0 + b1 uv1) + b2 uv2; albedo (before linearisation) = diffuse
 * texel or the material colour (no vertex colour here, unlike the raster
 * pass); emissive texel or 0; roughness / metallic = g / b of the
 * metallic-roughness texel or the material's factors. */
struct Surf {
    v3 albedo, emissive;
    float roughness, metallic;
};
static Surf hit_surface(const or_scene& s, const hr_material& m, const hr_vertex& x0, const hr_vertex& x1,
                        const hr_vertex& x2, float b0, float b1, float b2) {
    Surf f;
    f.albedo = ld3(m.diffuse_ior);
    f.emissive = mk(0.0f, 0.0f, 0.0f);
    f.roughness = m.shading_metal_rough_idx[2];
    f.metallic = m.shading_metal_rough_idx[1];
    if (!textured(m)) return f;
    float u = (x0.uv[0] * b0 + x1.uv[0] * b1) + x2.uv[0] * b2;
    float v = (x0.uv[1] * b0 + x1.uv[1] * b1) + x2.uv[1] * b2;
    if (m.textures[0] > -1.0f) f.albedo = tex_rgba8(s, m.textures[0], u, v);
    if (m.textures[2] > -1.0f) f.emissive = tex_rgba8(s, m.textures[2], u, v);
    if (m.textures[3] > -1.0f) {
        v3 rm = tex_rgba8(s, m.textures[3], u, v);
        f.roughness = rm.y;
        f.metallic = rm.z;
    }
    return f;
}

/* direction of one shadow sample towards light i (helpers.glsl:82-93 for
 * disks, stratified cell s of S for rectangles) */
static inline v3 light_sample_dir(const Frame& f, uint32_t i, v3 P, v3 L, Pcg& g, uint32_t s, uint32_t S) {
    const hr_light& lt = f.lights[i];
    float r1 = u01(g);
    float r2 = u01(g);
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    if (!isDir && f.shapes && f.shapes[i].type == HR_LIGHT_RECT)
        return sampleRect(ld3(lt.pos_maxdist), f.shapes[i], P, s, S, r1, r2);
    return sampleDisk(lt, P, L, r1, r2);
}

/* hybridHit.rchit:24-183 for one closest hit. Returns the payload colour,
 * sets dir/scatter/origin. */
struct HitOut {
    v3 color, dir, origin;
    bool scatter;
};
static HitOut shade_hit(const or_scene& s, const Frame& f, v3 ro, v3 rd, const Hit& hit, uint32_t seed,
                        uint32_t sample, uint32_t depth, Cnt& c) {
    HitOut out;
    float b1 = hit.u, b2 = hit.v;
    float b0 = 1.0f - b1 - b2;
    const Prim& p = s.prims[s.tri_prim[hit.tri]];
    uint32_t local = hit.tri - p.tri_offset;
    uint32_t ib = s.ibase[p.mesh] + p.first_index + 3 * local;
    const hr_vertex& x0 = s.verts[s.vbase[p.mesh] + s.idx[ib + 0]];
    const hr_vertex& x1 = s.verts[s.vbase[p.mesh] + s.idx[ib + 1]];
    const hr_vertex& x2 = s.verts[s.vbase[p.mesh] + s.idx[ib + 2]];
    v3 nrm = add(add(scl(ld3(x0.normal), b0), scl(ld3(x1.normal), b1)), scl(ld3(x2.normal), b2));
    const float* M = p.nm;
    v3 N = normalize(mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                        (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z));
    v3 V = normalize(neg(rd));
    float NdotV = clamp01(dot(N, V));
    v3 worldPos = add(ro, scl(rd, hit.t));

    v3 Lo = mk(0, 0, 0);
    const hr_material& mat = s.mats[p.material];
    int mode = (int)mat.shading_metal_rough_idx[0];
    const Surf sf = hit_surface(s, mat, x0, x1, x2, b0, b1, b2);
    v3 albedo = pow3(sf.albedo, 2.2f);
    float roughness = sf.roughness;
    float metallic = sf.metallic;
    v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
    v3 irradiance = env_lookup(s.irr, N, PI_H);

    v3 direction = mk(1, 1, 1);
    bool scatter = false;
    v3 rdn = normalize(rd);
    v3 reflected = mk(0, 0, 0);
    if (mode == 3) {
        v3 Nr = N;
        if (roughness > 0.0f) {
            Pcg g = ray_stream(seed, PURPOSE_HIT_REFLECT, sample * f.max_depth + depth);
            float r1 = u01(g), r2 = u01(g);
            Nr = sampleGGX(N, roughness, r1, r2);
        }
        reflected = reflect(rdn, Nr);
    }
    for (uint32_t i = 0; i < f.nl; ++i) {
        const hr_light& lt = f.lights[i];
        bool isDir = lt.pos_maxdist[3] < 0.0f;
        v3 L = isDir ? ld3(lt.pos_maxdist) : sub(ld3(lt.pos_maxdist), worldPos);
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
        v3 lc = ld3(lt.color_intensity);
        if (mode == 0) {
            if (NdotL > 0.0f) {
                bool shadowed = true;
                if (isDir || dist < maxd) {
                    Pcg g = ray_stream(seed, PURPOSE_HIT_SHADOW, (sample * f.max_depth + depth) * f.nl + i);
                    v3 dir = light_sample_dir(f, i, worldPos, L, g, 0, 1);
                    float tmax = isDir ? 1000.0f : dist + 1.0f;
                    ++c.hitshadow;
                    shadowed = occluded(s, add(worldPos, scl(dir, 1e-2f)), dir, 0.001f, tmax, c);
                    if (shadowed) ++c.occluded;
                }
                if (!shadowed) shadowFactor += 1.0f;
            }
            v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
            v3 F = FresnelSchlick(NdotH, F0);
            float D = DistributionGGX(N, H, roughness);
            float G = GeometrySmith(N, V, L, roughness);
            v3 numerator = scl(F, D * G);
            float denominator = fmaxf(4.0f * clamp01(dot(N, V)) * NdotL, 0.000001f);
            v3 specular = divs(numerator, denominator);
            v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
            v3 term = add(divs(mul(kD, albedo), PI_H), specular);
            Lo = add(Lo, scl(mul(term, radiance), NdotL));
            direction = mk(1, 1, 1);
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
            v3 add_t = (intensity > 0.0f) ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                                          : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, add_t);
            direction = glass_hit_dir(N, rd, mat.diffuse_ior[3]);
            scatter = true;
        }
    }
    v3 F = FresnelSchlick(NdotV, F0);
    v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
    v3 ambient = mul(mul(kD, albedo), irradiance);
    out.color = add(add(Lo, ambient), sf.emissive);
    out.dir = direction;
    out.scatter = scatter;
    out.origin = worldPos;
    return out;
}

/* closesthit.rchit:24-186 for one closest hit of the pure ray tracing mode:
 * `samples` shadow rays per light (closesthit.rchit:99-121; a light beyond
 * its range, and a directional light, stays shadowed), the factor is the
 * unshadowed fraction; attenuation without a directional case; the glossy
 * extension (GGX lobe) for shading model 3 with roughness > 0. */
static HitOut shade_rt(const or_scene& s, const Frame& f, v3 ro, v3 rd, const Hit& hit, uint32_t seed,
                       uint32_t sample, uint32_t depth, Cnt& c) {
    HitOut out;
    float b1 = hit.u, b2 = hit.v;
    float b0 = 1.0f - b1 - b2;
    const Prim& p = s.prims[s.tri_prim[hit.tri]];
    uint32_t local = hit.tri - p.tri_offset;
    uint32_t ib = s.ibase[p.mesh] + p.first_index + 3 * local;
    const hr_vertex& x0 = s.verts[s.vbase[p.mesh] + s.idx[ib + 0]];
    const hr_vertex& x1 = s.verts[s.vbase[p.mesh] + s.idx[ib + 1]];
    const hr_vertex& x2 = s.verts[s.vbase[p.mesh] + s.idx[ib + 2]];
    v3 nrm = add(add(scl(ld3(x0.normal), b0), scl(ld3(x1.normal), b1)), scl(ld3(x2.normal), b2));
    const float* M = p.nm;
    v3 N = normalize(mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                        (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z));
    v3 V = normalize(neg(rd));
    float NdotV = clamp01(dot(N, V));
    v3 worldPos = add(ro, scl(rd, hit.t));
    v3 Lo = mk(0, 0, 0);
    const hr_material& mat = s.mats[p.material];
    int mode = (int)mat.shading_metal_rough_idx[0];
    const Surf sf = hit_surface(s, mat, x0, x1, x2, b0, b1, b2);
    v3 albedo = pow3(sf.albedo, 2.2f);
    float roughness = sf.roughness;
    float metallic = sf.metallic;
    v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
    v3 irradiance = env_lookup(s.irr, N, PI_H);
    v3 direction = mk(1, 1, 1);
    bool scatter = false;
    v3 reflected = mk(0, 0, 0);
    if (mode == 3) {
        v3 Nr = N;
        if (roughness > 0.0f) {
            Pcg g = ray_stream(seed, PURPOSE_RT_REFLECT, sample * f.max_depth + depth);
            float r1 = u01(g), r2 = u01(g);
            Nr = sampleGGX(N, roughness, r1, r2);
        }
        reflected = reflect(normalize(rd), Nr);
    }
    uint32_t SS = f.samples;
    for (uint32_t i = 0; i < f.nl; ++i) {
        const hr_light& lt = f.lights[i];
        bool isDir = lt.pos_maxdist[3] < 0.0f;
        v3 L = isDir ? ld3(lt.pos_maxdist) : sub(ld3(lt.pos_maxdist), worldPos);
        float maxd = lt.pos_maxdist[3];
        float dist = length(L);
        L = normalize(L);
        float intensity = isDir ? 1.0f : (lt.color_intensity[3] / (dist * dist));
        v3 H = normalize(add(V, L));
        float NdotL = clamp01(dot(N, L));
        float NdotH = clamp01(dot(N, H));
        float shadowFactor = 0.0f;
        if (NdotL > 0.0f) {
            Pcg gs = ray_stream(seed, PURPOSE_RT_SHADOW, (sample * f.max_depth + depth) * f.nl + i);
            for (uint32_t a = 0; a < SS; ++a) {
                bool shadowed = true;
                if (dist < maxd) {
                    v3 dir = light_sample_dir(f, i, worldPos, L, gs, a, SS);
                    ++c.hitshadow;
                    shadowed = occluded(s, add(worldPos, scl(dir, 1e-2f)), dir, 0.001f, dist + 1.0f, c);
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
        v3 lc = ld3(lt.color_intensity);
        if (mode == 0) {
            v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
            v3 F = FresnelSchlick(NdotH, F0);
            float D = DistributionGGX(N, H, roughness);
            float G = GeometrySmith(N, V, L, roughness);
            v3 numerator = scl(F, D * G);
            float denominator = fmaxf(4.0f * clamp01(dot(N, V)) * NdotL, 0.000001f);
            v3 specular = divs(numerator, denominator);
            v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
            v3 term = add(divs(mul(kD, albedo), PI_H), specular);
            Lo = add(Lo, scl(mul(term, radiance), NdotL));
            direction = mk(1, 1, 1);
            scatter = false;
        } else if (mode == 3) {
            bool sc = dot(reflected, N) > 0.0f;
            v3 t = (NdotL > 0.0f && intensity > 0.0f)
                       ? mul(scl(scl(scl(scl(lc, intensity), attenuation), shadowFactor), metallic), albedo)
                       : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, t);
            direction = reflected;
            scatter = sc;
        } else if (mode == 4) {
            float ior = mat.diffuse_ior[3];
            float ndv = dot(N, normalize(rd));
            v3 refrNormal = ndv > 0.0f ? neg(N) : N;
            float refrEta = ndv > 0.0f ? 1.0f / ior : ior;
            v3 t = (intensity > 0.0f) ? scl(mul(scl(scl(lc, intensity), attenuation), albedo), metallic)
                                      : scl(mul(irradiance, albedo), metallic);
            Lo = add(Lo, t);
            float radicand = 1.0f + refrEta * refrEta * (ndv * ndv - 1.0f);
            direction = radicand < 0.0f ? reflect(rd, N) : refract(rd, refrNormal, refrEta);
            scatter = true;
        }
    }
    v3 F = FresnelSchlick(NdotV, F0);
    v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
    v3 ambient = mul(mul(kD, albedo), irradiance);
    out.color = add(add(Lo, ambient), sf.emissive);
    out.dir = direction;
    out.scatter = scatter;
    out.origin = worldPos;
    return out;
}

} // namespace

/* ======================================================================== */
/* C interface                                                               */
/* ======================================================================== */
extern "C" {

uint32_t or_tea(uint32_t v0, uint32_t v1) { return tea(v0, v1); }
uint32_t or_lcg(uint32_t* st) { return lcg(*st); }
float or_rnd(uint32_t* st) { return rnd(*st); }
void or_pcg32_seq(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Pcg g = pcg_init(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(g);
}
float or_u01_stream(uint32_t pixel_seed, uint32_t purpose, uint32_t index, uint32_t k) {
    Pcg g = ray_stream(pixel_seed, purpose, index);
    float u = 0.0f;
    for (uint32_t i = 0; i <= k; ++i) u = u01(g);
    return u;
}
uint32_t or_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
void or_set_treelets(int on) { g_treelets = on ? 1 : 0; }
void or_atan2(const float* y, const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = hr_atan2(y[i], x[i]);
}
void or_asin(const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = hr_asin(x[i]);
}
void or_sincos2pi(const float* u, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) hr_sincos2pi(u[i], &s[i], &c[i]);
}
float or_distribution_ggx(const float* N, const float* H, float r) { return DistributionGGX(ld3(N), ld3(H), r); }
float or_geometry_smith(const float* N, const float* V, const float* L, float r) {
    return GeometrySmith(ld3(N), ld3(V), ld3(L), r);
}
void or_fresnel_schlick(float ct, const float* F0, float* out3) {
    v3 f = FresnelSchlick(ct, ld3(F0));
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
void or_sample_disk(const hr_light* light, const float* pos, const float* L, float r1, float r2, float* d) {
    v3 v = sampleDisk(*light, ld3(pos), ld3(L), r1, r2);
    d[0] = v.x; d[1] = v.y; d[2] = v.z;
}

void or_set_plant(int plant) { g_plant = plant; }
void or_sample_ggx(const float* N, float roughness, float r1, float r2, float* out) {
    v3 d = sampleGGX(ld3(N), roughness, r1, r2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void or_sample_rect(const float* centre, const hr_light_shape* sh, const float* pos, uint32_t s, uint32_t S,
                    float r1, float r2, float* out) {
    v3 d = sampleRect(ld3(centre), *sh, ld3(pos), s, S, r1, r2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void or_glass_raygen(const float* N, const float* V, float ior, float* out) {
    v3 d = glass_raygen_dir(ld3(N), ld3(V), ior);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void or_glass_hit(const float* N, const float* rd, float ior, float* out) {
    v3 d = glass_hit_dir(ld3(N), ld3(rd), ior);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void or_nested_cell(uint32_t s, uint32_t S, uint32_t* cx, uint32_t* nx, uint32_t* cy, uint32_t* ny) {
    nested_cell(s, S, *cx, *nx, *cy, *ny);
}
int or_adaptive_stop(uint32_t k, uint32_t N, uint32_t S) { return adaptive_stop(k, N, S) ? 1 : 0; }
uint32_t or_adaptive_next(uint32_t N, uint32_t S) { return adaptive_next(N, S); }
int or_intersect_tri(const float* org, const float* dir, float tmin, float tmax, const float* v0,
                     const float* v1, const float* v2, float* t, float* u, float* v) {
    RayP r = prep_ray(ld3(org), ld3(dir), tmin, tmax);
    float tt, b1, b2;
    if (!intersect_tri(r, ld3(v0), ld3(v1), ld3(v2), tmin, tmax, tt, b1, b2)) return 0;
    *t = tt; *u = b1; *v = b2;
    return 1;
}

int or_scene_build(const hr_mesh* meshes, uint32_t mesh_count, const hr_primitive* prims, uint32_t prim_count,
                   const hr_material* materials, uint32_t material_count, uint32_t leaf_max, or_scene** out) {
    if (!out || (mesh_count && !meshes) || (prim_count && !prims) || !materials || material_count == 0)
        return HR_E_INVALID;
    if (leaf_max == 0) leaf_max = 4;
    if (leaf_max > 16) return HR_E_INVALID;
    or_scene* s = new or_scene();
    s->leaf_max = leaf_max;
    s->mats.assign(materials, materials + material_count);
    for (uint32_t m = 0; m < mesh_count; ++m) {
        s->vbase.push_back((uint32_t)s->verts.size());
        s->ibase.push_back((uint32_t)s->idx.size());
        s->verts.insert(s->verts.end(), meshes[m].vertices, meshes[m].vertices + meshes[m].vertex_count);
        s->idx.insert(s->idx.end(), meshes[m].indices, meshes[m].indices + meshes[m].index_count);
    }
    uint32_t tri = 0;
    for (uint32_t i = 0; i < prim_count; ++i) {
        const hr_primitive& hp = prims[i];
        if (hp.mesh >= mesh_count || hp.index_count % 3 || hp.first_index + hp.index_count > meshes[hp.mesh].index_count ||
            hp.material < 0 || (uint32_t)hp.material >= material_count) {
            delete s;
            return HR_E_INVALID;
        }
        const hr_mesh& mh = meshes[hp.mesh];
        for (uint32_t k = 0; k < hp.index_count; ++k) {
            if (mh.indices[hp.first_index + k] >= mh.vertex_count) {
                delete s;
                return HR_E_INVALID;
            }
        }
        Prim p;
        p.mesh = hp.mesh;
        p.first_index = hp.first_index;
        p.index_count = hp.index_count;
        p.material = hp.material;
        memcpy(p.model, hp.model, sizeof(p.model));
        normal_matrix(p.model, p.nm);
        p.tri_offset = tri;
        p.tri_count = hp.index_count / 3;
        tri += p.tri_count;
        s->prims.push_back(p);
    }
    s->ntri = tri;
    s->tri_prim.resize(tri);
    for (uint32_t i = 0; i < prim_count; ++i)
        for (uint32_t t = 0; t < s->prims[i].tri_count; ++t) s->tri_prim[s->prims[i].tri_offset + t] = i;
    s->wtri.assign((size_t)tri * 9, 0.0f);
    s->tbox.assign((size_t)tri * 6, 0.0f);
    transform_tris(*s);
    build_bvh(*s, g_treelets != 0);
    *out = s;
    return HR_OK;
}
int or_scene_refit(or_scene* s, const float* models) {
    if (!s || !models) return HR_E_INVALID;
    for (size_t i = 0; i < s->prims.size(); ++i) {
        memcpy(s->prims[i].model, models + 16 * i, 16 * sizeof(float));
        normal_matrix(s->prims[i].model, s->prims[i].nm);
    }
    transform_tris(*s);
    if (s->ntri >= 2) refit_boxes(*s);
    layout_nodes(*s);
    compute_depth(*s);
    return HR_OK;
}
int or_scene_set_textures(or_scene* s, const hr_texture* t, uint32_t n) {
    if (!s || (n && !t)) return HR_E_INVALID;
    s->texels.clear(); s->tex_offset.clear(); s->tex_w.clear(); s->tex_h.clear();
    for (uint32_t i = 0; i < n; ++i) {
        if (!t[i].rgba8 || !t[i].width || !t[i].height) return HR_E_INVALID;
        s->tex_offset.push_back(s->texels.size());
        s->tex_w.push_back(t[i].width);
        s->tex_h.push_back(t[i].height);
        size_t m = (size_t)t[i].width * t[i].height;
        for (size_t k = 0; k < m; ++k) {
            const uint8_t* c = t[i].rgba8 + 4 * k;
            s->texels.push_back((uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24));
        }
    }
    return HR_OK;
}

int or_scene_set_env(or_scene* s, const hr_env* e) {
    if (!s || !e) return HR_E_INVALID;
    s->sky.w = e->sky_width; s->sky.h = e->sky_height;
    s->sky.px.assign(e->skybox, e->skybox + (size_t)e->sky_width * e->sky_height * 4);
    s->irr.w = e->irr_width; s->irr.h = e->irr_height;
    s->irr.px.assign(e->irradiance, e->irradiance + (size_t)e->irr_width * e->irr_height * 4);
    return HR_OK;
}
void or_scene_destroy(or_scene* s) { delete s; }
uint32_t or_scene_tri_count(const or_scene* s) { return s->ntri; }
uint32_t or_scene_node_count(const or_scene* s) { return (uint32_t)s->nodes.size(); }
uint32_t or_scene_depth(const or_scene* s) { return s->depth; }
int or_scene_export(const or_scene* s, uint32_t which, void* dst, uint64_t bytes) {
    const void* src = nullptr;
    size_t n = 0;
    std::vector<uint32_t> tmp;
    switch (which) {
        case 0: src = s->nodes.data(); n = s->nodes.size() * sizeof(TNode); break;
        case 1: src = s->stris.data(); n = s->stris.size() * sizeof(STri); break;
        case 2: src = s->keys.data(); n = s->keys.size() * 4; break;
        case 3: src = s->ids.data(); n = s->ids.size() * 4; break;
        case 4:
            if (s->ntri >= 2) {
                tmp.resize((size_t)(s->ntri - 1) * 4);
                for (uint32_t i = 0; i + 1 < s->ntri; ++i) {
                    tmp[4 * i] = s->kl[i]; tmp[4 * i + 1] = s->kr[i];
                    tmp[4 * i + 2] = s->kfirst[i]; tmp[4 * i + 3] = s->klast[i];
                }
            }
            src = tmp.data(); n = tmp.size() * 4;
            break;
        default: return HR_E_INVALID;
    }
    if (bytes < n) return HR_E_INVALID;
    if (n) memcpy(dst, src, n);
    return HR_OK;
}

void or_trace_rays(const or_scene* s, const hr_ray* rays, uint32_t n, int any_hit, int use_brute, hr_hit* hits,
                   or_counters* cnt, uint32_t threads) {
    Cnt total;
    uint32_t rows = (n + 255) / 256;
    parallel_rows(rows, 0, 1, threads, [&](uint32_t blk, Cnt& c) {
        for (uint32_t i = blk * 256; i < std::min(n, blk * 256 + 256); ++i) {
            RayP r = prep_ray(ld3(rays[i].org), ld3(rays[i].dir), rays[i].tmin, rays[i].tmax);
            Hit h;
            if (use_brute) brute(*s, r, any_hit != 0, h);
            else traverse(*s, r, any_hit != 0, h, c);
            hits[i].t = h.t; hits[i].u = h.u; hits[i].v = h.v; hits[i].tri = h.tri;
        }
    }, total);
    to_counters(total, cnt);
}

/* World-space direction of the primary ray through the centre of pixel
 * (x, y) (raygen.rgen:29-40); x, y may lie outside the image. */
static v3 primary_dir(const hr_camera* cam, uint32_t W, uint32_t H, uint32_t x, uint32_t y) {
    float ux = ((float)x + 0.5f) / (float)W;
    float uy = ((float)y + 0.5f) / (float)H;
    float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    float tgt[4], dir[4];
    mul4(cam->inv_proj, dx, dy, 1.0f, 1.0f, tgt);
    v3 tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
    mul4(cam->inv_view, tn.x, tn.y, tn.z, 0.0f, dir);
    return mk(dir[0], dir[1], dir[2]);
}
/* The world position a rasteriser interpolates at pixel (x, y) for the
 * triangle's plane (v0, ng): where the ray from o through that pixel meets it. */
static v3 plane_point(const hr_camera* cam, uint32_t W, uint32_t H, v3 o, v3 v0, v3 ng, uint32_t x, uint32_t y) {
    v3 d = primary_dir(cam, W, H, x, y);
    float t = dot(sub(v0, o), ng) / dot(d, ng);
    return add(o, scl(d, t));
}
/* perturbNormal + cotangent_frame (geometry_shader.frag:28-49). dFdx(p) and
 * dFdy(p) are fine derivatives: differences across the pixel's 2x2 quad
 * (columns x & ~1 and x | 1, rows y & ~1 and y | 1) of the positions on the
 * hit triangle's plane. The shader's duv1 / duv2 are dFdx(p).xy / dFdy(p).xy
 * (it differentiates p, not uv) and are kept so. Nin: interpolated, transformed,
 * unnormalised normal; tx: normal-map texel. A zero or non-finite frame keeps
 * normalize(Nin). */
static v3 perturb_normal(const or_scene& s, const hr_camera* cam, uint32_t W, uint32_t H, v3 o, uint32_t tri,
                         v3 Nin, v3 tx, uint32_t x, uint32_t y) {
    v3 np = mk(tx.x * (255.0f / 127.0f) - 128.0f / 127.0f, tx.y * (255.0f / 127.0f) - 128.0f / 127.0f,
               tx.z * (255.0f / 127.0f) - 128.0f / 127.0f);
    const float* w = &s.wtri[(size_t)tri * 9];
    v3 v0 = mk(w[0], w[1], w[2]), v1 = mk(w[3], w[4], w[5]), v2 = mk(w[6], w[7], w[8]);
    v3 ng = cross(sub(v1, v0), sub(v2, v0));
    uint32_t x0 = x & ~1u, x1 = x | 1u, y0 = y & ~1u, y1 = y | 1u;
    v3 dp1 = sub(plane_point(cam, W, H, o, v0, ng, x1, y), plane_point(cam, W, H, o, v0, ng, x0, y));
    v3 dp2 = sub(plane_point(cam, W, H, o, v0, ng, x, y1), plane_point(cam, W, H, o, v0, ng, x, y0));
    v3 dp2perp = cross(dp2, Nin);
    v3 dp1perp = cross(Nin, dp1);
    v3 T = add(scl(dp2perp, dp1.x), scl(dp1perp, dp2.x));
    v3 B = add(scl(dp2perp, dp1.y), scl(dp1perp, dp2.y));
    float tt = dot(T, T), bb = dot(B, B);
    float m = (bb > tt) ? bb : tt;
    if (!(std::isfinite(tt) && std::isfinite(bb) && m > 0.0f)) return normalize(Nin);
    float invmax = 1.0f / sqrtf(m);
    T = scl(T, invmax);
    B = scl(B, invmax);
    return normalize(add(add(scl(T, np.x), scl(B, np.y)), scl(Nin, np.z)));
}

/* raygen.rgen:29-40 primary ray, then the G-buffer written by
 * geometry_shader.frag:51-76 (hits) and skybox.frag:16-21 (background). */
void or_gbuffer_primary(const or_scene* s, const hr_camera* cam, uint32_t W, uint32_t H, float* position,
                        float* normal, float* albedo, float* material, float* emissive, uint32_t row_start,
                        uint32_t row_step, or_counters* cnt, uint32_t threads) {
    Cnt total;
    parallel_rows(H, row_start, row_step, threads, [&](uint32_t y, Cnt& c) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t p = ((size_t)y * W + x) * 4;
            float org[4];
            mul4(cam->inv_view, 0.0f, 0.0f, 0.0f, 1.0f, org);
            v3 o = mk(org[0], org[1], org[2]), d = primary_dir(cam, W, H, x, y);
            v3 ro = add(o, scl(d, 1e-2f));
            RayP r = prep_ray(ro, d, 0.001f, 10000.0f);
            Hit h;
            ++c.primary;
            bool hit = traverse(*s, r, false, h, c);
            float* P = position + p; float* Nn = normal + p; float* A = albedo + p;
            float* Mt = material + p; float* E = emissive + p;
            for (int k = 0; k < 4; ++k) { P[k] = 0; Nn[k] = 0; A[k] = 0; Mt[k] = 0; E[k] = 0; }
            if (!hit) {
                v3 sky = env_lookup(s->sky, normalize(d), PI_M);
                A[0] = sky.x; A[1] = sky.y; A[2] = sky.z; A[3] = 1.0f;
                continue;
            }
            const Prim& pr = s->prims[s->tri_prim[h.tri]];
            uint32_t local = h.tri - pr.tri_offset;
            uint32_t ib = s->ibase[pr.mesh] + pr.first_index + 3 * local;
            const hr_vertex& x0 = s->verts[s->vbase[pr.mesh] + s->idx[ib + 0]];
            const hr_vertex& x1 = s->verts[s->vbase[pr.mesh] + s->idx[ib + 1]];
            const hr_vertex& x2 = s->verts[s->vbase[pr.mesh] + s->idx[ib + 2]];
            float b1 = h.u, b2 = h.v, b0 = 1.0f - b1 - b2;
            v3 nrm = add(add(scl(ld3(x0.normal), b0), scl(ld3(x1.normal), b1)), scl(ld3(x2.normal), b2));
            const float* M = pr.nm;
            /* inNormal of basic.vert:44, interpolated, not yet normalised */
            v3 Nin = mk((M[0] * nrm.x + M[1] * nrm.y) + M[2] * nrm.z, (M[3] * nrm.x + M[4] * nrm.y) + M[5] * nrm.z,
                        (M[6] * nrm.x + M[7] * nrm.y) + M[8] * nrm.z);
            v3 N = normalize(Nin);
            v3 wp = add(ro, scl(d, h.t));
            const hr_material& mt = s->mats[pr.material];
            v3 alb = ld3(mt.diffuse_ior);
            v3 emi = mk(0.0f, 0.0f, 0.0f);
            v3 mtl = mk(0.0f, mt.shading_metal_rough_idx[2], mt.shading_metal_rough_idx[1]);
            if (textured(mt)) { /* geometry_shader.frag:54-64 */
                float u = (x0.uv[0] * b0 + x1.uv[0] * b1) + x2.uv[0] * b2;
                float v = (x0.uv[1] * b0 + x1.uv[1] * b1) + x2.uv[1] * b2;
                if (mt.textures[0] > -1.0f) {
                    v3 col = add(add(scl(ld3(x0.color), b0), scl(ld3(x1.color), b1)), scl(ld3(x2.color), b2));
                    alb = mul(tex_rgba8(*s, mt.textures[0], u, v), col);
                }
                if (mt.textures[2] > -1.0f) emi = tex_rgba8(*s, mt.textures[2], u, v);
                if (mt.textures[3] > -1.0f) mtl = tex_rgba8(*s, mt.textures[3], u, v);
                if (mt.textures[1] > -1.0f)
                    N = perturb_normal(*s, cam, W, H, o, h.tri, Nin, tex_rgba8(*s, mt.textures[1], u, v), x, y);
            }
            P[0] = wp.x; P[1] = wp.y; P[2] = wp.z; P[3] = mt.shading_metal_rough_idx[3];
            Nn[0] = N.x; Nn[1] = N.y; Nn[2] = N.z; Nn[3] = 1.0f;
            A[0] = alb.x; A[1] = alb.y; A[2] = alb.z; A[3] = 1.0f;
            Mt[0] = mtl.x; Mt[1] = mtl.y; Mt[2] = mtl.z; Mt[3] = 0.0f;
            E[0] = emi.x; E[1] = emi.y; E[2] = emi.z; E[3] = 0.0f;
        }
    }, total);
    to_counters(total, cnt);
}

/* shadowRaygen.rgen:28-93 */
void or_trace_shadows(const or_scene* s, uint32_t W, uint32_t H, const float* position, const float* normal,
                      const hr_light* lights, const hr_light_shape* shapes, uint32_t nl, const hr_camera* cam,
                      const hr_trace_params* prm, float* shadow, uint32_t row_start, uint32_t row_step,
                      or_counters* cnt, uint32_t threads) {
    Cnt total;
    Frame f{lights, shapes, nl, prm->shadow_samples, prm->refl_spp, prm->max_depth, W, H, frame_index(cam),
            (prm->flags & HR_FLAG_ADAPTIVE) != 0};
    size_t plane = (size_t)W * H;
    parallel_rows(H, row_start, row_step, threads, [&](uint32_t y, Cnt& c) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            uint32_t seed = tea(y * W + x, f.frame);
            const float* P4 = position + 4 * p;
            const float* N4 = normal + 4 * p;
            bool background = P4[3] == 0.0f && N4[3] == 0.0f;
            if (background) {
                for (uint32_t i = 0; i < nl; ++i) shadow[i * plane + p] = 0.0f;
                continue;
            }
            int mode = (int)s->mats[mat_index(*s, P4[3])].shading_metal_rough_idx[0];
            v3 P = ld3(P4);
            v3 N = normalize(ld3(N4));
            if (mode == 4) {
                for (uint32_t i = 0; i < nl; ++i) shadow[i * plane + p] = 1.0f;
                continue;
            }
            uint32_t S = f.samples;
            for (uint32_t i = 0; i < nl; ++i) {
                const hr_light& lt = lights[i];
                bool isDir = lt.pos_maxdist[3] < 0.0f;
                v3 L = isDir ? ld3(lt.pos_maxdist) : sub(ld3(lt.pos_maxdist), P);
                float maxd = lt.pos_maxdist[3];
                float dist = length(L);
                L = normalize(L);
                float NdotL = clamp01(dot(N, L));
                float factor = 0.0f;
                if (NdotL > 0.0f && f.adaptive && (isDir || dist < maxd)) {
                    /* 1 ray for a hard light; else the stages of the nested
                     * sequence, each adding to the samples before it */
                    const bool hard = hard_light(f, i);
                    uint32_t target = hard ? 1u : adaptive_first(S), unocc = 0, sI = 0;
                    for (;;) {
                        for (; sI < target; ++sI) {
                            Pcg g = ray_stream(seed, PURPOSE_ADAPTIVE_SHADOW, i * 128u + sI);
                            v3 dir = light_sample_dir_nested(f, i, P, L, g, sI, hard ? 1u : S);
                            float tmax = isDir ? 1000.0f : dist + 1.0f;
                            ++c.shadow;
                            bool sh = occluded(*s, add(P, scl(dir, 1e-1f)), dir, 0.001f, tmax, c);
                            if (sh) ++c.occluded; else ++unocc;
                        }
                        if (hard || adaptive_stop(unocc, target, S)) break;
                        target = adaptive_next(target, S);
                    }
                    factor = (float)unocc / (float)target;
                } else if (NdotL > 0.0f) {
                    for (uint32_t sI = 0; sI < S; ++sI) {
                        bool shadowed;
                        if (isDir || dist < maxd) {
                            Pcg g = ray_stream(seed, PURPOSE_SURFACE_SHADOW, i * S + sI);
                            v3 dir = light_sample_dir(f, i, P, L, g, sI, S);
                            float tmax = isDir ? 1000.0f : dist + 1.0f;
                            ++c.shadow;
                            shadowed = occluded(*s, add(P, scl(dir, 1e-1f)), dir, 0.001f, tmax, c);
                            if (shadowed) ++c.occluded;
                        } else {
                            shadowed = false;
                        }
                        if (!shadowed) factor += 1.0f;
                    }
                    factor /= (float)S;
                }
                shadow[i * plane + p] = factor;
            }
        }
    }, total);
    to_counters(total, cnt);
}

/* TESTS ONLY: the ray that or_trace_shadows traces for sample sI of light li
 * at pixel (x, y) with adaptive sampling (the same preconditions and set-up
 * as its adaptive branch). Returns 0 when the factor is set without rays
 * (background, glass, facing away, out of reach), 1 for a light sampled by
 * the nested sequence of shadow_samples, 2 for a hard light (one sample). */
int or_adaptive_shadow_ray(const or_scene* s, uint32_t W, uint32_t H, const float* position, const float* normal,
                           const hr_light* lights, const hr_light_shape* shapes, uint32_t nl, const hr_camera* cam,
                           const hr_trace_params* prm, uint32_t x, uint32_t y, uint32_t li, uint32_t sI, hr_ray* ray) {
    Frame f{lights, shapes, nl, prm->shadow_samples, prm->refl_spp, prm->max_depth, W, H, frame_index(cam), true};
    size_t p = (size_t)y * W + x;
    const float* P4 = position + 4 * p;
    const float* N4 = normal + 4 * p;
    if (P4[3] == 0.0f && N4[3] == 0.0f) return 0;
    if ((int)s->mats[mat_index(*s, P4[3])].shading_metal_rough_idx[0] == 4) return 0;
    v3 P = ld3(P4);
    v3 N = normalize(ld3(N4));
    const hr_light& lt = lights[li];
    bool isDir = lt.pos_maxdist[3] < 0.0f;
    v3 L = isDir ? ld3(lt.pos_maxdist) : sub(ld3(lt.pos_maxdist), P);
    float dist = length(L);
    L = normalize(L);
    if (!(clamp01(dot(N, L)) > 0.0f) || !(isDir || dist < lt.pos_maxdist[3])) return 0;
    const bool hard = hard_light(f, li);
    Pcg g = ray_stream(tea(y * W + x, f.frame), PURPOSE_ADAPTIVE_SHADOW, li * 128u + sI);
    v3 dir = light_sample_dir_nested(f, li, P, L, g, sI, hard ? 1u : f.samples);
    v3 o = add(P, scl(dir, 1e-1f));
    ray->org[0] = o.x; ray->org[1] = o.y; ray->org[2] = o.z;
    ray->dir[0] = dir.x; ray->dir[1] = dir.y; ray->dir[2] = dir.z;
    ray->tmin = 0.001f;
    ray->tmax = isDir ? 1000.0f : dist + 1.0f;
    return hard ? 2 : 1;
}

/* hybridRaygen.rgen:27-186 with the bounce chain of :153-170 */
void or_shade_hybrid(const or_scene* s, uint32_t W, uint32_t H, const float* position, const float* normal,
                     const float* albedo_p, const float* material, const float* emissive, const hr_light* lights,
                     const hr_light_shape* shapes, uint32_t nl, const hr_camera* cam, const hr_trace_params* prm,
                     const float* shadow, float* lit, float* refl, uint32_t row_start, uint32_t row_step,
                     or_counters* cnt, uint32_t threads) {
    Cnt total;
    Frame f{lights, shapes, nl, prm->shadow_samples, prm->refl_spp, prm->max_depth, W, H, frame_index(cam),
            (prm->flags & HR_FLAG_ADAPTIVE) != 0};
    size_t plane = (size_t)W * H;
    float cpos[4];
    mul4(cam->inv_view, 0.0f, 0.0f, 0.0f, 1.0f, cpos);
    v3 camPos = mk(cpos[0], cpos[1], cpos[2]);
    parallel_rows(H, row_start, row_step, threads, [&](uint32_t y, Cnt& c) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            uint32_t seed = tea(y * W + x, f.frame);
            const float* P4 = position + 4 * p;
            const float* N4 = normal + 4 * p;
            const hr_material& mat = s->mats[mat_index(*s, P4[3])];
            int mode = (int)mat.shading_metal_rough_idx[0];
            v3 albedo = pow3(ld3(albedo_p + 4 * p), 2.2f);
            v3 mtl = ld3(material + 4 * p);
            v3 emis = ld3(emissive + 4 * p);
            bool background = P4[3] == 0.0f && N4[3] == 0.0f;
            float* out = lit + 4 * p;
            float* rf = refl + 4 * p;
            rf[0] = rf[1] = rf[2] = rf[3] = 1.0f;
            if (background) {
                out[0] = albedo.x; out[1] = albedo.y; out[2] = albedo.z; out[3] = 1.0f;
                continue;
            }
            v3 position3 = ld3(P4);
            float roughness = mtl.y, metallic = mtl.z;
            v3 F0 = mixs(mk(0.04f, 0.04f, 0.04f), albedo, metallic);
            v3 V = normalize(sub(camPos, position3));
            v3 N = normalize(ld3(N4));
            float NdotV = clamp01(dot(N, V));
            v3 irradiance = env_lookup(s->irr, N, PI_H);
            float attenuation = 1.0f;
            v3 rayColor = mk(0, 0, 0);
            v3 origin = position3;
            v3 direction = mk(0, 0, 0);
            if (mode == 3) direction = reflect(normalize(neg(V)), N);
            if (mode == 4) direction = glass_raygen_dir(N, V, mat.diffuse_ior[3]);
            for (uint32_t i = 0; i < nl; ++i) {
                const hr_light& lt = lights[i];
                bool isDir = lt.pos_maxdist[3] < 0.0f;
                v3 L = isDir ? ld3(lt.pos_maxdist) : sub(ld3(lt.pos_maxdist), position3);
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
                v3 lc = ld3(lt.color_intensity);
                if (mode == 0) {
                    v3 radiance = scl(scl(scl(lc, intensity), attenuation), shadowFactor);
                    v3 Hh = normalize(add(V, L));
                    float D = DistributionGGX(N, Hh, roughness);
                    float G = GeometrySmith(N, V, L, roughness);
                    v3 F = FresnelSchlick(fmaxf(g_plant == PLANT_FRESNEL_NDOTH ? dot(N, Hh) : dot(Hh, V), 0.0f), F0);
                    v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
                    v3 numerator = scl(F, D * G);
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
            v3 rfac = mk(1, 1, 1);
            // hybridRaygen.rgen:66-67, 125/144: the bounce direction is set in
            // the light loop; with no lights it stays (0,0,0) and the bounce's
            // colour multiplies rayColor = 0 (decision: no bounce is traced)
            if ((mode == 3 || mode == 4) && f.max_depth > 0 && nl > 0) {
                uint32_t spp = pixel_spp(f, mode, roughness);
                v3 acc = mk(0, 0, 0);
                for (uint32_t sI = 0; sI < spp; ++sI) {
                    v3 d = direction;
                    v3 prod = mk(1, 1, 1);
                    bool alive = true;
                    if (mode == 3 && roughness > 0.0f) {
                        Pcg g = ray_stream(seed, PURPOSE_REFLECT, sI);
                        float r1 = u01(g), r2 = u01(g);
                        v3 Hm = sampleGGX(N, roughness, r1, r2);
                        d = reflect(normalize(neg(V)), Hm);
                        if (!(dot(d, N) > 0.0f)) {
                            prod = mk(0, 0, 0);
                            alive = false;
                        }
                    }
                    v3 o = origin;
                    for (uint32_t depth = 0; alive && depth < f.max_depth; ++depth) {
                        v3 ro = add(o, scl(d, 1e-2f));
                        RayP r = prep_ray(ro, d, 0.001f, 1000.0f);
                        Hit h;
                        ++c.refl;
                        if (!traverse(*s, r, false, h, c)) {
                            v3 sky = pow3(env_lookup(s->sky, normalize(d), PI_M), 2.2f);
                            prod = mul(prod, sky);
                            break;
                        }
                        HitOut ho = shade_hit(*s, f, ro, d, h, seed, sI, depth, c);
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
            v3 F = FresnelSchlick(NdotV, F0);
            v3 kD = scl(sub(mk(1, 1, 1), F), 1.0f - metallic);
            v3 ambient = mul(mul(kD, albedo), irradiance);
            v3 fc = add(add(rayColor, ambient), emis);
            out[0] = fc.x; out[1] = fc.y; out[2] = fc.z; out[3] = 1.0f;
            rf[0] = rfac.x; rf[1] = rfac.y; rf[2] = rfac.z; rf[3] = 1.0f;
        }
    }, total);
    to_counters(total, cnt);
}


/* raygen.rgen:20-81: camera samples, the closest-hit chain, accumulation */
void or_render_rt(const or_scene* s, const hr_light* lights, const hr_light_shape* shapes, uint32_t nl,
                  const hr_camera* cam, const hr_trace_params* prm, uint32_t W, uint32_t H, float* accum,
                  uint32_t row_start, uint32_t row_step, or_counters* cnt, uint32_t threads) {
    Cnt total;
    Frame f{lights, shapes, nl, prm->shadow_samples, prm->refl_spp, prm->max_depth, W, H, frame_index(cam), false};
    const int frame_signed = (int)cam->frame[0];
    parallel_rows(H, row_start, row_step, threads, [&](uint32_t y, Cnt& c) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            uint32_t seed = tea(y * W + x, f.frame);
            v3 finalColor = mk(0, 0, 0);
            uint32_t spp = std::max(f.spp, 1u);
            for (uint32_t nS = 0; nS < spp; ++nS) {
                float ox = 0.5f, oy = 0.5f;
                if (nS > 0) {
                    Pcg gp = ray_stream(seed, PURPOSE_RT_PIXEL, nS);
                    ox = u01(gp);
                    oy = u01(gp);
                }
                float ux = ((float)x + ox) / (float)W;
                float uy = ((float)y + oy) / (float)H;
                float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
                float org[4], tgt[4], dir[4];
                mul4(cam->inv_view, 0.0f, 0.0f, 0.0f, 1.0f, org);
                mul4(cam->inv_proj, dx, dy, 1.0f, 1.0f, tgt);
                v3 tn = normalize(mk(tgt[0], tgt[1], tgt[2]));
                mul4(cam->inv_view, tn.x, tn.y, tn.z, 0.0f, dir);
                v3 o = mk(org[0], org[1], org[2]), d = mk(dir[0], dir[1], dir[2]);
                v3 rayColor = mk(1, 1, 1);
                for (uint32_t depth = 0; depth < f.max_depth; ++depth) {
                    v3 ro = add(o, scl(d, 1e-2f));
                    RayP r = prep_ray(ro, d, 0.001f, 10000.0f);
                    Hit h;
                    if (depth == 0) ++c.primary; else ++c.refl;
                    if (!traverse(*s, r, false, h, c)) {
                        rayColor = mul(rayColor, pow3(env_lookup(s->sky, normalize(d), PI_M), 2.2f));
                        break;
                    }
                    HitOut hc = shade_rt(*s, f, ro, d, h, seed, nS, depth, c);
                    rayColor = mul(rayColor, hc.color);
                    if (!hc.scatter) break;
                    o = hc.origin;
                    d = hc.dir;
                }
                finalColor = add(finalColor, rayColor);
            }
            finalColor = divs(finalColor, (float)spp);
            float* out = accum + 4 * p;
            if (frame_signed < 0) {
                out[0] = finalColor.x; out[1] = finalColor.y; out[2] = finalColor.z;
            } else {
                float a = 1.0f / (1.0f + (float)frame_signed);
                v3 m = add(scl(mk(out[0], out[1], out[2]), 1.0f - a), scl(finalColor, a));
                out[0] = m.x; out[1] = m.y; out[2] = m.z;
            }
            out[3] = 1.0f;
        }
    }, total);
    to_counters(total, cnt);
}
} /* extern "C" */

/* postFragment.frag:11-20 over n float4 pixels: rgb = pow(max(rgb, 0.001),
 * 1 / 2.2), alpha kept; rgba8 (may be NULL) = rint(clamp(v, 0, 1) * 255). */
void or_post_gamma(const float* in, uint64_t n, float* out, uint8_t* rgba8) {
    const float g = 1.0f / 2.2f;
    for (uint64_t i = 0; i < n; ++i) {
        float o[4] = {powf(fmaxf(in[4 * i], 0.001f), g), powf(fmaxf(in[4 * i + 1], 0.001f), g),
                      powf(fmaxf(in[4 * i + 2], 0.001f), g), in[4 * i + 3]};
        for (int k = 0; k < 4; ++k) {
            if (out) out[4 * i + k] = o[k];
            if (rgba8) rgba8[4 * i + k] = (uint8_t)rintf(fminf(fmaxf(o[k], 0.0f), 1.0f) * 255.0f);
        }
    }
}
