// device_math.h - device-side algebra shared by the LBVH, traversal and
// shading kernels of libhr (gfx950).
//
// The library is compiled with -ffp-contract=off: every float expression is
// evaluated in the order written (no implicit FMA), and division and sqrt are
// correctly rounded. Explicit fmaf() is used only where it is part of the
// definition (the slab test). This is what makes integer results, hit/miss,
// triangle ids, hit distances and barycentrics reproducible bit for bit by a
// scalar CPU build of the same definitions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HR_DEV __device__ __forceinline__

namespace hrd {

struct v3 {
    float x, y, z;
};
HR_DEV v3 mk(float x, float y, float z) { return v3{x, y, z}; }
HR_DEV v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
HR_DEV v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
HR_DEV v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
HR_DEV v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
HR_DEV v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
HR_DEV v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
HR_DEV float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
HR_DEV v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
HR_DEV float length(v3 a) { return sqrtf(dot(a, a)); }
HR_DEV v3 normalize(v3 a) {
    float inv = 1.0f / sqrtf(dot(a, a));
    return scl(a, inv);
}
HR_DEV float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
HR_DEV float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
HR_DEV v3 ld3(float4 a) { return mk(a.x, a.y, a.z); }
HR_DEV v3 ld3p(const float* p) { return mk(p[0], p[1], p[2]); }
// GLSL reflect / refract / mix
HR_DEV v3 reflect(v3 I, v3 N) {
    float d2 = 2.0f * dot(N, I);
    return sub(I, scl(N, d2));
}
HR_DEV v3 refract(v3 I, v3 N, float eta) {
    float ndi = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float a = eta * ndi + sqrtf(k);
    return sub(scl(I, eta), scl(N, a));
}
HR_DEV v3 mixs(v3 x, v3 y, float a) {
    float oma = 1.0f - a;
    return add(scl(x, oma), scl(y, a));
}
HR_DEV v3 pow3(v3 a, float e) { return mk(powf(a.x, e), powf(a.y, e), powf(a.z, e)); }
// ordered selection: first operand on ties (box unions must not depend on the
// sign-of-zero conventions of a min instruction)
HR_DEV float minsel(float a, float b) { return (b < a) ? b : a; }
HR_DEV float maxsel(float a, float b) { return (b > a) ? b : a; }

// column-major 4x4 times (x, y, z, w), sums left to right
HR_DEV void mul4(const float* m, float x, float y, float z, float w, float* o) {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

// ---- atan2 / asin / sincos(2 pi u) by Taylor series -------------------------
HR_DEV float atan_small(float s) {  // |s| <= tan(pi/8), Taylor series to s^19
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
HR_DEV float atan01(float t) {  // atan t = pi/4 + atan((t-1)/(t+1)) above tan(pi/8)
    if (t > 0.41421356237309503f) {
        float s = (t - 1.0f) / (t + 1.0f);
        return 0.78539816339744831f + atan_small(s);
    }
    return atan_small(t);
}
HR_DEV float atan2_hr(float y, float x) {
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
HR_DEV float asin_hr(float x) {
    x = fminf(fmaxf(x, -1.0f), 1.0f);
    return atan2_hr(x, sqrtf((1.0f - x) * (1.0f + x)));
}
// quadrant from floor(4u); Taylor series of sin(pi/2 f) to f^11, cos to f^12
HR_DEV void sincos2pi_hr(float u, float* so, float* co) {
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
    float rs, rc;
    if (q == 0) { rs = s; rc = c; }
    else if (q == 1) { rs = c; rc = -s; }
    else if (q == 2) { rs = -s; rc = -c; }
    else { rs = -c; rc = s; }
    *so = rs;
    *co = rc;
}

// ---- random numbers -------------------------------------------------------
// random.glsl:3-17: 16 rounds of TEA, returns v0
HR_DEV uint32_t tea(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
    for (uint32_t n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// PCG32 XSH-RR 64/32 (O'Neill 2014)
struct Pcg {
    uint64_t state, inc;
};
HR_DEV uint32_t pcg_next(Pcg& g) {
    uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
HR_DEV Pcg pcg_init(uint64_t initstate, uint64_t initseq) {
    Pcg g;
    g.state = 0u;
    g.inc = (initseq << 1u) | 1u;
    pcg_next(g);
    g.state += initstate;
    pcg_next(g);
    return g;
}
HR_DEV float u01(Pcg& g) { return (float)(pcg_next(g) >> 8) * (1.0f / 16777216.0f); }
// the state `delta` steps ahead, in O(log delta) (F. Brown, "Random number
// generation with arbitrary strides", 1994): composes state' = a * state + c
// by repeated squaring of (a, c)
HR_DEV void pcg_advance(Pcg& g, uint32_t delta) {
    uint64_t cur_mult = 6364136223846793005ULL, cur_plus = g.inc;
    uint64_t acc_mult = 1u, acc_plus = 0u;
    while (delta > 0u) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    g.state = acc_mult * g.state + acc_plus;
}

// stream purposes: stream id = (purpose << 32) | index
enum : uint32_t {
    PURPOSE_SURFACE_SHADOW = 1,  // index = light * samples + sample
    PURPOSE_REFLECT = 2,         // index = spp sample
    PURPOSE_HIT_SHADOW = 3,      // index = (sample * max_depth + depth) * nl + light
    PURPOSE_HIT_REFLECT = 4,     // index = sample * max_depth + depth
    PURPOSE_ADAPTIVE_SHADOW = 5,  // index = light * 128 + s (s: sample of the nested sequence, < 64)
    PURPOSE_RT_PIXEL = 6,         // pure ray tracing: index = pixel sample (sub-pixel offset)
    PURPOSE_RT_SHADOW = 7,        // pure ray tracing: index = ((sample * max_depth + depth) * nl + light)
    PURPOSE_RT_REFLECT = 8        // pure ray tracing: index = sample * max_depth + depth
};
HR_DEV Pcg ray_stream(uint32_t pixel_seed, uint32_t purpose, uint32_t index) {
    return pcg_init((uint64_t)pixel_seed, ((uint64_t)purpose << 32) | (uint64_t)index);
}

// ---- Morton codes -----------------------------------------------------------
HR_DEV uint32_t expand_bits(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
HR_DEV uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand_bits(x) << 2) | (expand_bits(y) << 1) | expand_bits(z);
}
HR_DEV uint32_t quant10(float c, float lo, float scale) {
    float q = floorf((c - lo) * scale);
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}

// ---- ray setup, slab test, watertight triangle test -------------------------
struct RayP {
    v3 o, d;
    v3 idir, ood;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
HR_DEV float safe_inv(float d) { return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d); }
HR_DEV RayP prep_ray(v3 o, v3 d) {
    RayP r;
    r.o = o;
    r.d = d;
    r.idir = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    r.ood = mul(o, r.idir);
    int kz = 0;
    if (fabsf(d.y) > fabsf(comp(d, kz))) kz = 1;
    if (fabsf(d.z) > fabsf(comp(d, kz))) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    float dz = comp(d, kz);
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}
// Returns true when [max(entry, tmin), min(exit, tbest)] is non-empty, with
// the exit distance widened by 2^-22 relative.
HR_DEV bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz, const RayP& r, float tmin,
                    float tbest, float& tn) {
    float tx0 = fmaf(lx, r.idir.x, -r.ood.x), tx1 = fmaf(hx, r.idir.x, -r.ood.x);
    float ty0 = fmaf(ly, r.idir.y, -r.ood.y), ty1 = fmaf(hy, r.idir.y, -r.ood.y);
    float tz0 = fmaf(lz, r.idir.z, -r.ood.z), tz1 = fmaf(hz, r.idir.z, -r.ood.z);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tbest));
    tn = n;
    return n <= f * 1.0000002384185791f;
}
// Woop, Benthin, Wald: "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013
HR_DEV bool intersect_tri(const RayP& r, v3 v0, v3 v1, v3 v2, float tmin, float tmax, float& t_out,
                          float& b1_out, float& b2_out) {
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

}  // namespace hrd
