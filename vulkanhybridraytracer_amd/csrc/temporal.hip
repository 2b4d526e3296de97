// temporal.hip - the reference's temporal shadow filter (blur.comp:35-81,
// dispatched by renderer.cpp:2661-2675 between the shadow and hybrid passes)
// for any frame size and light count, and the motion vectors its
// reprojection reads (geometry_shader.frag:72-75 with basic.vert:38-46).
//
// One lane per pixel, looping over the lights (the pixel's motion vector is
// read once); rows are contiguous, so a wave's loads and stores of each
// light's factors are coalesced. Every expression is evaluated as written
// (no contraction), so results are bit-exact against the oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace hr {
namespace {

// imageLoad outside the image reads 0 (robust image access)
__device__ __forceinline__ float load_or_zero(const float* img, int x, int y, int W, int H) {
    return (x >= 0 && x < W && y >= 0 && y < H) ? img[(size_t)y * W + x] : 0.0f;
}

// GLSL mix(x, y, a) = x * (1 - a) + y * a
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// frame > 0: running mean of the traced factors (blur.comp:40-47)
// frame <= 0: reprojected history clamped to the 3x3 neighbourhood of the new
// factor, blended with weight 0.4 (blur.comp:48-79). prev holds the previous
// output (it may be out itself when frame > 0: each pixel reads only itself).
__global__ __launch_bounds__(256) void k_filter_shadows(uint32_t W, uint32_t H, uint32_t nl, int32_t frame,
                                                        const float* __restrict__ shadow, const float2* motion,
                                                        const float* prev, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const int x = (int)(i % W), y = (int)(i / W);
    const size_t plane = (size_t)W * H;
    if (frame > 0) {
        const float a = 1.0f / ((float)frame + 1.0f);
        for (uint32_t l = 0; l < nl; ++l) {
            const size_t o = l * plane + i;
            out[o] = mixf(prev[o], shadow[o], a);
        }
        return;
    }
    // reprojection; the y offset uses the x component of the motion vector,
    // as blur.comp:52 evaluates vec2(x + r.x * W, y + r * H)
    const float2 m = motion[i];
    const float rx = m.x * 2.0f - 1.0f;
    const float lx = (float)x + rx * (float)W;
    const float ly = (float)y + rx * (float)H;
    const int px = (int)lx, py = (int)ly;  // ivec2(): truncation toward zero
    for (uint32_t l = 0; l < nl; ++l) {
        const float* in = shadow + l * plane;
        const float center = in[i];
        float mn = center, mx = center;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const float c = load_or_zero(in, x + dx, y + dy, (int)W, (int)H);
                mn = fminf(mn, c);
                mx = fmaxf(mx, c);
            }
        float old = load_or_zero(prev + l * plane, px, py, (int)W, (int)H);
        old = fmaxf(mn, old);
        old = fminf(mx, old);
        out[l * plane + i] = mixf(old, center, 0.4f);
    }
}

// motion = (ndc - ndc_prev).xy * 0.5 + 0.5 of the pixel's world position;
// background pixels (position.w == 0 && normal.w == 0) get 0, as the skybox
// pass writes (skybox.frag:19)
__global__ __launch_bounds__(256) void k_motion(uint32_t W, uint32_t H, const float4* pos, const float4* nrm,
                                                MotionArgs a, float2* motion) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const float4 P = pos[i], N = nrm[i];
    if (P.w == 0.0f && N.w == 0.0f) {
        motion[i] = make_float2(0.0f, 0.0f);
        return;
    }
    float c[4], q[4];
    const float v[4] = {P.x, P.y, P.z, 1.0f};
    for (int r = 0; r < 4; ++r) {
        c[r] = ((a.vp[r] * v[0] + a.vp[4 + r] * v[1]) + a.vp[8 + r] * v[2]) + a.vp[12 + r] * v[3];
        q[r] = ((a.vp_prev[r] * v[0] + a.vp_prev[4 + r] * v[1]) + a.vp_prev[8 + r] * v[2]) + a.vp_prev[12 + r] * v[3];
    }
    const float nx = c[0] / c[3], ny = c[1] / c[3];
    const float px = q[0] / q[3], py = q[1] / q[3];
    motion[i] = make_float2((nx - px) * 0.5f + 0.5f, (ny - py) * 0.5f + 0.5f);
}

}  // namespace

hipError_t launch_filter_shadows(uint32_t W, uint32_t H, uint32_t nl, int32_t frame, const float* shadow,
                                 const float* motion, const float* prev, float* out, hipStream_t st) {
    const uint32_t n = W * H;
    k_filter_shadows<<<(n + 255) / 256, 256, 0, st>>>(W, H, nl, frame, shadow,
                                                      reinterpret_cast<const float2*>(motion), prev, out);
    return hipGetLastError();
}

hipError_t launch_motion(uint32_t W, uint32_t H, const float4* pos, const float4* nrm, const MotionArgs& a,
                         float* motion, hipStream_t st) {
    const uint32_t n = W * H;
    k_motion<<<(n + 255) / 256, 256, 0, st>>>(W, H, pos, nrm, a, reinterpret_cast<float2*>(motion));
    return hipGetLastError();
}

}  // namespace hr
