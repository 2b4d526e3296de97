// gbuffer_io.hip - the raster G-buffer formats of renderer.cpp:252-261 in and
// out of the FP32 planes the trace kernels read.
//
//   k_ingest    RGBA16F position / normal, RGBA8 albedo / material / emissive,
//               RG16F motion -> FP32 planes, decoded as the shaders decode
//               them (shadowRaygen.rgen:36-41, hybridRaygen.rgen:40-47):
//               normal = n * 2 - 1, UNORM8 = c / 255
//   k_quantize  FP32 planes -> the raster formats, as geometry_shader.frag:66-75
//               writes them (normal * 0.5 + 0.5, material (0, rough, metal))
//               and the attachments store them: half by round to nearest
//               even, UNORM8 = rint(clamp(v, 0, 1) * 255)
//   k_post_gamma  the display pass of postFragment.frag
// One lane per pixel; every plane is read or written as one 8-byte (RGBA16F),
// 4-byte (RGBA8, RG16F) or 16-byte (FP32) access, so a wave's accesses are
// contiguous.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace hr {
namespace {

__device__ __forceinline__ float h2f(uint16_t b) { return __half2float(__ushort_as_half(b)); }
__device__ __forceinline__ uint16_t f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }
__device__ __forceinline__ float u8f(uint32_t c) { return (float)c / 255.0f; }
__device__ __forceinline__ uint32_t fu8(float v) {
    float c = fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f;
    return (uint32_t)rintf(c);
}
__device__ __forceinline__ float4 rgba8(uint32_t w) {
    return make_float4(u8f(w & 0xFFu), u8f((w >> 8) & 0xFFu), u8f((w >> 16) & 0xFFu), u8f(w >> 24));
}
__device__ __forceinline__ uint32_t pack8(float a, float b, float c, float d) {
    return fu8(a) | (fu8(b) << 8) | (fu8(c) << 16) | (fu8(d) << 24);
}

__global__ __launch_bounds__(256) void k_ingest(uint32_t n, const ushort4* pos, const ushort4* nrm,
                                                const uint32_t* alb, const ushort2* mot, const uint32_t* mat,
                                                const uint32_t* emi, GBufView G, float2* motion_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ushort4 p = pos[i], q = nrm[i];
    G.pos[i] = make_float4(h2f(p.x), h2f(p.y), h2f(p.z), h2f(p.w));
    G.nrm[i] = make_float4(h2f(q.x) * 2.0f - 1.0f, h2f(q.y) * 2.0f - 1.0f, h2f(q.z) * 2.0f - 1.0f, h2f(q.w));
    G.alb[i] = rgba8(alb[i]);
    G.mat[i] = rgba8(mat[i]);
    const float4 e = rgba8(emi[i]);
    G.emi[i] = make_float4(e.x, e.y, e.z, 0.0f);
    if (mot && motion_out) {
        const ushort2 m = mot[i];
        motion_out[i] = make_float2(h2f(m.x), h2f(m.y));
    }
}

__global__ __launch_bounds__(256) void k_quantize(uint32_t n, GBufView G, const float2* motion_in, ushort4* pos,
                                                  ushort4* nrm, uint32_t* alb, ushort2* mot, uint32_t* mat,
                                                  uint32_t* emi) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = G.pos[i], N = G.nrm[i], A = G.alb[i], M = G.mat[i], E = G.emi[i];
    pos[i] = make_ushort4(f2h(P.x), f2h(P.y), f2h(P.z), f2h(P.w));
    // background normals are written as 0 (skybox.frag:17), surfaces as n * 0.5 + 0.5
    const bool bg = N.w == 0.0f;
    nrm[i] = bg ? make_ushort4(0, 0, 0, 0)
                : make_ushort4(f2h(N.x * 0.5f + 0.5f), f2h(N.y * 0.5f + 0.5f), f2h(N.z * 0.5f + 0.5f), f2h(N.w));
    alb[i] = pack8(A.x, A.y, A.z, A.w);
    mat[i] = pack8(M.x, M.y, M.z, M.w);
    emi[i] = pack8(E.x, E.y, E.z, 1.0f);
    if (mot) {
        const float2 m = motion_in ? motion_in[i] : make_float2(0.0f, 0.0f);
        mot[i] = make_ushort2(f2h(m.x), f2h(m.y));
    }
}

// postFragment.frag:11-20: rgb = pow(max(rgb, 0.001), 1 / 2.2), alpha kept;
// optionally stored to an RGBA8 UNORM image
__global__ __launch_bounds__(256) void k_post_gamma(uint32_t n, const float4* in, float4* out, uint32_t* rgba8) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = in[i];
    const float g = 1.0f / 2.2f;
    const float4 o = make_float4(powf(fmaxf(c.x, 0.001f), g), powf(fmaxf(c.y, 0.001f), g),
                                 powf(fmaxf(c.z, 0.001f), g), c.w);
    if (out) out[i] = o;
    if (rgba8) rgba8[i] = pack8(o.x, o.y, o.z, o.w);
}

}  // namespace

hipError_t launch_post_gamma(uint32_t n, const float4* in, float4* out, uint32_t* rgba8, hipStream_t st) {
    k_post_gamma<<<(n + 255) / 256, 256, 0, st>>>(n, in, out, rgba8);
    return hipGetLastError();
}

hipError_t launch_gbuffer_ingest(uint32_t n, const void* pos, const void* nrm, const void* alb, const void* mot,
                                 const void* mat, const void* emi, GBufView G, float* motion_out, hipStream_t st) {
    k_ingest<<<(n + 255) / 256, 256, 0, st>>>(n, static_cast<const ushort4*>(pos), static_cast<const ushort4*>(nrm),
                                              static_cast<const uint32_t*>(alb), static_cast<const ushort2*>(mot),
                                              static_cast<const uint32_t*>(mat), static_cast<const uint32_t*>(emi), G,
                                              reinterpret_cast<float2*>(motion_out));
    return hipGetLastError();
}

hipError_t launch_gbuffer_quantize(uint32_t n, GBufView G, const float* motion_in, void* pos, void* nrm, void* alb,
                                   void* mot, void* mat, void* emi, hipStream_t st) {
    k_quantize<<<(n + 255) / 256, 256, 0, st>>>(n, G, reinterpret_cast<const float2*>(motion_in),
                                                static_cast<ushort4*>(pos), static_cast<ushort4*>(nrm),
                                                static_cast<uint32_t*>(alb), static_cast<ushort2*>(mot),
                                                static_cast<uint32_t*>(mat), static_cast<uint32_t*>(emi));
    return hipGetLastError();
}

}  // namespace hr
