// denoise.hip - kernels of the kernel-predicting denoiser of the low-sample
// hybrid pass (no counterpart in the reference, whose only "denoise" is the
// temporal shadow filter of blur.comp).
//
// Activations and weights are BF16 in NHWC, accumulation is FP32:
//   k_conv   implicit-GEMM 3x3 convolution (zero padding) on
//            v_mfma_f32_16x16x32_bf16: GEMM M = output channels, N = pixels,
//            K = 9 * Cin padded to 32. A lane's 8 K values are 8 contiguous
//            channels of one tap (one 16-byte load); weights are pre-arranged
//            in fragment order so a wave's weight load is one contiguous
//            1 KiB. A wave computes 64 channels x 64 pixels (4 x 4 tiles).
//            Epilogue: bias, ReLU, 4 channels per lane stored as one 8-byte
//            BF16 write (or FP32), into a channel slice. The input gradient
//            is the same kernel on the output gradient with flipped,
//            transposed weights.
//   k_wgrad  weight gradient, GEMM M = output channels, N = 9 * Cin, K =
//            pixels: 32-pixel tiles of the output gradient and of the input
//            patches are loaded with 16-byte loads and transposed through LDS
//            so each MFMA fragment is one 16-byte LDS read; one block = 64 x
//            64 (co, k) over a chunk of pixels, FP32 partial sums per chunk
//            added by k_reduce in a fixed order (deterministic). Bias
//            gradients come from the same LDS tiles.
//   pooling, upsampling, ReLU masks, the input features, the
//   kernel-prediction softmax and its backward with the tone-mapped L1 loss,
//   and Adam (FP32 master weights, BF16 copies in both fragment layouts) are
//   elementwise kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

#include <type_traits>

namespace hr {
namespace dn {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(__bf16 h) { return (float)h; }
__device__ __forceinline__ __bf16 f2bf(float f) { return (__bf16)f; }

// index of GEMM element (co, k) in fragment order
__host__ __device__ __forceinline__ size_t frag_index(uint32_t co, uint32_t k, uint32_t cout_pad) {
    return ((size_t)((k >> 5) * (cout_pad >> 4) + (co >> 4)) * 64u + (co & 15u) + ((k & 31u) >> 3) * 16u) * 8u + (k & 7u);
}

// kernel prediction: out = sum_k softmax(logits)_k * noisy[clamp(p + off_k)]
// over the 5x5 neighbourhood, off_k = (k % 5 - 2, k / 5 - 2)
// lg: a pixel's first 28 logits (16-byte aligned), read as 16-byte chunks
__device__ __forceinline__ void softmax25(const float* lg, float* w) {
    const float4* l4 = reinterpret_cast<const float4*>(lg);
    float v[28];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const float4 t = l4[q];
        v[4 * q] = t.x;
        v[4 * q + 1] = t.y;
        v[4 * q + 2] = t.z;
        v[4 * q + 3] = t.w;
    }
    float m = -INFINITY;
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        w[k] = v[k];
        m = fmaxf(m, w[k]);
    }
    float s = 0.0f;
    for (int k = 0; k < 25; ++k) {
        w[k] = expf(w[k] - m);
        s += w[k];
    }
    float inv = 1.0f / s;
    for (int k = 0; k < 25; ++k) w[k] *= inv;
}
__device__ __forceinline__ float4 tap_color(const float4* img, int x, int y, int k, uint32_t W, uint32_t H) {
    int xx = min(max(x + k % 5 - 2, 0), (int)W - 1), yy = min(max(y + k / 5 - 2, 0), (int)H - 1);
    return img[(size_t)yy * W + xx];
}
// the denoised colour of pixel (x, y) of a W x H frame from its logits
__device__ __forceinline__ void kpcn_pixel(const float* lg, const float4* noisy, float4* out, int x, int y,
                                           uint32_t W, uint32_t H) {
    float w[25];
    softmax25(lg, w);
    float r = 0.f, g = 0.f, b = 0.f;
    for (int k = 0; k < 25; ++k) {
        float4 c = tap_color(noisy, x, y, k, W, H);
        r += w[k] * c.x;
        g += w[k] * c.y;
        b += w[k] * c.z;
    }
    out[(size_t)y * W + x] = make_float4(r, g, b, 1.0f);
}

// ---- convolution (forward and input gradient) ----------------------------------
constexpr int CONV_MT = 4;  // 16-pixel tiles per wave
constexpr int CONV_NT = 4;  // 16-channel tiles per wave

// Epilogue of both convolution kernels: bias, ReLU, and 4 channels per lane
// stored into the output's channel slice (BF16 8-byte or FP32 16-byte writes).
template <int NS>
__device__ __forceinline__ void conv_store(const ConvArgs& a, const f32x4 (&acc)[CONV_MT][NS],
                                           const uint32_t (&pix)[CONV_MT], const bool (&pok)[CONV_MT],
                                           uint32_t s0, uint32_t lane) {
    // C layout: row (channel) = (lane / 16) * 4 + r, column (pixel) = lane % 16
    const bool vec_ok = ((a.out_C | a.out_off) & 3u) == 0;  // 4-channel groups aligned
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const uint32_t co = (s0 + s) * 16u + (lane >> 4) * 4u;
        if (co >= a.Cout) continue;
        float b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) b[r] = (a.bias && co + r < a.Cout) ? a.bias[co + r] : 0.0f;
#pragma unroll
        for (int m = 0; m < CONV_MT; ++m) {
            if (!pok[m]) continue;
            const uint32_t p = pix[m];
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[m][s][r] + b[r];
                if (a.relu) v[r] = fmaxf(v[r], 0.0f);
            }
            const size_t o = (size_t)p * a.out_C + a.out_off + co;
            if (a.out_f32) {
                if (vec_ok && co + 4u <= a.Cout) {
                    *reinterpret_cast<float4*>(a.out_f32 + o) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
                    for (int r = 0; r < 4; ++r)
                        if (co + r < a.Cout) a.out_f32[o + r] = v[r];
                }
            } else {
                if (vec_ok && co + 4u <= a.Cout) {
                    bf16x4 q = {f2bf(v[0]), f2bf(v[1]), f2bf(v[2]), f2bf(v[3])};
                    *reinterpret_cast<bf16x4*>(a.out + o) = q;
                } else {
                    for (int r = 0; r < 4; ++r)
                        if (co + r < a.Cout) a.out[o + r] = f2bf(v[r]);
                }
            }
        }
    }
}

// NS 16-channel output tiles per wave (a compile-time count, so the
// accumulators stay in AGPRs): tiles s_base + blockIdx.y * NS onwards.
template <int NS>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a, uint32_t s_base) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t P = a.N * a.H * a.W;
    const uint32_t pix0 = wave * (16u * CONV_MT);
    if (pix0 >= P) return;
    const uint32_t s0 = s_base + blockIdx.y * NS;  // first 16-channel tile
    const uint32_t ns_total = a.Cout_pad >> 4;
    f32x4 acc[CONV_MT][NS];
#pragma unroll
    for (int m = 0; m < CONV_MT; ++m)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[m][s] = f32x4{0.f, 0.f, 0.f, 0.f};
    int py[CONV_MT], px[CONV_MT];
    long pbase[CONV_MT];
    bool pv[CONV_MT];
#pragma unroll
    for (int m = 0; m < CONV_MT; ++m) {
        uint32_t p = pix0 + m * 16u + (lane & 15u);
        pv[m] = p < P;
        uint32_t pp = pv[m] ? p : 0u;
        px[m] = (int)(pp % a.W);
        py[m] = (int)((pp / a.W) % a.H);
        pbase[m] = (long)pp * a.in_C + a.in_off;
    }
    // this lane's K position: k0 = kk + (lane / 16) * 8 -> (tap, ci)
    uint32_t ci = (lane >> 4) * 8u, tap = 0;
    while (ci >= a.Cin) {
        ci -= a.Cin;
        ++tap;
    }
    const bf16x8 zero = {};
    const bf16x8* wfrag = reinterpret_cast<const bf16x8*>(a.wt) + (size_t)s0 * 64u + lane;
    const size_t wstep = (size_t)ns_total * 64u;  // fragments per 32-wide K step
    for (uint32_t kk = 0; kk < a.Kpad; kk += 32u) {
        bf16x8 av[CONV_MT];
        const bool kval = tap < 9u;
        const int dy = (int)(tap / 3u) - 1, dx = (int)(tap % 3u) - 1;
        const long doff = ((long)dy * (long)a.W + dx) * (long)a.in_C + (long)ci;
#pragma unroll
        for (int m = 0; m < CONV_MT; ++m) {
            const int yy = py[m] + dy, xx = px[m] + dx;
            const bool ok = kval && pv[m] && yy >= 0 && yy < (int)a.H && xx >= 0 && xx < (int)a.W;
            av[m] = ok ? *reinterpret_cast<const bf16x8*>(a.in + pbase[m] + doff) : zero;
        }
        const bf16x8* wk = wfrag + (size_t)(kk >> 5) * wstep;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const bf16x8 bv = wk[s * 64];
#pragma unroll
            for (int m = 0; m < CONV_MT; ++m)
                acc[m][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv, av[m], acc[m][s], 0, 0, 0);
        }
        ci += 32u;
        while (ci >= a.Cin && tap < 9u) {
            ci -= a.Cin;
            ++tap;
        }
    }
    uint32_t pix[CONV_MT];
    bool pok[CONV_MT];
#pragma unroll
    for (int m = 0; m < CONV_MT; ++m) {
        pix[m] = pix0 + m * 16u + (lane & 15u);
        pok[m] = pix[m] < P;
    }
    conv_store<NS>(a, acc, pix, pok, s0, lane);
}

// The same convolution with the input staged in LDS: a workgroup computes an
// 8 x 32 pixel tile (wave w: rows 2w, 2w + 1) and first copies the 10 x 34
// input halo, all Cin channels, into LDS (zero outside the image; pixel
// stride Cin + 8 elements against bank conflicts), so the 9 taps read their
// fragments from LDS instead of re-reading global memory. Same K walk,
// fragments and MFMA order per K step as k_conv. TWO: the halo's channels from
// cin1 on come from in2 at half resolution (ConvArgs); LDS holds them once per
// half-resolution pixel (a 6 x 18 halo after the 10 x 34 one), and a tap reads
// the half-resolution pixel under its full-resolution one. KP: the kernel
// prediction in the epilogue (ConvArgs::kp_out): the logits pass through LDS,
// KP_STRIDE floats per pixel, where the halo tile was.
constexpr int LT_H = 8, LT_W = 32, HALO_H = LT_H + 2, HALO_W = LT_W + 2;
constexpr int HB_H = LT_H / 2 + 2, HB_W = LT_W / 2 + 2;  // half-resolution halo of a two-input layer
constexpr int HALO_BATCH = 8;
constexpr uint32_t KP_STRIDE = 28;  // logits 0..27 of a pixel (25 used), 16-byte aligned rows
constexpr size_t KP_LDS = 4u * 64u * KP_STRIDE * sizeof(float);

template <int NS, bool TWO, bool KP = false, bool DEEP = (NS > 2 || TWO)>
__global__ __launch_bounds__(256) void k_conv_lds(ConvArgs a, uint32_t tiles_x, uint32_t tiles_y, uint32_t s_base) {
    static_assert(!KP || (NS == 2 && !TWO), "kernel prediction: one launch of 32 logits, one input");
    extern __shared__ __align__(16) __bf16 tile[];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    // TWO: the halo of in (cin1 channels, stride ps) is followed at
    // base2 by the half-resolution halo of in2 (HB_H x HB_W pixels, Cin - cin1
    // channels, stride ps2)
    const uint32_t cin1 = TWO ? a.cin1 : a.Cin;
    const uint32_t ps = cin1 + 8u, ps2 = a.Cin - cin1 + 8u;
    const uint32_t base2 = (uint32_t)(HALO_H * HALO_W) * ps;
    uint32_t b = blockIdx.x;
    const uint32_t tx = b % tiles_x;
    b /= tiles_x;
    const uint32_t ty = b % tiles_y;
    const uint32_t n = b / tiles_y;
    const int x0 = (int)(tx * LT_W), y0 = (int)(ty * LT_H);
    const int H = (int)a.H, W = (int)a.W;
    const uint32_t cpp = cin1 >> 3, cpp2 = (a.Cin - cin1) >> 3;  // 16-byte chunks per pixel
    const uint32_t total1 = (uint32_t)(HALO_H * HALO_W) * cpp;
    const uint32_t total = total1 + (TWO ? (uint32_t)(HB_H * HB_W) * cpp2 : 0u);
    const bf16x8 zero = {};
    // HALO_BATCH chunks per thread in flight before they are written to LDS,
    // so the global latency is paid once per batch, not once per chunk
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += 256u * HALO_BATCH) {
        bf16x8 v[HALO_BATCH];
        uint32_t dst[HALO_BATCH];
#pragma unroll
        for (int j = 0; j < HALO_BATCH; ++j) {
            const uint32_t i = i0 + (uint32_t)j * 256u;
            v[j] = zero;
            dst[j] = ~0u;
            if (i < total1) {
                const uint32_t hp = i / cpp, ch = i - hp * cpp;
                const int y = y0 + (int)(hp / HALO_W) - 1, x = x0 + (int)(hp % HALO_W) - 1;
                if (y >= 0 && y < H && x >= 0 && x < W)
                    v[j] = *reinterpret_cast<const bf16x8*>(
                        a.in + ((size_t)(n * a.H + (uint32_t)y) * a.W + (uint32_t)x) * a.in_C + a.in_off + ch * 8u);
                dst[j] = hp * ps + ch * 8u;
            } else if (TWO && i < total) {
                // half-resolution pixel (y0 / 2 - 1 + row, x0 / 2 - 1 + col):
                // rows and columns of in2 that the nearest 2x upsampling of
                // the halo reads; zero outside in2's H / 2 x W / 2
                const uint32_t i2 = i - total1, hq = i2 / cpp2, ch = i2 - hq * cpp2;
                const int Y = (y0 >> 1) - 1 + (int)(hq / HB_W), X = (x0 >> 1) - 1 + (int)(hq % HB_W);
                if (Y >= 0 && Y < (H >> 1) && X >= 0 && X < (W >> 1))
                    v[j] = *reinterpret_cast<const bf16x8*>(
                        a.in2 + ((size_t)(n * (a.H >> 1) + (uint32_t)Y) * (a.W >> 1) + (uint32_t)X) * a.in2_C +
                        ch * 8u);
                dst[j] = base2 + hq * ps2 + ch * 8u;
            }
        }
#pragma unroll
        for (int j = 0; j < HALO_BATCH; ++j)
            if (dst[j] != ~0u) *reinterpret_cast<bf16x8*>(tile + dst[j]) = v[j];
    }
    __syncthreads();
    const uint32_t s0 = s_base + blockIdx.y * NS;
    const uint32_t ns_total = a.Cout_pad >> 4;
    f32x4 acc[CONV_MT][NS];
#pragma unroll
    for (int m = 0; m < CONV_MT; ++m)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[m][s] = f32x4{0.f, 0.f, 0.f, 0.f};
    int rbase[CONV_MT];  // LDS element offset of this lane's pixel (tap (1, 1) is the pixel itself)
    uint32_t pix[CONV_MT];
    bool pok[CONV_MT];
#pragma unroll
    for (int m = 0; m < CONV_MT; ++m) {
        const int ry = (int)(2u * w) + (m >> 1), rx = (m & 1) * 16 + (int)(lane & 15u);
        rbase[m] = ((ry + 1) * HALO_W + rx + 1) * (int)ps;
        const int y = y0 + ry, x = x0 + rx;
        pok[m] = y < H && x < W;
        pix[m] = pok[m] ? (n * a.H + (uint32_t)y) * a.W + (uint32_t)x : 0u;
    }
    uint32_t ci = (lane >> 4) * 8u, tap = 0;
    while (ci >= a.Cin) {
        ci -= a.Cin;
        ++tap;
    }
    const bf16x8* wfrag = reinterpret_cast<const bf16x8*>(a.wt) + (size_t)s0 * 64u + lane;
    const size_t wstep = (size_t)ns_total * 64u;
    const uint32_t nk = a.Kpad >> 5;  // 32-wide K steps
    // One K step: this lane's 8 input channels of its 4 pixels from LDS, then
    // CONV_MT x NS MFMAs with the step's weight fragments. Only the last step
    // can hold padding K (Kpad - 9 Cin < 32), so only it checks the tap.
    auto kstep = [&](const bf16x8 (&wv)[NS], auto last) {
        const bool kval = !decltype(last)::value || tap < 9u;
        const uint32_t tc = kval ? tap : 4u;  // padding K: any in-tile address, value unused
        const int dy = (int)(tc / 3u) - 1, dx = (int)(tc % 3u) - 1;
        const int doff = (dy * HALO_W + dx) * (int)ps + (int)ci;
        bf16x8 av[CONV_MT];
        if constexpr (TWO) {
            // channels from cin1 on: tap pixel (ry + dy, rx + dx) of the tile
            // reads half-resolution halo pixel ((ry + dy + 2) / 2, (rx + dx + 2) / 2)
            const bool second = ci >= cin1;
            const int rw = (int)(2u * w) + dy + 2, cl = (int)(lane & 15u) + dx + 2;
            const int b2 = (int)base2 + (int)ci - (int)cin1 + (cl >> 1) * (int)ps2;
            const int r0 = (rw >> 1) * HB_W * (int)ps2, r1 = ((rw + 1) >> 1) * HB_W * (int)ps2;
#pragma unroll
            for (int m = 0; m < CONV_MT; ++m) {
                const int o2 = b2 + ((m >> 1) ? r1 : r0) + (m & 1) * 8 * (int)ps2;
                const bf16x8 v = *reinterpret_cast<const bf16x8*>(tile + (second ? o2 : rbase[m] + doff));
                av[m] = kval ? v : zero;
            }
        } else {
#pragma unroll
            for (int m = 0; m < CONV_MT; ++m) {
                const bf16x8 v = *reinterpret_cast<const bf16x8*>(tile + (rbase[m] + doff));
                av[m] = kval ? v : zero;
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int m = 0; m < CONV_MT; ++m)
                acc[m][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv[s], av[m], acc[m][s], 0, 0, 0);
        ci += 32u;
        while (ci >= a.Cin && tap < 9u) {
            ci -= a.Cin;
            ++tap;
        }
    };
    // Weight fragments one K step ahead in two buffers used in turn, or two
    // steps ahead in three (DEEP: wider tiles, two-input layers and small
    // grids, where few waves per SIMD hide the load latency); no copy between
    // them, so waiting for one does not wait for the others. The
    // prefetch index is clamped to the last step: always in bounds, and the
    // value loaded past the end is not used.
    auto wload = [&](bf16x8 (&wv)[NS], uint32_t k) {
        const bf16x8* wk = wfrag + (size_t)min(k, nk - 1u) * wstep;
#pragma unroll
        for (int s = 0; s < NS; ++s) wv[s] = wk[s * 64];
    };
    using full = std::false_type;  // every K of the step is a real tap
    using last = std::true_type;   // the last step
    const uint32_t nmain = nk - 1u;
    uint32_t k = 0;
    bf16x8 wa[NS], wb[NS];
    wload(wa, 0u);
    if constexpr (DEEP) {
        bf16x8 wc[NS];
        wload(wb, 1u);
        for (; k + 3u <= nmain; k += 3u) {
            wload(wc, k + 2u);
            kstep(wa, full{});
            wload(wa, k + 3u);
            kstep(wb, full{});
            wload(wb, k + 4u);
            kstep(wc, full{});
        }
        // 0, 1 or 2 full steps left before the last (wa holds step k, wb step k + 1)
        if (k + 2u == nmain) {
            wload(wc, k + 2u);
            kstep(wa, full{});
            kstep(wb, full{});
            kstep(wc, last{});
        } else if (k + 1u == nmain) {
            kstep(wa, full{});
            kstep(wb, last{});
        } else {
            kstep(wa, last{});
        }
    } else {
        for (; k + 2u <= nmain; k += 2u) {
            wload(wb, k + 1u);
            kstep(wa, full{});
            wload(wa, k + 2u);
            kstep(wb, full{});
        }
        if (k < nmain) {  // one full step left before the last
            wload(wb, k + 1u);
            kstep(wa, full{});
            kstep(wb, last{});
        } else {
            kstep(wa, last{});
        }
    }
    if constexpr (KP) {
        // every wave is past its last halo read before the logits overwrite it
        __syncthreads();
        float* lg = reinterpret_cast<float*>(tile) + w * (64u * KP_STRIDE);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const uint32_t co = (s0 + s) * 16u + (lane >> 4) * 4u;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (co + r >= KP_STRIDE) continue;
                const float b = (a.bias && co + r < a.Cout) ? a.bias[co + r] : 0.0f;
#pragma unroll
                for (int m = 0; m < CONV_MT; ++m) {
                    float v = acc[m][s][r] + b;
                    if (a.relu) v = fmaxf(v, 0.0f);
                    lg[(m * 16u + (lane & 15u)) * KP_STRIDE + co + r] = v;
                }
            }
        }
        __syncthreads();
        // lane L: slot L of the wave = tile m = L / 16, column L % 16
        const uint32_t m = lane >> 4;
        const int y = y0 + (int)(2u * w + (m >> 1)), x = x0 + (int)((m & 1u) * 16u + (lane & 15u));
        if (y < (int)a.kp_H && x < (int)a.kp_W)
            kpcn_pixel(lg + lane * KP_STRIDE, a.kp_noisy, a.kp_out, x, y, a.kp_W, a.kp_H);
    } else {
        conv_store<NS>(a, acc, pix, pok, s0, lane);
    }
}

// ---- weight gradient --------------------------------------------------------------
// partial[chunk][co][k] = sum over the chunk's pixels p of dout[p][co] *
// in[p + offset(tap)][ci], k = tap * Cin + ci (zero outside the image).
// Block: 4 waves, tile co [64 by, +64) x k [64 bz, +64); wave w owns k columns
// [16 w, +16) for all 64 channels. LDS holds both 32-pixel tiles transposed
// (channel-major), row stride 40 elements (80 B) to spread the banks.
constexpr uint32_t WG_LDS = 40;

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[64 * WG_LDS];
    __shared__ __attribute__((aligned(16))) __bf16 sB[64 * WG_LDS];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t chunk = blockIdx.x;
    const uint32_t co_base = blockIdx.y * 64u, k_base = blockIdx.z * 64u;
    const uint32_t P = a.N * a.H * a.W;
    const uint32_t p_begin = chunk * a.chunk, p_end = min(P, p_begin + a.chunk);
    // loader roles: pixel row t / 8 of the tile, 8 channels at (t % 8) * 8
    const uint32_t lp = t >> 3, l8 = (t & 7u) * 8u;
    const uint32_t lco = co_base + l8;
    const bool co_ok = lco < a.Cout_pad;
    const uint32_t lk = k_base + l8;
    const bool k_ok = lk < a.K;
    const uint32_t ltap = k_ok ? lk / a.Cin : 0u, lci = k_ok ? lk - ltap * a.Cin : 0u;
    const int ldy = (int)(ltap / 3u) - 1, ldx = (int)(ltap % 3u) - 1;
    f32x4 acc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[s] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.0f;
    const bool do_bias = blockIdx.z == 0 && t < 64u && co_base + t < a.Cout_pad;
    const bf16x8 zero = {};
    for (uint32_t p0 = p_begin; p0 < p_end; p0 += 32u) {
        const uint32_t p = p0 + lp;
        bf16x8 va = zero, vb = zero;
        if (p < p_end) {
            if (co_ok) va = *reinterpret_cast<const bf16x8*>(a.dout + (size_t)p * a.dout_C + a.dout_off + lco);
            if (k_ok) {
                const uint32_t x = p % a.W, r = p / a.W, y = r % a.H;
                const int yy = (int)y + ldy, xx = (int)x + ldx;
                if (yy >= 0 && yy < (int)a.H && xx >= 0 && xx < (int)a.W)
                    vb = *reinterpret_cast<const bf16x8*>(
                        a.in + ((long)p + (long)ldy * (long)a.W + ldx) * (long)a.in_C + a.in_off + lci);
            }
        }
        __syncthreads();  // previous tile consumed
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sA[(l8 + j) * WG_LDS + lp] = va[j];
            sB[(l8 + j) * WG_LDS + lp] = vb[j];
        }
        __syncthreads();
        const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(&sB[(w * 16u + (lane & 15u)) * WG_LDS + (lane >> 4) * 8u]);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 afr = *reinterpret_cast<const bf16x8*>(&sA[(s * 16u + (lane & 15u)) * WG_LDS + (lane >> 4) * 8u]);
            acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[s], 0, 0, 0);
        }
        if (do_bias) {
#pragma unroll
            for (int j = 0; j < 32; ++j) bsum += bf2f(sA[t * WG_LDS + j]);
        }
    }
    float* out = a.partial + (size_t)chunk * a.Cout_pad * a.Kpad;
    const uint32_t col = k_base + w * 16u + (lane & 15u);
    if (col < a.Kpad) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = co_base + s * 16u + (lane >> 4) * 4u + r;
                if (row < a.Cout_pad) out[(size_t)row * a.Kpad + col] = acc[s][r];
            }
        }
    }
    if (do_bias) a.bpartial[(size_t)chunk * a.Cout_pad + co_base + t] = bsum;
}

// out[i] = sum over chunks (in order) of partial[c][i]
__global__ void k_reduce(const float* partial, uint32_t chunks, uint32_t n, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) s += partial[(size_t)c * n + i];
    out[i] = s;
}

// ---- elementwise -------------------------------------------------------------------
// g[p][c] = 0 where act[p][c] <= 0 (ReLU backward on channel slices)
__global__ void k_relu_bwd(__bf16* g, uint32_t g_C, uint32_t g_off, const __bf16* act, uint32_t a_C, uint32_t a_off,
                           uint32_t C, uint32_t P) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * C) return;
    size_t p = i / C, c = i - p * C;
    if (!(bf2f(act[p * a_C + a_off + c]) > 0.0f)) g[p * g_C + g_off + c] = f2bf(0.0f);
}

// 2x2 average pool of a channel slice (H, W) -> dense (H/2, W/2) x C
__global__ void k_pool(const __bf16* in, uint32_t in_C, uint32_t in_off, __bf16* out, uint32_t C, uint32_t N,
                       uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)N * Ho * Wo * C;
    if (i >= total) return;
    uint32_t c = i % C;
    size_t q = i / C;
    uint32_t x = q % Wo, y = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    uint32_t Wi = Wo * 2, Hi = Ho * 2;
    const __bf16* b = in + (((size_t)n * Hi + 2 * y) * Wi + 2 * x) * in_C + in_off + c;
    float s = (bf2f(b[0]) + bf2f(b[in_C])) + (bf2f(b[(size_t)Wi * in_C]) + bf2f(b[(size_t)Wi * in_C + in_C]));
    out[i] = f2bf(0.25f * s);
}
// gradient of the pool, added into a channel slice of the input gradient
__global__ void k_pool_bwd_add(const __bf16* gout, __bf16* gin, uint32_t gin_C, uint32_t gin_off, uint32_t C,
                               uint32_t N, uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t Hi = Ho * 2, Wi = Wo * 2;
    size_t total = (size_t)N * Hi * Wi * C;
    if (i >= total) return;
    uint32_t c = i % C;
    size_t q = i / C;
    uint32_t x = q % Wi, y = (q / Wi) % Hi, n = q / ((size_t)Wi * Hi);
    float g = 0.25f * bf2f(gout[(((size_t)n * Ho + y / 2) * Wo + x / 2) * C + c]);
    size_t o = q * gin_C + gin_off + c;
    gin[o] = f2bf(bf2f(gin[o]) + g);
}
// nearest 2x upsample: dense (H/2, W/2) x C -> channel slice of (H, W)
__global__ void k_up(const __bf16* in, __bf16* out, uint32_t out_C, uint32_t out_off, uint32_t C, uint32_t N,
                     uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)N * Ho * Wo * C;
    if (i >= total) return;
    uint32_t c = i % C;
    size_t q = i / C;
    uint32_t x = q % Wo, y = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    uint32_t Hi = Ho / 2, Wi = Wo / 2;
    out[q * out_C + out_off + c] = in[(((size_t)n * Hi + y / 2) * Wi + x / 2) * C + c];
}
// gradient of the upsample: sum of the 2x2 block of the slice gradient
__global__ void k_up_bwd(const __bf16* gout, uint32_t g_C, uint32_t g_off, __bf16* gin, uint32_t C, uint32_t N,
                         uint32_t Hi, uint32_t Wi) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = (size_t)N * Hi * Wi * C;
    if (i >= total) return;
    uint32_t c = i % C;
    size_t q = i / C;
    uint32_t x = q % Wi, y = (q / Wi) % Hi, n = q / ((size_t)Wi * Hi);
    uint32_t Wo = Wi * 2, Ho = Hi * 2;
    const __bf16* b = gout + (((size_t)n * Ho + 2 * y) * Wo + 2 * x) * g_C + g_off + c;
    float s = (bf2f(b[0]) + bf2f(b[g_C])) + (bf2f(b[(size_t)Wo * g_C]) + bf2f(b[(size_t)Wo * g_C + g_C]));
    gin[i] = f2bf(s);
}

// The same four operations on 8 channels per thread (16-byte loads and
// stores; per element the arithmetic above), for channel counts and offsets
// that are multiples of 8.
__global__ void k_pool8(const __bf16* in, uint32_t in_C, uint32_t in_off, __bf16* out, uint32_t C, uint32_t N,
                        uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t C8 = C >> 3;
    size_t total = (size_t)N * Ho * Wo * C8;
    if (i >= total) return;
    uint32_t c = (uint32_t)(i % C8) * 8u;
    size_t q = i / C8;
    uint32_t x = q % Wo, y = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    uint32_t Wi = Wo * 2, Hi = Ho * 2;
    const __bf16* b = in + (((size_t)n * Hi + 2 * y) * Wi + 2 * x) * in_C + in_off + c;
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(b), a1 = *reinterpret_cast<const bf16x8*>(b + in_C);
    const bf16x8 a2 = *reinterpret_cast<const bf16x8*>(b + (size_t)Wi * in_C);
    const bf16x8 a3 = *reinterpret_cast<const bf16x8*>(b + (size_t)Wi * in_C + in_C);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f2bf(0.25f * ((bf2f(a0[k]) + bf2f(a1[k])) + (bf2f(a2[k]) + bf2f(a3[k]))));
    *reinterpret_cast<bf16x8*>(out + q * C + c) = o;
}
__global__ void k_pool_bwd_add8(const __bf16* gout, __bf16* gin, uint32_t gin_C, uint32_t gin_off, uint32_t C,
                                uint32_t N, uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t Hi = Ho * 2, Wi = Wo * 2;
    const uint32_t C8 = C >> 3;
    size_t total = (size_t)N * Hi * Wi * C8;
    if (i >= total) return;
    uint32_t c = (uint32_t)(i % C8) * 8u;
    size_t q = i / C8;
    uint32_t x = q % Wi, y = (q / Wi) % Hi, n = q / ((size_t)Wi * Hi);
    const bf16x8 g = *reinterpret_cast<const bf16x8*>(gout + (((size_t)n * Ho + y / 2) * Wo + x / 2) * C + c);
    bf16x8* d = reinterpret_cast<bf16x8*>(gin + q * gin_C + gin_off + c);
    bf16x8 v = *d;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = f2bf(bf2f(v[k]) + 0.25f * bf2f(g[k]));
    *d = v;
}
__global__ void k_up8(const __bf16* in, __bf16* out, uint32_t out_C, uint32_t out_off, uint32_t C, uint32_t N,
                      uint32_t Ho, uint32_t Wo) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t C8 = C >> 3;
    size_t total = (size_t)N * Ho * Wo * C8;
    if (i >= total) return;
    uint32_t c = (uint32_t)(i % C8) * 8u;
    size_t q = i / C8;
    uint32_t x = q % Wo, y = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    uint32_t Hi = Ho / 2, Wi = Wo / 2;
    *reinterpret_cast<bf16x8*>(out + q * out_C + out_off + c) =
        *reinterpret_cast<const bf16x8*>(in + (((size_t)n * Hi + y / 2) * Wi + x / 2) * C + c);
}
__global__ void k_up_bwd8(const __bf16* gout, uint32_t g_C, uint32_t g_off, __bf16* gin, uint32_t C, uint32_t N,
                          uint32_t Hi, uint32_t Wi) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t C8 = C >> 3;
    size_t total = (size_t)N * Hi * Wi * C8;
    if (i >= total) return;
    uint32_t c = (uint32_t)(i % C8) * 8u;
    size_t q = i / C8;
    uint32_t x = q % Wi, y = (q / Wi) % Hi, n = q / ((size_t)Wi * Hi);
    uint32_t Wo = Wi * 2, Ho = Hi * 2;
    const __bf16* b = gout + (((size_t)n * Ho + 2 * y) * Wo + 2 * x) * g_C + g_off + c;
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(b), a1 = *reinterpret_cast<const bf16x8*>(b + g_C);
    const bf16x8 a2 = *reinterpret_cast<const bf16x8*>(b + (size_t)Wo * g_C);
    const bf16x8 a3 = *reinterpret_cast<const bf16x8*>(b + (size_t)Wo * g_C + g_C);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f2bf((bf2f(a0[k]) + bf2f(a1[k])) + (bf2f(a2[k]) + bf2f(a3[k])));
    *reinterpret_cast<bf16x8*>(gin + q * C + c) = o;
}

// input features of image n (16 channels, 9 used; zero outside the frame):
// albedo, normal, distance to the eye / 100, roughness, log(1 + luminance)
__global__ void k_features(FeatArgs a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Hp * a.Wp) return;
    uint32_t x = i % a.Wp, y = i / a.Wp;
    float f[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) f[k] = 0.0f;
    if (x < a.W && y < a.H) {
        size_t p = (size_t)y * a.W + x;
        float4 pos = a.pos[p], nrm = a.nrm[p], alb = a.alb[p], mat = a.mat[p], col = a.noisy[p];
        f[0] = alb.x;
        f[1] = alb.y;
        f[2] = alb.z;
        f[3] = nrm.x;
        f[4] = nrm.y;
        f[5] = nrm.z;
        float dx = pos.x - a.cam[0], dy = pos.y - a.cam[1], dz = pos.z - a.cam[2];
        f[6] = (nrm.w == 0.0f) ? 0.0f : sqrtf((dx * dx + dy * dy) + dz * dz) * 0.01f;  // background: 0
        f[7] = mat.y;
        float lum = (0.2126f * col.x + 0.7152f * col.y) + 0.0722f * col.z;
        f[8] = logf(1.0f + fmaxf(lum, 0.0f));
    }
    bf16x8 lo, hi;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        lo[k] = f2bf(f[k]);
        hi[k] = f2bf(f[8 + k]);
    }
    bf16x8* o = reinterpret_cast<bf16x8*>(a.x0 + ((size_t)n * a.Hp * a.Wp + i) * 16);
    o[0] = lo;
    o[1] = hi;
}

__global__ void k_kpcn(KpcnArgs a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.W * a.H) return;
    int x = (int)(i % a.W), y = (int)(i / a.W);
    kpcn_pixel(a.logits + ((size_t)n * a.Hp * a.Wp + (size_t)y * a.Wp + x) * 32, a.noisy, a.out, x, y, a.W, a.H);
}
// loss = sum over pixels and colour channels of |T(out) - T(target)|, T(v) =
// max(v, 0) / (1 + max(v, 0)) (a.loss 0), or of (out - target)^2 / (target^2
// + 0.01) (a.loss 1, relative MSE); writes the logits gradient times scale
// and one partial loss sum per block
__global__ __launch_bounds__(256) void k_kpcn_bwd(KpcnArgs a, uint32_t n, float scale, float* loss_partial) {
    __shared__ float red[256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.0f;
    if (i < a.W * a.H) {
        int x = (int)(i % a.W), y = (int)(i / a.W);
        size_t lp = ((size_t)n * a.Hp * a.Wp + (size_t)y * a.Wp + x) * 32;
        float w[25];
        softmax25(a.logits + lp, w);
        float o[3] = {0.f, 0.f, 0.f};
        for (int k = 0; k < 25; ++k) {
            float4 c = tap_color(a.noisy, x, y, k, a.W, a.H);
            o[0] += w[k] * c.x;
            o[1] += w[k] * c.y;
            o[2] += w[k] * c.z;
        }
        float4 t = a.target[i];
        float tv[3] = {t.x, t.y, t.z};
        float go[3];
        for (int ch = 0; ch < 3; ++ch) {
            if (a.loss == 1u) {
                // relative MSE: (o - t)^2 / (t^2 + 0.01)
                float den = tv[ch] * tv[ch] + 0.01f;
                float d = o[ch] - tv[ch];
                lsum += d * d / den;
                go[ch] = 2.0f * d / den * scale;
            } else {
                float ov = fmaxf(o[ch], 0.0f), tt = fmaxf(tv[ch], 0.0f);
                float d = ov / (1.0f + ov) - tt / (1.0f + tt);
                lsum += fabsf(d);
                float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
                go[ch] = o[ch] > 0.0f ? sg * scale / ((1.0f + ov) * (1.0f + ov)) : 0.0f;
            }
        }
        float dw[25], dot = 0.0f;
        for (int k = 0; k < 25; ++k) {
            float4 c = tap_color(a.noisy, x, y, k, a.W, a.H);
            dw[k] = (go[0] * c.x + go[1] * c.y) + go[2] * c.z;
            dot += w[k] * dw[k];
        }
        __bf16* gl = a.dlogits + lp;
        for (int k = 0; k < 25; ++k) gl[k] = f2bf(w[k] * (dw[k] - dot));
        for (int k = 25; k < 32; ++k) gl[k] = f2bf(0.0f);
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}
__global__ void k_fill_bf16(__bf16* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = f2bf(0.0f);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_adam(AdamArgs a) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    float g = a.grad[i];
    float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    float v = a.b2 * a.v[i] + (1.0f - a.b2) * (g * g);
    a.m[i] = m;
    a.v[i] = v;
    float mh = m / a.bc1, vh = v / a.bc2;
    a.w[i] = a.w[i] - a.lr * mh / (sqrtf(vh) + a.eps);
}
// BF16 copies of one layer: forward weights, and input-gradient weights
// W'[ci][tap' * Cout_pad + co] = W[co][(8 - tap') * Cin + ci]
__global__ void k_pack_weights(PackArgs a) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)a.Cout_pad * a.Kpad) return;
    uint32_t co = i / a.Kpad, k = i % a.Kpad;
    __bf16 w = f2bf(a.w[i]);
    a.wf[frag_index(co, k, a.Cout_pad)] = w;
    if (k < 9 * a.Cin) {
        uint32_t tap = k / a.Cin, ci = k % a.Cin;
        a.wd[frag_index(ci, (8 - tap) * a.Cout_pad + co, a.Cin)] = w;
    }
}

}  // namespace dn

// ---- launchers ------------------------------------------------------------------------
static inline uint32_t nblk(size_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// One launch per tile count: CONV_NT tiles per wave over the first
// (Cout_pad / 16) / CONV_NT groups, the remaining 1..3 tiles in a second launch.
template <int NS>
static void conv_launch(const ConvArgs& a, uint32_t ny, uint32_t s_base, bool lds_path, size_t lds, hipStream_t st) {
    if (lds_path) {
        const uint32_t tx = (a.W + dn::LT_W - 1) / dn::LT_W, ty = (a.H + dn::LT_H - 1) / dn::LT_H;
        if (a.in2) dn::k_conv_lds<NS, true><<<dim3(a.N * tx * ty, ny), 256, lds, st>>>(a, tx, ty, s_base);
        else dn::k_conv_lds<NS, false><<<dim3(a.N * tx * ty, ny), 256, lds, st>>>(a, tx, ty, s_base);
    } else {
        const size_t P = (size_t)a.N * a.H * a.W;
        dn::k_conv<NS><<<dim3(nblk(P, 4 * 16 * dn::CONV_MT), ny), 256, 0, st>>>(a, s_base);
    }
}

hipError_t dn_conv(const ConvArgs& a, hipStream_t st) {
    // LDS-staged halo when the tile fits in 96 KiB (up to 128 input channels;
    // gfx950 has 160 KiB per CU) and 16-byte chunks stay aligned
    // (two inputs: the halo of in's cin1 channels, then in2's half-resolution
    // halo of the other Cin - cin1)
    if (a.in2 && (a.cin1 == 0u || a.cin1 >= a.Cin)) return hipErrorInvalidValue;
    const uint32_t cin1 = a.in2 ? a.cin1 : a.Cin;
    const size_t lds = ((size_t)dn::HALO_H * dn::HALO_W * (cin1 + 8u) +
                        (a.in2 ? (size_t)dn::HB_H * dn::HB_W * (a.Cin - cin1 + 8u) : 0u)) *
                       sizeof(__bf16);
    const bool lds_path = lds <= 98304u && (a.Cin & 7u) == 0 && (a.in_C & 7u) == 0 && (a.in_off & 7u) == 0;
    if (a.in2 && (!lds_path || (a.cin1 & 7u) || (a.in2_C & 7u) || (a.H & 1u) || (a.W & 1u)))
        return hipErrorInvalidValue;  // two inputs: LDS path, 8-channel chunks, even sizes
    const uint32_t ns_total = a.Cout_pad >> 4;
    if (a.kp_out) {
        // fused kernel prediction: one image, 25..32 logits, the LDS path
        if (!lds_path || a.in2 || a.N != 1u || ns_total != 2u || a.Cout < 25u || !a.kp_noisy || a.kp_W > a.W ||
            a.kp_H > a.H)
            return hipErrorInvalidValue;
        const uint32_t tx = (a.W + dn::LT_W - 1) / dn::LT_W, ty = (a.H + dn::LT_H - 1) / dn::LT_H;
        dn::k_conv_lds<2, false, true><<<tx * ty, 256, (lds > dn::KP_LDS ? lds : dn::KP_LDS), st>>>(a, tx, ty, 0u);
        return hipGetLastError();
    }
    if (lds_path && !a.in2 && ns_total >= 4u && (ns_total & 1u) == 0u) {
        // small grids (the 1/8 and 1/16 resolution layers: 36-136 tiles for
        // 256 CUs): two 16-channel tiles per wave over twice the workgroups
        // (not for two inputs, whose halo copy costs more than it gains)
        const uint32_t tx = (a.W + dn::LT_W - 1) / dn::LT_W, ty = (a.H + dn::LT_H - 1) / dn::LT_H;
        if (a.N * tx * ty < 256u) {
            dn::k_conv_lds<2, false, false, true><<<dim3(a.N * tx * ty, ns_total / 2u), 256, lds, st>>>(a, tx, ty, 0u);
            return hipGetLastError();
        }
    }
    const uint32_t full = ns_total / dn::CONV_NT, rem = ns_total % dn::CONV_NT;
    static_assert(dn::CONV_NT == 4, "remainder launches below cover 1..3 tiles");
    if (full) conv_launch<4>(a, full, 0u, lds_path, lds, st);
    const uint32_t sb = full * dn::CONV_NT;
    if (rem == 1) conv_launch<1>(a, 1u, sb, lds_path, lds, st);
    if (rem == 2) conv_launch<2>(a, 1u, sb, lds_path, lds, st);
    if (rem == 3) conv_launch<3>(a, 1u, sb, lds_path, lds, st);
    return hipGetLastError();
}
uint32_t dn_wgrad_chunks(uint32_t P, uint32_t* chunk) {
    uint32_t c = (P + 255u) / 256u;
    c = (c + 31u) & ~31u;
    if (c < 512u) c = 512u;
    *chunk = c;
    return (P + c - 1) / c;
}
hipError_t dn_wgrad(const WgradArgs& a, float* wgrad, float* bgrad, hipStream_t st) {
    uint32_t P = a.N * a.H * a.W;
    uint32_t chunks = (P + a.chunk - 1) / a.chunk;
    dim3 grid(chunks, (a.Cout_pad + 63) / 64, (a.Kpad + 63) / 64);
    dn::k_wgrad<<<grid, 256, 0, st>>>(a);
    dn::k_reduce<<<nblk((size_t)a.Cout_pad * a.Kpad, 256), 256, 0, st>>>(a.partial, chunks, a.Cout_pad * a.Kpad, wgrad);
    dn::k_reduce<<<nblk(a.Cout_pad, 256), 256, 0, st>>>(a.bpartial, chunks, a.Cout_pad, bgrad);
    return hipGetLastError();
}
hipError_t dn_relu_bwd(__bf16* g, uint32_t g_C, uint32_t g_off, const __bf16* act, uint32_t a_C, uint32_t a_off,
                       uint32_t C, uint32_t P, hipStream_t st) {
    dn::k_relu_bwd<<<nblk((size_t)P * C, 256), 256, 0, st>>>(g, g_C, g_off, act, a_C, a_off, C, P);
    return hipGetLastError();
}
hipError_t dn_pool(const __bf16* in, uint32_t in_C, uint32_t in_off, __bf16* out, uint32_t C, uint32_t N, uint32_t Ho,
                   uint32_t Wo, hipStream_t st) {
    if (((C | in_C | in_off) & 7u) == 0)
        dn::k_pool8<<<nblk((size_t)N * Ho * Wo * (C >> 3), 256), 256, 0, st>>>(in, in_C, in_off, out, C, N, Ho, Wo);
    else
        dn::k_pool<<<nblk((size_t)N * Ho * Wo * C, 256), 256, 0, st>>>(in, in_C, in_off, out, C, N, Ho, Wo);
    return hipGetLastError();
}
hipError_t dn_pool_bwd_add(const __bf16* gout, __bf16* gin, uint32_t gin_C, uint32_t gin_off, uint32_t C, uint32_t N,
                           uint32_t Ho, uint32_t Wo, hipStream_t st) {
    if (((C | gin_C | gin_off) & 7u) == 0)
        dn::k_pool_bwd_add8<<<nblk((size_t)N * Ho * Wo * 4 * (C >> 3), 256), 256, 0, st>>>(gout, gin, gin_C, gin_off, C,
                                                                                          N, Ho, Wo);
    else
        dn::k_pool_bwd_add<<<nblk((size_t)N * Ho * Wo * 4 * C, 256), 256, 0, st>>>(gout, gin, gin_C, gin_off, C, N, Ho,
                                                                                  Wo);
    return hipGetLastError();
}
hipError_t dn_up(const __bf16* in, __bf16* out, uint32_t out_C, uint32_t out_off, uint32_t C, uint32_t N, uint32_t Ho,
                 uint32_t Wo, hipStream_t st) {
    if (((C | out_C | out_off) & 7u) == 0)
        dn::k_up8<<<nblk((size_t)N * Ho * Wo * (C >> 3), 256), 256, 0, st>>>(in, out, out_C, out_off, C, N, Ho, Wo);
    else
        dn::k_up<<<nblk((size_t)N * Ho * Wo * C, 256), 256, 0, st>>>(in, out, out_C, out_off, C, N, Ho, Wo);
    return hipGetLastError();
}
hipError_t dn_up_bwd(const __bf16* gout, uint32_t g_C, uint32_t g_off, __bf16* gin, uint32_t C, uint32_t N, uint32_t Hi,
                     uint32_t Wi, hipStream_t st) {
    if (((C | g_C | g_off) & 7u) == 0)
        dn::k_up_bwd8<<<nblk((size_t)N * Hi * Wi * (C >> 3), 256), 256, 0, st>>>(gout, g_C, g_off, gin, C, N, Hi, Wi);
    else
        dn::k_up_bwd<<<nblk((size_t)N * Hi * Wi * C, 256), 256, 0, st>>>(gout, g_C, g_off, gin, C, N, Hi, Wi);
    return hipGetLastError();
}
hipError_t dn_features(const FeatArgs& a, uint32_t n, hipStream_t st) {
    dn::k_features<<<nblk((size_t)a.Hp * a.Wp, 256), 256, 0, st>>>(a, n);
    return hipGetLastError();
}
hipError_t dn_kpcn(const KpcnArgs& a, uint32_t n, hipStream_t st) {
    dn::k_kpcn<<<nblk((size_t)a.W * a.H, 256), 256, 0, st>>>(a, n);
    return hipGetLastError();
}
uint32_t dn_kpcn_blocks(uint32_t W, uint32_t H) { return nblk((size_t)W * H, 256); }
hipError_t dn_kpcn_bwd(const KpcnArgs& a, uint32_t n, float scale, float* loss_partial, hipStream_t st) {
    dn::k_kpcn_bwd<<<dn_kpcn_blocks(a.W, a.H), 256, 0, st>>>(a, n, scale, loss_partial);
    return hipGetLastError();
}
hipError_t dn_fill_bf16(__bf16* p, size_t n, hipStream_t st) {
    if (n) dn::k_fill_bf16<<<nblk(n, 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}
hipError_t dn_adam(const AdamArgs& a, hipStream_t st) {
    dn::k_adam<<<nblk(a.n, 256), 256, 0, st>>>(a);
    return hipGetLastError();
}
hipError_t dn_pack(const PackArgs& a, hipStream_t st) {
    dn::k_pack_weights<<<nblk((size_t)a.Cout_pad * a.Kpad, 256), 256, 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace hr
