// internal.h - data layouts in HBM and launcher interfaces of libhr.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "hr/hr.h"

namespace hr {

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t MAX_DEPTH = 64;       // traversal stack bound (checked at build)
constexpr uint32_t STACK_LDS = 16;       // entries per lane kept in LDS
constexpr uint32_t STACK_SPILL = MAX_DEPTH - STACK_LDS;  // entries per lane in private memory
constexpr uint32_t LEAF_MAX_LIMIT = 16;  // leaf code holds count - 1 in 4 bits
constexpr uint32_t MAX_TRIS = 1u << 27;  // leaf code holds the first triangle in 27 bits
constexpr uint32_t SHADOW_QUEUES = 8;    // shadow work bands, one per XCD
// Band b's counter is word b * BAND_STRIDE of its buffer: 256 B apart, since
// device-scope atomics are performed at memory and serialise per cache line.
constexpr uint32_t BAND_STRIDE = 64;
constexpr size_t BAND_COUNTER_BYTES = (size_t)SHADOW_QUEUES * BAND_STRIDE * 4u;

__host__ __device__ inline uint32_t leaf_code(uint32_t first, uint32_t count) { return LEAF_BIT | ((count - 1u) << 27) | first; }

// Traversal node, 64 B, one per Karras internal node:
//   a = (child 0 box min, child 0 code)   b = (child 0 box max, child 1 code)
//   c = (child 1 box min, 0)              d = (child 1 box max, 0)
// A code below LEAF_BIT is an internal node index; a leaf code is
// LEAF_BIT | (count - 1) << 27 | first sorted triangle.
struct __align__(64) TNode {
    float4 a, b, c, d;
};
// 4-wide traversal node, 128 B, one per wide root (slot = the binary node's
// index): child k's box in component k of lo/hi x/y/z and its code in code[k]
// (an internal slot, a leaf code, or EMPTY4 with an inverted box). Built from
// the binary nodes top-down: a wide node starts from its binary node's two
// children and twice opens the internal child with the largest box area. The
// timed traversal kernels walk it, the counting kernels walk the binary one.
constexpr uint32_t EMPTY4 = 0xFFFFFFFFu;
// depth buckets of the level-by-level box refit (a tree that passes the depth
// check has internal nodes at most MAX_DEPTH + 16 edges deep)
constexpr uint32_t REFIT_LEVELS = 128;
// 3 pushes per 4-wide level; a wide level descends at least one binary level
constexpr uint32_t MAX_STACK4 = 3u * MAX_DEPTH + 1u;
struct __align__(64) TNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    uint4 code;
    uint4 pad;
};
// Sorted world-space triangle, 48 B: (v0, triangle id), (v1, 0), (v2, 0).
struct __align__(16) STri {
    float4 v0, v1, v2;
};
// Per primitive constants for shading, 64 B.
struct __align__(16) PrimInfo {
    uint32_t tri_offset;   // first global triangle id of the primitive
    uint32_t index_base;   // offset of its first index in the concatenated index array
    uint32_t vertex_base;  // offset of its mesh's first vertex in the concatenated vertex arrays
    int32_t material;
    float nm[9];           // transpose(inverse(mat3(model))), row-major
    uint32_t _pad[3];
};
static_assert(sizeof(TNode) == 64, "TNode");
static_assert(sizeof(TNode4) == 128, "TNode4");
static_assert(sizeof(STri) == 48, "STri");
static_assert(sizeof(PrimInfo) == 64, "PrimInfo");

// One material texture inside the scene's texel array (RGBA8 packed in a
// uint32, red in the low byte).
struct TexInfo {
    uint64_t offset;  // first texel
    uint32_t w, h;
};
static_assert(sizeof(TexInfo) == 16, "TexInfo");

// Read-only view of a built scene, passed to kernels by value.
struct DevScene {
    const TNode* nodes;
    const TNode4* nodes4;
    const STri* tris;
    const uint32_t* tri_prim;   // global triangle id -> primitive
    const PrimInfo* prims;
    const uint32_t* indices;    // concatenated mesh indices
    const float4* vnormal;      // concatenated per-vertex normals (xyz, 0)
    const uint8_t* verts;       // concatenated hr_vertex records (44 B: colour at 24, uv at 36)
    const float4* wtri;         // world-space triangle vertices, 3 per triangle id
    const hr_material* mats;
    const float4* sky;
    const float4* irr;
    const uint32_t* texels;     // material textures, RGBA8
    const TexInfo* tex;
    uint32_t node_count, tri_count, nmat, prim_count;  // node_count: binary internal nodes (1 for one triangle)
    uint32_t wide_count;                               // 4-wide nodes
    uint32_t sky_w, sky_h, irr_w, irr_h;
    uint32_t ntex;
};

// Per-frame uniforms, passed to kernels by value (replaces the camera, light,
// samples UBO/SSBOs of renderer.cpp:1943-2028 and 2946-3123).
struct FrameArgs {
    hr_light lights[HR_MAX_LIGHTS];
    hr_light_shape shapes[HR_MAX_LIGHTS];
    float inv_view[16];
    float inv_proj[16];
    uint32_t nl;
    uint32_t frame;
    uint32_t width, height;
    uint32_t samples, spp, max_depth;
    uint32_t has_shapes;
    uint32_t adaptive;  // HR_FLAG_ADAPTIVE
    uint32_t unorm8;    // UNORM8_SHADOW | UNORM8_LIT: the reference's 8-bit image formats
};
constexpr uint32_t UNORM8_SHADOW = 1u;  // shadow buffers are R8_UNORM
constexpr uint32_t UNORM8_LIT = 2u;     // lit is B8G8R8A8_UNORM

struct GBufView {
    float4* pos;
    float4* nrm;
    float4* alb;
    float4* mat;
    float4* emi;
};

// counters slots (uint64)
enum : int {
    C_PRIMARY = 0,
    C_SHADOW = 1,
    C_REFL = 2,
    C_HITSHADOW = 3,
    C_OCCLUDED = 4,
    C_NODES = 5,
    C_TRIS = 6,
    C_OVERFLOW = 7,
    C_COUNT = 8
};
// Counter k is word k * C_STRIDE of the counter buffer (256 B apart): every
// wave of every kernel adds its totals at its end, and device-scope atomics
// to one cache line serialise.
constexpr uint32_t C_STRIDE = 32;

// ---- launchers (trace.hip) ---------------------------------------------------
hipError_t launch_gbuffer(const DevScene& s, const FrameArgs& f, GBufView gb, unsigned long long* counters,
                          bool count, hipStream_t st);
// One frame of a shadow launch, in device memory: the launch's groups
// (pixel, light) are numbered frame after frame from group_base.
struct ShadowJob {
    DevScene S;
    FrameArgs F;
    GBufView G;
    float* shadow;
    uint32_t group_base;
    uint32_t tile_w;   // pixel tile of one wave: tile_w x (64 / tile_w)
    uint32_t tiles_x;  // tiles per row
    uint32_t _pad;
};
uint32_t shadow_groups(const FrameArgs& f, uint32_t tile_w);
// writes `job` to device memory at dst (16-byte aligned) from a kernel argument
hipError_t launch_store_job(const ShadowJob& job, ShadowJob* dst, hipStream_t st);
// Frames with power-of-two sample counts in 4..64 (with adaptive sampling the
// penumbra counts), all adaptive or none: one lane per (group, sample) item,
// 2^lg2s = the largest count. scratch: device memory of
// shadows_spl_scratch() bytes.
size_t shadows_spl_scratch(uint32_t ngroups, uint32_t lg2s, bool adaptive);
// node_count: the 4-wide node slots of a single frame's scene (to choose
// 32-bit node addressing), ignored for several frames
hipError_t launch_shadows_spl(const ShadowJob* jobs_dev, uint32_t njobs, uint32_t ngroups, uint32_t lg2s,
                              bool adaptive, unsigned long long* counters, hipStream_t st, uint32_t* next_item,
                              uint32_t* scratch, uint32_t resident_blocks, uint32_t wide_count);
// jobs_dev: njobs ShadowJob in device memory; next_group: BAND_COUNTER_BYTES
// of device memory holding the SHADOW_QUEUES band counters, BAND_STRIDE words
// apart (reset by the launch); resident_blocks: the persistent grid size
hipError_t launch_shadows(const ShadowJob* jobs_dev, uint32_t njobs, uint32_t ngroups, unsigned long long* counters,
                          bool count, hipStream_t st, uint32_t* next_group, uint32_t resident_blocks);
// Workspace of the wavefront reflection pipeline for one slab of pixels
// [p0, p1): path queues (ping-pong), hits, hit shadow rays, visibility masks,
// pending list, per (pixel, sample) results and per-pixel direct / ambient.
struct WaveArgs {
    float4* paths0;
    float4* paths1;
    float4* hits;
    float4* srays;
    uint32_t* vis;
    uint32_t* pending;
    float4* pathcol;
    float4* pix;
    uint32_t* counts;  // queue counters, word i * BAND_STRIDE: 0, 1 path queues, 2 shadow rays, 3 pending
    uint32_t* takes;   // band take counters of the two trace stages: 2 x BAND_COUNTER_BYTES
    uint32_t p0, p1, spp_max, slab_pixels;
    uint32_t blocks;   // persistent grid of the trace stages
};
size_t wave_bytes_per_slot(uint32_t nl);
hipError_t launch_shade(const DevScene& s, const FrameArgs& f, GBufView gb, const float* shadow, float4* lit,
                        float4* refl, unsigned long long* counters, bool count, hipStream_t st, WaveArgs w);
hipError_t launch_trace_rays(const DevScene& s, const hr_ray* rays, uint32_t n, int any_hit, hr_hit* hits,
                             unsigned long long* counters, bool count, hipStream_t st);
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st);
// pure ray tracing mode (rt.hip): accum float4[H][W], progressive when frame_signed >= 0
// Wavefront pure ray tracing (1..64 shadow samples): workspace bytes, 0 when
// the frame does not fit it (then launch_rt).
size_t rt_work_bytes(const FrameArgs& f);
hipError_t launch_rt_wavefront(const DevScene& s, const FrameArgs& f, float4* accum, int32_t frame_signed,
                               unsigned long long* counters, hipStream_t st, void* work, uint32_t* next_item,
                               uint32_t resident_blocks);
hipError_t launch_rt(const DevScene& s, const FrameArgs& f, float4* accum, int32_t frame_signed,
                     unsigned long long* counters, bool count, hipStream_t st);

// ---- temporal shadow filter and motion vectors (temporal.hip) ----------------
struct MotionArgs {
    float vp[16];       // projection * view of this frame, column-major
    float vp_prev[16];  // of the previous frame
};
// shadow, prev, out: float[nl][H][W]; motion: float2[H][W] (read when frame <= 0)
hipError_t launch_filter_shadows(uint32_t W, uint32_t H, uint32_t nl, int32_t frame, const float* shadow,
                                 const float* motion, const float* prev, float* out, hipStream_t st);
hipError_t launch_motion(uint32_t W, uint32_t H, const float4* pos, const float4* nrm, const MotionArgs& a,
                         float* motion, hipStream_t st);

// ---- raster G-buffer formats (gbuffer_io.hip) --------------------------------
hipError_t launch_gbuffer_ingest(uint32_t n, const void* pos, const void* nrm, const void* alb, const void* mot,
                                 const void* mat, const void* emi, GBufView G, float* motion_out, hipStream_t st);
hipError_t launch_gbuffer_quantize(uint32_t n, GBufView G, const float* motion_in, void* pos, void* nrm, void* alb,
                                   void* mot, void* mat, void* emi, hipStream_t st);
// post pass (postFragment.frag): float4 in -> float4 out and/or RGBA8
hipError_t launch_post_gamma(uint32_t n, const float4* in, float4* out, uint32_t* rgba8, hipStream_t st);

// ---- LBVH build (lbvh.hip) ------------------------------------------------------
struct BuildBuffers {
    // inputs
    const uint8_t* verts;       // concatenated 44-byte vertices
    const uint32_t* indices;    // concatenated indices
    const PrimInfo* prims;
    const float* models;        // prim_count x 16
    const uint32_t* prim_tri_offset;  // prim_count + 1
    uint32_t prim_count, tri_count, leaf_max;
    // outputs / scratch
    float4* wtri;       // tri_count x 3 (primitive order)
    float4* tbox;       // tri_count x 2
    uint32_t* keys;     // sorted Morton codes
    uint32_t* ids;      // sorted triangle ids
    uint32_t* keys_tmp;
    uint32_t* ids_tmp;
    uint32_t* hist;     // radix histogram, 256 x blocks
    uint32_t* bounds;   // 6 ordered-uint centroid bounds
    uint32_t* kl;       // Karras children
    uint32_t* kr;
    uint32_t* kfirst;
    uint32_t* klast;
    uint32_t* kparent;  // internal node -> parent
    uint32_t* lparent;  // sorted leaf -> parent
    uint32_t* flags;    // internal-node depth (edges from the root)
    uint32_t* lvl;      // refit order: internal nodes by depth (node_count), then REFIT_LEVELS + 1 offsets and
                        // REFIT_LEVELS counts
    uint32_t refit_levels;  // host side: depth buckets in use, set by the build
    float4* kbox;       // internal node boxes x 2
    TNode* nodes;
    TNode4* nodes4;     // wide_count 4-wide nodes, dense (the root first), allocated between
                        // lbvh_build and lbvh_build_wide
    uint4* wsrc;        // per 4-wide node, the binary source of each child slot: (node << 1) | side, or EMPTY4
    uint32_t wide_count;  // host side: set by lbvh_build
    STri* stris;
    uint32_t* tri_prim;
    uint32_t* depth;    // 1 uint
};
size_t radix_hist_words(uint32_t n);
// full build in two calls: lbvh_build makes the binary tree and chooses the
// 4-wide nodes (sets wide_count, synchronising the stream); the caller then
// allocates nodes4 and wsrc with wide_count entries, and lbvh_build_wide fills them
// treelets: restructure the Karras tree (kept only if the result passes the
// traversal depth bound MAX_DEPTH; else the plain tree is rebuilt)
hipError_t lbvh_build(BuildBuffers& b, hipStream_t st, bool treelets);
hipError_t lbvh_build_wide(BuildBuffers& b, hipStream_t st);
hipError_t lbvh_refit(BuildBuffers& b, hipStream_t st);   // new models, same topology (binary and 4-wide)

// ---- denoiser kernels (denoise.hip) ------------------------------------------------
// Activations are BF16 NHWC images of N x H x W pixels; a tensor is a channel
// slice [off, off + count) of a buffer with C channels per pixel.
// Weights of a 3x3 layer, GEMM form W[co][k], k = tap * Cin + ci, tap =
// (dy + 1) * 3 + (dx + 1), are stored in MFMA fragment order: element (co, k)
// at ((k / 32 * (Cout_pad / 16) + co / 16) * 64 + (co % 16) + (k % 32) / 8 * 16) * 8 + k % 8.
struct ConvArgs {
    const __bf16* in;
    const __bf16* wt;      // fragment order, Cout_pad x Kpad
    const float* bias;     // Cout or null
    __bf16* out;           // BF16 output (when out_f32 is null)
    float* out_f32;        // FP32 output instead
    uint32_t N, H, W;
    uint32_t in_C, in_off, Cin, K, Kpad;  // K = 9 * Cin, Kpad = K rounded up to 32
    uint32_t out_C, out_off, Cout, Cout_pad;
    uint32_t relu;
    // two inputs (in2 not null; the LDS path only): channels [0, cin1) from
    // in (at in_off, stride in_C), channels [cin1, Cin) from in2 (stride
    // in2_C) at half resolution, pixel (x >> 1, y >> 1): a concatenation with
    // a nearest 2x upsampling, without either written out
    const __bf16* in2;
    uint32_t in2_C, cin1;
    // kernel prediction fused into the epilogue (kp_out not null; the LDS
    // path, N = 1, 2 output tiles): no logits are written; every pixel of the
    // kp_W x kp_H frame gets out = the 5x5 softmax-weighted sum of kp_noisy
    // its logits (bias added) predict, as dn_kpcn computes it
    const float4* kp_noisy;
    float4* kp_out;
    uint32_t kp_W, kp_H;
};
struct WgradArgs {
    const __bf16* dout;    // gradient of the layer output (after the ReLU mask)
    const __bf16* in;      // layer input
    float* partial;        // chunks x Cout_pad x Kpad
    float* bpartial;       // chunks x Cout_pad
    uint32_t N, H, W;
    uint32_t dout_C, dout_off, in_C, in_off;
    uint32_t Cin, K, Kpad, Cout_pad, chunk;
};
struct FeatArgs {
    const float4 *pos, *nrm, *alb, *mat, *noisy;
    __bf16* x0;            // N x Hp x Wp x 16
    float cam[3];
    uint32_t W, H, Wp, Hp;
};
struct KpcnArgs {
    const float* logits;   // N x Hp x Wp x 32 (25 used)
    __bf16* dlogits;
    const float4* noisy;   // image n, W x H
    const float4* target;
    float4* out;
    uint32_t W, H, Wp, Hp;
    uint32_t loss;         // 0 tone-mapped L1, 1 relative MSE
};
struct AdamArgs {
    float *w, *grad, *m, *v;
    size_t n;
    float lr, b1, b2, eps, bc1, bc2;
};
struct PackArgs {
    const float* w;        // FP32 master, Cout_pad x Kpad, row-major
    __bf16* wf;            // forward weights, fragment order
    __bf16* wd;            // input-gradient weights (flipped taps), fragment order, Cin x Kpad_d
    uint32_t Cin, Cout_pad, Kpad, Kpad_d;
};
hipError_t dn_conv(const ConvArgs& a, hipStream_t st);
uint32_t dn_wgrad_chunks(uint32_t P, uint32_t* chunk);
hipError_t dn_wgrad(const WgradArgs& a, float* wgrad, float* bgrad, hipStream_t st);
hipError_t dn_relu_bwd(__bf16* g, uint32_t g_C, uint32_t g_off, const __bf16* act, uint32_t a_C, uint32_t a_off,
                       uint32_t C, uint32_t P, hipStream_t st);
hipError_t dn_pool(const __bf16* in, uint32_t in_C, uint32_t in_off, __bf16* out, uint32_t C, uint32_t N, uint32_t Ho,
                   uint32_t Wo, hipStream_t st);
hipError_t dn_pool_bwd_add(const __bf16* gout, __bf16* gin, uint32_t gin_C, uint32_t gin_off, uint32_t C, uint32_t N,
                           uint32_t Ho, uint32_t Wo, hipStream_t st);
hipError_t dn_up(const __bf16* in, __bf16* out, uint32_t out_C, uint32_t out_off, uint32_t C, uint32_t N, uint32_t Ho,
                 uint32_t Wo, hipStream_t st);
hipError_t dn_up_bwd(const __bf16* gout, uint32_t g_C, uint32_t g_off, __bf16* gin, uint32_t C, uint32_t N, uint32_t Hi,
                     uint32_t Wi, hipStream_t st);
hipError_t dn_features(const FeatArgs& a, uint32_t n, hipStream_t st);
hipError_t dn_kpcn(const KpcnArgs& a, uint32_t n, hipStream_t st);
uint32_t dn_kpcn_blocks(uint32_t W, uint32_t H);
hipError_t dn_kpcn_bwd(const KpcnArgs& a, uint32_t n, float scale, float* loss_partial, hipStream_t st);
hipError_t dn_fill_bf16(__bf16* p, size_t n, hipStream_t st);
hipError_t dn_adam(const AdamArgs& a, hipStream_t st);
hipError_t dn_pack(const PackArgs& a, hipStream_t st);

}  // namespace hr
