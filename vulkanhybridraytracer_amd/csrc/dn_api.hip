// dn_api.hip - extern "C" entry points of the learned denoiser: the network
// topology, its device buffers, forward and backward passes and the Adam
// step as kernel launches on the context's stream.
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.h"

using namespace hr;

namespace {

struct LayerDesc {
    uint32_t cin, cout, level;
};
constexpr LayerDesc LAYERS[HR_DN_LAYERS] = {
    {9, 32, 0},   {32, 32, 0},  {32, 48, 1}, {48, 48, 1},  {48, 64, 2},  {64, 64, 2}, {64, 64, 3},
    {64, 64, 3},  {64, 64, 4},  {64, 64, 4}, {128, 64, 3}, {64, 64, 3},  {128, 64, 2}, {64, 64, 2},
    {112, 48, 1}, {48, 48, 1},  {80, 32, 0}, {32, 32, 0},  {32, 25, 0}};
enum {
    L_E0A, L_E0B, L_E1A, L_E1B, L_E2A, L_E2B, L_E3A, L_E3B, L_E4A, L_E4B,
    L_D3A, L_D3B, L_D2A, L_D2B, L_D1A, L_D1B, L_D0A, L_D0B, L_HEAD
};

// activation buffers: pyramid level and channels per pixel
enum {
    B_X0, B_E0A, B_CAT0, B_D0A, B_D0B, B_P1, B_E1A, B_CAT1, B_D1A, B_D1B, B_P2, B_E2A,
    B_CAT2, B_D2A, B_D2B, B_P3, B_E3A, B_CAT3, B_D3A, B_D3B, B_P4, B_E4A, B_E4B, B_COUNT
};
struct BufDesc {
    uint32_t level, C;
};
constexpr BufDesc BUFS[B_COUNT] = {{0, 16}, {0, 32}, {0, 80}, {0, 32}, {0, 32}, {1, 32}, {1, 48}, {1, 112},
                                   {1, 48}, {1, 48}, {2, 48}, {2, 64}, {2, 128}, {2, 64}, {2, 64}, {3, 64},
                                   {3, 64}, {3, 128}, {3, 64}, {3, 64}, {4, 64}, {4, 64}, {4, 64}};
constexpr int B_LOGITS = -1;

// The forward program. A concatenation buffer CATl holds the encoder output
// of level l in its first channels and the upsampled decoder input after it.
enum OpKind { OP_CONV, OP_POOL, OP_UP };
struct Op {
    OpKind kind;
    int layer;
    int src;
    uint32_t src_off, C;  // source slice (C = input channels of a convolution)
    int dst;
    uint32_t dst_off;
};
constexpr Op OPS[] = {
    {OP_CONV, L_E0A, B_X0, 0, 16, B_E0A, 0},     {OP_CONV, L_E0B, B_E0A, 0, 32, B_CAT0, 0},
    {OP_POOL, -1, B_CAT0, 0, 32, B_P1, 0},       {OP_CONV, L_E1A, B_P1, 0, 32, B_E1A, 0},
    {OP_CONV, L_E1B, B_E1A, 0, 48, B_CAT1, 0},   {OP_POOL, -1, B_CAT1, 0, 48, B_P2, 0},
    {OP_CONV, L_E2A, B_P2, 0, 48, B_E2A, 0},     {OP_CONV, L_E2B, B_E2A, 0, 64, B_CAT2, 0},
    {OP_POOL, -1, B_CAT2, 0, 64, B_P3, 0},       {OP_CONV, L_E3A, B_P3, 0, 64, B_E3A, 0},
    {OP_CONV, L_E3B, B_E3A, 0, 64, B_CAT3, 0},   {OP_POOL, -1, B_CAT3, 0, 64, B_P4, 0},
    {OP_CONV, L_E4A, B_P4, 0, 64, B_E4A, 0},     {OP_CONV, L_E4B, B_E4A, 0, 64, B_E4B, 0},
    {OP_UP, -1, B_E4B, 0, 64, B_CAT3, 64},       {OP_CONV, L_D3A, B_CAT3, 0, 128, B_D3A, 0},
    {OP_CONV, L_D3B, B_D3A, 0, 64, B_D3B, 0},    {OP_UP, -1, B_D3B, 0, 64, B_CAT2, 64},
    {OP_CONV, L_D2A, B_CAT2, 0, 128, B_D2A, 0},  {OP_CONV, L_D2B, B_D2A, 0, 64, B_D2B, 0},
    {OP_UP, -1, B_D2B, 0, 64, B_CAT1, 48},       {OP_CONV, L_D1A, B_CAT1, 0, 112, B_D1A, 0},
    {OP_CONV, L_D1B, B_D1A, 0, 48, B_D1B, 0},    {OP_UP, -1, B_D1B, 0, 48, B_CAT0, 32},
    {OP_CONV, L_D0A, B_CAT0, 0, 80, B_D0A, 0},   {OP_CONV, L_D0B, B_D0A, 0, 32, B_D0B, 0},
    {OP_CONV, L_HEAD, B_D0B, 0, 32, B_LOGITS, 0},
};
constexpr int NOPS = sizeof(OPS) / sizeof(OPS[0]);
constexpr uint32_t LEVELS = 5;
constexpr uint32_t ALIGN = 1u << (LEVELS - 1);  // padded frame size multiple: 16

inline uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// padded shapes and parameter offsets of one layer
struct LayerInt {
    uint32_t cin_pad, cout_pad, Kpad, Kpad_d;
    size_t w_off, b_off;         // FP32 master: W[cout_pad][Kpad] then b[cout_pad]
    __bf16* wf = nullptr;        // forward weights, fragment order
    __bf16* wd = nullptr;        // input-gradient weights, fragment order
};
void layer_shapes(uint32_t cin, uint32_t cout, LayerInt& li) {
    li.cin_pad = round_up(cin, 16);
    li.cout_pad = round_up(cout, 16);
    li.Kpad = round_up(9 * li.cin_pad, 32);
    li.Kpad_d = round_up(9 * li.cout_pad, 32);
}

struct Arena {
    uint32_t N = 0, Hp = 0, Wp = 0, loss_cap = 0;
    bool grads = false;
    void* block = nullptr;
    __bf16* act[B_COUNT] = {};
    __bf16* grad[B_COUNT] = {};
    float* logits = nullptr;
    __bf16* dlogits = nullptr;
    float* partial = nullptr;
    float* bpartial = nullptr;
    float* loss = nullptr;
};

}  // namespace

struct hr_denoiser {
    hr_context* ctx = nullptr;
    LayerInt L[HR_DN_LAYERS];
    size_t nparam = 0;
    void* pblock = nullptr;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint64_t step = 0;
    uint32_t loss = HR_DN_LOSS_L1_TONEMAPPED;
    Arena train, infer;
};

namespace {

uint64_t public_count() {
    uint64_t n = 0;
    for (const LayerDesc& d : LAYERS) n += 9ull * d.cin * d.cout + d.cout;
    return n;
}

// public W[cout][3][3][cin], b[cout] <-> internal padded FP32 master
void public_to_internal(const hr_denoiser* dn, const float* pub, std::vector<float>& in) {
    in.assign(dn->nparam, 0.0f);
    size_t o = 0;
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        const LayerDesc& d = LAYERS[l];
        const LayerInt& li = dn->L[l];
        for (uint32_t co = 0; co < d.cout; ++co)
            for (uint32_t t = 0; t < 9; ++t)
                for (uint32_t ci = 0; ci < d.cin; ++ci) in[li.w_off + (size_t)co * li.Kpad + t * li.cin_pad + ci] = pub[o++];
        for (uint32_t co = 0; co < d.cout; ++co) in[li.b_off + co] = pub[o++];
    }
}
void internal_to_public(const hr_denoiser* dn, const float* in, float* pub) {
    size_t o = 0;
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        const LayerDesc& d = LAYERS[l];
        const LayerInt& li = dn->L[l];
        for (uint32_t co = 0; co < d.cout; ++co)
            for (uint32_t t = 0; t < 9; ++t)
                for (uint32_t ci = 0; ci < d.cin; ++ci) pub[o++] = in[li.w_off + (size_t)co * li.Kpad + t * li.cin_pad + ci];
        for (uint32_t co = 0; co < d.cout; ++co) pub[o++] = in[li.b_off + co];
    }
}

hr_status pack_all(hr_denoiser* dn) {
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        const LayerInt& li = dn->L[l];
        PackArgs p{dn->w + li.w_off, li.wf, li.wd, li.cin_pad, li.cout_pad, li.Kpad, li.Kpad_d};
        HR_HIP(dn->ctx, dn_pack(p, dn->ctx->stream));
    }
    return HR_OK;
}

hr_status upload_weights(hr_denoiser* dn, const float* pub) {
    hr_context* ctx = dn->ctx;
    std::vector<float> in;
    public_to_internal(dn, pub, in);
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HR_HIP(ctx, hipMemcpy(dn->w, in.data(), dn->nparam * sizeof(float), hipMemcpyHostToDevice));
    HR_HIP(ctx, hipMemset(dn->m, 0, dn->nparam * sizeof(float)));
    HR_HIP(ctx, hipMemset(dn->v, 0, dn->nparam * sizeof(float)));
    dn->step = 0;
    return pack_all(dn);
}

void arena_free(Arena& A) {
    if (A.block) (void)hipFree(A.block);
    A = Arena{};
}

hr_status ensure_arena(hr_denoiser* dn, Arena& A, uint32_t N, uint32_t Hp, uint32_t Wp, bool grads, uint32_t loss_blocks) {
    hr_context* ctx = dn->ctx;
    if (A.block && A.N == N && A.Hp == Hp && A.Wp == Wp && A.grads == grads && A.loss_cap >= loss_blocks) return HR_OK;
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    arena_free(A);
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t px0 = (size_t)N * Hp * Wp, total = 0;
    size_t act_b[B_COUNT];
    for (int b = 0; b < B_COUNT; ++b) {
        act_b[b] = al(((size_t)N * (Hp >> BUFS[b].level) * (Wp >> BUFS[b].level)) * BUFS[b].C * sizeof(__bf16));
        total += act_b[b] * (grads ? 2 : 1);
    }
    size_t logits_b = al(px0 * 32 * sizeof(float)), dlogits_b = grads ? al(px0 * 32 * sizeof(__bf16)) : 0;
    size_t part_b = 0, bpart_b = 0;
    if (grads) {
        for (const Op& op : OPS) {
            if (op.kind != OP_CONV) continue;
            const LayerInt& li = dn->L[op.layer];
            uint32_t lvl = LAYERS[op.layer].level, chunk = 0;
            uint32_t chunks = dn_wgrad_chunks(N * (Hp >> lvl) * (Wp >> lvl), &chunk);
            part_b = std::max(part_b, al((size_t)chunks * li.cout_pad * li.Kpad * sizeof(float)));
            bpart_b = std::max(bpart_b, al((size_t)chunks * li.cout_pad * sizeof(float)));
        }
    }
    size_t loss_b = al((size_t)std::max(loss_blocks, 1u) * sizeof(float));
    total += logits_b + dlogits_b + part_b + bpart_b + loss_b;
    void* blk = nullptr;
    HR_HIP(ctx, hipMalloc(&blk, total));
    A.block = blk;
    char* c = static_cast<char*>(blk);
    for (int b = 0; b < B_COUNT; ++b) {
        A.act[b] = reinterpret_cast<__bf16*>(c);
        c += act_b[b];
        if (grads) {
            A.grad[b] = reinterpret_cast<__bf16*>(c);
            c += act_b[b];
        }
    }
    A.logits = reinterpret_cast<float*>(c);
    c += logits_b;
    if (grads) {
        A.dlogits = reinterpret_cast<__bf16*>(c);
        c += dlogits_b;
        A.partial = reinterpret_cast<float*>(c);
        c += part_b;
        A.bpartial = reinterpret_cast<float*>(c);
        c += bpart_b;
    }
    A.loss = reinterpret_cast<float*>(c);
    A.N = N;
    A.Hp = Hp;
    A.Wp = Wp;
    A.grads = grads;
    A.loss_cap = std::max(loss_blocks, 1u);
    return HR_OK;
}

hr_status check_images(hr_context* ctx, const hr_dn_image* im, uint32_t n, bool need_target, bool need_out) {
    if (!im || n == 0) return fail(ctx, HR_E_INVALID, "no images");
    for (uint32_t i = 0; i < n; ++i) {
        const hr_gbuffer& g = im[i].gbuffer;
        if (g.width == 0 || g.height == 0 || g.width > 16384 || g.height > 16384)
            return fail(ctx, HR_E_INVALID, "image %u: bad size %ux%u", i, g.width, g.height);
        if (g.width != im[0].gbuffer.width || g.height != im[0].gbuffer.height)
            return fail(ctx, HR_E_INVALID, "image %u: all images of a call must have one size", i);
        if (!g.position || !g.normal || !g.albedo || !g.material || !im[i].noisy)
            return fail(ctx, HR_E_INVALID, "image %u: missing G-buffer plane or noisy colour", i);
        if (need_target && !im[i].target) return fail(ctx, HR_E_INVALID, "image %u: missing target", i);
        if (need_out && !im[i].out) return fail(ctx, HR_E_INVALID, "image %u: missing output", i);
    }
    return HR_OK;
}

FeatArgs feat_args(const Arena& A, const hr_dn_image& im) {
    FeatArgs f{};
    f.pos = reinterpret_cast<const float4*>(im.gbuffer.position);
    f.nrm = reinterpret_cast<const float4*>(im.gbuffer.normal);
    f.alb = reinterpret_cast<const float4*>(im.gbuffer.albedo);
    f.mat = reinterpret_cast<const float4*>(im.gbuffer.material);
    f.noisy = reinterpret_cast<const float4*>(im.noisy);
    f.x0 = A.act[B_X0];
    for (int k = 0; k < 3; ++k) f.cam[k] = im.eye[k];
    f.W = im.gbuffer.width;
    f.H = im.gbuffer.height;
    f.Wp = A.Wp;
    f.Hp = A.Hp;
    return f;
}
KpcnArgs kpcn_args(const Arena& A, const hr_dn_image& im, uint32_t loss) {
    KpcnArgs k{};
    k.loss = loss;
    k.logits = A.logits;
    k.dlogits = A.dlogits;
    k.noisy = reinterpret_cast<const float4*>(im.noisy);
    k.target = reinterpret_cast<const float4*>(im.target);
    k.out = reinterpret_cast<float4*>(im.out);
    k.W = im.gbuffer.width;
    k.H = im.gbuffer.height;
    k.Wp = A.Wp;
    k.Hp = A.Hp;
    return k;
}

// The upsampling that fills the second part of concatenation buffer `buf`
// (false: `buf` is not one).
bool concat_of(int buf, Op& up) {
    for (const Op& o : OPS)
        if (o.kind == OP_UP && o.dst == buf) {
            up = o;
            return true;
        }
    return false;
}

// Inference (no gradients) writes no concatenation: a concatenation buffer
// holds only its encoder output, compact (channel stride = the encoder's
// channels, dst_off of the upsampling), the pooling reads it there, and the
// decoder layer that reads the concatenation takes its second part straight
// from the upsampling's half-resolution source (two-input convolution). The
// values are those of the written-out concatenation. With `image` (inference
// of one image), the head layer's epilogue also applies the predicted kernels
// and writes image->out; the logits are not written (kpcn_done: it did).
hr_status forward(hr_denoiser* dn, const Arena& A, const hr_dn_image* image = nullptr, bool* kpcn_done = nullptr) {
    hr_context* ctx = dn->ctx;
    hipStream_t st = ctx->stream;
    const bool fused = !A.grads;
    for (const Op& op : OPS) {
        Op up{};
        if (fused && op.kind == OP_UP) continue;
        if (op.kind == OP_CONV) {
            const LayerDesc& d = LAYERS[op.layer];
            const LayerInt& li = dn->L[op.layer];
            ConvArgs a{};
            a.in = A.act[op.src];
            a.wt = li.wf;
            a.bias = dn->w + li.b_off;
            a.N = A.N;
            a.H = A.Hp >> d.level;
            a.W = A.Wp >> d.level;
            a.in_C = BUFS[op.src].C;
            a.in_off = op.src_off;
            a.Cin = li.cin_pad;
            a.K = 9 * li.cin_pad;
            a.Kpad = li.Kpad;
            if (op.dst == B_LOGITS) {
                a.out_f32 = A.logits;
                a.out_C = 32;
            } else {
                a.out = A.act[op.dst];
                a.out_C = BUFS[op.dst].C;
            }
            a.out_off = op.dst_off;
            a.Cout = d.cout;
            a.Cout_pad = li.cout_pad;
            a.relu = op.layer != L_HEAD;
            if (fused && image && op.dst == B_LOGITS && A.N == 1u) {
                a.kp_noisy = reinterpret_cast<const float4*>(image->noisy);
                a.kp_out = reinterpret_cast<float4*>(image->out);
                a.kp_W = image->gbuffer.width;
                a.kp_H = image->gbuffer.height;
                if (kpcn_done) *kpcn_done = true;
            }
            if (fused && op.dst != B_LOGITS && concat_of(op.dst, up)) a.out_C = up.dst_off;
            if (fused && concat_of(op.src, up)) {
                if (op.src_off != 0u || up.dst_off + up.C != li.cin_pad) return fail(ctx, HR_E_INVALID, "concat layout");
                a.in_C = up.dst_off;
                a.in2 = A.act[up.src];
                a.in2_C = BUFS[up.src].C;
                a.cin1 = up.dst_off;
            }
            HR_HIP(ctx, dn_conv(a, st));
        } else if (op.kind == OP_POOL) {
            uint32_t lvl = BUFS[op.dst].level;
            const uint32_t in_C = fused && concat_of(op.src, up) ? up.dst_off : BUFS[op.src].C;
            HR_HIP(ctx, dn_pool(A.act[op.src], in_C, op.src_off, A.act[op.dst], op.C, A.N, A.Hp >> lvl, A.Wp >> lvl,
                                st));
        } else {
            uint32_t lvl = BUFS[op.dst].level;
            HR_HIP(ctx, dn_up(A.act[op.src], A.act[op.dst], BUFS[op.dst].C, op.dst_off, op.C, A.N, A.Hp >> lvl,
                              A.Wp >> lvl, st));
        }
    }
    return HR_OK;
}

// Backward program (the forward program reversed). Every slice gradient is
// first written whole (input gradient of the consuming convolution, or the
// upsample backward) before the pool backward adds to it.
hr_status backward(hr_denoiser* dn, const Arena& A) {
    hr_context* ctx = dn->ctx;
    hipStream_t st = ctx->stream;
    for (int i = NOPS - 1; i >= 0; --i) {
        const Op& op = OPS[i];
        if (op.kind == OP_CONV) {
            const LayerDesc& d = LAYERS[op.layer];
            const LayerInt& li = dn->L[op.layer];
            const uint32_t H = A.Hp >> d.level, W = A.Wp >> d.level, P = A.N * H * W;
            __bf16* gout;
            uint32_t gC, goff;
            if (op.dst == B_LOGITS) {
                gout = A.dlogits;
                gC = 32;
                goff = 0;
            } else {
                gout = A.grad[op.dst];
                gC = BUFS[op.dst].C;
                goff = op.dst_off;
            }
            if (op.layer != L_HEAD)
                HR_HIP(ctx, dn_relu_bwd(gout, gC, goff, A.act[op.dst], BUFS[op.dst].C, op.dst_off, d.cout, P, st));
            WgradArgs w{};
            w.dout = gout;
            w.in = A.act[op.src];
            w.partial = A.partial;
            w.bpartial = A.bpartial;
            w.N = A.N;
            w.H = H;
            w.W = W;
            w.dout_C = gC;
            w.dout_off = goff;
            w.in_C = BUFS[op.src].C;
            w.in_off = op.src_off;
            w.Cin = li.cin_pad;
            w.K = 9 * li.cin_pad;
            w.Kpad = li.Kpad;
            w.Cout_pad = li.cout_pad;
            dn_wgrad_chunks(P, &w.chunk);
            HR_HIP(ctx, dn_wgrad(w, dn->g + li.w_off, dn->g + li.b_off, st));
            if (op.layer != L_E0A) {
                ConvArgs a{};
                a.in = gout;
                a.wt = li.wd;
                a.N = A.N;
                a.H = H;
                a.W = W;
                a.in_C = gC;
                a.in_off = goff;
                a.Cin = li.cout_pad;
                a.K = 9 * li.cout_pad;
                a.Kpad = li.Kpad_d;
                a.out = A.grad[op.src];
                a.out_C = BUFS[op.src].C;
                a.out_off = op.src_off;
                a.Cout = li.cin_pad;
                a.Cout_pad = li.cin_pad;
                a.relu = 0;
                HR_HIP(ctx, dn_conv(a, st));
            }
        } else if (op.kind == OP_POOL) {
            uint32_t lvl = BUFS[op.dst].level;
            HR_HIP(ctx, dn_pool_bwd_add(A.grad[op.dst], A.grad[op.src], BUFS[op.src].C, op.src_off, op.C, A.N,
                                        A.Hp >> lvl, A.Wp >> lvl, st));
        } else {
            uint32_t lvl = BUFS[op.src].level;
            HR_HIP(ctx, dn_up_bwd(A.grad[op.dst], BUFS[op.dst].C, op.dst_off, A.grad[op.src], op.C, A.N, A.Hp >> lvl,
                                  A.Wp >> lvl, st));
        }
    }
    return HR_OK;
}

// forward + loss + backward of n images into dn->g; loss partial sums stay on
// the device (A.loss, n x blocks) until read
hr_status loss_and_gradients(hr_denoiser* dn, const hr_dn_image* im, uint32_t n, uint32_t* loss_blocks) {
    hr_context* ctx = dn->ctx;
    const uint32_t W = im[0].gbuffer.width, H = im[0].gbuffer.height;
    const uint32_t blocks = dn_kpcn_blocks(W, H);
    HR_TRY(ensure_arena(dn, dn->train, n, round_up(H, ALIGN), round_up(W, ALIGN), true, n * blocks));
    Arena& A = dn->train;
    for (uint32_t i = 0; i < n; ++i) HR_HIP(ctx, dn_features(feat_args(A, im[i]), i, ctx->stream));
    HR_TRY(forward(dn, A));
    HR_HIP(ctx, dn_fill_bf16(A.dlogits, (size_t)A.N * A.Hp * A.Wp * 32, ctx->stream));
    const float scale = (float)(1.0 / (3.0 * n * (double)W * H));
    for (uint32_t i = 0; i < n; ++i)
        HR_HIP(ctx, dn_kpcn_bwd(kpcn_args(A, im[i], dn->loss), i, scale, A.loss + (size_t)i * blocks, ctx->stream));
    HR_TRY(backward(dn, A));
    *loss_blocks = n * blocks;
    return HR_OK;
}

// mean loss of the last loss_and_gradients call (waits for the stream)
hr_status read_loss(hr_denoiser* dn, uint32_t blocks, const hr_dn_image* im, uint32_t n, float* loss) {
    hr_context* ctx = dn->ctx;
    std::vector<float> part(blocks);
    HR_HIP(ctx, hipMemcpyAsync(part.data(), dn->train.loss, blocks * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float f : part) s += f;
    *loss = (float)(s / (3.0 * n * (double)im[0].gbuffer.width * im[0].gbuffer.height));
    return HR_OK;
}

// He-normal initialisation (std sqrt(2 / fan_in)); the head is scaled by 0.1
// so the initial kernels are close to a uniform 5x5 filter. splitmix64 +
// Box-Muller.
void init_weights(uint64_t seed, std::vector<float>& pub) {
    pub.resize(public_count());
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
    auto next = [&s]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    size_t o = 0;
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        const LayerDesc& d = LAYERS[l];
        double sd = std::sqrt(2.0 / (9.0 * d.cin)) * (l == L_HEAD ? 0.1 : 1.0);
        size_t nw = 9ull * d.cin * d.cout;
        for (size_t i = 0; i < nw; ++i) {
            double u1 = next(), u2 = next();
            pub[o++] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
        for (uint32_t i = 0; i < d.cout; ++i) pub[o++] = 0.0f;
    }
}

// temporary device copies of one layer's weights for the building-block calls
struct TmpLayer {
    LayerInt li;
    void* block = nullptr;
    float* w = nullptr;
    ~TmpLayer() {
        if (block) (void)hipFree(block);
    }
};
hr_status tmp_layer(hr_context* ctx, const float* weight, const float* bias, uint32_t cin, uint32_t cout, TmpLayer& t) {
    layer_shapes(cin, cout, t.li);
    LayerInt& li = t.li;
    size_t nw = (size_t)li.cout_pad * li.Kpad;
    li.w_off = 0;
    li.b_off = nw;
    size_t nmaster = nw + li.cout_pad;
    size_t nf = (size_t)li.cout_pad * li.Kpad, nd = (size_t)li.cin_pad * li.Kpad_d;
    size_t bytes = nmaster * 4 + (nf + nd) * 2 + 512;
    HR_HIP(ctx, hipMalloc(&t.block, bytes));
    HR_HIP(ctx, hipMemset(t.block, 0, bytes));
    t.w = static_cast<float*>(t.block);
    li.wf = reinterpret_cast<__bf16*>(static_cast<char*>(t.block) + ((nmaster * 4 + 255) & ~(size_t)255));
    li.wd = li.wf + ((nf + 127) & ~(size_t)127);
    std::vector<float> m(nmaster, 0.0f);
    for (uint32_t co = 0; co < cout; ++co) {
        for (uint32_t tp = 0; tp < 9; ++tp)
            for (uint32_t ci = 0; ci < cin; ++ci)
                m[(size_t)co * li.Kpad + tp * li.cin_pad + ci] = weight[((size_t)co * 9 + tp) * cin + ci];
        if (bias) m[nw + co] = bias[co];
    }
    HR_HIP(ctx, hipMemcpy(t.w, m.data(), nmaster * 4, hipMemcpyHostToDevice));
    PackArgs p{t.w, li.wf, li.wd, li.cin_pad, li.cout_pad, li.Kpad, li.Kpad_d};
    HR_HIP(ctx, dn_pack(p, ctx->stream));
    return HR_OK;
}

}  // namespace

extern "C" {

uint64_t hr_denoiser_param_count(void) { return public_count(); }

hr_status hr_denoiser_layer(uint32_t i, uint32_t* cin, uint32_t* cout, uint32_t* level) {
    if (i >= HR_DN_LAYERS) return HR_E_INVALID;
    if (cin) *cin = LAYERS[i].cin;
    if (cout) *cout = LAYERS[i].cout;
    if (level) *level = LAYERS[i].level;
    return HR_OK;
}

hr_status hr_denoiser_create(hr_context* ctx, const float* weights, uint64_t count, uint64_t seed, hr_denoiser** out) {
    if (!ctx || !out) return fail(ctx, HR_E_INVALID, "null argument");
    *out = nullptr;
    if (weights && count != public_count())
        return fail(ctx, HR_E_INVALID, "weights: %llu values given, %llu expected", (unsigned long long)count,
                    (unsigned long long)public_count());
    HR_HIP(ctx, hipSetDevice(ctx->device));
    hr_denoiser* dn = new hr_denoiser;
    dn->ctx = ctx;
    size_t n = 0, nbf = 0;
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        LayerInt& li = dn->L[l];
        layer_shapes(LAYERS[l].cin, LAYERS[l].cout, li);
        li.w_off = n;
        n += (size_t)li.cout_pad * li.Kpad;
        li.b_off = n;
        n += li.cout_pad;
        nbf += ((size_t)li.cout_pad * li.Kpad + 127) / 128 * 128 + ((size_t)li.cin_pad * li.Kpad_d + 127) / 128 * 128;
    }
    dn->nparam = n;
    size_t fb = (n * sizeof(float) + 255) & ~(size_t)255;
    size_t bytes = 4 * fb + nbf * sizeof(__bf16);
    hipError_t e = hipMalloc(&dn->pblock, bytes);
    if (e == hipSuccess) e = hipMemset(dn->pblock, 0, bytes);
    if (e != hipSuccess) {
        hr_status s = hip_fail(ctx, e, "denoiser parameters");
        hr_denoiser_destroy(dn);
        return s;
    }
    char* c = static_cast<char*>(dn->pblock);
    dn->w = reinterpret_cast<float*>(c);
    dn->g = reinterpret_cast<float*>(c + fb);
    dn->m = reinterpret_cast<float*>(c + 2 * fb);
    dn->v = reinterpret_cast<float*>(c + 3 * fb);
    __bf16* b = reinterpret_cast<__bf16*>(c + 4 * fb);
    for (int l = 0; l < HR_DN_LAYERS; ++l) {
        LayerInt& li = dn->L[l];
        li.wf = b;
        b += ((size_t)li.cout_pad * li.Kpad + 127) / 128 * 128;
        li.wd = b;
        b += ((size_t)li.cin_pad * li.Kpad_d + 127) / 128 * 128;
    }
    std::vector<float> init;
    if (!weights) {
        init_weights(seed, init);
        weights = init.data();
    }
    hr_status s = upload_weights(dn, weights);
    if (s != HR_OK) {
        hr_denoiser_destroy(dn);
        return s;
    }
    *out = dn;
    return HR_OK;
}

void hr_denoiser_destroy(hr_denoiser* dn) {
    if (!dn) return;
    (void)hipSetDevice(dn->ctx->device);
    (void)hipStreamSynchronize(dn->ctx->stream);
    arena_free(dn->train);
    arena_free(dn->infer);
    if (dn->pblock) (void)hipFree(dn->pblock);
    delete dn;
}

hr_status hr_denoiser_get_weights(hr_denoiser* dn, float* weights, uint64_t count) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    if (!weights || count != public_count()) return fail(ctx, HR_E_INVALID, "weights: %llu values expected",
                                                         (unsigned long long)public_count());
    HR_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<float> in(dn->nparam);
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HR_HIP(ctx, hipMemcpy(in.data(), dn->w, dn->nparam * sizeof(float), hipMemcpyDeviceToHost));
    internal_to_public(dn, in.data(), weights);
    return HR_OK;
}

hr_status hr_denoiser_set_weights(hr_denoiser* dn, const float* weights, uint64_t count) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    if (!weights || count != public_count()) return fail(ctx, HR_E_INVALID, "weights: %llu values expected",
                                                         (unsigned long long)public_count());
    HR_HIP(ctx, hipSetDevice(ctx->device));
    return upload_weights(dn, weights);
}

hr_status hr_denoiser_get_state(hr_denoiser* dn, float* weights, float* adam_m, float* adam_v, uint64_t count,
                                uint64_t* step) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    if (!weights || count != public_count()) return fail(ctx, HR_E_INVALID, "state: %llu values expected",
                                                         (unsigned long long)public_count());
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float> in(dn->nparam);
    float* dst[3] = {weights, adam_m, adam_v};
    const float* src[3] = {dn->w, dn->m, dn->v};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        HR_HIP(ctx, hipMemcpy(in.data(), src[k], dn->nparam * sizeof(float), hipMemcpyDeviceToHost));
        internal_to_public(dn, in.data(), dst[k]);
    }
    if (step) *step = dn->step;
    return HR_OK;
}

hr_status hr_denoiser_set_state(hr_denoiser* dn, const float* weights, const float* adam_m, const float* adam_v,
                                uint64_t count, uint64_t step) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    if (!weights || !adam_m || !adam_v || count != public_count())
        return fail(ctx, HR_E_INVALID, "state: weights, adam_m and adam_v of %llu values expected",
                    (unsigned long long)public_count());
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i]) || !std::isfinite(adam_m[i]) || !(adam_v[i] >= 0.0f) ||
            !std::isfinite(adam_v[i]))
            return fail(ctx, HR_E_INVALID, "state: value %llu is not finite (or a negative second moment)",
                        (unsigned long long)i);
    HR_HIP(ctx, hipSetDevice(ctx->device));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float> in;
    float* dst[3] = {dn->w, dn->m, dn->v};
    const float* src[3] = {weights, adam_m, adam_v};
    for (int k = 0; k < 3; ++k) {
        public_to_internal(dn, src[k], in);
        HR_HIP(ctx, hipMemcpy(dst[k], in.data(), dn->nparam * sizeof(float), hipMemcpyHostToDevice));
    }
    dn->step = step;
    return pack_all(dn);
}

hr_status hr_denoiser_set_loss(hr_denoiser* dn, uint32_t loss) {
    if (!dn) return HR_E_INVALID;
    if (loss != HR_DN_LOSS_L1_TONEMAPPED && loss != HR_DN_LOSS_REL_MSE)
        return fail(dn->ctx, HR_E_INVALID, "unknown loss %u", loss);
    dn->loss = loss;
    return HR_OK;
}

hr_status hr_denoise(hr_denoiser* dn, const hr_dn_image* image, float* ms) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    HR_TRY(check_images(ctx, image, 1, false, true));
    HR_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t W = image->gbuffer.width, H = image->gbuffer.height;
    HR_TRY(ensure_arena(dn, dn->infer, 1, round_up(H, ALIGN), round_up(W, ALIGN), false, 1));
    const Arena& A = dn->infer;
    if (ms) HR_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    HR_HIP(ctx, dn_features(feat_args(A, *image), 0, ctx->stream));
    bool kpcn_done = false;
    HR_TRY(forward(dn, A, image, &kpcn_done));
    if (!kpcn_done) HR_HIP(ctx, dn_kpcn(kpcn_args(A, *image, dn->loss), 0, ctx->stream));
    if (ms) {
        HR_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
        HR_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
        HR_HIP(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    }
    return HR_OK;
}

hr_status hr_denoiser_train_step(hr_denoiser* dn, const hr_dn_image* images, uint32_t n, float lr, float* loss) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    HR_TRY(check_images(ctx, images, n, true, false));
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(ctx, HR_E_INVALID, "learning rate must be positive");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t blocks = 0;
    HR_TRY(loss_and_gradients(dn, images, n, &blocks));
    dn->step += 1;
    AdamArgs a{};
    a.w = dn->w;
    a.grad = dn->g;
    a.m = dn->m;
    a.v = dn->v;
    a.n = dn->nparam;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.bc1 = (float)(1.0 - std::pow(0.9, (double)dn->step));
    a.bc2 = (float)(1.0 - std::pow(0.999, (double)dn->step));
    HR_HIP(ctx, dn_adam(a, ctx->stream));
    HR_TRY(pack_all(dn));
    if (loss) HR_TRY(read_loss(dn, blocks, images, n, loss));
    return HR_OK;
}

hr_status hr_denoiser_gradients(hr_denoiser* dn, const hr_dn_image* images, uint32_t n, float* loss, float* grad,
                                uint64_t count) {
    if (!dn) return HR_E_INVALID;
    hr_context* ctx = dn->ctx;
    HR_TRY(check_images(ctx, images, n, true, false));
    if (grad && count != public_count()) return fail(ctx, HR_E_INVALID, "gradient: %llu values expected",
                                                     (unsigned long long)public_count());
    HR_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t blocks = 0;
    HR_TRY(loss_and_gradients(dn, images, n, &blocks));
    float l = 0.0f;
    HR_TRY(read_loss(dn, blocks, images, n, &l));
    if (loss) *loss = l;
    if (grad) {
        std::vector<float> in(dn->nparam);
        HR_HIP(ctx, hipMemcpy(in.data(), dn->g, dn->nparam * sizeof(float), hipMemcpyDeviceToHost));
        internal_to_public(dn, in.data(), grad);
    }
    return HR_OK;
}

hr_status hr_dn_conv3x3_forward(hr_context* ctx, const void* x, uint32_t n, uint32_t h, uint32_t w, uint32_t cin,
                                const float* weight, const float* bias, uint32_t cout, int relu, void* y) {
    if (!ctx) return HR_E_INVALID;
    if (!x || !y || !weight || n == 0 || h == 0 || w == 0 || cout == 0 || cout > 1024)
        return fail(ctx, HR_E_INVALID, "conv3x3: bad argument");
    if (cin == 0 || cin % 16 || cin > 1024) return fail(ctx, HR_E_INVALID, "conv3x3: cin must be a multiple of 16");
    if ((uint64_t)n * h * w >= (1ull << 31)) return fail(ctx, HR_E_INVALID, "conv3x3: too many pixels");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    TmpLayer t;
    HR_TRY(tmp_layer(ctx, weight, bias, cin, cout, t));
    ConvArgs a{};
    a.in = static_cast<const __bf16*>(x);
    a.wt = t.li.wf;
    a.bias = bias ? t.w + t.li.b_off : nullptr;
    a.out = static_cast<__bf16*>(y);
    a.N = n;
    a.H = h;
    a.W = w;
    a.in_C = cin;
    a.in_off = 0;
    a.Cin = cin;
    a.K = 9 * cin;
    a.Kpad = t.li.Kpad;
    a.out_C = cout;
    a.out_off = 0;
    a.Cout = cout;
    a.Cout_pad = t.li.cout_pad;
    a.relu = relu ? 1u : 0u;
    HR_HIP(ctx, dn_conv(a, ctx->stream));
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HR_OK;
}

hr_status hr_dn_conv3x3_backward(hr_context* ctx, const void* x, const void* dy, uint32_t n, uint32_t h, uint32_t w,
                                 uint32_t cin, const float* weight, uint32_t cout, void* dx, float* dweight,
                                 float* dbias) {
    if (!ctx) return HR_E_INVALID;
    if (!x || !dy || !weight || n == 0 || h == 0 || w == 0) return fail(ctx, HR_E_INVALID, "conv3x3: bad argument");
    if (cin == 0 || cin % 16 || cin > 1024 || cout == 0 || cout % 16 || cout > 1024)
        return fail(ctx, HR_E_INVALID, "conv3x3 backward: cin and cout must be multiples of 16");
    if ((uint64_t)n * h * w >= (1ull << 31)) return fail(ctx, HR_E_INVALID, "conv3x3: too many pixels");
    HR_HIP(ctx, hipSetDevice(ctx->device));
    TmpLayer t;
    HR_TRY(tmp_layer(ctx, weight, nullptr, cin, cout, t));
    const LayerInt& li = t.li;
    const uint32_t P = n * h * w;
    if (dx) {
        ConvArgs a{};
        a.in = static_cast<const __bf16*>(dy);
        a.wt = li.wd;
        a.out = static_cast<__bf16*>(dx);
        a.N = n;
        a.H = h;
        a.W = w;
        a.in_C = cout;
        a.Cin = li.cout_pad;
        a.K = 9 * li.cout_pad;
        a.Kpad = li.Kpad_d;
        a.out_C = cin;
        a.Cout = cin;
        a.Cout_pad = cin;
        HR_HIP(ctx, dn_conv(a, ctx->stream));
    }
    if (dweight || dbias) {
        WgradArgs g{};
        g.dout = static_cast<const __bf16*>(dy);
        g.in = static_cast<const __bf16*>(x);
        g.N = n;
        g.H = h;
        g.W = w;
        g.dout_C = cout;
        g.in_C = cin;
        g.Cin = cin;
        g.K = 9 * cin;
        g.Kpad = li.Kpad;
        g.Cout_pad = li.cout_pad;
        uint32_t chunks = dn_wgrad_chunks(P, &g.chunk);
        size_t nw = (size_t)li.cout_pad * li.Kpad;
        size_t bytes = ((size_t)chunks * nw + (size_t)chunks * li.cout_pad + nw + li.cout_pad) * sizeof(float);
        void* blk = nullptr;
        HR_HIP(ctx, hipMalloc(&blk, bytes));
        float* f = static_cast<float*>(blk);
        g.partial = f;
        g.bpartial = f + (size_t)chunks * nw;
        float* gw = g.bpartial + (size_t)chunks * li.cout_pad;
        float* gb = gw + nw;
        hipError_t e = dn_wgrad(g, gw, gb, ctx->stream);
        std::vector<float> hw(nw + li.cout_pad);
        if (e == hipSuccess) e = hipMemcpyAsync(hw.data(), gw, hw.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(blk);
        if (e != hipSuccess) return hip_fail(ctx, e, "conv3x3 weight gradient");
        for (uint32_t co = 0; co < cout; ++co) {
            if (dweight)
                for (uint32_t tp = 0; tp < 9; ++tp)
                    for (uint32_t ci = 0; ci < cin; ++ci)
                        dweight[((size_t)co * 9 + tp) * cin + ci] = hw[(size_t)co * li.Kpad + tp * cin + ci];
            if (dbias) dbias[co] = hw[nw + co];
        }
    }
    HR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HR_OK;
}

}  // extern "C"
