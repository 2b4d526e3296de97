// image_io.cpp - host-side image decoding for the asset loaders.
//
// The reference loads every texture with stbi_load(path, &w, &h, &c,
// STBI_rgb_alpha) (vk_textures.cpp:15) from stb_image (a dependency that is
// not vendored in the reference tree; its version is not pinned there). This
// file restates what that call returns, as RGBA8, for the formats the
// reference's assets use:
//   PNG   every colour type and bit depth, Adam7 interlacing, tRNS; 16-bit
//         samples keep their high byte, 1/2/4-bit grey is scaled to 8 bits
//         (x 255, 85, 17), palettes are expanded; zlib inflate (RFC 1950/1951)
//         from the system zlib
//   JPEG  baseline and progressive Huffman (ITU T.81), 1 or 3 components,
//         any sampling factors; the integer IDCT of the IJG "islow" method
//         (Loeffler, Ligtenberg, Moschytz) with 12-bit constants, 2x chroma
//         upsampling by the 3:1 triangle filter (other factors replicate),
//         YCbCr -> RGB in 20-bit fixed point
//   HDR   Radiance RGBE, flat and run-length scanlines; floats converted to
//         8 bits as stbi_load does for HDR files:
//         (uint8) clamp(pow(v, 1 / 2.2) * 255 + 0.5, 0, 255), alpha 255
// Files are untrusted: every size comes from the file, so images above
// kMaxPixels are refused, and no buffer is sized from a header before the
// data that would fill it is known to be there (inflate output grows with
// what the stream yields; HDR and JPEG check the bytes left first).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hr/hr.h"
#include "image_io.h"

namespace hr {
namespace img {

namespace {

struct Fail {
    std::string msg;
};
[[noreturn]] void fail(const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{buf};
}

// 16384 x 16384: RGBA8 output of at most 1 GiB
constexpr uint64_t kMaxPixels = 1ull << 28;
void check_size(const char* fmt, uint64_t w, uint64_t h) {
    if (w * h > kMaxPixels) fail("%s: %llu x %llu pixels exceed the %llu-pixel limit", fmt, (unsigned long long)w,
                                 (unsigned long long)h, (unsigned long long)kMaxPixels);
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
uint32_t be16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }

// ---- PNG -----------------------------------------------------------------------
// The first `expect` bytes of the zlib stream `in`. The output buffer grows
// (doubling, capped at expect) as the stream yields data, so a header that
// claims a large image costs memory only if the stream delivers it.
std::vector<uint8_t> inflate_all(const std::vector<uint8_t>& in, size_t expect) {
    if (in.size() > 0xFFFFFFFFu) fail("png: image data too large");
    std::vector<uint8_t> out(std::min<size_t>(std::max<size_t>(expect, 1), std::max<size_t>(in.size() * 8, 1u << 16)));
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) fail("png: zlib init failed");
    zs.next_in = const_cast<Bytef*>(in.data());
    zs.avail_in = (uInt)in.size();
    int rc = Z_OK;
    size_t got = 0;
    for (;;) {
        const size_t room = std::min<size_t>(out.size() - got, 0x40000000u);
        zs.next_out = out.data() + got;
        zs.avail_out = (uInt)room;
        rc = inflate(&zs, Z_NO_FLUSH);
        got += room - zs.avail_out;
        if (rc != Z_OK || got >= expect) break;
        if (zs.avail_out == 0 && got == out.size()) {  // full: grow towards expect
            try {
                out.resize(std::min<size_t>(std::max<size_t>(out.size(), 1) * 2, expect));
            } catch (...) {
                inflateEnd(&zs);
                throw;
            }
        } else if (zs.avail_in == 0) {
            break;  // input exhausted
        }
    }
    inflateEnd(&zs);
    if (rc != Z_STREAM_END && rc != Z_BUF_ERROR && rc != Z_OK) fail("png: corrupt image data (zlib %d)", rc);
    if (got < expect) fail("png: image data too short (%zu of %zu bytes)", got, expect);
    out.resize(expect);
    return out;
}

uint8_t paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

// Undo the per-row filters of one (sub)image in place: raw holds h rows of
// 1 + stride bytes; returns the h x stride unfiltered bytes.
std::vector<uint8_t> unfilter(const uint8_t* raw, uint32_t w, uint32_t h, uint32_t bits_pp) {
    size_t stride = ((size_t)w * bits_pp + 7) / 8;
    size_t bpp = bits_pp >= 8 ? bits_pp / 8 : 1;
    std::vector<uint8_t> out((size_t)h * stride);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* src = raw + (size_t)y * (stride + 1);
        uint8_t ft = src[0];
        ++src;
        uint8_t* cur = &out[(size_t)y * stride];
        const uint8_t* prev = y ? &out[(size_t)(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= bpp ? cur[i - bpp] : 0;
            int b = prev ? prev[i] : 0;
            int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int x = src[i];
            switch (ft) {
                case 0: cur[i] = (uint8_t)x; break;
                case 1: cur[i] = (uint8_t)(x + a); break;
                case 2: cur[i] = (uint8_t)(x + b); break;
                case 3: cur[i] = (uint8_t)(x + ((a + b) >> 1)); break;
                case 4: cur[i] = (uint8_t)(x + paeth(a, b, c)); break;
                default: fail("png: bad filter type %u", ft);
            }
        }
    }
    return out;
}

struct PngInfo {
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    uint8_t pal[256][4];
    uint32_t npal = 0;
    bool has_key = false;
    uint16_t key[3] = {0, 0, 0};
};

uint32_t channels_of(uint32_t ctype) {
    switch (ctype) {
        case 0: return 1;
        case 2: return 3;
        case 3: return 1;
        case 4: return 2;
        case 6: return 4;
    }
    fail("png: bad colour type %u", ctype);
}

// sample s of pixel x in an unfiltered row
uint32_t sample(const uint8_t* row, uint32_t idx, uint32_t depth) {
    if (depth == 8) return row[idx];
    if (depth == 16) return be16(row + 2 * (size_t)idx);
    uint32_t per = 8 / depth;
    uint32_t byte = row[idx / per];
    uint32_t shift = 8 - depth * (idx % per + 1);
    return (byte >> shift) & ((1u << depth) - 1u);
}

// Convert one unfiltered (sub)image to RGBA8 at output pixels
// (x0 + x * dx, y0 + y * dy).
void emit(const PngInfo& I, const std::vector<uint8_t>& px, uint32_t w, uint32_t h, uint32_t x0, uint32_t y0,
          uint32_t dx, uint32_t dy, uint8_t* out) {
    uint32_t ch = channels_of(I.ctype);
    size_t stride = ((size_t)w * ch * I.depth + 7) / 8;
    static const uint8_t scale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = &px[(size_t)y * stride];
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t* o = out + (((size_t)(y0 + y * dy) * I.w) + (x0 + x * dx)) * 4;
            uint32_t s[4];
            for (uint32_t k = 0; k < ch; ++k) s[k] = sample(row, x * ch + k, I.depth);
            auto to8 = [&](uint32_t v) -> uint8_t {
                if (I.depth == 16) return (uint8_t)(v >> 8);
                if (I.depth < 8) return (uint8_t)(v * scale[I.depth]);
                return (uint8_t)v;
            };
            switch (I.ctype) {
                case 0: {
                    uint8_t g = to8(s[0]);
                    bool t = I.has_key && (I.depth == 16 ? s[0] == I.key[0] : g == (uint8_t)(I.key[0] * (I.depth < 8 ? scale[I.depth] : 1)));
                    o[0] = o[1] = o[2] = g;
                    o[3] = t ? 0 : 255;
                    break;
                }
                case 2: {
                    bool t = I.has_key;
                    for (int k = 0; k < 3; ++k) {
                        o[k] = to8(s[k]);
                        if (I.depth == 16 ? s[k] != I.key[k] : o[k] != (uint8_t)I.key[k]) t = false;
                    }
                    o[3] = t ? 0 : 255;
                    break;
                }
                case 3: {
                    if (s[0] >= I.npal) fail("png: palette index %u out of range", s[0]);
                    memcpy(o, I.pal[s[0]], 4);
                    break;
                }
                case 4: o[0] = o[1] = o[2] = to8(s[0]); o[3] = to8(s[1]); break;
                case 6: for (int k = 0; k < 4; ++k) o[k] = to8(s[k]); break;
            }
        }
    }
}

void decode_png(const uint8_t* d, size_t n, Image& out) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (n < 8 || memcmp(d, sig, 8) != 0) fail("png: bad signature");
    PngInfo I;
    for (int i = 0; i < 256; ++i) I.pal[i][3] = 255;
    std::vector<uint8_t> idat;
    size_t p = 8;
    bool ihdr = false;
    while (p + 8 <= n) {
        uint32_t len = be32(d + p);
        const uint8_t* type = d + p + 4;
        if (p + 12 + (size_t)len > n) fail("png: truncated chunk");
        const uint8_t* c = d + p + 8;
        if (!memcmp(type, "IHDR", 4)) {
            if (len < 13) fail("png: bad IHDR");
            I.w = be32(c);
            I.h = be32(c + 4);
            I.depth = c[8];
            I.ctype = c[9];
            I.interlace = c[12];
            if (c[10] != 0 || c[11] != 0 || I.interlace > 1) fail("png: unsupported compression/filter/interlace");
            if (!I.w || !I.h || I.w > (1u << 24) || I.h > (1u << 24)) fail("png: bad size");
            check_size("png", I.w, I.h);
            uint32_t ch = channels_of(I.ctype);
            bool ok = I.depth == 8 || I.depth == 16 || ((I.ctype == 0 || I.ctype == 3) && I.depth < 8 &&
                                                        (I.depth == 1 || I.depth == 2 || I.depth == 4));
            if (!ok || (I.ctype == 3 && I.depth == 16)) fail("png: bad bit depth %u", I.depth);
            (void)ch;
            ihdr = true;
        } else if (!memcmp(type, "PLTE", 4)) {
            I.npal = len / 3;
            if (I.npal > 256 || len % 3) fail("png: bad PLTE");
            for (uint32_t i = 0; i < I.npal; ++i) {
                I.pal[i][0] = c[3 * i];
                I.pal[i][1] = c[3 * i + 1];
                I.pal[i][2] = c[3 * i + 2];
            }
        } else if (!memcmp(type, "tRNS", 4)) {
            if (I.ctype == 3) {
                for (uint32_t i = 0; i < len && i < 256; ++i) I.pal[i][3] = c[i];
            } else if (I.ctype == 0 && len >= 2) {
                I.has_key = true;
                I.key[0] = (uint16_t)be16(c);
            } else if (I.ctype == 2 && len >= 6) {
                I.has_key = true;
                for (int k = 0; k < 3; ++k) I.key[k] = (uint16_t)be16(c + 2 * k);
            }
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), c, c + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (!ihdr || idat.empty()) fail("png: missing IHDR or IDAT");
    // an 8-bit key for 8-bit images is compared in the 8-bit domain
    uint32_t ch = channels_of(I.ctype);
    uint32_t bpp = ch * I.depth;
    out.w = I.w;
    out.h = I.h;
    if (!I.interlace) {
        size_t raw = (size_t)I.h * (((size_t)I.w * bpp + 7) / 8 + 1);
        std::vector<uint8_t> z = inflate_all(idat, raw);
        out.px.assign((size_t)I.w * I.h * 4, 0);
        std::vector<uint8_t> px = unfilter(z.data(), I.w, I.h, bpp);
        emit(I, px, I.w, I.h, 0, 0, 1, 1, out.px.data());
        return;
    }
    static const uint32_t xo[7] = {0, 4, 0, 2, 0, 1, 0}, yo[7] = {0, 0, 4, 0, 2, 0, 1};
    static const uint32_t xs[7] = {8, 8, 4, 4, 2, 2, 1}, ys[7] = {8, 8, 8, 4, 4, 2, 2};
    size_t total = 0;
    uint32_t pw[7], ph[7];
    for (int k = 0; k < 7; ++k) {
        pw[k] = I.w > xo[k] ? (I.w - xo[k] + xs[k] - 1) / xs[k] : 0;
        ph[k] = I.h > yo[k] ? (I.h - yo[k] + ys[k] - 1) / ys[k] : 0;
        if (pw[k] && ph[k]) total += (size_t)ph[k] * (((size_t)pw[k] * bpp + 7) / 8 + 1);
    }
    std::vector<uint8_t> z = inflate_all(idat, total);
    out.px.assign((size_t)I.w * I.h * 4, 0);
    size_t off = 0;
    for (int k = 0; k < 7; ++k) {
        if (!pw[k] || !ph[k]) continue;
        std::vector<uint8_t> px = unfilter(z.data() + off, pw[k], ph[k], bpp);
        emit(I, px, pw[k], ph[k], xo[k], yo[k], xs[k], ys[k], out.px.data());
        off += (size_t)ph[k] * (((size_t)pw[k] * bpp + 7) / 8 + 1);
    }
}

// ---- JPEG ----------------------------------------------------------------------
const uint8_t kZigzag[64 + 16] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                  35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                  58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
                                  // extra entries so a corrupt run past 63 stays in bounds
                                  63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct Huff {
    // canonical Huffman code: for each length, the first code and the index
    // of its first symbol
    uint8_t sym[256];
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    bool set = false;
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int bw = 0, bh = 0;  // blocks per line / column (padded to whole MCUs)
    std::vector<int16_t> coef;  // bw * bh * 64, natural order
    std::vector<uint8_t> pix;   // (bw * 8) x (bh * 8)
    int dc_pred = 0;
};

struct Jpeg {
    const uint8_t* d;
    size_t n, p = 0;
    uint16_t q[4][64];
    Huff hdc[4], hac[4];
    std::vector<Comp> c;
    int w = 0, h = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false;
    int restart = 0;
    // bit reader
    uint32_t acc = 0;
    int nbits = 0;
    bool marker_hit = false;
    int eobrun = 0;
    bool adobe = false;
    int adobe_transform = -1;

    int byte() {
        if (p >= n) fail("jpeg: truncated");
        return d[p++];
    }
    int u16() {
        int a = byte();
        return (a << 8) | byte();
    }
    void fill() {
        while (nbits <= 24) {
            int b = 0;
            if (!marker_hit && p < n) {
                b = d[p];
                if (b == 0xFF) {
                    int b2 = p + 1 < n ? d[p + 1] : 0;
                    if (b2 == 0) {
                        p += 2;
                    } else {
                        marker_hit = true;  // a marker: feed zeros from here
                        b = 0;
                    }
                } else {
                    ++p;
                }
            }
            acc |= (uint32_t)b << (24 - nbits);
            nbits += 8;
        }
    }
    int bits(int k) {
        if (k == 0) return 0;
        fill();
        int v = (int)(acc >> (32 - k));
        acc <<= k;
        nbits -= k;
        return v;
    }
    int bit() { return bits(1); }
    int decode(const Huff& hf) {
        if (!hf.set) fail("jpeg: missing Huffman table");
        fill();
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | (int)(acc >> 31);
            acc <<= 1;
            --nbits;
            if (hf.maxcode[l] >= 0 && code <= hf.maxcode[l] && code >= hf.mincode[l])
                return hf.sym[hf.valptr[l] + code - hf.mincode[l]];
        }
        fail("jpeg: bad Huffman code");
    }
    static int extend(int v, int t) { return t && v < (1 << (t - 1)) ? v - (1 << t) + 1 : v; }
    int receive_extend(int s) { return s ? extend(bits(s), s) : 0; }
    // a DC difference category (T.81 F.1.2.1: at most 11 for 8-bit samples;
    // 15 is the most the bit reader takes at once)
    int dc_category(const Huff& hf) {
        const int t = decode(hf);
        if (t > 15) fail("jpeg: bad DC category %d", t);
        return t;
    }
    void reset_bits() {
        acc = 0;
        nbits = 0;
        marker_hit = false;
    }
};

void read_dht(Jpeg& J, int len) {
    int end = (int)J.p + len - 2;
    while ((int)J.p < end) {
        int tc_th = J.byte();
        int tc = tc_th >> 4, th = tc_th & 15;
        if (tc > 1 || th > 3) fail("jpeg: bad DHT");
        int counts[17] = {0};
        int total = 0;
        for (int l = 1; l <= 16; ++l) total += counts[l] = J.byte();
        if (total > 256) fail("jpeg: bad DHT");
        Huff& hf = tc ? J.hac[th] : J.hdc[th];
        for (int i = 0; i < total; ++i) hf.sym[i] = (uint8_t)J.byte();
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            hf.valptr[l] = k;
            hf.mincode[l] = code;
            code += counts[l];
            k += counts[l];
            hf.maxcode[l] = counts[l] ? code - 1 : -1;
            code <<= 1;
        }
        hf.set = true;
    }
}

// progressive and baseline block decoding (ITU T.81 F.2.2, G.1.2)
// The DC predictor sums the differences modulo 2^32 (a corrupt file may
// drive it past the int range; valid files stay far inside it).
inline int add_wrap(int a, int b) { return (int)(int32_t)((uint32_t)a + (uint32_t)b); }

void decode_block_baseline(Jpeg& J, Comp& C, int16_t* blk) {
    int t = J.dc_category(J.hdc[C.td]);
    int diff = J.receive_extend(t);
    C.dc_pred = add_wrap(C.dc_pred, diff);
    blk[0] = (int16_t)C.dc_pred;
    for (int k = 1; k < 64;) {
        int rs = J.decode(J.hac[C.ta]);
        int r = rs >> 4, s = rs & 15;
        if (s == 0) {
            if (r != 15) break;
            k += 16;
            continue;
        }
        k += r;
        if (k > 63) fail("jpeg: bad AC run");
        blk[kZigzag[k]] = (int16_t)J.receive_extend(s);
        ++k;
    }
}

void decode_dc_prog(Jpeg& J, Comp& C, int16_t* blk, int ah, int al) {
    if (ah == 0) {
        int t = J.dc_category(J.hdc[C.td]);
        int diff = J.receive_extend(t);
        C.dc_pred = add_wrap(C.dc_pred, diff);
        blk[0] = (int16_t)((uint32_t)C.dc_pred << al);
    } else if (J.bit()) {
        blk[0] = (int16_t)(blk[0] | (1 << al));
    }
}

void decode_ac_first(Jpeg& J, Comp& C, int16_t* blk, int ss, int se, int al) {
    if (J.eobrun > 0) {
        --J.eobrun;
        return;
    }
    for (int k = ss; k <= se;) {
        int rs = J.decode(J.hac[C.ta]);
        int r = rs >> 4, s = rs & 15;
        if (s == 0) {
            if (r < 15) {
                J.eobrun = (1 << r) - 1;
                if (r) J.eobrun += J.bits(r);
                break;
            }
            k += 16;
            continue;
        }
        k += r;
        if (k > 63) fail("jpeg: bad AC run");
        blk[kZigzag[k]] = (int16_t)(J.receive_extend(s) * (1 << al));
        ++k;
    }
}

void decode_ac_refine(Jpeg& J, Comp& C, int16_t* blk, int ss, int se, int al) {
    int p1 = 1 << al, m1 = -1 * (1 << al);
    int k = ss;
    if (J.eobrun <= 0) {
        for (; k <= se;) {
            int rs = J.decode(J.hac[C.ta]);
            int r = rs >> 4, s = rs & 15;
            int val = 0;
            if (s == 0) {
                if (r < 15) {
                    J.eobrun = (1 << r);
                    if (r) J.eobrun += J.bits(r);
                    break;  // the rest of the band is refined below
                }
                // r == 15: skip 16 zero coefficients (refining nonzero ones on the way)
            } else {
                if (s != 1) fail("jpeg: bad refinement code");
                val = J.bit() ? p1 : m1;
            }
            while (k <= se) {
                int16_t* coef = &blk[kZigzag[k]];
                if (*coef != 0) {
                    if (J.bit() && (*coef & p1) == 0) *coef = (int16_t)(*coef >= 0 ? *coef + p1 : *coef + m1);
                } else {
                    if (r == 0) {
                        if (val) *coef = (int16_t)val;
                        ++k;
                        break;
                    }
                    --r;
                }
                ++k;
            }
        }
    }
    if (J.eobrun > 0) {
        for (; k <= se; ++k) {
            int16_t* coef = &blk[kZigzag[k]];
            if (*coef != 0 && J.bit() && (*coef & p1) == 0)
                *coef = (int16_t)(*coef >= 0 ? *coef + p1 : *coef + m1);
        }
        --J.eobrun;
    }
}

// IJG "islow" integer IDCT with 12-bit constants; input dequantized
// coefficients in natural order, output 8x8 samples (level shift +128).
// The sums and products are taken modulo 2^32 (unsigned), which is what
// 32-bit int arithmetic gives whenever it does not overflow, and stays
// defined for the out-of-range coefficients of a corrupt file; the shifts
// are arithmetic shifts of the 32-bit two's-complement value.
inline uint32_t f2f(double x) { return (uint32_t)(int)(x * 4096 + 0.5); }
inline int asr(uint32_t v, int n) { return (int)(int32_t)v >> n; }
inline uint8_t clamp8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

#define HR_IDCT_1D(s0, s1, s2, s3, s4, s5, s6, s7)               \
    uint32_t t0, t1, t2, t3, p1, p2, p3, p4, p5, x0, x1, x2, x3; \
    p2 = (uint32_t)(s2);                                         \
    p3 = (uint32_t)(s6);                                         \
    p1 = (p2 + p3) * f2f(0.5411961);                             \
    t2 = p1 + p3 * f2f(-1.847759065);                            \
    t3 = p1 + p2 * f2f(0.765366865);                             \
    p2 = (uint32_t)(s0);                                         \
    p3 = (uint32_t)(s4);                                         \
    t0 = (p2 + p3) * 4096u;                                      \
    t1 = (p2 - p3) * 4096u;                                      \
    x0 = t0 + t3;                                                \
    x3 = t0 - t3;                                                \
    x1 = t1 + t2;                                                \
    x2 = t1 - t2;                                                \
    t0 = (uint32_t)(s7);                                         \
    t1 = (uint32_t)(s5);                                         \
    t2 = (uint32_t)(s3);                                         \
    t3 = (uint32_t)(s1);                                         \
    p3 = t0 + t2;                                                \
    p4 = t1 + t3;                                                \
    p1 = t0 + t3;                                                \
    p2 = t1 + t2;                                                \
    p5 = (p3 + p4) * f2f(1.175875602);                           \
    t0 = t0 * f2f(0.298631336);                                  \
    t1 = t1 * f2f(2.053119869);                                  \
    t2 = t2 * f2f(3.072711026);                                  \
    t3 = t3 * f2f(1.501321110);                                  \
    p1 = p5 + p1 * f2f(-0.899976223);                            \
    p2 = p5 + p2 * f2f(-2.562915447);                            \
    p3 = p3 * f2f(-1.961570560);                                 \
    p4 = p4 * f2f(-0.390180644);                                 \
    t3 += p1 + p4;                                               \
    t2 += p2 + p3;                                               \
    t1 += p2 + p4;                                               \
    t0 += p1 + p3;

void idct_block(uint8_t* out, int out_stride, const int16_t* d) {
    int v[64];
    for (int i = 0; i < 8; ++i) {  // columns
        const int16_t* s = d + i;
        int* o = v + i;
        if (s[8] == 0 && s[16] == 0 && s[24] == 0 && s[32] == 0 && s[40] == 0 && s[48] == 0 && s[56] == 0) {
            int dc = s[0] * 4;
            for (int k = 0; k < 8; ++k) o[8 * k] = dc;
        } else {
            HR_IDCT_1D(s[0], s[8], s[16], s[24], s[32], s[40], s[48], s[56])
            x0 += 512;
            x1 += 512;
            x2 += 512;
            x3 += 512;
            o[0] = asr(x0 + t3, 10);
            o[56] = asr(x0 - t3, 10);
            o[8] = asr(x1 + t2, 10);
            o[48] = asr(x1 - t2, 10);
            o[16] = asr(x2 + t1, 10);
            o[40] = asr(x2 - t1, 10);
            o[24] = asr(x3 + t0, 10);
            o[32] = asr(x3 - t0, 10);
        }
    }
    for (int i = 0; i < 8; ++i) {  // rows
        const int* s = v + 8 * i;
        uint8_t* o = out + (size_t)i * out_stride;
        HR_IDCT_1D(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7])
        x0 += 65536u + (128u << 17);
        x1 += 65536u + (128u << 17);
        x2 += 65536u + (128u << 17);
        x3 += 65536u + (128u << 17);
        o[0] = clamp8(asr(x0 + t3, 17));
        o[7] = clamp8(asr(x0 - t3, 17));
        o[1] = clamp8(asr(x1 + t2, 17));
        o[6] = clamp8(asr(x1 - t2, 17));
        o[2] = clamp8(asr(x2 + t1, 17));
        o[5] = clamp8(asr(x2 - t1, 17));
        o[3] = clamp8(asr(x3 + t0, 17));
        o[4] = clamp8(asr(x3 - t0, 17));
    }
}
#undef HR_IDCT_1D

void read_sof(Jpeg& J, int len) {
    if (J.byte() != 8) fail("jpeg: only 8-bit samples are supported");
    J.h = J.u16();
    J.w = J.u16();
    int nc = J.byte();
    if (!J.w || !J.h) fail("jpeg: zero size (DNL is not supported)");
    check_size("jpeg", (uint64_t)J.w, (uint64_t)J.h);
    if (nc != 1 && nc != 3) fail("jpeg: %d components (1 or 3 supported)", nc);
    if (len != 8 + 3 * nc) fail("jpeg: bad SOF");
    J.c.resize(nc);
    for (auto& C : J.c) {
        C.id = J.byte();
        int hv = J.byte();
        C.h = hv >> 4;
        C.v = hv & 15;
        C.tq = J.byte();
        if (C.h < 1 || C.h > 4 || C.v < 1 || C.v > 4 || C.tq > 3) fail("jpeg: bad component");
        J.hmax = std::max(J.hmax, C.h);
        J.vmax = std::max(J.vmax, C.v);
    }
    J.mcux = (J.w + 8 * J.hmax - 1) / (8 * J.hmax);
    J.mcuy = (J.h + 8 * J.vmax - 1) / (8 * J.vmax);
    // every block inside the image codes its DC coefficient with at least one
    // bit in a scan after this frame header
    uint64_t blocks = 0;
    for (const auto& C : J.c) {
        const uint64_t cw = (((uint64_t)J.w * C.h + J.hmax - 1) / J.hmax + 7) / 8;
        const uint64_t ch = (((uint64_t)J.h * C.v + J.vmax - 1) / J.vmax + 7) / 8;
        blocks += cw * ch;
    }
    if (blocks > 8 * (uint64_t)(J.n - J.p)) fail("jpeg: truncated (%llu blocks in %zu bytes)",
                                                 (unsigned long long)blocks, J.n - J.p);
    for (auto& C : J.c) {
        C.bw = J.mcux * C.h;
        C.bh = J.mcuy * C.v;
        C.coef.assign((size_t)C.bw * C.bh * 64, 0);
    }
}

void read_scan(Jpeg& J) {
    int ns = J.byte();
    if (ns < 1 || ns > (int)J.c.size()) fail("jpeg: bad SOS");
    std::vector<Comp*> sc;
    for (int i = 0; i < ns; ++i) {
        int id = J.byte(), t = J.byte();
        Comp* found = nullptr;
        for (auto& C : J.c)
            if (C.id == id) found = &C;
        if (!found) fail("jpeg: SOS names an unknown component");
        found->td = t >> 4;
        found->ta = t & 15;
        if (found->td > 3 || found->ta > 3) fail("jpeg: bad table index");
        sc.push_back(found);
    }
    int ss = J.byte(), se = J.byte(), a = J.byte();
    int ah = a >> 4, al = a & 15;
    if (!J.progressive) {
        ss = 0;
        se = 63;
        ah = al = 0;
    } else {
        if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1)) fail("jpeg: bad progressive scan");
    }
    J.reset_bits();
    J.eobrun = 0;
    for (auto* C : sc) C->dc_pred = 0;
    // MCU iteration: interleaved scans walk MCUs, single-component scans walk
    // the component's own blocks (only those inside the image)
    int mcus;
    int cw = 0, ch = 0;
    if (ns == 1) {
        Comp& C = *sc[0];
        cw = (((J.w * C.h + J.hmax - 1) / J.hmax) + 7) / 8;
        ch = (((J.h * C.v + J.vmax - 1) / J.vmax) + 7) / 8;
        mcus = cw * ch;
    } else {
        mcus = J.mcux * J.mcuy;
    }
    int todo = J.restart ? J.restart : 0x7fffffff;
    for (int m = 0; m < mcus; ++m) {
        if (ns == 1) {
            Comp& C = *sc[0];
            int bx = m % cw, by = m / cw;
            int16_t* blk = &C.coef[((size_t)by * C.bw + bx) * 64];
            if (!J.progressive) decode_block_baseline(J, C, blk);
            else if (ss == 0) decode_dc_prog(J, C, blk, ah, al);
            else if (ah == 0) decode_ac_first(J, C, blk, ss, se, al);
            else decode_ac_refine(J, C, blk, ss, se, al);
        } else {
            int mx = m % J.mcux, my = m / J.mcux;
            for (auto* Cp : sc) {
                Comp& C = *Cp;
                for (int yy = 0; yy < C.v; ++yy)
                    for (int xx = 0; xx < C.h; ++xx) {
                        int bx = mx * C.h + xx, by = my * C.v + yy;
                        int16_t* blk = &C.coef[((size_t)by * C.bw + bx) * 64];
                        if (!J.progressive) decode_block_baseline(J, C, blk);
                        else decode_dc_prog(J, C, blk, ah, al);
                    }
            }
        }
        if (--todo == 0 && m + 1 < mcus) {
            // expect RSTn: skip to the marker and reset predictors
            while (J.p + 1 < J.n && !(J.d[J.p] == 0xFF && J.d[J.p + 1] >= 0xD0 && J.d[J.p + 1] <= 0xD7)) ++J.p;
            if (J.p + 1 >= J.n) fail("jpeg: missing restart marker");
            J.p += 2;
            J.reset_bits();
            J.eobrun = 0;
            for (auto* C : sc) C->dc_pred = 0;
            todo = J.restart;
        }
    }
    // move to the next marker
    while (J.p + 1 < J.n && !(J.d[J.p] == 0xFF && J.d[J.p + 1] != 0 && !(J.d[J.p + 1] >= 0xD0 && J.d[J.p + 1] <= 0xD7)))
        ++J.p;
}

// 2x upsampling by the 3:1 triangle filter (each output sample is 3/4 of the
// nearer and 1/4 of the farther input sample), horizontal then vertical as
// one 2-D filter for 2x2; other factors replicate samples.
void upsample(const Comp& C, const Jpeg& J, std::vector<uint8_t>& full) {
    int sw = C.bw * 8, sh = C.bh * 8;
    int fx = J.hmax / C.h, fy = J.vmax / C.v;
    full.assign((size_t)J.w * J.h, 0);
    if (fx == 1 && fy == 1) {
        for (int y = 0; y < J.h; ++y) memcpy(&full[(size_t)y * J.w], &C.pix[(size_t)y * sw], J.w);
        return;
    }
    int lw = (J.w + fx - 1) / fx, lh = (J.h + fy - 1) / fy;  // low-res samples that matter
    auto at = [&](int x, int y) -> int { return C.pix[(size_t)y * sw + x]; };
    (void)sh;
    if (fx == 2 && fy == 2) {
        for (int y = 0; y < J.h; ++y) {
            int yl = y >> 1;
            int yf = (y & 1) ? std::min(yl + 1, lh - 1) : std::max(yl - 1, 0);
            for (int x = 0; x < J.w; ++x) {
                int xl = x >> 1;
                int xf = (x & 1) ? std::min(xl + 1, lw - 1) : std::max(xl - 1, 0);
                int tn = 3 * at(xl, yl) + at(xl, yf);  // column sums at the near and far columns
                int tf = 3 * at(xf, yl) + at(xf, yf);
                int v = lw == 1 ? (4 * tn + 8) >> 4 : (3 * tn + tf + 8) >> 4;
                full[(size_t)y * J.w + x] = (uint8_t)v;
            }
        }
        return;
    }
    if (fx == 2 && fy == 1) {
        for (int y = 0; y < J.h; ++y)
            for (int x = 0; x < J.w; ++x) {
                int xl = x >> 1;
                int xf = (x & 1) ? std::min(xl + 1, lw - 1) : std::max(xl - 1, 0);
                full[(size_t)y * J.w + x] = (uint8_t)((3 * at(xl, y) + at(xf, y) + 2) >> 2);
            }
        return;
    }
    if (fx == 1 && fy == 2) {
        for (int y = 0; y < J.h; ++y) {
            int yl = y >> 1;
            int yf = (y & 1) ? std::min(yl + 1, lh - 1) : std::max(yl - 1, 0);
            for (int x = 0; x < J.w; ++x)
                full[(size_t)y * J.w + x] = (uint8_t)((3 * at(x, yl) + at(x, yf) + 2) >> 2);
        }
        return;
    }
    for (int y = 0; y < J.h; ++y)
        for (int x = 0; x < J.w; ++x) full[(size_t)y * J.w + x] = (uint8_t)at(x / fx, y / fy);
}

void decode_jpeg(const uint8_t* d, size_t n, Image& out) {
    Jpeg J;
    J.d = d;
    J.n = n;
    memset(J.q, 0, sizeof(J.q));
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) fail("jpeg: bad signature");
    J.p = 2;
    bool frame = false, done = false;
    while (!done) {
        // find the next marker
        while (J.p < n && d[J.p] != 0xFF) ++J.p;
        while (J.p < n && d[J.p] == 0xFF) ++J.p;
        if (J.p >= n) break;
        int m = d[J.p++];
        switch (m) {
            case 0xD8: break;
            case 0xD9: done = true; break;
            case 0xC0:
            case 0xC1:
            case 0xC2: {
                int len = J.u16();
                J.progressive = m == 0xC2;
                read_sof(J, len);
                frame = true;
                break;
            }
            case 0xC4: read_dht(J, J.u16()); break;
            case 0xDB: {
                int len = J.u16();
                int end = (int)J.p + len - 2;
                while ((int)J.p < end) {
                    int pq_tq = J.byte();
                    int pq = pq_tq >> 4, tq = pq_tq & 15;
                    if (tq > 3 || pq > 1) fail("jpeg: bad DQT");
                    for (int k = 0; k < 64; ++k) J.q[tq][kZigzag[k]] = (uint16_t)(pq ? J.u16() : J.byte());
                }
                break;
            }
            case 0xDD: J.u16(); J.restart = J.u16(); break;
            case 0xDA: {
                if (!frame) fail("jpeg: scan before frame");
                J.u16();
                read_scan(J);
                break;
            }
            case 0xEE: {  // Adobe APP14: colour transform flag
                int len = J.u16();
                if (len >= 14 && J.p + 12 <= n && !memcmp(d + J.p, "Adobe", 5)) {
                    J.adobe = true;
                    J.adobe_transform = d[J.p + 11];
                }
                J.p += len - 2;
                break;
            }
            default:
                if ((m >= 0xC3 && m <= 0xCF)) fail("jpeg: unsupported coding process (SOF%d)", m - 0xC0);
                if (m >= 0xD0 && m <= 0xD7) break;
                J.p += J.u16() - 2;  // APPn, COM and the rest
                break;
        }
    }
    if (!frame) fail("jpeg: no frame");
    for (auto& C : J.c) {
        C.pix.assign((size_t)C.bw * 8 * C.bh * 8, 0);
        int16_t tmp[64];
        for (int by = 0; by < C.bh; ++by)
            for (int bx = 0; bx < C.bw; ++bx) {
                const int16_t* blk = &C.coef[((size_t)by * C.bw + bx) * 64];
                for (int k = 0; k < 64; ++k) tmp[k] = (int16_t)(blk[k] * J.q[C.tq][k]);
                idct_block(&C.pix[((size_t)by * 8) * C.bw * 8 + bx * 8], C.bw * 8, tmp);
            }
    }
    out.w = (uint32_t)J.w;
    out.h = (uint32_t)J.h;
    out.px.assign((size_t)J.w * J.h * 4, 255);
    std::vector<uint8_t> ch[3];
    for (size_t k = 0; k < J.c.size(); ++k) upsample(J.c[k], J, ch[k]);
    size_t np = (size_t)J.w * J.h;
    if (J.c.size() == 1) {
        for (size_t i = 0; i < np; ++i) out.px[4 * i] = out.px[4 * i + 1] = out.px[4 * i + 2] = ch[0][i];
        return;
    }
    bool rgb = J.adobe && J.adobe_transform == 0;
    if (!J.adobe && J.c[0].id == 'R' && J.c[1].id == 'G' && J.c[2].id == 'B') rgb = true;
    for (size_t i = 0; i < np; ++i) {
        if (rgb) {
            out.px[4 * i] = ch[0][i];
            out.px[4 * i + 1] = ch[1][i];
            out.px[4 * i + 2] = ch[2][i];
            continue;
        }
        // JFIF YCbCr -> RGB in 20-bit fixed point
        auto fx20 = [](double x) { return ((int)(x * 4096.0 + 0.5)) << 8; };
        int yf = ((int)ch[0][i] << 20) + (1 << 19);
        int cb = (int)ch[1][i] - 128, cr = (int)ch[2][i] - 128;
        int r = yf + cr * fx20(1.40200);
        int g = yf + (cr * -fx20(0.71414)) + ((cb * -fx20(0.34414)) & (int)0xffff0000);
        int b = yf + cb * fx20(1.77200);
        out.px[4 * i] = clamp8(r >> 20);
        out.px[4 * i + 1] = clamp8(g >> 20);
        out.px[4 * i + 2] = clamp8(b >> 20);
    }
}

// ---- Radiance HDR ----------------------------------------------------------------
void decode_hdr(const uint8_t* d, size_t n, Image& out, std::vector<float>* fout) {
    size_t p = 0;
    auto line = [&]() {
        std::string s;
        while (p < n && d[p] != '\n') s.push_back((char)d[p++]);
        if (p < n) ++p;
        return s;
    };
    std::string first = line();
    if (first != "#?RADIANCE" && first != "#?RGBE") fail("hdr: bad signature");
    bool fmt = false;
    for (;;) {
        if (p >= n) fail("hdr: truncated header");
        std::string s = line();
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    if (!fmt) fail("hdr: unsupported format");
    std::string res = line();
    // "-Y <height> +X <width>" (the only orientation stbi_load reads)
    auto field = [&](size_t& k, const char* tag) -> long {
        const size_t t = strlen(tag);
        if (res.compare(k, t, tag) != 0) fail("hdr: unsupported orientation");
        k += t;
        long v = 0;
        size_t digits = 0;
        while (k < res.size() && res[k] >= '0' && res[k] <= '9' && digits < 9) v = v * 10 + (res[k++] - '0'), ++digits;
        if (!digits || v <= 0) fail("hdr: unsupported orientation");
        return v;
    };
    size_t k = 0;
    const long hl = field(k, "-Y ");
    const long wl = field(k, " +X ");
    if (k != res.size()) fail("hdr: unsupported orientation");
    check_size("hdr", (uint64_t)wl, (uint64_t)hl);
    const int h = (int)hl, w = (int)wl;
    // the least data the scanlines can take: flat RGBE, or per run-length
    // line a 4-byte header and per channel 2 bytes per 127 values
    const bool rle_ok = !(w < 8 || w >= 32768);
    const uint64_t least = rle_ok ? (uint64_t)h * (4u + 8u * (((uint64_t)w + 126u) / 127u)) : (uint64_t)w * h * 4u;
    if (least > n - p) fail("hdr: truncated data");
    std::vector<uint8_t> rgbe((size_t)w * h * 4);
    auto flat = [&](size_t first_px) {
        size_t need = ((size_t)w * h - first_px) * 4;
        if (p + need > n) fail("hdr: truncated data");
        memcpy(&rgbe[first_px * 4], d + p, need);
        p += need;
    };
    if (!rle_ok) {
        flat(0);
    } else {
        std::vector<uint8_t> line_buf((size_t)w * 4);
        for (int y = 0; y < h; ++y) {
            if (p + 4 > n) fail("hdr: truncated data");
            uint8_t c1 = d[p], c2 = d[p + 1], c3 = d[p + 2];
            if (c1 != 2 || c2 != 2 || (c3 & 0x80)) {  // not run-length encoded: flat from here
                flat((size_t)y * w);
                break;
            }
            int len = (d[p + 2] << 8) | d[p + 3];
            p += 4;
            if (len != w) fail("hdr: bad scanline length");
            for (int k = 0; k < 4; ++k) {
                int x = 0;
                while (x < w) {
                    if (p >= n) fail("hdr: truncated data");
                    int count = d[p++];
                    if (count > 128) {
                        count -= 128;
                        if (p >= n || x + count > w) fail("hdr: bad run");
                        uint8_t v = d[p++];
                        for (int i = 0; i < count; ++i) line_buf[(size_t)(x++) * 4 + k] = v;
                    } else {
                        if (count == 0 || x + count > w || p + count > n) fail("hdr: bad literal run");
                        for (int i = 0; i < count; ++i) line_buf[(size_t)(x++) * 4 + k] = d[p++];
                    }
                }
            }
            memcpy(&rgbe[(size_t)y * w * 4], line_buf.data(), line_buf.size());
        }
    }
    out.w = (uint32_t)w;
    out.h = (uint32_t)h;
    out.px.assign((size_t)w * h * 4, 255);
    if (fout) fout->assign((size_t)w * h * 4, 1.0f);
    const float gamma = 1.0f / 2.2f, scale = 1.0f;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* e = &rgbe[4 * i];
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (e[3]) {
            float m = (float)ldexp(1.0f, (int)e[3] - (int)(128 + 8));
            f[0] = e[0] * m;
            f[1] = e[1] * m;
            f[2] = e[2] * m;
        }
        for (int k = 0; k < 3; ++k) {
            if (fout) (*fout)[4 * i + k] = f[k];
            float z = (float)pow((double)(f[k] * scale), (double)gamma) * 255 + 0.5f;
            if (z < 0) z = 0;
            if (z > 255) z = 255;
            out.px[4 * i + k] = (uint8_t)(int)z;
        }
    }
}

}  // namespace

bool decode(const uint8_t* data, size_t size, Image& out, std::string& err, std::vector<float>* hdr_float) {
    try {
        if (size >= 8 && data[0] == 0x89 && data[1] == 'P' && data[2] == 'N' && data[3] == 'G') {
            decode_png(data, size, out);
        } else if (size >= 3 && data[0] == 0xFF && data[1] == 0xD8) {
            decode_jpeg(data, size, out);
        } else if (size >= 6 && data[0] == '#' && data[1] == '?') {
            decode_hdr(data, size, out, hdr_float);
        } else {
            err = "unknown image format (PNG, JPEG and Radiance HDR are supported)";
            return false;
        }
    } catch (const Fail& f) {
        err = f.msg;
        return false;
    } catch (const std::bad_alloc&) {
        err = "out of memory";
        return false;
    } catch (const std::exception& e) {
        err = std::string("decode failed: ") + e.what();
        return false;
    }
    return true;
}

bool read_file(const std::string& path, std::vector<uint8_t>& bytes, std::string& err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        fclose(f);
        err = "cannot read " + path;
        return false;
    }
    bytes.resize((size_t)sz);
    size_t got = sz ? fread(bytes.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if (got != (size_t)sz) {
        err = "short read on " + path;
        return false;
    }
    return true;
}

}  // namespace img
}  // namespace hr

namespace {
void set_err(char* err, size_t len, const std::string& m) {
    if (err && len) {
        snprintf(err, len, "%s", m.c_str());
    }
}
hr_status to_image(hr::img::Image& im, hr_image* out) {
    out->width = im.w;
    out->height = im.h;
    out->rgba8 = static_cast<uint8_t*>(malloc(im.px.size()));
    if (!out->rgba8) return HR_E_OOM;
    memcpy(out->rgba8, im.px.data(), im.px.size());
    return HR_OK;
}
}  // namespace

extern "C" {

hr_status hr_image_decode(const uint8_t* data, size_t size, hr_image* out, char* err, size_t err_len) {
    if (!out || (!data && size)) return HR_E_INVALID;
    memset(out, 0, sizeof(*out));
    hr::img::Image im;
    std::string e;
    if (!hr::img::decode(data, size, im, e)) {
        set_err(err, err_len, e);
        return HR_E_INVALID;
    }
    return to_image(im, out);
}

hr_status hr_image_load(const char* path, hr_image* out, char* err, size_t err_len) {
    if (!path || !out) return HR_E_INVALID;
    memset(out, 0, sizeof(*out));
    std::vector<uint8_t> bytes;
    std::string e;
    if (!hr::img::read_file(path, bytes, e)) {
        set_err(err, err_len, e);
        return HR_E_INVALID;
    }
    return hr_image_decode(bytes.data(), bytes.size(), out, err, err_len);
}

void hr_image_free(hr_image* img) {
    if (!img) return;
    free(img->rgba8);
    img->rgba8 = nullptr;
    img->width = img->height = 0;
}

}  // extern "C"
