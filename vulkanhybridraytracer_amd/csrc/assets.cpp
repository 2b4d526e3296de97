// assets.cpp - host-side asset ingest: OBJ and glTF 2.0 (.gltf / .glb) files
// to the arrays hr_scene_build and hr_scene_set_textures take.
//
// Restates the reference's loaders:
//   OBJ   Mesh::load_from_obj (vk_mesh.cpp:96-150) over tinyobjloader's
//         LoadObj (triangulate = true): faces fan-triangulated, every face
//         vertex (position, normal, colour (1, 1, 1), uv = (u, 1 - v) when the
//         file has texture coordinates) deduplicated by value in first-seen
//         order; one primitive over all indices with the default material
//         (Prefab::createOBJprefab, vk_mesh.cpp:832-841; Material defaults,
//         material.h:17-31).
//   glTF  Prefab::GET / loadNode / loadMaterial / loadTextures
//         (vk_mesh.cpp:646-901): scene 0's root nodes in order; a node's
//         children are loaded before its own primitives (so their vertices
//         come first); local matrix T * R * S or `matrix`; POSITION, NORMAL
//         (normalised; missing -> normalize(0)), TEXCOORD_0 (missing -> 0);
//         invertNormals applies to the root nodes only (children are loaded
//         with the default false); indices u8/u16/u32 offset by the
//         primitive's first vertex; materials deduplicated by value
//         (Material::operator==, material.cpp:51-62) before their texture
//         indices are remapped to texture ids in first-use order
//         (Texture::get_id, vk_textures.cpp:154-167). A material's texture
//         index selects images[index] as the reference does
//         (vk_mesh.cpp:809-827); HR_ASSET_TEXTURE_SOURCE selects
//         images[textures[index].source] instead.
// The primitive list is in the order the reference instances them
// (Node::node_to_instance, vk_mesh.cpp:494-515: a node's primitives, then
// its children), each with model = the node's global matrix
// (Node::getGlobalMatrix, vk_mesh.cpp:416-422).
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "hr/hr.h"
#include "image_io.h"

struct hr_asset {
    std::vector<hr_vertex> verts;
    std::vector<uint32_t> idx;
    std::vector<hr_primitive> prims;
    std::vector<hr_material> mats;
    std::vector<hr::img::Image> images;
    std::vector<hr_texture> tex;
};

namespace {

struct Fail {
    std::string msg;
};
[[noreturn]] void fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{buf};
}

// ---- column-major 4x4 float matrices, products as glm evaluates them ----------
struct M4 {
    float m[16];  // m[c * 4 + r]
};
M4 ident() {
    M4 a;
    for (int i = 0; i < 16; ++i) a.m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    return a;
}
// result column i = ((A0 * B[i][0] + A1 * B[i][1]) + A2 * B[i][2]) + A3 * B[i][3]
M4 mul(const M4& A, const M4& B) {
    M4 R;
    for (int i = 0; i < 4; ++i)
        for (int r = 0; r < 4; ++r) {
            float s = A.m[0 * 4 + r] * B.m[i * 4 + 0];
            s = s + A.m[1 * 4 + r] * B.m[i * 4 + 1];
            s = s + A.m[2 * 4 + r] * B.m[i * 4 + 2];
            s = s + A.m[3 * 4 + r] * B.m[i * 4 + 3];
            R.m[i * 4 + r] = s;
        }
    return R;
}

// ---- JSON -------------------------------------------------------------------------
struct JVal {
    enum Type { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal* get(const char* k) const {
        if (t != OBJ) return nullptr;
        for (auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    const JVal& at(size_t i) const {
        if (t != ARR || i >= a.size()) fail("gltf: index %zu out of range", i);
        return a[i];
    }
    size_t size() const { return t == ARR ? a.size() : 0; }
};

struct JParser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    void expect(char c) {
        ws();
        if (p >= e || *p != c) fail("gltf: JSON syntax error (expected '%c')", c);
        ++p;
    }
    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s.push_back((char)cp);
        } else if (cp < 0x800) {
            s.push_back((char)(0xC0 | (cp >> 6)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            s.push_back((char)(0xE0 | (cp >> 12)));
            s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            s.push_back((char)(0xF0 | (cp >> 18)));
            s.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
            s.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            s.push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    uint32_t hex4() {
        if (e - p < 4) fail("gltf: bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("gltf: bad \\u escape");
        }
        return v;
    }
    std::string str() {
        expect('"');
        std::string s;
        while (p < e && *p != '"') {
            char c = *p++;
            if (c != '\\') {
                s.push_back(c);
                continue;
            }
            if (p >= e) break;
            char x = *p++;
            switch (x) {
                case '"': s.push_back('"'); break;
                case '\\': s.push_back('\\'); break;
                case '/': s.push_back('/'); break;
                case 'b': s.push_back('\b'); break;
                case 'f': s.push_back('\f'); break;
                case 'n': s.push_back('\n'); break;
                case 'r': s.push_back('\r'); break;
                case 't': s.push_back('\t'); break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(s, cp);
                    break;
                }
                default: fail("gltf: bad escape");
            }
        }
        if (p >= e) fail("gltf: unterminated string");
        ++p;
        return s;
    }
    JVal val() {
        if (++depth > 256) fail("gltf: JSON nested too deeply");
        ws();
        if (p >= e) fail("gltf: unexpected end of JSON");
        JVal v;
        if (*p == '{') {
            ++p;
            v.t = JVal::OBJ;
            ws();
            if (p < e && *p == '}') {
                ++p;
            } else {
                for (;;) {
                    std::string k = str();
                    expect(':');
                    v.o.emplace_back(std::move(k), val());
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    expect('}');
                    break;
                }
            }
        } else if (*p == '[') {
            ++p;
            v.t = JVal::ARR;
            ws();
            if (p < e && *p == ']') {
                ++p;
            } else {
                for (;;) {
                    v.a.push_back(val());
                    ws();
                    if (p < e && *p == ',') {
                        ++p;
                        continue;
                    }
                    expect(']');
                    break;
                }
            }
        } else if (*p == '"') {
            v.t = JVal::STR;
            v.s = str();
        } else if (e - p >= 4 && !strncmp(p, "true", 4)) {
            v.t = JVal::BOOL;
            v.b = true;
            p += 4;
        } else if (e - p >= 5 && !strncmp(p, "false", 5)) {
            v.t = JVal::BOOL;
            p += 5;
        } else if (e - p >= 4 && !strncmp(p, "null", 4)) {
            p += 4;
        } else {
            std::string num;
            while (p < e && (isdigit((unsigned char)*p) || *p == '-' || *p == '+' || *p == '.' || *p == 'e' || *p == 'E'))
                num.push_back(*p++);
            if (num.empty()) fail("gltf: JSON syntax error");
            char* end = nullptr;
            v.t = JVal::NUM;
            v.n = strtod(num.c_str(), &end);
            if (!end || *end) fail("gltf: bad number '%s'", num.c_str());
        }
        --depth;
        return v;
    }
};

double num(const JVal* v, double dflt) {
    if (!v) return dflt;
    if (v->t != JVal::NUM) fail("gltf: expected a number");
    return v->n;
}
long inum(const JVal* v, long dflt) {
    double d = num(v, (double)dflt);
    if (d != std::floor(d) || d < -1e15 || d > 1e15) fail("gltf: expected an integer");
    return (long)d;
}
long req_inum(const JVal& o, const char* k) {
    const JVal* v = o.get(k);
    if (!v) fail("gltf: missing '%s'", k);
    return inum(v, 0);
}
// byte offsets, lengths, strides and counts: whole numbers >= 0 (below 1e15)
size_t unum(const JVal* v, long dflt, const char* what) {
    const long x = inum(v, dflt);
    if (x < 0) fail("gltf: negative %s", what);
    return (size_t)x;
}
size_t req_unum(const JVal& o, const char* k) {
    const long x = req_inum(o, k);
    if (x < 0) fail("gltf: negative %s", k);
    return (size_t)x;
}
// count elements of elem bytes, stride bytes apart from byte off, all inside
// a buffer of size bytes (no sum or product below can wrap: each operand is
// at most 1e15 and the division keeps the last one in range)
bool span_ok(size_t size, size_t off, size_t count, size_t stride, size_t elem) {
    if (off > size || elem > size - off) return count == 0 && off <= size;
    if (count == 0) return true;
    return stride != 0 && count - 1 <= (size - off - elem) / stride;
}

// ---- paths, URIs, base64 ------------------------------------------------------------
std::string dirname_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}
std::string percent_decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && isxdigit((unsigned char)s[i + 1]) && isxdigit((unsigned char)s[i + 2])) {
            o.push_back((char)strtol(s.substr(i + 1, 2).c_str(), nullptr, 16));
            i += 2;
        } else {
            o.push_back(s[i]);
        }
    }
    return o;
}
std::vector<uint8_t> base64(const std::string& s, size_t from) {
    static int8_t T[256];
    static bool init = false;
    if (!init) {
        memset(T, -1, sizeof(T));
        const char* a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; ++i) T[(uint8_t)a[i]] = (int8_t)i;
        init = true;
    }
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        uint8_t c = (uint8_t)s[i];
        if (c == '=') break;
        if (T[c] < 0) {
            if (c == '\n' || c == '\r' || c == ' ') continue;
            fail("gltf: bad base64 data");
        }
        acc = (acc << 6) | (uint32_t)T[c];
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}
// bytes of a uri: data URI (base64) or a file relative to the asset
std::vector<uint8_t> uri_bytes(const std::string& base_dir, const std::string& uri) {
    if (uri.compare(0, 5, "data:") == 0) {
        size_t k = uri.find(";base64,");
        if (k == std::string::npos) fail("gltf: only base64 data URIs are supported");
        return base64(uri, k + 8);
    }
    std::string path = uri.size() && uri[0] == '/' ? uri : base_dir + percent_decode(uri);
    std::vector<uint8_t> b;
    std::string err;
    if (!hr::img::read_file(path, b, err)) fail("%s", err.c_str());
    return b;
}

// ---- materials as the reference keeps them (material.h:13-37) -----------------------
struct Mat {
    int shading = 0;
    float diffuse[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float metallic = 1.0f, roughness = 1.0f, ior = 1.0f;
    int tex[4] = {-1, -1, -1, -1};  // diffuse, metallicRoughness, emissive, normal (field order of Material)
    bool operator==(const Mat& m) const {
        return shading == m.shading && diffuse[0] == m.diffuse[0] && diffuse[1] == m.diffuse[1] &&
               diffuse[2] == m.diffuse[2] && diffuse[3] == m.diffuse[3] && tex[0] == m.tex[0] &&
               tex[3] == m.tex[3] && tex[2] == m.tex[2] && tex[1] == m.tex[1] && ior == m.ior &&
               metallic == m.metallic && roughness == m.roughness;
    }
};
// Material::materialToShader (material.cpp:19-26)
hr_material to_shader(const Mat& m, int index) {
    hr_material g;
    g.diffuse_ior[0] = m.diffuse[0];
    g.diffuse_ior[1] = m.diffuse[1];
    g.diffuse_ior[2] = m.diffuse[2];
    g.diffuse_ior[3] = m.ior;
    g.textures[0] = (float)m.tex[0];  // diffuse
    g.textures[1] = (float)m.tex[3];  // normal
    g.textures[2] = (float)m.tex[2];  // emissive
    g.textures[3] = (float)m.tex[1];  // metallic-roughness
    g.shading_metal_rough_idx[0] = (float)m.shading;
    g.shading_metal_rough_idx[1] = m.metallic;
    g.shading_metal_rough_idx[2] = m.roughness;
    g.shading_metal_rough_idx[3] = (float)index;
    return g;
}

// ---- OBJ ------------------------------------------------------------------------------
struct VKey {
    hr_vertex v;
    bool operator==(const VKey& o) const {
        const float* a = &v.position[0];
        const float* b = &o.v.position[0];
        for (int i = 0; i < 11; ++i)
            if (!(a[i] == b[i])) return false;
        return true;
    }
};
struct VHash {
    size_t operator()(const VKey& k) const {
        const float* a = &k.v.position[0];
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < 11; ++i) {
            float f = a[i] == 0.0f ? 0.0f : a[i];  // +0 and -0 compare equal
            uint32_t u;
            memcpy(&u, &f, 4);
            h = (h ^ u) * 1099511628211ull;
        }
        return (size_t)h;
    }
};

bool parse_floats(const char* s, float* out, int nmax, int& got) {
    got = 0;
    while (got < nmax) {
        while (*s == ' ' || *s == '\t') ++s;
        if (!*s || *s == '#') break;
        char* end = nullptr;
        double d = strtod(s, &end);
        if (end == s) return false;
        out[got++] = (float)d;
        s = end;
    }
    return true;
}

void load_obj(const std::string& path, hr_asset& A) {
    std::vector<uint8_t> bytes;
    std::string err;
    if (!hr::img::read_file(path, bytes, err)) fail("%s", err.c_str());
    std::vector<float> V, N, T;
    struct FV {
        long v, t, n;
    };
    std::vector<FV> corners;  // triangulated face vertices in file order
    size_t line_no = 0;
    std::string line;
    size_t p = 0;
    auto fix = [&](long i, size_t count, const char* what) -> long {
        long r;
        if (i > 0) r = i - 1;
        else if (i < 0) r = (long)count + i;
        else fail("obj line %zu: %s index 0", line_no, what);
        if (r < 0 || (size_t)r >= count) fail("obj line %zu: %s index out of range", line_no, what);
        return r;
    };
    while (p < bytes.size()) {
        size_t q = p;
        while (q < bytes.size() && bytes[q] != '\n') ++q;
        line.assign((const char*)&bytes[p], q - p);
        p = q + 1;
        ++line_no;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') ++s;
        if (!*s || *s == '#') continue;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            float f[3] = {0, 0, 0};
            int got;
            if (!parse_floats(s + 2, f, 3, got) || got < 3) fail("obj line %zu: bad vertex", line_no);
            V.insert(V.end(), f, f + 3);
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            float f[3] = {0, 0, 0};
            int got;
            if (!parse_floats(s + 3, f, 3, got) || got < 3) fail("obj line %zu: bad normal", line_no);
            N.insert(N.end(), f, f + 3);
        } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
            float f[2] = {0, 0};
            int got;
            if (!parse_floats(s + 3, f, 2, got) || got < 1) fail("obj line %zu: bad texture coordinate", line_no);
            T.insert(T.end(), f, f + 2);
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            std::vector<FV> face;
            const char* c = s + 2;
            for (;;) {
                while (*c == ' ' || *c == '\t') ++c;
                if (!*c || *c == '#') break;
                FV fv{0, 0, 0};
                char* end = nullptr;
                fv.v = strtol(c, &end, 10);
                if (end == c) fail("obj line %zu: bad face", line_no);
                c = end;
                if (*c == '/') {
                    ++c;
                    if (*c != '/') {
                        fv.t = strtol(c, &end, 10);
                        if (end == c) fail("obj line %zu: bad face", line_no);
                        c = end;
                    }
                    if (*c == '/') {
                        ++c;
                        fv.n = strtol(c, &end, 10);
                        if (end == c) fail("obj line %zu: bad face", line_no);
                        c = end;
                    }
                }
                if (*c && *c != ' ' && *c != '\t') fail("obj line %zu: bad face", line_no);
                // resolve relative indices against the counts so far
                FV r;
                r.v = fix(fv.v, V.size() / 3, "vertex");
                r.t = fv.t ? fix(fv.t, T.size() / 2, "texture coordinate") : -1;
                r.n = fv.n ? fix(fv.n, N.size() / 3, "normal") : -1;
                face.push_back(r);
            }
            if (face.size() < 3) continue;  // as tinyobjloader: faces need 3+ vertices
            for (size_t k = 2; k < face.size(); ++k) {
                corners.push_back(face[0]);
                corners.push_back(face[k - 1]);
                corners.push_back(face[k]);
            }
        }
        // o, g, s, usemtl, mtllib, l, p and the rest do not change the mesh
    }
    bool has_uv = !T.empty();
    std::unordered_map<VKey, uint32_t, VHash> unique;
    for (const FV& f : corners) {
        if (f.n < 0) fail("obj: a face vertex has no normal (the loader reads one for every vertex)");
        if (has_uv && f.t < 0) fail("obj: a face vertex has no texture coordinate while the file has some");
        VKey k;
        memset(&k.v, 0, sizeof(k.v));
        for (int i = 0; i < 3; ++i) {
            k.v.position[i] = V[3 * f.v + i];
            k.v.normal[i] = N[3 * f.n + i];
            k.v.color[i] = 1.0f;
        }
        if (has_uv) {
            k.v.uv[0] = T[2 * f.t];
            k.v.uv[1] = 1.0f - T[2 * f.t + 1];
        }
        auto it = unique.find(k);
        uint32_t id;
        if (it == unique.end()) {
            id = (uint32_t)A.verts.size();
            unique.emplace(k, id);
            A.verts.push_back(k.v);
        } else {
            id = it->second;
        }
        A.idx.push_back(id);
    }
    hr_primitive pr;
    memset(&pr, 0, sizeof(pr));
    pr.mesh = 0;
    pr.first_index = 0;
    pr.index_count = (uint32_t)A.idx.size();
    pr.material = 0;
    M4 I = ident();
    memcpy(pr.model, I.m, sizeof(pr.model));
    A.prims.push_back(pr);
    A.mats.push_back(to_shader(Mat(), 0));
}

// ---- glTF -------------------------------------------------------------------------------
struct GNode {
    M4 local;
    int parent = -1;
    std::vector<int> children;
    std::vector<int> prims;  // indices into the loader's primitive list
};

struct Gltf {
    JVal root;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;
    uint32_t flags = 0;
    std::vector<GNode> nodes;  // loader nodes (one per visit)
    struct Prim {
        uint32_t first_index, index_count;
        int material;
    };
    std::vector<Prim> prims;
    std::vector<Mat> mats;
    std::vector<std::string> tex_keys;  // texture ids in first-use order
    hr_asset* A = nullptr;

    const JVal& arr(const char* k) {
        static const JVal empty = [] {
            JVal v;
            v.t = JVal::ARR;
            return v;
        }();
        const JVal* v = root.get(k);
        return v ? *v : empty;
    }

    // element e, component c of a float accessor with `comps` components
    struct View {
        const uint8_t* base;
        size_t stride, count;
    };
    View float_accessor(long ai, int comps, const char* what) {
        const JVal& acc = arr("accessors").at((size_t)ai);
        if (inum(acc.get("componentType"), 0) != 5126) fail("gltf: %s accessor is not float", what);
        long bv = inum(acc.get("bufferView"), -1);
        if (bv < 0) fail("gltf: %s accessor without a buffer view", what);
        const JVal& view = arr("bufferViews").at((size_t)bv);
        long bi = req_inum(view, "buffer");
        if (bi < 0 || (size_t)bi >= buffers.size()) fail("gltf: buffer %ld out of range", bi);
        size_t off = unum(view.get("byteOffset"), 0, "byteOffset") + unum(acc.get("byteOffset"), 0, "byteOffset");
        size_t count = req_unum(acc, "count");
        size_t stride = unum(view.get("byteStride"), 0, "byteStride");
        if (stride == 0) stride = 4 * (size_t)comps;
        const auto& buf = buffers[(size_t)bi];
        if (!span_ok(buf.size(), off, count, stride, 4 * (size_t)comps))
            fail("gltf: %s accessor exceeds its buffer", what);
        return View{buf.data() + off, stride, count};
    }
    static float rd(const View& v, size_t e, int c) {
        float f;
        memcpy(&f, v.base + e * v.stride + 4 * (size_t)c, 4);
        return f;
    }

    M4 local_matrix(const JVal& n) {
        M4 m = ident();
        const JVal* t = n.get("translation");
        if (t && t->size() == 3) {  // glm::translate(I, t)
            m.m[12] = (float)num(&t->at(0), 0);
            m.m[13] = (float)num(&t->at(1), 0);
            m.m[14] = (float)num(&t->at(2), 0);
        }
        const JVal* r = n.get("rotation");
        if (r && r->size() == 4) {  // matrix *= mat4(quat(x, y, z, w))
            double x = num(&r->at(0), 0), y = num(&r->at(1), 0), z = num(&r->at(2), 0), w = num(&r->at(3), 1);
            M4 R = ident();
            R.m[0] = (float)(1 - 2 * (y * y + z * z));
            R.m[1] = (float)(2 * (x * y + w * z));
            R.m[2] = (float)(2 * (x * z - w * y));
            R.m[4] = (float)(2 * (x * y - w * z));
            R.m[5] = (float)(1 - 2 * (x * x + z * z));
            R.m[6] = (float)(2 * (y * z + w * x));
            R.m[8] = (float)(2 * (x * z + w * y));
            R.m[9] = (float)(2 * (y * z - w * x));
            R.m[10] = (float)(1 - 2 * (x * x + y * y));
            m = mul(m, R);
        }
        const JVal* s = n.get("scale");
        if (s && s->size() == 3) {  // glm::scale(m, s): columns 0..2 scaled
            float sv[3] = {(float)num(&s->at(0), 1), (float)num(&s->at(1), 1), (float)num(&s->at(2), 1)};
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 4; ++k) m.m[c * 4 + k] = m.m[c * 4 + k] * sv[c];
        }
        const JVal* mm = n.get("matrix");
        if (mm && mm->size() == 16)
            for (int i = 0; i < 16; ++i) m.m[i] = (float)num(&mm->at((size_t)i), 0);
        return m;
    }

    int load_material(long index) {
        Mat m;
        if (index > -1) {
            const JVal& tm = arr("materials").at((size_t)index);
            const JVal* pbr = tm.get("pbrMetallicRoughness");
            const JVal* bc = pbr ? pbr->get("baseColorFactor") : nullptr;
            if (bc && bc->size() == 4)
                for (int k = 0; k < 4; ++k) m.diffuse[k] = (float)num(&bc->at((size_t)k), 1);
            m.metallic = (float)num(pbr ? pbr->get("metallicFactor") : nullptr, 1.0);
            m.roughness = (float)num(pbr ? pbr->get("roughnessFactor") : nullptr, 1.0);
            auto tex_index = [&](const JVal* o, const char* k) -> int {
                const JVal* t = o ? o->get(k) : nullptr;
                return t ? (int)inum(t->get("index"), -1) : -1;
            };
            m.tex[0] = tex_index(pbr, "baseColorTexture");
            m.tex[1] = tex_index(pbr, "metallicRoughnessTexture");
            m.tex[2] = tex_index(&tm, "emissiveTexture");
            m.tex[3] = tex_index(&tm, "normalTexture");
        }
        for (size_t i = 0; i < mats.size(); ++i)
            if (mats[i] == m) return (int)i;
        mats.push_back(m);
        return (int)mats.size() - 1;
    }

    // Texture::get_id for images[image]: its id in first-use order
    int texture_id(long tex_index) {
        long image = tex_index;
        if (flags & HR_ASSET_TEXTURE_SOURCE) {
            const JVal& t = arr("textures").at((size_t)tex_index);
            image = req_inum(t, "source");
        }
        const JVal& im = arr("images").at((size_t)image);
        std::string key;
        std::vector<uint8_t> bytes;
        const JVal* uri = im.get("uri");
        if (uri && uri->t == JVal::STR) {
            key = uri->s.compare(0, 5, "data:") == 0 ? "#data" + std::to_string(image) : dir + percent_decode(uri->s);
        } else {
            key = "#image" + std::to_string(image);
        }
        for (size_t i = 0; i < tex_keys.size(); ++i)
            if (tex_keys[i] == key) return (int)i;
        if (uri && uri->t == JVal::STR) {
            bytes = uri_bytes(dir, uri->s);
        } else {  // embedded in a buffer view
            long bv = inum(im.get("bufferView"), -1);
            if (bv < 0) fail("gltf: image %ld has neither a uri nor a buffer view", image);
            const JVal& view = arr("bufferViews").at((size_t)bv);
            long bi = req_inum(view, "buffer");
            if (bi < 0 || (size_t)bi >= buffers.size()) fail("gltf: buffer %ld out of range", bi);
            size_t off = unum(view.get("byteOffset"), 0, "byteOffset"), len = req_unum(view, "byteLength");
            if (!span_ok(buffers[(size_t)bi].size(), off, 1, 1, len))
                fail("gltf: image buffer view exceeds its buffer");
            bytes.assign(buffers[(size_t)bi].begin() + (long)off, buffers[(size_t)bi].begin() + (long)(off + len));
        }
        hr::img::Image img;
        std::string err;
        if (!hr::img::decode(bytes.data(), bytes.size(), img, err)) fail("image %s: %s", key.c_str(), err.c_str());
        A->images.push_back(std::move(img));
        tex_keys.push_back(key);
        return (int)tex_keys.size() - 1;
    }

    void load_textures(int mi) {
        Mat& m = mats[(size_t)mi];
        // field order of Prefab::loadTextures: diffuse, normal, emissive, metallicRoughness
        static const int order[4] = {0, 3, 2, 1};
        for (int k : order)
            if (m.tex[k] > -1) m.tex[k] = texture_id(m.tex[k]);
    }

    int load_node(long ni, int parent, bool invert, int depth) {
        if (depth > 128) fail("gltf: node hierarchy deeper than 128 (cycle?)");
        // a node listed as a child more than once is visited once per listing
        // (the reference instances it each time); a shared subtree repeated
        // at every level would multiply without bound
        if (nodes.size() >= (1u << 20)) fail("gltf: more than %u node visits", 1u << 20);
        const JVal& tn = arr("nodes").at((size_t)ni);
        int id = (int)nodes.size();
        nodes.push_back(GNode());
        nodes[(size_t)id].local = local_matrix(tn);
        nodes[(size_t)id].parent = parent;
        const JVal* ch = tn.get("children");
        if (ch)
            for (size_t i = 0; i < ch->size(); ++i) {
                int c = load_node(inum(&ch->at(i), 0), id, false, depth + 1);  // children: invertNormals = false
                nodes[(size_t)id].children.push_back(c);
            }
        long mesh = inum(tn.get("mesh"), -1);
        if (mesh > -1) {
            const JVal& tm = arr("meshes").at((size_t)mesh);
            const JVal* ps = tm.get("primitives");
            for (size_t i = 0; ps && i < ps->size(); ++i) load_primitive(ps->at(i), id, invert);
        }
        return id;
    }

    void load_primitive(const JVal& tp, int node, bool invert) {
        uint32_t first_index = (uint32_t)A->idx.size();
        uint32_t first_vertex = (uint32_t)A->verts.size();
        const JVal* at = tp.get("attributes");
        long pa = at ? inum(at->get("POSITION"), -1) : -1;
        long na = at ? inum(at->get("NORMAL"), -1) : -1;
        long ta = at ? inum(at->get("TEXCOORD_0"), -1) : -1;
        size_t nv = 0;
        View P{nullptr, 0, 0}, Nn{nullptr, 0, 0}, Tc{nullptr, 0, 0};
        if (pa > -1) {
            P = float_accessor(pa, 3, "POSITION");
            nv = P.count;
        }
        if (na > -1) Nn = float_accessor(na, 3, "NORMAL");
        if (ta > -1) Tc = float_accessor(ta, 2, "TEXCOORD_0");
        if ((na > -1 && Nn.count < nv) || (ta > -1 && Tc.count < nv)) fail("gltf: attribute shorter than POSITION");
        for (size_t v = 0; v < nv; ++v) {
            hr_vertex x;
            memset(&x, 0, sizeof(x));
            float n[3] = {0.0f, 0.0f, 0.0f};
            if (na > -1)
                for (int k = 0; k < 3; ++k) n[k] = rd(Nn, v, k);
            // glm::normalize: v * inversesqrt(dot(v, v))
            float inv = 1.0f / sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
            for (int k = 0; k < 3; ++k) {
                x.position[k] = rd(P, v, k);
                float nk = n[k] * inv;
                x.normal[k] = invert ? nk * -1.0f : nk;
                x.color[k] = 1.0f;
            }
            if (ta > -1) {
                x.uv[0] = rd(Tc, v, 0);
                x.uv[1] = rd(Tc, v, 1);
            }
            A->verts.push_back(x);
        }
        long ia = inum(tp.get("indices"), -1);
        uint32_t count = 0;
        if (ia > -1) {
            const JVal& acc = arr("accessors").at((size_t)ia);
            long bv = req_inum(acc, "bufferView");
            const JVal& view = arr("bufferViews").at((size_t)bv);
            long bi = req_inum(view, "buffer");
            if (bi < 0 || (size_t)bi >= buffers.size()) fail("gltf: buffer %ld out of range", bi);
            size_t off = unum(view.get("byteOffset"), 0, "byteOffset") + unum(acc.get("byteOffset"), 0, "byteOffset");
            size_t n = req_unum(acc, "count");
            long ct = inum(acc.get("componentType"), 0);
            size_t es = ct == 5125 ? 4 : ct == 5123 ? 2 : ct == 5121 ? 1 : 0;
            if (!es) fail("gltf: index component type %ld not supported", ct);
            const auto& buf = buffers[(size_t)bi];
            if (!span_ok(buf.size(), off, n, es, es)) fail("gltf: index accessor exceeds its buffer");
            for (size_t i = 0; i < n; ++i) {
                const uint8_t* q = buf.data() + off + i * es;
                uint32_t v = es == 4 ? (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24)
                           : es == 2 ? (uint32_t)q[0] | ((uint32_t)q[1] << 8) : q[0];
                A->idx.push_back(v + first_vertex);
            }
            count = (uint32_t)n;
        } else {  // non-indexed: vertices in order
            for (size_t i = 0; i < nv; ++i) A->idx.push_back((uint32_t)(first_vertex + i));
            count = (uint32_t)nv;
        }
        Prim pr;
        pr.first_index = first_index;
        pr.index_count = count;
        pr.material = load_material(inum(tp.get("material"), -1));
        load_textures(pr.material);
        nodes[(size_t)node].prims.push_back((int)prims.size());
        prims.push_back(pr);
    }

    M4 global(int n) {
        const GNode& g = nodes[(size_t)n];
        if (g.parent < 0) return g.local;
        return mul(global(g.parent), g.local);
    }

    void instance(int n) {  // Node::node_to_instance order
        if (!nodes[(size_t)n].prims.empty()) {
            M4 m = global(n);
            for (int pi : nodes[(size_t)n].prims) {
                hr_primitive hp;
                memset(&hp, 0, sizeof(hp));
                hp.mesh = 0;
                hp.first_index = prims[(size_t)pi].first_index;
                hp.index_count = prims[(size_t)pi].index_count;
                hp.material = prims[(size_t)pi].material;
                memcpy(hp.model, m.m, sizeof(hp.model));
                A->prims.push_back(hp);
            }
        }
        for (int c : nodes[(size_t)n].children) instance(c);
    }
};

void load_gltf(const std::string& path, uint32_t flags, hr_asset& A) {
    std::vector<uint8_t> bytes;
    std::string err;
    if (!hr::img::read_file(path, bytes, err)) fail("%s", err.c_str());
    Gltf G;
    G.flags = flags;
    G.dir = dirname_of(path);
    G.A = &A;
    std::string json;
    std::vector<uint8_t> bin;
    bool glb = bytes.size() >= 12 && !memcmp(bytes.data(), "glTF", 4);
    if (glb) {
        auto le32 = [&](size_t o) {
            return (uint32_t)bytes[o] | ((uint32_t)bytes[o + 1] << 8) | ((uint32_t)bytes[o + 2] << 16) |
                   ((uint32_t)bytes[o + 3] << 24);
        };
        if (le32(4) != 2) fail("glb: version %u (2 supported)", le32(4));
        size_t total = std::min<size_t>(le32(8), bytes.size());
        size_t o = 12;
        while (o + 8 <= total) {
            uint32_t len = le32(o), type = le32(o + 4);
            if (o + 8 + (size_t)len > total) fail("glb: truncated chunk");
            if (type == 0x4E4F534Au) json.assign((const char*)&bytes[o + 8], len);
            else if (type == 0x004E4942u && bin.empty()) bin.assign(bytes.begin() + (long)(o + 8), bytes.begin() + (long)(o + 8 + len));
            o += 8 + (size_t)len;
        }
        if (json.empty()) fail("glb: no JSON chunk");
    } else {
        json.assign((const char*)bytes.data(), bytes.size());
    }
    JParser jp{json.data(), json.data() + json.size()};
    G.root = jp.val();
    if (G.root.t != JVal::OBJ) fail("gltf: the document is not a JSON object");
    const JVal& bufs = G.arr("buffers");
    for (size_t i = 0; i < bufs.size(); ++i) {
        const JVal* uri = bufs.at(i).get("uri");
        if (uri && uri->t == JVal::STR) {
            G.buffers.push_back(uri_bytes(G.dir, uri->s));
        } else if (glb && i == 0) {
            G.buffers.push_back(bin);
        } else {
            fail("gltf: buffer %zu has no data", i);
        }
        long bl = inum(bufs.at(i).get("byteLength"), 0);
        if ((size_t)bl > G.buffers.back().size()) fail("gltf: buffer %zu is shorter than its byteLength", i);
    }
    const JVal& scenes = G.arr("scenes");
    if (scenes.size() == 0) fail("gltf: no scene");
    const JVal* roots = scenes.at(0).get("nodes");
    std::vector<int> root_ids;
    bool invert = (flags & HR_ASSET_INVERT_NORMALS) != 0;
    for (size_t i = 0; roots && i < roots->size(); ++i) root_ids.push_back(G.load_node(inum(&roots->at(i), 0), -1, invert, 0));
    for (int r : root_ids) G.instance(r);
    for (size_t i = 0; i < G.mats.size(); ++i) A.mats.push_back(to_shader(G.mats[i], (int)i));
    if (A.mats.empty()) A.mats.push_back(to_shader(Mat(), 0));
}

bool ends_with(const std::string& s, const char* suf) {
    size_t n = strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

void set_err(char* err, size_t len, const std::string& m) {
    if (err && len) snprintf(err, len, "%s", m.c_str());
}

}  // namespace

extern "C" {

hr_status hr_asset_load(const char* path, uint32_t flags, hr_asset** out, char* err, size_t err_len) {
    if (!path || !out) return HR_E_INVALID;
    *out = nullptr;
    std::unique_ptr<hr_asset> A(new hr_asset());
    std::string p = path;
    try {
        if (ends_with(p, ".obj")) load_obj(p, *A);
        else if (ends_with(p, ".gltf") || ends_with(p, ".glb")) load_gltf(p, flags, *A);
        else fail("%s: not an .obj, .gltf or .glb file", path);
    } catch (const Fail& f) {
        set_err(err, err_len, f.msg);
        return HR_E_INVALID;
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "out of memory");
        return HR_E_OOM;
    } catch (const std::exception& e) {
        set_err(err, err_len, std::string("asset load failed: ") + e.what());
        return HR_E_INVALID;
    }
    for (const auto& im : A->images) A->tex.push_back(hr_texture{im.px.data(), im.w, im.h});
    *out = A.release();
    return HR_OK;
}

hr_status hr_asset_view_get(const hr_asset* a, hr_asset_view* v) {
    if (!a || !v) return HR_E_INVALID;
    memset(v, 0, sizeof(*v));
    v->mesh.vertices = a->verts.data();
    v->mesh.vertex_count = (uint32_t)a->verts.size();
    v->mesh.indices = a->idx.data();
    v->mesh.index_count = (uint32_t)a->idx.size();
    v->primitives = a->prims.data();
    v->primitive_count = (uint32_t)a->prims.size();
    v->materials = a->mats.data();
    v->material_count = (uint32_t)a->mats.size();
    v->textures = a->tex.data();
    v->texture_count = (uint32_t)a->tex.size();
    return HR_OK;
}

void hr_asset_destroy(hr_asset* a) { delete a; }

}  // extern "C"
