// glTF 2.0 scenes (.gltf with external or data-URI buffers, and binary .glb): host code.
//
// Plasma reads glTF through tinygltf and translates it in glTFLoader.cpp:26-386; this
// is the same translation over a parser of its own:
//   - every scene's node tree is walked from its roots; a node's world transform is
//     parent x local, local = its matrix (column-major) or T x R x S with the rotation
//     quaternion (x, y, z, w) turned into a matrix unnormalized (glTFLoader.cpp:26-51),
//     in float;
//   - each TRIANGLES primitive (mode 4, the default) of a node's mesh becomes an
//     instance; primitives without a float POSITION accessor, or of other modes, are
//     skipped.  Vertex data are POSITION, NORMAL and TEXCOORD_0 as stored (FLOAT only:
//     accessors of other component types read as absent, :96-124); indices are
//     UNSIGNED_BYTE / SHORT / INT, and a primitive without them, or with another
//     index type, gets sequential indices (:128-176, :213-222).  A primitive without
//     normals gets Foundation's calculateNormals (:226-232).  The same (mesh,
//     primitive) under several nodes is one mesh with several instances (Plasma
//     copies it per node; the images are the same);
//   - materials: a Standard Surface with base_color = baseColorFactor.rgb, metalness =
//     metallicFactor, specular_roughness = roughnessFactor (glTF defaults 1, 1, 1) and
//     Aurora's defaults otherwise (:400-414); a primitive without a material gets the
//     default material; a node's AURORA_materials_materialX file is reported (the
//     MaterialX library is not here to translate it, :290-341);
//   - cameras: each camera node's view matrix is the inverse of its local transform
//     (the reference does not apply the parents', :37-85), with the perspective or
//     orthographic parameters.
// A .glb holds a JSON chunk and an optional BIN chunk that a buffer without a uri
// refers to; a .gltf's buffers are data: URIs (base64) or files next to it.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/aph.h"
#include "assets.h"

namespace {

// ---- JSON ----
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    const Json* get(const char* key) const
    {
        if (kind != Obj) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num(const char* key, double dflt) const
    {
        const Json* v = get(key);
        return v && v->kind == Num ? v->n : dflt;
    }
    // an int field; a value outside int's range (or NaN) reads as INT_MIN, which every
    // index check rejects
    int integer(const char* key, int dflt) const
    {
        const double v = num(key, dflt);
        return v >= -2147483648.0 && v <= 2147483647.0 ? (int)v : INT_MIN;
    }
    // a byte count or offset: absent reads as 0; false unless a whole number in [0, 2^53]
    bool size(const char* key, size_t& out) const
    {
        const double v = num(key, 0.0);
        if (!(v >= 0.0 && v <= 9007199254740992.0) || v != std::floor(v)) return false;
        out = (size_t)v;
        return true;
    }
    const std::vector<Json>& arr(const char* key) const
    {
        static const std::vector<Json> none;
        const Json* v = get(key);
        return v && v->kind == Arr ? v->a : none;
    }
    std::string str(const char* key) const
    {
        const Json* v = get(key);
        return v && v->kind == Str ? v->s : std::string();
    }
};

class JsonParser {
public:
    JsonParser(const char* p, const char* end) : p_(p), end_(end) {}
    bool parse(Json& out)
    {
        if (!value(out, 0)) return false;
        ws();
        return p_ == end_;
    }

private:
    const char* p_;
    const char* end_;

    void ws()
    {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if ((size_t)(end_ - p_) < n || std::strncmp(p_, w, n) != 0) return false;
        p_ += n;
        return true;
    }
    static void utf8(std::string& s, uint32_t c)
    {
        if (c < 0x80) {
            s += (char)c;
        } else if (c < 0x800) {
            s += (char)(0xC0 | (c >> 6));
            s += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            s += (char)(0xE0 | (c >> 12));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        } else {
            s += (char)(0xF0 | (c >> 18));
            s += (char)(0x80 | ((c >> 12) & 0x3F));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        }
    }
    bool hex4(uint32_t& v)
    {
        if (end_ - p_ < 4) return false;
        v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string& s)
    {
        if (p_ >= end_ || *p_ != '"') return false;
        p_++;
        while (p_ < end_ && *p_ != '"') {
            const char c = *p_++;
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= end_) return false;
            const char e = *p_++;
            switch (e) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                uint32_t c1 = 0;
                if (!hex4(c1)) return false;
                if (c1 >= 0xD800 && c1 < 0xDC00) {  // a surrogate pair
                    uint32_t c2 = 0;
                    if (!lit("\\u") || !hex4(c2) || c2 < 0xDC00 || c2 >= 0xE000) return false;
                    c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                }
                utf8(s, c1);
                break;
            }
            default: return false;
            }
        }
        if (p_ >= end_) return false;
        p_++;
        return true;
    }
    bool value(Json& v, int depth)
    {
        if (depth > 256) return false;
        ws();
        if (p_ >= end_) return false;
        const char c = *p_;
        if (c == '{') {
            p_++;
            v.kind = Json::Obj;
            ws();
            if (p_ < end_ && *p_ == '}') {
                p_++;
                return true;
            }
            for (;;) {
                ws();
                std::pair<std::string, Json> kv;
                if (!string(kv.first)) return false;
                ws();
                if (p_ >= end_ || *p_++ != ':') return false;
                if (!value(kv.second, depth + 1)) return false;
                v.o.push_back(std::move(kv));
                ws();
                if (p_ >= end_) return false;
                if (*p_ == ',') { p_++; continue; }
                if (*p_ == '}') { p_++; return true; }
                return false;
            }
        }
        if (c == '[') {
            p_++;
            v.kind = Json::Arr;
            ws();
            if (p_ < end_ && *p_ == ']') {
                p_++;
                return true;
            }
            for (;;) {
                Json e;
                if (!value(e, depth + 1)) return false;
                v.a.push_back(std::move(e));
                ws();
                if (p_ >= end_) return false;
                if (*p_ == ',') { p_++; continue; }
                if (*p_ == ']') { p_++; return true; }
                return false;
            }
        }
        if (c == '"') {
            v.kind = Json::Str;
            return string(v.s);
        }
        if (lit("true")) { v.kind = Json::Bool; v.b = true; return true; }
        if (lit("false")) { v.kind = Json::Bool; v.b = false; return true; }
        if (lit("null")) { v.kind = Json::Null; return true; }
        // number: strtod on a bounded copy (the buffer need not be NUL-terminated)
        const char* q = p_;
        while (q < end_ && (std::strchr("+-0123456789.eE", *q) != nullptr)) q++;
        if (q == p_ || q - p_ > 64) return false;
        char tmp[72];
        std::memcpy(tmp, p_, (size_t)(q - p_));
        tmp[q - p_] = 0;
        char* e = nullptr;
        v.kind = Json::Num;
        v.n = std::strtod(tmp, &e);
        if (e != tmp + (q - p_)) return false;
        p_ = q;
        return true;
    }
};

// ---- buffers ----
bool base64(const std::string& in, size_t from, std::vector<uint8_t>& out)
{
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < in.size(); i++) {
        const char c = in[i];
        if (c == '=') break;
        const int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFF));
        }
    }
    return true;
}

std::string dirOf(const std::string& path)
{
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

// a relative URI as a file path: %XX escapes decoded
std::string uriPath(const std::string& uri, const std::string& base)
{
    std::string s;
    for (size_t i = 0; i < uri.size(); i++) {
        if (uri[i] == '%' && i + 2 < uri.size()) {
            s += (char)std::strtol(uri.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            s += uri[i];
        }
    }
    return !s.empty() && s[0] == '/' ? s : base + s;
}

bool readFile(const std::string& path, std::vector<uint8_t>& out)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    out.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return true;
}

uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

// ---- transforms (column-major 4x4 float, m[col][row]) ----
struct Mat4 {
    float m[4][4];
};
Mat4 identity()
{
    Mat4 r{};
    for (int k = 0; k < 4; k++) r.m[k][k] = 1.0f;
    return r;
}
// column c of a x b = sum over k of a's column k times b[c][k], added left to right
Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r;
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++)
            r.m[c][row] = ((a.m[0][row] * b.m[c][0] + a.m[1][row] * b.m[c][1]) + a.m[2][row] * b.m[c][2]) +
                          a.m[3][row] * b.m[c][3];
    return r;
}

Mat4 nodeTransform(const Json& node)
{
    const std::vector<Json>& mat = node.arr("matrix");
    if (mat.size() == 16) {
        Mat4 r;
        for (int k = 0; k < 16; k++) r.m[k / 4][k % 4] = (float)mat[k].n;
        return r;
    }
    float t[3] = {0, 0, 0}, s[3] = {1, 1, 1}, q[4] = {0, 0, 0, 1};  // q = (x, y, z, w)
    const std::vector<Json>& tr = node.arr("translation");
    const std::vector<Json>& ro = node.arr("rotation");
    const std::vector<Json>& sc = node.arr("scale");
    if (tr.size() == 3)
        for (int k = 0; k < 3; k++) t[k] = (float)tr[k].n;
    if (ro.size() == 4)
        for (int k = 0; k < 4; k++) q[k] = (float)ro[k].n;
    if (sc.size() == 3)
        for (int k = 0; k < 3; k++) s[k] = (float)sc[k].n;
    // rotation matrix of the quaternion (the standard formula, unnormalized input)
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z, xz = x * z, xy = x * y, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    float R[3][3];  // R[col][row]
    R[0][0] = 1.0f - 2.0f * (yy + zz);
    R[0][1] = 2.0f * (xy + wz);
    R[0][2] = 2.0f * (xz - wy);
    R[1][0] = 2.0f * (xy - wz);
    R[1][1] = 1.0f - 2.0f * (xx + zz);
    R[1][2] = 2.0f * (yz + wx);
    R[2][0] = 2.0f * (xz + wy);
    R[2][1] = 2.0f * (yz - wx);
    R[2][2] = 1.0f - 2.0f * (xx + yy);
    // T x R x S: the products with the identity's zeros and ones are exact
    Mat4 r{};
    for (int c = 0; c < 3; c++)
        for (int row = 0; row < 3; row++) r.m[c][row] = R[c][row] * s[c];
    for (int row = 0; row < 3; row++) r.m[3][row] = t[row];
    r.m[3][3] = 1.0f;
    return r;
}

// inverse of a 4x4 (camera view matrices), in double then rounded
bool inverse(const Mat4& a, float out[16])
{
    double m[16], inv[16];
    for (int k = 0; k < 16; k++) m[k] = a.m[k / 4][k % 4];  // m[col * 4 + row]
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    // out is row-major: out[row * 4 + col] = inv[col * 4 + row] / det
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) out[row * 4 + col] = (float)(inv[col * 4 + row] / det);
    return true;
}

struct GltfFile {
    aph_gltf_file pub{};
    std::vector<aph_gltf_mesh> meshes;
    std::vector<aph_gltf_instance> instances;
    std::vector<aph_material> materials;
    std::vector<aph_gltf_camera> cameras;
    std::vector<std::vector<float>> pos, nrm, uv;
    std::vector<std::vector<uint32_t>> idx;
    std::vector<std::unique_ptr<std::string>> strings;
};

class Loader {
public:
    Json doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string base;
    GltfFile* out = nullptr;
    std::map<std::pair<int, int>, int> meshOf;  // (mesh, primitive) -> mesh index

    int componentCount(const std::string& type)
    {
        if (type == "SCALAR") return 1;
        if (type == "VEC2") return 2;
        if (type == "VEC3") return 3;
        if (type == "VEC4" || type == "MAT2") return 4;
        if (type == "MAT3") return 9;
        if (type == "MAT4") return 16;
        return 0;
    }
    // the bytes of accessor i's elements: base pointer, stride, count, component type
    bool view(int i, const uint8_t*& ptr, size_t& stride, size_t& count, int& ctype, int& ncomp, size_t csize)
    {
        const std::vector<Json>& acc = doc.arr("accessors");
        if (i < 0 || i >= (int)acc.size()) return false;
        const Json& a = acc[i];
        ctype = a.integer("componentType", 0);
        ncomp = componentCount(a.str("type"));
        if (!a.size("count", count)) return false;
        const int bvi = a.integer("bufferView", -1);
        const std::vector<Json>& bvs = doc.arr("bufferViews");
        if (ncomp == 0 || bvi < 0 || bvi >= (int)bvs.size()) return false;
        const Json& bv = bvs[bvi];
        const int bi = bv.integer("buffer", -1);
        if (bi < 0 || bi >= (int)buffers.size()) return false;
        size_t bvOff, aOff, bvLen, bs;
        if (!bv.size("byteOffset", bvOff) || !a.size("byteOffset", aOff) || !bv.size("byteLength", bvLen) ||
            !bv.size("byteStride", bs))
            return false;
        // each term is at most 2^53, so the sums cannot wrap
        const size_t off = bvOff + aOff, bvEnd = bvOff + bvLen;
        stride = bs ? bs : (size_t)ncomp * csize;
        const size_t elem = (size_t)ncomp * csize;
        if (off > bvEnd || bvEnd > buffers[bi].size()) return false;
        // the last element ends inside the view: off + (count - 1) stride + elem <= bvEnd,
        // tested without forming the product
        if (count > 0 && (elem > bvEnd - off || (count - 1) > (bvEnd - off - elem) / stride)) return false;
        ptr = buffers[bi].data() + off;
        return true;
    }
    // FLOAT accessors of `want` components only; other component types read as absent
    // (an empty array), another element type is an error
    bool floats(int i, int want, std::vector<float>& v)
    {
        v.clear();
        if (i < 0) return true;
        const std::vector<Json>& acc = doc.arr("accessors");
        if (i >= (int)acc.size()) return false;
        if (acc[i].integer("componentType", 0) != 5126) return true;
        if (componentCount(acc[i].str("type")) != want) return false;
        const uint8_t* p;
        size_t stride, count;
        int ct, nc;
        if (!view(i, p, stride, count, ct, nc, 4)) return false;
        v.resize(count * nc);
        for (size_t e = 0; e < count; e++) std::memcpy(&v[e * nc], p + e * stride, 4 * (size_t)nc);
        return true;
    }
    // UNSIGNED_BYTE / SHORT / INT; other types read as absent
    bool indices(int i, std::vector<uint32_t>& v)
    {
        v.clear();
        if (i < 0) return true;
        const std::vector<Json>& acc = doc.arr("accessors");
        if (i >= (int)acc.size()) return false;
        const int ct = acc[i].integer("componentType", 0);
        const size_t cs = ct == 5121 ? 1 : ct == 5123 ? 2 : ct == 5125 ? 4 : 0;
        if (cs == 0) return true;
        const uint8_t* p;
        size_t stride, count;
        int ct2, nc;
        if (!view(i, p, stride, count, ct2, nc, cs)) return false;
        v.resize(count);
        for (size_t e = 0; e < count; e++) {
            const uint8_t* q = p + e * stride;
            v[e] = cs == 1 ? q[0] : cs == 2 ? (uint32_t)(q[0] | (q[1] << 8)) : le32(q);
        }
        return true;
    }

    int primitiveMesh(int meshIdx, int primIdx, const Json& prim, int& material)
    {
        const auto key = std::make_pair(meshIdx, primIdx);
        material = prim.integer("material", -1);
        if (material >= (int)doc.arr("materials").size()) material = -1;
        auto it = meshOf.find(key);
        if (it != meshOf.end()) return it->second;
        int result = -1;
        const Json* attrs = prim.get("attributes");
        if (prim.integer("mode", 4) == 4 && attrs) {
            std::vector<float> pos, nrm, uv;
            std::vector<uint32_t> idx;
            if (!floats(attrs->integer("POSITION", -1), 3, pos) || !floats(attrs->integer("NORMAL", -1), 3, nrm) ||
                !floats(attrs->integer("TEXCOORD_0", -1), 2, uv) || !indices(prim.integer("indices", -1), idx))
                return -2;
            if (!pos.empty()) {
                const uint32_t nv = (uint32_t)(pos.size() / 3);
                if (idx.empty()) {
                    idx.resize(nv);
                    for (uint32_t k = 0; k < nv; k++) idx[k] = k;
                }
                idx.resize(idx.size() / 3 * 3);
                for (uint32_t k : idx)
                    if (k >= nv) return -2;
                if (nrm.size() != pos.size()) {
                    if (!nrm.empty()) return -2;
                    aph::angleWeightedNormals(pos, idx, nrm);
                }
                if (!uv.empty() && uv.size() != (size_t)nv * 2) return -2;
                result = (int)out->pos.size();
                out->pos.push_back(std::move(pos));
                out->nrm.push_back(std::move(nrm));
                out->uv.push_back(std::move(uv));
                out->idx.push_back(std::move(idx));
            }
        }
        meshOf[key] = result;
        return result;
    }

    // node visits over all scenes: a file whose nodes share children (not a tree) could
    // otherwise ask for exponentially many
    size_t visits = 0;
    static constexpr size_t kMaxVisits = size_t(1) << 24;
    static int index(const Json& v)
    {
        return v.kind == Json::Num && v.n >= 0.0 && v.n <= 2147483647.0 ? (int)v.n : -1;
    }

    bool node(int ni, const Mat4& parent, int depth)
    {
        const std::vector<Json>& nodes = doc.arr("nodes");
        if (ni < 0 || ni >= (int)nodes.size() || depth > 64 || ++visits > kMaxVisits) return false;
        const Json& nd = nodes[ni];
        const Mat4 world = mul(parent, nodeTransform(nd));
        const int mi = nd.integer("mesh", -1);
        if (mi >= 0) {
            const std::vector<Json>& meshes = doc.arr("meshes");
            if (mi >= (int)meshes.size()) return false;
            std::string mtlx;
            if (const Json* ext = nd.get("extensions"))
                if (const Json* mx = ext->get("AURORA_materials_materialX")) {
                    const std::string f = mx->str("materialXFile");
                    if (!f.empty()) mtlx = f[0] == '/' ? f : base + f;
                }
            const std::vector<Json>& prims = meshes[mi].arr("primitives");
            for (size_t pi = 0; pi < prims.size(); pi++) {
                int material = -1;
                const int m = primitiveMesh(mi, (int)pi, prims[pi], material);
                if (m == -2) return false;
                if (m < 0) continue;
                aph_gltf_instance inst{};
                inst.mesh = (uint32_t)m;
                inst.material = material;
                inst.node = ni;
                for (int row = 0; row < 3; row++)
                    for (int col = 0; col < 4; col++) inst.o2w[row * 4 + col] = world.m[col][row];
                if (!mtlx.empty()) {
                    out->strings.emplace_back(new std::string(mtlx));
                    inst.materialx = out->strings.back()->c_str();
                }
                out->instances.push_back(inst);
            }
        }
        for (const Json& c : nd.arr("children"))
            if (!node(index(c), world, depth + 1)) return false;
        return true;
    }

    bool cameras(const std::vector<Json>& roots, int depth)
    {
        const std::vector<Json>& nodes = doc.arr("nodes");
        const std::vector<Json>& cams = doc.arr("cameras");
        if (depth > 64) return false;
        for (const Json& r : roots) {
            const int ni = index(r);
            if (ni < 0 || ni >= (int)nodes.size() || ++visits > kMaxVisits) return false;
            const Json& nd = nodes[ni];
            const int ci = nd.integer("camera", -1);
            if (ci >= 0 && ci < (int)cams.size()) {
                const Json& cam = cams[ci];
                aph_gltf_camera c{};
                c.node = ni;
                if (!inverse(nodeTransform(nd), c.view)) return false;
                if (cam.str("type") == "orthographic") {
                    const Json* o = cam.get("orthographic");
                    c.type = 1;
                    if (o) {
                        c.znear = (float)o->num("znear", 0);
                        c.zfar = (float)o->num("zfar", 0);
                        c.xmag = (float)o->num("xmag", 0);
                        c.ymag = (float)o->num("ymag", 0);
                    }
                } else {
                    const Json* pr = cam.get("perspective");
                    c.type = 0;
                    if (pr) {
                        c.yfov = (float)pr->num("yfov", 0);
                        c.aspect = (float)pr->num("aspectRatio", 0);
                        c.znear = (float)pr->num("znear", 0);
                        c.zfar = (float)pr->num("zfar", 0);
                    }
                }
                out->cameras.push_back(c);
            }
            if (!cameras(nd.arr("children"), depth + 1)) return false;
        }
        return true;
    }
};

}  // namespace

extern "C" {

static int gltfLoad(const char* path, aph_gltf_file** out);

int aph_gltf_load(const char* path, aph_gltf_file** out)
{
    if (!path || !out) return APH_E_INVALID;
    *out = nullptr;
    // no exception leaves the C boundary: a file whose arrays do not fit in host memory
    // reports an allocation failure
    try {
        return gltfLoad(path, out);
    } catch (const std::bad_alloc&) {
        return APH_E_MEMORY;
    } catch (...) {
        return APH_E_INVALID;
    }
}

static int gltfLoad(const char* path, aph_gltf_file** out)
{
    std::vector<uint8_t> bytes;
    if (!readFile(path, bytes)) return APH_E_INVALID;
    Loader L;
    L.base = dirOf(path);
    const uint8_t* json = bytes.data();
    size_t jsonLen = bytes.size();
    std::vector<uint8_t> bin;
    bool haveBin = false;
    if (bytes.size() >= 12 && std::memcmp(bytes.data(), "glTF", 4) == 0) {
        // binary container: header (magic, version 2, length), then chunks of (length, type, data)
        if (le32(&bytes[4]) != 2 || le32(&bytes[8]) > bytes.size()) return APH_E_INVALID;
        const size_t total = le32(&bytes[8]);
        size_t off = 12;
        json = nullptr;
        while (off + 8 <= total) {
            const size_t len = le32(&bytes[off]);
            const uint32_t type = le32(&bytes[off + 4]);
            if (off + 8 + len > total) return APH_E_INVALID;
            if (type == 0x4E4F534Au && !json) {  // "JSON"
                json = &bytes[off + 8];
                jsonLen = len;
            } else if (type == 0x004E4942u && !haveBin) {  // "BIN\0"
                bin.assign(bytes.begin() + (long)(off + 8), bytes.begin() + (long)(off + 8 + len));
                haveBin = true;
            }
            off += 8 + ((len + 3) & ~(size_t)3);
        }
        if (!json) return APH_E_INVALID;
    }
    JsonParser jp((const char*)json, (const char*)json + jsonLen);
    if (!jp.parse(L.doc) || L.doc.kind != Json::Obj) return APH_E_INVALID;
    for (const Json& b : L.doc.arr("buffers")) {
        std::vector<uint8_t> data;
        const std::string uri = b.str("uri");
        if (uri.empty()) {
            if (!haveBin) return APH_E_INVALID;
            data = bin;
        } else if (uri.compare(0, 5, "data:") == 0) {
            const size_t comma = uri.find(',');
            if (comma == std::string::npos || uri.find(";base64") == std::string::npos ||
                !base64(uri, comma + 1, data))
                return APH_E_INVALID;
        } else if (!readFile(uriPath(uri, L.base), data)) {
            return APH_E_INVALID;
        }
        size_t len;
        if (!b.size("byteLength", len) || data.size() < len) return APH_E_INVALID;
        L.buffers.push_back(std::move(data));
    }
    std::unique_ptr<GltfFile> f(new GltfFile());
    L.out = f.get();
    for (const Json& m : L.doc.arr("materials")) {
        aph_material a;
        aph_material_defaults(&a);
        const Json* pbr = m.get("pbrMetallicRoughness");
        float bc[3] = {1, 1, 1}, metal = 1.0f, rough = 1.0f;
        if (pbr) {
            const std::vector<Json>& c = pbr->arr("baseColorFactor");
            if (c.size() >= 3)
                for (int k = 0; k < 3; k++) bc[k] = (float)c[k].n;
            metal = (float)pbr->num("metallicFactor", 1.0);
            rough = (float)pbr->num("roughnessFactor", 1.0);
        }
        std::memcpy(a.base_color, bc, sizeof(bc));
        a.metalness = metal;
        a.specular_roughness = rough;
        f->materials.push_back(a);
    }
    for (const Json& sc : L.doc.arr("scenes")) {
        if (!L.cameras(sc.arr("nodes"), 0)) return APH_E_INVALID;
        for (const Json& r : sc.arr("nodes"))
            if (!L.node(Loader::index(r), identity(), 0)) return APH_E_INVALID;
    }
    for (size_t k = 0; k < f->pos.size(); k++) {
        aph_gltf_mesh m{};
        m.nv = (uint32_t)(f->pos[k].size() / 3);
        m.nt = (uint32_t)(f->idx[k].size() / 3);
        m.positions = f->pos[k].data();
        m.normals = f->nrm[k].data();
        m.uvs = f->uv[k].empty() ? nullptr : f->uv[k].data();
        m.indices = f->idx[k].data();
        f->meshes.push_back(m);
    }
    f->pub.nmesh = (uint32_t)f->meshes.size();
    f->pub.ninst = (uint32_t)f->instances.size();
    f->pub.nmat = (uint32_t)f->materials.size();
    f->pub.ncam = (uint32_t)f->cameras.size();
    f->pub.meshes = f->meshes.data();
    f->pub.instances = f->instances.data();
    f->pub.materials = f->materials.data();
    f->pub.cameras = f->cameras.data();
    static_assert(offsetof(GltfFile, pub) == 0, "aph_gltf_file is the first member");
    *out = &f.release()->pub;
    return APH_OK;
}

void aph_gltf_free(aph_gltf_file* f) { delete reinterpret_cast<GltfFile*>(f); }

}  // extern "C"
