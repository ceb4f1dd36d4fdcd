// On-disk inputs: Wavefront OBJ + MTL scenes and Radiance HDR lat-long images (host code).
//
// Plasma loads OBJ files through tinyobjloader and translates them into Aurora geometry
// and Standard Surface materials (Applications/Plasma/OBJLoader.cpp:77-470), and loads
// lat-long environments from .hdr files (Plasma.cpp:69, :1341) through stb_image.
// Here both are parsed directly:
//   - OBJ: v / vn / vt / f (1-based or negative indices; v, v/t, v//n, v/t/n), o / g
//     start a shape, usemtl, mtllib.  Quads are split along their shorter diagonal (0-2
//     when strictly shorter, else 1-3), as tinyobjloader v2.0.0rc13 triangulates them;
//     larger polygons are fanned from their first corner, which covers the same surface
//     as that library's ear clipping for convex polygons.  A corner whose index lies
//     outside the arrays read so far rejects the file (the reference warns and drops it).  Each shape
//     becomes one mesh whose vertices are the distinct (position, normal, uv) triples in
//     first-use order (OBJLoader.cpp:266-356); normals are normalized, V is flipped, and
//     a shape without normals gets the angle-weighted normals of Foundation's
//     calculateNormals (Geometry.cpp:29-101, including its use of pi minus the second
//     vertex's angle).  A shape's material is its first face's (OBJLoader.cpp:443-449).
//   - MTL: newmtl, Kd, Ks, Ke, Tf, Ni, d, Tr, Pr, Pm and the texture paths map_Kd,
//     map_Pr, map_Ke, map_d, norm / bump / map_bump; tinyobjloader's defaults (colours 0,
//     Ni 1, d 1); the Standard Surface mapping of OBJLoader.cpp:162-229 (sRGB colours
//     linearized with Plasma's polynomial, pch.h:125-128; IOR 1 -> 1.5; transmission =
//     1 - d; a black Tf -> white; emission 1 when Ke is non-zero or an emission map
//     exists; opacity 1).
//   - HDR: "#?RADIANCE" / "#?RGBE" header, 32-bit_rle_rgbe, "-Y h +X w" orientation,
//     new-style run-length scanlines or flat pixels; RGBE to float as stb_image does
//     (mantissa * 2^(exponent - 136), exponent 0 -> black).
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
#include <tuple>
#include <vector>

#include "../../include/aph.h"
#include "assets.h"

namespace {

struct ObjMaterial {
    std::string name;
    float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0}, tf[3] = {0, 0, 0};
    float ni = 1.0f, d = 1.0f, pr = 0.0f, pm = 0.0f;
    bool hasD = false;
    std::string mapKd, mapPr, mapKe, mapD, mapNorm;
};

// Plasma's sRGB approximation (pch.h:125-128).
float srgbToLinearApprox(float v) { return v * (v * (v * 0.305306011f + 0.682171111f) + 0.012522878f); }

std::string dirOf(const std::string& path)
{
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

// The last token(s) of a texture statement: the file name after any options.
std::string texturePath(std::istringstream& ls, const std::string& base)
{
    std::string tok, last;
    while (ls >> tok) last = tok;
    if (last.empty()) return std::string();
    return last[0] == '/' ? last : base + last;
}

bool parseMtl(const std::string& path, std::vector<ObjMaterial>& mats, std::map<std::string, int>& byName)
{
    std::ifstream in(path);
    if (!in) return false;
    const std::string base = dirOf(path);
    std::string line;
    ObjMaterial* cur = nullptr;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            ObjMaterial m;
            ls >> m.name;
            byName[m.name] = (int)mats.size();
            mats.push_back(m);
            cur = &mats.back();
            continue;
        }
        if (!cur) continue;
        auto rgb = [&](float* c) { ls >> c[0] >> c[1] >> c[2]; };
        if (key == "Kd") rgb(cur->kd);
        else if (key == "Ks") rgb(cur->ks);
        else if (key == "Ke") rgb(cur->ke);
        else if (key == "Tf") rgb(cur->tf);
        else if (key == "Ni") ls >> cur->ni;
        else if (key == "d") {
            ls >> cur->d;
            cur->hasD = true;
        } else if (key == "Tr" && !cur->hasD) {  // d wins when both are given
            float tr = 0.0f;
            ls >> tr;
            cur->d = 1.0f - tr;
        } else if (key == "Pr") ls >> cur->pr;
        else if (key == "Pm") ls >> cur->pm;
        else if (key == "map_Kd") cur->mapKd = texturePath(ls, base);
        else if (key == "map_Pr") cur->mapPr = texturePath(ls, base);
        else if (key == "map_Ke") cur->mapKe = texturePath(ls, base);
        else if (key == "map_d") cur->mapD = texturePath(ls, base);
        else if (key == "norm") cur->mapNorm = texturePath(ls, base);
        else if ((key == "bump" || key == "map_bump") && cur->mapNorm.empty()) cur->mapNorm = texturePath(ls, base);
    }
    return true;
}

struct ShapeBuild {
    std::string name;
    int material = -1;
    bool materialSet = false;
    std::vector<std::tuple<int, int, int>> corners;  // (v, vt, vn), 0-based, -1 = none; 3 per triangle
};

struct ObjFile {
    aph_obj_file pub{};
    std::vector<aph_obj_shape> shapes;
    std::vector<aph_material> materials;
    std::vector<std::vector<float>> pos, nrm, uv;
    std::vector<std::vector<uint32_t>> idx;
    std::vector<std::string> names, textures;
    std::vector<const char*> namePtr, texturePtr;
};

}  // namespace

namespace aph {

// Foundation::calculateNormals (Geometry.cpp:29-101): per face, the normal of
// cross(p2 - p1, p3 - p1) weighted at each corner by acos of the edge directions' dot
// product as that code forms it, summed per vertex and normalized.
void angleWeightedNormals(const std::vector<float>& pos, const std::vector<uint32_t>& idx, std::vector<float>& out)
{
    out.assign(pos.size(), 0.0f);
    auto P = [&](uint32_t i) { return std::make_tuple(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); };
    auto norm = [](float& x, float& y, float& z) {
        const float l = std::sqrt(x * x + y * y + z * z);
        x /= l;
        y /= l;
        z /= l;
    };
    for (size_t f = 0; f + 2 < idx.size(); f += 3) {
        const uint32_t i1 = idx[f], i2 = idx[f + 1], i3 = idx[f + 2];
        float x1, y1, z1, x2, y2, z2, x3, y3, z3;
        std::tie(x1, y1, z1) = P(i1);
        std::tie(x2, y2, z2) = P(i2);
        std::tie(x3, y3, z3) = P(i3);
        float ax = x3 - x1, ay = y3 - y1, az = z3 - z1;  // vA
        float bx = x2 - x1, by = y2 - y1, bz = z2 - z1;  // vB
        float nx = by * az - bz * ay, ny = bz * ax - bx * az, nz = bx * ay - by * ax;  // cross(vB, vA)
        norm(nx, ny, nz);
        float v31x = ax, v31y = ay, v31z = az, v21x = bx, v21y = by, v21z = bz;
        float v32x = x3 - x2, v32y = y3 - y2, v32z = z3 - z2;
        norm(v31x, v31y, v31z);
        norm(v21x, v21y, v21z);
        norm(v32x, v32y, v32z);
        const float d1 = v31x * v21x + v31y * v21y + v31z * v21z;
        const float d2 = -(v32x * v21x + v32y * v21y + v32z * v21z);
        const float d3 = v31x * v32x + v31y * v32y + v31z * v32z;
        const float w[3] = {std::fabs(std::acos(d1)), std::fabs(std::acos(-d2)), std::fabs(std::acos(d3))};
        const uint32_t ids[3] = {i1, i2, i3};
        for (int k = 0; k < 3; k++) {
            out[3 * ids[k]] += nx * w[k];
            out[3 * ids[k] + 1] += ny * w[k];
            out[3 * ids[k] + 2] += nz * w[k];
        }
    }
    for (size_t v = 0; v < out.size(); v += 3) norm(out[v], out[v + 1], out[v + 2]);
}

}  // namespace aph

using aph::angleWeightedNormals;

extern "C" {

static int objLoad(const char* path, aph_obj_file** out);
static int hdrLoad(const char* path, float** rgb, uint32_t* width, uint32_t* height);

// No exception leaves the C boundary: input whose arrays do not fit in host memory
// reports an allocation failure.
int aph_obj_load(const char* path, aph_obj_file** out)
{
    try {
        return objLoad(path, out);
    } catch (const std::bad_alloc&) {
        return APH_E_MEMORY;
    } catch (...) {
        return APH_E_INVALID;
    }
}

int aph_hdr_load(const char* path, float** rgb, uint32_t* width, uint32_t* height)
{
    try {
        return hdrLoad(path, rgb, width, height);
    } catch (const std::bad_alloc&) {
        return APH_E_MEMORY;
    } catch (...) {
        return APH_E_INVALID;
    }
}

static int objLoad(const char* path, aph_obj_file** out)
{
    if (!path || !out) return APH_E_INVALID;
    *out = nullptr;
    std::ifstream in(path);
    if (!in) return APH_E_INVALID;
    const std::string base = dirOf(path);
    std::vector<float> V, VT, VN;
    std::vector<ObjMaterial> mats;
    std::map<std::string, int> byName;
    std::vector<ShapeBuild> shapes(1);
    int curMat = -1;
    std::string line;
    auto resolve = [](long i, size_t n) -> int {
        // 0 is not a valid OBJ index; neither is one beyond the arrays' int range
        if (i > 0) return i <= INT_MAX ? (int)(i - 1) : -2;
        if (i < 0) return i >= -(long)n ? (int)((long)n + i) : -2;
        return -2;
    };
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "v") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            V.insert(V.end(), {x, y, z});
        } else if (key == "vt") {
            float u = 0, v = 0;
            ls >> u >> v;
            VT.insert(VT.end(), {u, v});
        } else if (key == "vn") {
            float x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            VN.insert(VN.end(), {x, y, z});
        } else if (key == "o" || key == "g") {
            ShapeBuild s;
            std::getline(ls, s.name);
            if (!shapes.back().corners.empty()) shapes.push_back(s);
            else shapes.back().name = s.name;
        } else if (key == "usemtl") {
            std::string n;
            ls >> n;
            auto it = byName.find(n);
            curMat = it == byName.end() ? -1 : it->second;
        } else if (key == "mtllib") {
            std::string f;
            while (ls >> f) parseMtl(f[0] == '/' ? f : base + f, mats, byName);
        } else if (key == "f") {
            std::vector<std::tuple<int, int, int>> poly;
            std::string tok;
            while (ls >> tok) {
                long a = 0, b = 0, c = 0;
                int have = 0;
                const char* s = tok.c_str();
                char* e = nullptr;
                a = std::strtol(s, &e, 10);
                have = 1;
                if (*e == '/') {
                    s = e + 1;
                    if (*s != '/') {
                        b = std::strtol(s, &e, 10);
                        have |= 2;
                    } else {
                        e = (char*)s;
                    }
                    if (*e == '/') {
                        c = std::strtol(e + 1, &e, 10);
                        have |= 4;
                    }
                }
                const int vi = resolve(a, V.size() / 3);
                const int ti = (have & 2) ? resolve(b, VT.size() / 2) : -1;
                const int ni = (have & 4) ? resolve(c, VN.size() / 3) : -1;
                if (vi < 0 || vi >= (int)(V.size() / 3) || ti == -2 || ti >= (int)(VT.size() / 2) || ni == -2 ||
                    ni >= (int)(VN.size() / 3))
                    return APH_E_INVALID;  // an index outside the arrays read so far
                poly.emplace_back(vi, ti, ni);
            }
            ShapeBuild& sh = shapes.back();
            if (!sh.materialSet) {
                sh.material = curMat;
                sh.materialSet = true;
            }
            if (poly.size() == 4) {  // split along the shorter diagonal, 0-2 on a strict win
                auto d2 = [&](int a, int b) {
                    const int ia = std::get<0>(poly[a]), ib = std::get<0>(poly[b]);
                    float s = 0.0f;
                    for (int k = 0; k < 3; k++) {
                        const float e = V[3 * ib + k] - V[3 * ia + k];
                        s += e * e;
                    }
                    return s;
                };
                const bool along02 = d2(0, 2) < d2(1, 3);
                const int tri[2][3] = {{0, 1, along02 ? 2 : 3}, {along02 ? 0 : 1, 2, 3}};
                for (const auto& t : tri)
                    for (int k : t) sh.corners.push_back(poly[k]);
            } else {
                for (size_t k = 1; k + 1 < poly.size(); k++) {  // fan triangulation
                    sh.corners.push_back(poly[0]);
                    sh.corners.push_back(poly[k]);
                    sh.corners.push_back(poly[k + 1]);
                }
            }
        }
    }
    std::unique_ptr<ObjFile> f(new ObjFile());
    for (const ObjMaterial& m : mats) {
        aph_material a;
        aph_material_defaults(&a);
        float ks[3], ke[3], tf[3];
        for (int k = 0; k < 3; k++) {
            a.base_color[k] = srgbToLinearApprox(m.kd[k]);
            ks[k] = srgbToLinearApprox(m.ks[k]);
            ke[k] = srgbToLinearApprox(m.ke[k]);
            tf[k] = srgbToLinearApprox(m.tf[k]);
        }
        a.metalness = m.pm;
        std::memcpy(a.specular_color, ks, sizeof(ks));
        a.specular_roughness = m.pr;
        a.specular_ior = m.ni == 1.0f ? 1.5f : m.ni;
        a.transmission = 1.0f - m.d;
        if (std::sqrt(tf[0] * tf[0] + tf[1] * tf[1] + tf[2] * tf[2]) == 0.0f) tf[0] = tf[1] = tf[2] = 1.0f;
        std::memcpy(a.transmission_color, tf, sizeof(tf));
        std::memcpy(a.emission_color, ke, sizeof(ke));
        const bool anyEmission = ke[0] != 0.0f || ke[1] != 0.0f || ke[2] != 0.0f;
        a.emission = (!m.mapKe.empty() || anyEmission) ? 1.0f : 0.0f;
        a.opacity[0] = a.opacity[1] = a.opacity[2] = 1.0f;
        f->materials.push_back(a);
        f->names.push_back(m.name);
        for (const std::string* t : {&m.mapKd, &m.mapPr, &m.mapKe, &m.mapD, &m.mapNorm}) f->textures.push_back(*t);
    }
    for (ShapeBuild& sb : shapes) {
        if (sb.corners.empty()) continue;
        const bool hasN = std::get<2>(sb.corners[0]) >= 0, hasT = std::get<1>(sb.corners[0]) >= 0;
        std::map<std::tuple<int, int, int>, uint32_t> seen;
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx;
        for (const auto& c : sb.corners) {
            auto it = seen.find(c);
            if (it != seen.end()) {
                idx.push_back(it->second);
                continue;
            }
            const int vi = std::get<0>(c), ti = std::get<1>(c), ni = std::get<2>(c);
            if ((hasN && ni < 0) || (hasT && ti < 0)) return APH_E_INVALID;  // channels must agree within a shape
            const uint32_t n = (uint32_t)(pos.size() / 3);
            pos.insert(pos.end(), {V[3 * vi], V[3 * vi + 1], V[3 * vi + 2]});
            if (hasN) {
                float x = VN[3 * ni], y = VN[3 * ni + 1], z = VN[3 * ni + 2];
                const float l = std::sqrt(x * x + y * y + z * z);
                nrm.insert(nrm.end(), {x / l, y / l, z / l});
            }
            if (hasT) uv.insert(uv.end(), {VT[2 * ti], 1.0f - VT[2 * ti + 1]});
            seen[c] = n;
            idx.push_back(n);
        }
        if (!hasN) angleWeightedNormals(pos, idx, nrm);
        int mat = sb.material;
        if (mat >= (int)mats.size()) mat = 0;
        f->pos.push_back(std::move(pos));
        f->nrm.push_back(std::move(nrm));
        f->uv.push_back(std::move(uv));
        f->idx.push_back(std::move(idx));
        aph_obj_shape s{};
        s.material = mat;
        f->shapes.push_back(s);
    }
    for (size_t k = 0; k < f->shapes.size(); k++) {
        aph_obj_shape& s = f->shapes[k];
        s.nv = (uint32_t)(f->pos[k].size() / 3);
        s.nt = (uint32_t)(f->idx[k].size() / 3);
        s.positions = f->pos[k].data();
        s.normals = f->nrm[k].data();
        s.uvs = f->uv[k].empty() ? nullptr : f->uv[k].data();
        s.indices = f->idx[k].data();
    }
    for (const std::string& n : f->names) f->namePtr.push_back(n.c_str());
    for (const std::string& t : f->textures) f->texturePtr.push_back(t.empty() ? nullptr : t.c_str());
    f->pub.nshape = (uint32_t)f->shapes.size();
    f->pub.nmat = (uint32_t)f->materials.size();
    f->pub.shapes = f->shapes.data();
    f->pub.materials = f->materials.data();
    f->pub.material_names = f->namePtr.data();
    f->pub.textures = f->texturePtr.data();
    aph_obj_file* p = &f->pub;
    static_assert(offsetof(ObjFile, pub) == 0, "aph_obj_file is the first member");
    f.release();
    *out = p;
    return APH_OK;
}

void aph_obj_free(aph_obj_file* f) { delete reinterpret_cast<ObjFile*>(f); }

static int hdrLoad(const char* path, float** rgb, uint32_t* width, uint32_t* height)
{
    if (!path || !rgb || !width || !height) return APH_E_INVALID;
    *rgb = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return APH_E_INVALID;
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    char line[256];
    if (!std::fgets(line, sizeof(line), f) ||
        (std::strncmp(line, "#?RADIANCE", 10) != 0 && std::strncmp(line, "#?RGBE", 6) != 0))
        return APH_E_INVALID;
    bool rgbe = false;
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '\n') break;
        if (std::strncmp(line, "FORMAT=32-bit_rle_rgbe", 22) == 0) rgbe = true;
    }
    int h = 0, w = 0;
    if (!rgbe || std::fscanf(f, "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0 || w > 65535 || h > 65535)
        return APH_E_UNSUPPORTED;  // only the standard orientation and pixel format
    std::fgetc(f);  // the line end
    // the image must be there before it is allocated: a run-length coded scanline takes
    // at least its 4-byte header and 2 bytes per run of up to 127 values in each of the
    // four channels, a flat one 4 bytes per pixel
    const long at = std::ftell(f);
    if (at < 0 || std::fseek(f, 0, SEEK_END) != 0) return APH_E_INVALID;
    const long end = std::ftell(f);
    if (end < at || std::fseek(f, at, SEEK_SET) != 0) return APH_E_INVALID;
    const uint64_t minLine = (w >= 8 && w < 32768) ? 4 + 8 * (uint64_t)((w + 126) / 127) : 4 * (uint64_t)w;
    if ((uint64_t)(end - at) < minLine * (uint64_t)h) return APH_E_INVALID;
    std::vector<float> out((size_t)w * h * 3);
    std::vector<uint8_t> scan((size_t)w * 4);
    for (int y = 0; y < h; y++) {
        int c0 = std::fgetc(f), c1 = std::fgetc(f), c2 = std::fgetc(f), c3 = std::fgetc(f);
        if (c3 == EOF) return APH_E_INVALID;
        if (c0 == 2 && c1 == 2 && !(c2 & 0x80) && w >= 8 && w < 32768) {
            if (((c2 << 8) | c3) != w) return APH_E_INVALID;
            for (int ch = 0; ch < 4; ch++) {  // the four channel planes, run-length coded
                int x = 0;
                while (x < w) {
                    int n = std::fgetc(f);
                    if (n == EOF) return APH_E_INVALID;
                    if (n > 128) {
                        n -= 128;
                        const int v = std::fgetc(f);
                        if (v == EOF || x + n > w) return APH_E_INVALID;
                        for (int k = 0; k < n; k++) scan[(size_t)(x++) * 4 + ch] = (uint8_t)v;
                    } else {
                        if (n == 0 || x + n > w) return APH_E_INVALID;
                        for (int k = 0; k < n; k++) {
                            const int v = std::fgetc(f);
                            if (v == EOF) return APH_E_INVALID;
                            scan[(size_t)(x++) * 4 + ch] = (uint8_t)v;
                        }
                    }
                }
            }
        } else {  // flat scanline: the four bytes just read are the first pixel
            scan[0] = (uint8_t)c0;
            scan[1] = (uint8_t)c1;
            scan[2] = (uint8_t)c2;
            scan[3] = (uint8_t)c3;
            if (w > 1 && std::fread(scan.data() + 4, 4, (size_t)w - 1, f) != (size_t)w - 1) return APH_E_INVALID;
        }
        for (int x = 0; x < w; x++) {
            const uint8_t* p = &scan[(size_t)x * 4];
            float* o = &out[((size_t)y * w + x) * 3];
            if (p[3] == 0) {
                o[0] = o[1] = o[2] = 0.0f;
            } else {
                const float s = std::ldexp(1.0f, (int)p[3] - 136);
                o[0] = p[0] * s;
                o[1] = p[1] * s;
                o[2] = p[2] * s;
            }
        }
    }
    float* buf = (float*)std::malloc(out.size() * sizeof(float));
    if (!buf) return APH_E_MEMORY;
    std::memcpy(buf, out.data(), out.size() * sizeof(float));
    *rgb = buf;
    *width = (uint32_t)w;
    *height = (uint32_t)h;
    return APH_OK;
}

void aph_host_free(void* p) { std::free(p); }

}  // extern "C"
