// aph_plasma: a batch renderer over the C-ABI, shaped like Plasma's command-line capture
// (Applications/Plasma/Plasma.cpp): load a scene file, set up the camera, light and
// environment as Plasma does, render, and write the image.  Built with g++ against
// include/aph.h alone.
//
// usage: aph_plasma <scene.obj|scene.gltf|scene.glb> [options]
//   --env FILE.hdr            lat-long light and background (Plasma.cpp:1341)
//   --dims W,H                image size (default 1280,720, Plasma.h:177)
//   --eye X,Y,Z --target X,Y,Z --up X,Y,Z
//                             camera view; without them the camera is fitted to the scene
//                             bounds along (0, -0.5, -1) (Plasma.cpp:953-955, Camera.cpp:145-178)
//   --fov DEG                 vertical field of view (default 45, Camera.h:75)
//   --camera FILE.gltf --camera_id N
//                             a camera read from a glTF file (glTFLoader.cpp:37-85)
//   --lightDir X,Y,Z --lightColor R,G,B --lightIntensity I
//                             the distant light (defaults (1,-0.5,0), (1,1,1), 2: Plasma.h:226-228)
//   --trace-depth N           bounces, clamped to [1,10] (default 5)
//   --russian_roulette 0|1    (default 1)
//   --output FILE             .png (post-processed 8-bit, gamma on), .hdr (Radiance RGBE of the
//                             accumulation, as stbi_write_hdr) or .pfm (float, bottom row first)
//   --output_spp N            samples in one render(0, N) call (default: render(0, 1000),
//                             Plasma.cpp:1500-1509)
//   --output_frames K         K frames of render(0, 1) (the frame-sequence capture, :1466-1497)
//   --dump_frames_dir DIR     with --output_frames: frame_NNN.png per frame
//   --rotate_deg_per_frame D  with --output_frames: orbit the camera about its target
//   --progressive MS          instead, Plasma's interactive loop: each frame takes the sample
//                             count the SampleCounter gives for an MS-millisecond frame
//                             (Plasma.cpp:1044-1066, Timer.h:97-281) until --output_spp (or
//                             1000) samples are in; one line per frame on stdout
//   --performance_output FILE the render time in milliseconds (Plasma.cpp:1630-1637)
//   --mode wavefront|megakernel  launch form (default wavefront)
//   --frame_out FILE          the aph_frame used, as raw bytes (for tests)
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "aph.h"
#include "sample_counter.h"

namespace {

using Vec3 = std::array<double, 3>;

struct Args {
    std::string scene;
    std::map<std::string, std::string> opt;
    bool has(const char* k) const { return opt.count(k) != 0; }
    std::string str(const char* k, const std::string& d = std::string()) const
    {
        auto it = opt.find(k);
        return it == opt.end() ? d : it->second;
    }
    double num(const char* k, double d) const { return has(k) ? std::atof(str(k).c_str()) : d; }
    // a whole-number option in [lo, hi]; anything else is an error
    static uint32_t checked(const char* k, double v, uint32_t lo, uint32_t hi)
    {
        if (!(v >= lo && v <= hi) || v != std::floor(v))
            throw std::runtime_error(std::string("--") + k + " takes a whole number in [" + std::to_string(lo) +
                                     ", " + std::to_string(hi) + "]");
        return (uint32_t)v;
    }
    uint32_t count(const char* k, uint32_t d, uint32_t lo, uint32_t hi) const { return checked(k, num(k, d), lo, hi); }
    bool vec(const char* k, std::vector<double>& v, size_t n) const
    {
        if (!has(k)) return false;
        v.clear();
        std::string s = str(k);
        size_t p = 0;
        while (p <= s.size()) {
            size_t q = s.find(',', p);
            if (q == std::string::npos) q = s.size();
            v.push_back(std::atof(s.substr(p, q - p).c_str()));
            p = q + 1;
        }
        if (v.size() != n) throw std::runtime_error(std::string("--") + k + " takes " + std::to_string(n) + " values");
        return true;
    }
};

Args parse(int argc, char** argv)
{
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        if (s.rfind("--", 0) == 0) {
            std::string k = s.substr(2), v;
            const size_t eq = k.find('=');
            if (eq != std::string::npos) {
                v = k.substr(eq + 1);
                k = k.substr(0, eq);
            } else if (i + 1 < argc) {
                v = argv[++i];
            } else {
                throw std::runtime_error("option --" + k + " needs a value");
            }
            a.opt[k] = v;
        } else if (a.scene.empty()) {
            a.scene = s;
        } else {
            throw std::runtime_error("unexpected argument " + s);
        }
    }
    return a;
}

bool endsWith(const std::string& s, const char* e)
{
    const size_t n = std::strlen(e);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; i++)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != e[i]) return false;
    return true;
}

Vec3 sub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec3 cross(const Vec3& a, const Vec3& b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec3 normalize(const Vec3& a)
{
    const double l = std::sqrt(dot(a, a));
    return {a[0] / l, a[1] / l, a[2] / l};
}

// Scene contents on the host: meshes, instances, materials and world bounds.
struct HostScene {
    struct Mesh {
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx;
    };
    std::vector<Mesh> meshes;
    std::vector<aph_instance> instances;
    std::vector<aph_material> materials;
    std::vector<aph_gltf_camera> cameras;
    Vec3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};

    void addBounds(const Mesh& m, const float* o2w)
    {
        for (size_t v = 0; v + 2 < m.pos.size(); v += 3) {
            for (int r = 0; r < 3; r++) {
                const double w = o2w[4 * r] * m.pos[v] + o2w[4 * r + 1] * m.pos[v + 1] + o2w[4 * r + 2] * m.pos[v + 2] +
                                 o2w[4 * r + 3];
                lo[r] = std::min(lo[r], w);
                hi[r] = std::max(hi[r], w);
            }
        }
    }
};

// Materials of the file, then Aurora's default for parts without one (appended once).
uint32_t defaultMaterial(HostScene& s, int& slot)
{
    if (slot < 0) {
        aph_material m;
        aph_material_defaults(&m);
        slot = (int)s.materials.size();
        s.materials.push_back(m);
    }
    return (uint32_t)slot;
}

HostScene loadScene(const std::string& path)
{
    HostScene s;
    int dflt = -1;
    if (endsWith(path, ".obj")) {
        aph_obj_file* f = nullptr;
        if (aph_obj_load(path.c_str(), &f) != APH_OK) throw std::runtime_error("cannot read OBJ " + path);
        s.materials.assign(f->materials, f->materials + f->nmat);
        for (uint32_t k = 0; k < f->nshape; k++) {
            const aph_obj_shape& sh = f->shapes[k];
            HostScene::Mesh m;
            m.pos.assign(sh.positions, sh.positions + 3 * sh.nv);
            m.nrm.assign(sh.normals, sh.normals + 3 * sh.nv);
            if (sh.uvs) m.uv.assign(sh.uvs, sh.uvs + 2 * sh.nv);
            m.idx.assign(sh.indices, sh.indices + 3 * sh.nt);
            aph_instance in{};
            in.o2w[0] = in.o2w[5] = in.o2w[10] = 1.0f;
            in.mesh = k;  // mesh index until uploaded
            in.material = sh.material < 0 ? defaultMaterial(s, dflt) : (uint32_t)sh.material;
            in.mask = 0xFF;
            s.addBounds(m, in.o2w);
            s.meshes.push_back(std::move(m));
            s.instances.push_back(in);
        }
        aph_obj_free(f);
    } else if (endsWith(path, ".gltf") || endsWith(path, ".glb")) {
        aph_gltf_file* f = nullptr;
        if (aph_gltf_load(path.c_str(), &f) != APH_OK) throw std::runtime_error("cannot read glTF " + path);
        s.materials.assign(f->materials, f->materials + f->nmat);
        for (uint32_t k = 0; k < f->nmesh; k++) {
            const aph_gltf_mesh& gm = f->meshes[k];
            HostScene::Mesh m;
            m.pos.assign(gm.positions, gm.positions + 3 * gm.nv);
            m.nrm.assign(gm.normals, gm.normals + 3 * gm.nv);
            if (gm.uvs) m.uv.assign(gm.uvs, gm.uvs + 2 * gm.nv);
            m.idx.assign(gm.indices, gm.indices + 3 * gm.nt);
            s.meshes.push_back(std::move(m));
        }
        for (uint32_t k = 0; k < f->ninst; k++) {
            const aph_gltf_instance& gi = f->instances[k];
            aph_instance in{};
            std::memcpy(in.o2w, gi.o2w, sizeof(in.o2w));
            in.mesh = gi.mesh;
            in.material = gi.material < 0 ? defaultMaterial(s, dflt) : (uint32_t)gi.material;
            in.mask = 0xFF;
            s.addBounds(s.meshes[gi.mesh], in.o2w);
            s.instances.push_back(in);
        }
        s.cameras.assign(f->cameras, f->cameras + f->ncam);
        aph_gltf_free(f);
    } else {
        throw std::runtime_error("scene files are .obj, .gltf or .glb: " + path);
    }
    if (s.instances.empty()) throw std::runtime_error("the scene has no triangles: " + path);
    return s;
}

// glm::lookAt as the rows (right, up, back, eye) of FrameData's inverse view.
void setView(aph_frame& fr, const Vec3& eye, const Vec3& target, const Vec3& up)
{
    const Vec3 f = normalize(sub(target, eye));
    const Vec3 s = normalize(cross(f, up));
    const Vec3 u = cross(s, f);
    const double rows[16] = {s[0], s[1], s[2], 0, u[0], u[1], u[2], 0, -f[0], -f[1], -f[2], 0, eye[0], eye[1], eye[2], 1};
    for (int i = 0; i < 16; i++) fr.inv_view[i] = (float)rows[i];
}

// glm::perspective's (2 / P00, 2 / P11).
void setPerspective(aph_frame& fr, double fovY, double aspect)
{
    const double t = 1.0 / std::tan(fovY / 2.0);
    const float p00 = (float)(t / aspect), p11 = (float)t;
    fr.is_ortho = 0;
    fr.view_size[0] = 2.0f / p00;
    fr.view_size[1] = 2.0f / p11;
}

// Camera::fit (Camera.cpp:145-178): target at the bounds' centre, the view box of the
// bounds' corners, and the distance at which that box fills the field of view.
void fitCamera(const HostScene& s, const Vec3& direction, double fovY, double aspect, Vec3& eye, Vec3& target)
{
    const Vec3 dir = normalize(direction);
    target = {(s.lo[0] + s.hi[0]) / 2, (s.lo[1] + s.hi[1]) / 2, (s.lo[2] + s.hi[2]) / 2};
    const Vec3 e0 = sub(target, dir);
    const Vec3 f = normalize(sub(target, e0));
    const Vec3 r = normalize(cross(f, Vec3{0, 1, 0}));
    const Vec3 u = cross(r, f);
    Vec3 vlo{1e300, 1e300, 1e300}, vhi{-1e300, -1e300, -1e300};
    for (int c = 0; c < 8; c++) {
        const Vec3 p{(c & 1) ? s.hi[0] : s.lo[0], (c & 2) ? s.hi[1] : s.lo[1], (c & 4) ? s.hi[2] : s.lo[2]};
        const Vec3 d = sub(p, e0);
        const Vec3 v{dot(r, d), dot(u, d), -dot(f, d)};
        for (int k = 0; k < 3; k++) {
            vlo[k] = std::min(vlo[k], v[k]);
            vhi[k] = std::max(vhi[k], v[k]);
        }
    }
    const Vec3 dims = sub(vhi, vlo);
    const double boxAspect = dims[0] / dims[1];
    const double halfFov = boxAspect > aspect ? std::atan(std::tan(fovY * 0.5) * aspect) : fovY * 0.5;
    const double size = boxAspect > aspect ? dims[0] : dims[1];
    const double dist = size / (2.0 * std::tan(halfFov)) + dims[2] / 2.0;
    eye = {target[0] - dist * dir[0], target[1] - dist * dir[1], target[2] - dist * dir[2]};
}

// ---- image files ----
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

void put32(std::vector<uint8_t>& b, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
}

// PNG, RGBA8, rows unfiltered, zlib stream of stored (uncompressed) deflate blocks.
void writePng(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t w, uint32_t h)
{
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (4 * w + 1));
    for (uint32_t y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba.begin() + (long)((size_t)y * w * 4), rgba.begin() + (long)((size_t)(y + 1) * w * 4));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back((uint8_t)(n & 0xFF));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xFF));
        z.push_back((uint8_t)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)off, raw.begin() + (long)(off + n));
        if (raw.empty()) break;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
        put32(out, (uint32_t)data.size());
        const size_t start = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), data.begin(), data.end());
        put32(out, crc32(&out[start], out.size() - start));
    };
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, no filter method, no interlace
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    std::ofstream f(path, std::ios::binary);
    f.write((const char*)out.data(), (std::streamsize)out.size());
    if (!f) throw std::runtime_error("cannot write " + path);
}

// Radiance RGBE as stbi_write_hdr writes it (the largest channel's exponent; alpha dropped).
void writeHdr(const std::string& path, const std::vector<float>& rgba, uint32_t w, uint32_t h)
{
    std::ofstream f(path, std::ios::binary);
    const std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y " + std::to_string(h) + " +X " +
                             std::to_string(w) + "\n";
    f.write(head.data(), (std::streamsize)head.size());
    std::vector<uint8_t> px((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const float* c = &rgba[4 * i];
        const float m = std::max(c[0], std::max(c[1], c[2]));
        if (m < 1e-32f) {
            px[4 * i] = px[4 * i + 1] = px[4 * i + 2] = px[4 * i + 3] = 0;
        } else {
            int e = 0;
            const float norm = std::frexp(m, &e) * 256.0f / m;
            for (int k = 0; k < 3; k++) px[4 * i + k] = (uint8_t)(c[k] * norm);
            px[4 * i + 3] = (uint8_t)(e + 128);
        }
    }
    f.write((const char*)px.data(), (std::streamsize)px.size());
    if (!f) throw std::runtime_error("cannot write " + path);
}

// Portable float map, RGB, little-endian, bottom row first.
void writePfm(const std::string& path, const std::vector<float>& rgba, uint32_t w, uint32_t h)
{
    std::ofstream f(path, std::ios::binary);
    const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    f.write(head.data(), (std::streamsize)head.size());
    std::vector<float> row((size_t)w * 3);
    for (uint32_t y = h; y-- > 0;) {
        for (uint32_t x = 0; x < w; x++)
            for (int k = 0; k < 3; k++) row[3 * x + k] = rgba[((size_t)y * w + x) * 4 + k];
        f.write((const char*)row.data(), (std::streamsize)(row.size() * sizeof(float)));
    }
    if (!f) throw std::runtime_error("cannot write " + path);
}

class Renderer {
public:
    Renderer()
    {
        if (aph_create(0, &ctx_) != APH_OK) throw std::runtime_error("no HIP device for the renderer");
    }
    ~Renderer() { aph_destroy(ctx_); }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    void check(int rc, const char* what)
    {
        if (rc != APH_OK) throw std::runtime_error(std::string(what) + ": " + aph_last_error(ctx_));
    }
    aph_ctx_t ctx() const { return ctx_; }

    void setScene(const HostScene& s)
    {
        std::vector<uint64_t> handles;
        for (const HostScene::Mesh& m : s.meshes) {
            uint64_t h = 0;
            check(aph_mesh_upload(ctx_, m.pos.data(), (uint32_t)(m.pos.size() / 3), m.idx.data(),
                                  (uint32_t)(m.idx.size() / 3), m.nrm.data(), m.uv.empty() ? nullptr : m.uv.data(), &h),
                  "mesh upload");
            check(aph_blas_build(ctx_, h, APH_BUILDER_SAH), "BLAS build");
            handles.push_back(h);
        }
        std::vector<aph_instance> inst = s.instances;
        for (aph_instance& in : inst) in.mesh = (uint32_t)handles[in.mesh];
        check(aph_materials_upload(ctx_, s.materials.data(), (uint32_t)s.materials.size()), "materials");
        check(aph_tlas_build(ctx_, inst.data(), (uint32_t)inst.size(), &tlas_), "TLAS build");
    }
    void setEnvironment(const aph_environment& e) { check(aph_environment_upload(ctx_, &e), "environment"); }
    void setTarget(uint32_t w, uint32_t h)
    {
        check(aph_target_create(ctx_, w, h, &target_), "render buffer");
        w_ = w;
        h_ = h;
    }
    void render(const aph_frame& fr, uint32_t start, uint32_t count, int mode)
    {
        aph_request r{};
        r.target = target_;
        r.tlas = tlas_;
        r.frame = fr;
        r.sample_start = start;
        r.sample_count = count;
        check(aph_render(ctx_, &r, 1, mode), "render");
    }
    void wait() { check(aph_wait(ctx_), "wait"); }
    std::vector<float> accumulation()
    {
        std::vector<float> px((size_t)w_ * h_ * 4);
        check(aph_readback(ctx_, target_, px.data(), px.size() * sizeof(float)), "readback");
        return px;
    }
    // the Integer_RGBA render buffer: post-processing (brightness 1, no tone mapping, gamma
    // on: RendererBase.cpp:64-72) stored as unorm8
    std::vector<uint8_t> display()
    {
        std::vector<float> px((size_t)w_ * h_ * 4);
        const float one[3] = {1.0f, 1.0f, 1.0f};
        check(aph_post(ctx_, target_, one, 0, 1, px.data(), px.size() * sizeof(float)), "post");
        std::vector<uint8_t> out(px.size());
        for (size_t i = 0; i < px.size(); i++) {
            const float v = std::min(std::max(px[i], 0.0f), 1.0f);
            out[i] = (uint8_t)std::lround(v * 255.0f);
        }
        return out;
    }

private:
    aph_ctx_t ctx_ = nullptr;
    uint64_t tlas_ = 0, target_ = 0;
    uint32_t w_ = 0, h_ = 0;
};

void writeImage(Renderer& r, const std::string& path, uint32_t w, uint32_t h)
{
    if (endsWith(path, ".hdr")) writeHdr(path, r.accumulation(), w, h);
    else if (endsWith(path, ".pfm")) writePfm(path, r.accumulation(), w, h);
    else writePng(path, r.display(), w, h);
}

int run(const Args& a)
{
    if (a.scene.empty()) throw std::runtime_error("no scene file given");
    uint32_t w = 1280, h = 720;
    std::vector<double> v;
    if (a.vec("dims", v, 2)) {
        w = Args::checked("dims", v[0], 1, 16384);
        h = Args::checked("dims", v[1], 1, 16384);
    }
    const HostScene scene = loadScene(a.scene);
    Renderer r;
    r.setScene(scene);
    r.setTarget(w, h);

    // environment: the file's lat-long image as light and background, else Aurora's gradients
    // (EnvironmentBase.cpp: white-to-black light, the blue background gradient)
    aph_environment env{};
    const float lt[3] = {1, 1, 1}, lb[3] = {0, 0, 0}, bt[3] = {0.02f, 0.25f, 0.60f}, bb[3] = {0.32f, 0.79f, 1.0f};
    std::memcpy(env.light_top, lt, sizeof(lt));
    std::memcpy(env.light_bottom, lb, sizeof(lb));
    std::memcpy(env.background_top, bt, sizeof(bt));
    std::memcpy(env.background_bottom, bb, sizeof(bb));
    for (int i = 0; i < 4; i++) env.light_transform[5 * i] = env.background_transform[5 * i] = 1.0f;
    float* hdr = nullptr;
    if (a.has("env")) {
        uint32_t ew = 0, eh = 0;
        if (aph_hdr_load(a.str("env").c_str(), &hdr, &ew, &eh) != APH_OK)
            throw std::runtime_error("cannot read " + a.str("env"));
        env.light_rgb = env.background_rgb = hdr;
        env.light_width = env.background_width = ew;
        env.light_height = env.background_height = eh;
    }
    r.setEnvironment(env);
    aph_host_free(hdr);

    // camera (Plasma.cpp:674-777)
    aph_frame fr{};
    const double aspect = (double)w / (double)h;
    double fov = a.num("fov", 45.0) * M_PI / 180.0;
    Vec3 eye, target, up{0, 1, 0};
    fitCamera(scene, Vec3{0.0, -0.5, -1.0}, fov, aspect, eye, target);
    if (a.has("eye") || a.has("target")) {
        // either given: the other keeps Plasma's default (eye (0, 0, 1), target the origin)
        eye = {0.0, 0.0, 1.0};
        target = {0.0, 0.0, 0.0};
        if (a.vec("eye", v, 3)) eye = {v[0], v[1], v[2]};
        if (a.vec("target", v, 3)) target = {v[0], v[1], v[2]};
    }
    if (a.vec("up", v, 3)) up = {v[0], v[1], v[2]};
    setView(fr, eye, target, up);
    setPerspective(fr, fov, aspect);
    if (a.has("camera")) {
        aph_gltf_file* f = nullptr;
        if (aph_gltf_load(a.str("camera").c_str(), &f) != APH_OK) throw std::runtime_error("cannot read " + a.str("camera"));
        const int id = (int)a.num("camera_id", 0);
        if (id < 0 || id >= (int)f->ncam) {
            aph_gltf_free(f);
            throw std::runtime_error("no camera " + std::to_string(id) + " in " + a.str("camera"));
        }
        const aph_gltf_camera c = f->cameras[id];
        aph_gltf_free(f);
        // camera-to-world = inverse(view); FrameData's rows are its columns
        float inv[12];
        float vm[12];
        for (int k = 0; k < 12; k++) vm[k] = c.view[k];
        aph_inverse3x4(vm, inv);
        for (int col = 0; col < 4; col++)
            for (int row = 0; row < 3; row++) fr.inv_view[4 * col + row] = inv[4 * row + col];
        fr.inv_view[3] = fr.inv_view[7] = fr.inv_view[11] = 0.0f;
        fr.inv_view[15] = 1.0f;
        if (c.type == 1) {
            fr.is_ortho = 1;
            fr.view_size[0] = 2.0f * c.xmag;
            fr.view_size[1] = 2.0f * c.ymag;
        } else {
            setPerspective(fr, c.yfov, c.aspect > 0.0f ? c.aspect : aspect);
        }
    }
    fr.trace_depth = std::min(std::max((int)a.num("trace-depth", 5), 1), 10);
    fr.max_luminance = 1000.0f;
    fr.rr_enabled = (int)a.num("russian_roulette", 1) != 0;
    fr.rr_start_depth = 3;
    fr.force_opaque_shadows = 1;

    // distant light (Plasma::updateLighting, Plasma.cpp:958-986; FrameData takes the direction
    // toward the light and the cosine of half the default 0.1-radian angular diameter,
    // HGIScene.cpp:120-128)
    Vec3 ldir{1.0, -0.5, 0.0}, lcol{1, 1, 1};
    if (a.vec("lightDir", v, 3)) ldir = {v[0], v[1], v[2]};
    if (a.vec("lightColor", v, 3)) lcol = {v[0], v[1], v[2]};
    const double lint = a.num("lightIntensity", 2.0);
    ldir = normalize(ldir);
    auto srgbToLinear = [](double c) { return c * (c * (c * 0.305306011 + 0.682171111) + 0.012522878); };
    fr.light_count = 1;
    for (int k = 0; k < 3; k++) {
        fr.lights[0].color_intensity[k] = (float)srgbToLinear(lcol[k]);
        fr.lights[0].direction[k] = (float)-ldir[k];
    }
    fr.lights[0].color_intensity[3] = (float)lint;
    fr.lights[0].cos_radius = std::cos(0.5f * 0.1f);

    const int mode = a.str("mode", "wavefront") == "megakernel" ? APH_MODE_MEGAKERNEL : APH_MODE_WAVEFRONT;
    const uint32_t spp = a.count("output_spp", 1000, 1, 1u << 20);
    aph_app::SteadyTimer timer;
    if (a.has("progressive")) {
        // Plasma's interactive update: as many samples per frame as fit the frame time
        const double frameMs = a.num("progressive", 33.3);
        if (!(frameMs > 0.0 && frameMs <= 1e6)) throw std::runtime_error("--progressive takes a frame time in (0, 1e6] ms");
        aph_app::SampleCounter<> counter((float)frameMs, 200.0f, spp);
        uint32_t frame = 0;
        while (!counter.isComplete()) {
            uint32_t start = 0;
            const uint32_t n = counter.update(start, false);
            if (n == 0) break;
            r.render(fr, start, n, mode);
            r.wait();
            std::printf("frame %u: samples %u..%u\n", frame++, start, start + n);
        }
    } else if (a.has("output_frames")) {
        const int frames = (int)a.count("output_frames", 32, 1, 100000);
        const double deg = a.num("rotate_deg_per_frame", 0.0);
        const std::string dump = a.str("dump_frames_dir");
        for (int k = 0; k < frames; k++) {
            if (deg != 0.0 && k > 0) {
                // orbit about the target around the view's up axis (Plasma.cpp:1476-1486)
                const Vec3 u = normalize(up);
                const Vec3 off = sub(eye, target);
                const double t = deg * M_PI / 180.0, c = std::cos(t), s = std::sin(t);
                const Vec3 kx = cross(u, off);
                const double kd = dot(u, off);
                for (int i = 0; i < 3; i++) eye[i] = target[i] + off[i] * c + kx[i] * s + u[i] * kd * (1 - c);
                setView(fr, eye, target, up);
            }
            r.render(fr, 0, 1, mode);
            if (!dump.empty()) {
                r.wait();
                char name[64];
                std::snprintf(name, sizeof(name), "/frame_%03d.png", k);
                writePng(dump + name, r.display(), w, h);
            }
        }
    } else {
        r.render(fr, 0, spp, mode);
    }
    r.wait();
    const float ms = timer.elapsed();
    if (a.has("performance_output")) {
        std::ofstream f(a.str("performance_output"));
        f << ms << "\n";
    }
    std::printf("Rendering completed in %d ms.\n", (int)ms);
    if (a.has("output")) writeImage(r, a.str("output"), w, h);
    if (a.has("frame_out")) {
        std::ofstream f(a.str("frame_out"), std::ios::binary);
        f.write((const char*)&fr, sizeof(fr));
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return run(parse(argc, argv));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "aph_plasma: %s\n", e.what());
        return 1;
    }
}
