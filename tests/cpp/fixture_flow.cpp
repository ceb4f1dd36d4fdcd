// A C++ host of the backend, built with g++ against include/aph.h alone: the way an
// Aurora-side HIPRenderer drives the C-ABI.  It runs the render flow of the reference's
// test fixture (Tests/Helpers/AuroraTestHelpers.cpp:574-626: a 45-degree perspective
// camera at (0, 0, -5) looking at the origin; :826-852: render(start, sampleCount) called
// `iterations` times with start advancing by sampleCount) on one mesh with the default
// Standard Surface material and environment, and writes the linear RGBA32F image.
//
// usage: fixture_flow mesh.bin out.rgba32f [width height iterations samples]
//   mesh.bin: uint32 nv, nt; float positions[3 nv]; float normals[3 nv]; uint32 indices[3 nt]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "aph.h"

namespace {

// Status codes become exceptions at this boundary, as Aurora's AU_FAIL would report them
// (Libraries/Foundation/API/Aurora/Foundation/Log.h:79-113).
class HipRenderer {
public:
    HipRenderer()
    {
        if (aph_create(0, &ctx_) != APH_OK) throw std::runtime_error("no HIP device for the renderer");
    }
    ~HipRenderer() { aph_destroy(ctx_); }
    HipRenderer(const HipRenderer&) = delete;
    HipRenderer& operator=(const HipRenderer&) = delete;

    void check(int rc, const char* what)
    {
        if (rc != APH_OK) throw std::runtime_error(std::string(what) + ": " + aph_last_error(ctx_));
    }

    // IScene: one instance of one mesh with the default material (base colour 0.8).
    void setMesh(const std::vector<float>& pos, const std::vector<float>& nrm, const std::vector<uint32_t>& idx)
    {
        uint64_t mesh = 0;
        check(aph_mesh_upload(ctx_, pos.data(), (uint32_t)(pos.size() / 3), idx.data(), (uint32_t)(idx.size() / 3),
                              nrm.empty() ? nullptr : nrm.data(), nullptr, &mesh),
              "mesh upload");
        check(aph_blas_build(ctx_, mesh, APH_BUILDER_SAH), "BLAS build");
        aph_instance inst{};
        inst.o2w[0] = inst.o2w[5] = inst.o2w[10] = 1.0f;  // identity, row-major 3x4
        inst.mesh = (uint32_t)mesh;
        inst.material = 0;
        inst.mask = 0xFF;
        check(aph_tlas_build(ctx_, &inst, 1, &tlas_), "TLAS build");
        aph_material m;
        aph_material_defaults(&m);
        m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.8f;
        check(aph_materials_upload(ctx_, &m, 1), "materials");
    }

    // IRenderer::createRenderBuffer + setTargets.
    void setTarget(uint32_t w, uint32_t h)
    {
        check(aph_target_create(ctx_, w, h, &target_), "render buffer");
        w_ = w;
        h_ = h;
    }

    // IRenderer::setCamera with glm::lookAt / glm::perspective, reduced to the FrameData
    // fields the path reads (RendererBase::updateFrameDataGPUStruct, RendererBase.cpp:225-295).
    void setCamera(const double eye[3], const double target[3], const double up[3], double fovYDeg)
    {
        double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
        normalize(f);
        double s[3];
        cross(f, up, s);
        normalize(s);
        double u[3];
        cross(s, f, u);
        const double rows[16] = {s[0], s[1], s[2], 0.0, u[0], u[1], u[2], 0.0, -f[0], -f[1], -f[2], 0.0,
                                 eye[0], eye[1], eye[2], 1.0};
        for (int i = 0; i < 16; i++) frame_.inv_view[i] = (float)rows[i];
        const double t = 1.0 / std::tan(fovYDeg * M_PI / 180.0 / 2.0);
        const float p00 = (float)(t / ((double)w_ / (double)h_)), p11 = (float)t;
        frame_.view_size[0] = 2.0f / p00;
        frame_.view_size[1] = 2.0f / p11;
        frame_.trace_depth = 5;
        frame_.max_luminance = 1000.0f;
        frame_.rr_enabled = 1;
        frame_.rr_start_depth = 3;
        frame_.force_opaque_shadows = 1;  // HGI's setting (HGIRenderer.cpp:490-491)
    }

    // IRenderer::render(sampleStart, sampleCount) and waitForTask.
    void render(uint32_t sampleStart, uint32_t sampleCount)
    {
        aph_request r{};
        r.target = target_;
        r.tlas = tlas_;
        r.frame = frame_;
        r.sample_start = sampleStart;
        r.sample_count = sampleCount;
        check(aph_render(ctx_, &r, 1, APH_MODE_WAVEFRONT), "render");
    }
    void waitForTask() { check(aph_wait(ctx_), "wait"); }

    // IRenderBuffer::data: the linear accumulation.
    std::vector<float> data()
    {
        std::vector<float> px((size_t)w_ * h_ * 4);
        check(aph_readback(ctx_, target_, px.data(), px.size() * sizeof(float)), "readback");
        return px;
    }

private:
    static void normalize(double v[3])
    {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int i = 0; i < 3; i++) v[i] /= l;
    }
    static void cross(const double a[3], const double b[3], double o[3])
    {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    }

    aph_ctx_t ctx_ = nullptr;
    uint64_t tlas_ = 0, target_ = 0;
    uint32_t w_ = 0, h_ = 0;
    aph_frame frame_{};
};

template <class T> void readArray(FILE* f, std::vector<T>& v, size_t n)
{
    v.resize(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("truncated mesh file");
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s mesh.bin out.rgba32f [width height iterations samples]\n", argv[0]);
        return 2;
    }
    const uint32_t w = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 128, h = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 128;
    const uint32_t iterations = argc > 5 ? (uint32_t)std::atoi(argv[5]) : 32;
    const uint32_t samples = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 4;
    try {
        FILE* f = std::fopen(argv[1], "rb");
        if (!f) throw std::runtime_error(std::string("cannot open ") + argv[1]);
        uint32_t nv = 0, nt = 0;
        if (std::fread(&nv, 4, 1, f) != 1 || std::fread(&nt, 4, 1, f) != 1) throw std::runtime_error("bad mesh file");
        std::vector<float> pos, nrm;
        std::vector<uint32_t> idx;
        readArray(f, pos, (size_t)nv * 3);
        readArray(f, nrm, (size_t)nv * 3);
        readArray(f, idx, (size_t)nt * 3);
        std::fclose(f);

        HipRenderer r;
        r.setMesh(pos, nrm, idx);
        r.setTarget(w, h);
        const double eye[3] = {0.0, 0.0, -5.0}, at[3] = {0.0, 0.0, 0.0}, up[3] = {0.0, 1.0, 0.0};
        r.setCamera(eye, at, up, 45.0);
        // FixtureBase::renderAndCheckBaselineImage: start advances by the sample count
        uint32_t start = 0;
        for (uint32_t i = 0; i < iterations; i++) {
            r.render(start, samples);
            start += samples;
        }
        r.waitForTask();
        const std::vector<float> px = r.data();
        FILE* o = std::fopen(argv[2], "wb");
        if (!o || std::fwrite(px.data(), sizeof(float), px.size(), o) != px.size()) throw std::runtime_error("write");
        std::fclose(o);
        std::printf("fixture_flow: %ux%u, %u x render(start, %u): %u samples per pixel\n", w, h, iterations, samples,
                    start);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fixture_flow: %s\n", e.what());
        return 1;
    }
    return 0;
}
