// C-ABI of the gfx950 path-tracing backend (include/aph.h): context,
// device-resident scene objects, ray queries, render requests, readback.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <string>
#include <vector>

#include "../../include/aph.h"
#include "alias.h"
#include "bvh_build.h"
#include "kernels.h"
#include "wavefront.h"
#include "context.h"

using namespace aph;

namespace {

int uploadMaterialTextures(aph_ctx* c);

constexpr uint32_t kSahMaxLeaf = 4;  // triangles per leaf of an SAH BLAS

// Double-precision 4x4 inverse (row-major), rounded to float.
bool inverse4(const float* mf, float* out)
{
    double m[16], inv[16];
    for (int i = 0; i < 16; i++) m[i] = mf[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return false;
    for (int i = 0; i < 16; i++) out[i] = (float)(inv[i] / det);
    return true;
}

void identity4(float* m)
{
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

void setDefaultEnv(aph_ctx* c)
{
    // EnvironmentBase.cpp:33-45 defaults.
    c->env.reset(new Environment());
    DevEnv& e = c->env->env;
    std::memset(&e, 0, sizeof(e));
    e.lightTop = f3{1.0f, 1.0f, 1.0f};
    e.lightBottom = f3{0.0f, 0.0f, 0.0f};
    e.bgTop = f3{0.02f, 0.25f, 0.60f};
    e.bgBottom = f3{0.32f, 0.79f, 1.0f};
    identity4(e.lightXf);
    identity4(e.lightXfInv);
    identity4(e.bgXf);
}

DevMaterial toDev(const aph_material& a)
{
    DevMaterial m;
    std::memset(&m, 0, sizeof(m));
    auto v = [](const float* p) { return f3{p[0], p[1], p[2]}; };
    m.base = a.base; m.baseColor = v(a.base_color);
    m.diffuseRoughness = a.diffuse_roughness; m.metalness = a.metalness; m.specular = a.specular;
    m.specularRoughness = a.specular_roughness; m.specularColor = v(a.specular_color); m.specularIOR = a.specular_ior;
    m.specularAnisotropy = a.specular_anisotropy; m.specularRotation = a.specular_rotation;
    m.transmission = a.transmission; m.subsurface = a.subsurface; m.transmissionColor = v(a.transmission_color);
    m.sheen = a.sheen; m.subsurfaceColor = v(a.subsurface_color); m.sheenRoughness = a.sheen_roughness;
    m.sheenColor = v(a.sheen_color); m.coat = a.coat; m.coatColor = v(a.coat_color);
    m.coatRoughness = a.coat_roughness; m.coatAnisotropy = a.coat_anisotropy; m.coatRotation = a.coat_rotation;
    m.coatIOR = a.coat_ior; m.coatAffectColor = a.coat_affect_color; m.coatAffectRoughness = a.coat_affect_roughness;
    m.emission = a.emission; m.thinWalled = a.thin_walled; m.emissionColor = v(a.emission_color);
    m.opacity = v(a.opacity);
    return m;
}


// Recomputes the child boxes of two-child node `idx` and its subtree from the
// instance boxes (leaf references carry instance indices); returns the node's box.
// Children with an inverted (empty) box are left as they are.
void refitTree(std::vector<BvhNode>& nodes, uint32_t idx, const std::vector<float4>& lo,
               const std::vector<float4>& hi, float4* outLo = nullptr, float4* outHi = nullptr)
{
    BvhNode& nd = nodes[idx];
    float4* clo[2] = {&nd.a, &nd.c};
    float4* chi[2] = {&nd.b, &nd.d};
    int32_t ref[2];
    std::memcpy(&ref[0], &nd.a.w, 4);
    std::memcpy(&ref[1], &nd.b.w, 4);
    float4 ulo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), uhi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    for (int k = 0; k < 2; k++) {
        if (clo[k]->x > chi[k]->x) continue;  // empty slot (a single-instance tree's second child)
        float4 l, h;
        if ((uint32_t)ref[k] & kLeafBit) {
            l = lo[leafFirst(ref[k])];
            h = hi[leafFirst(ref[k])];
        } else {
            refitTree(nodes, (uint32_t)ref[k], lo, hi, &l, &h);
        }
        clo[k]->x = l.x; clo[k]->y = l.y; clo[k]->z = l.z;
        chi[k]->x = h.x; chi[k]->y = h.y; chi[k]->z = h.z;
        ulo = make_float4(std::fmin(ulo.x, l.x), std::fmin(ulo.y, l.y), std::fmin(ulo.z, l.z), 0.0f);
        uhi = make_float4(std::fmax(uhi.x, h.x), std::fmax(uhi.y, h.y), std::fmax(uhi.z, h.z), 0.0f);
    }
    if (outLo) *outLo = ulo;
    if (outHi) *outHi = uhi;
}

// Validated device records of render requests (targets, TLASes, sample ranges, lights).
int buildRequests(aph_ctx* c, const aph_request* reqs, uint32_t nreq, std::vector<WavefrontRequest>& out)
{
    out.clear();
    out.reserve(nreq);
    for (uint32_t i = 0; i < nreq; i++) {
        const aph_request& r = reqs[i];
        Target* tg = lookup(c->targets, r.target);
        Tlas* tl = lookup(c->tlases, r.tlas);
        if (!tg) return fail(c, APH_E_INVALID, "request %u: unknown target", i);
        if (!tl) return fail(c, APH_E_INVALID, "request %u: unknown tlas", i);
        int rc = ensureTransmit(c, tl);
        if (rc) return rc;
        if ((uint64_t)r.sample_start + r.sample_count > 0xFFFFFFFFull)
            return fail(c, APH_E_INVALID, "request %u: sample range overflows", i);
        if (r.frame.light_count < 0 || r.frame.light_count > 4)
            return fail(c, APH_E_INVALID, "request %u: light_count must be in [0,4]", i);
        WavefrontRequest w;
        w.rs = renderScene(c, tl);
        w.frame = clampFrame(r.frame);
        w.width = tg->w;
        w.height = tg->h;
        w.sampleStart = r.sample_start;
        w.sampleCount = r.sample_count;
        w.seedOffset = r.seed_offset;
        w.accum = tg->accum.as<float4>();
        out.push_back(w);
    }
    return APH_OK;
}

}  // namespace

extern "C" {

int aph_version(void) { return APH_VERSION; }

int aph_create(int device, aph_ctx_t* out)
{
    if (!out) return APH_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return APH_E_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return APH_E_DEVICE;
    std::unique_ptr<aph_ctx> c(new aph_ctx());
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return APH_E_DEVICE;
    if (c->counters.alloc(sizeof(DevCounters)) != hipSuccess) return APH_E_MEMORY;
    if (hipMemset(c->counters.p, 0, sizeof(DevCounters)) != hipSuccess) return APH_E_DEVICE;
    setDefaultEnv(c.get());
    aph_material dm;
    aph_material_defaults(&dm);
    DevMaterial d = toDev(dm);
    if (c->mats.alloc(sizeof(DevMaterial)) != hipSuccess) return APH_E_MEMORY;
    if (hipMemcpy(c->mats.p, &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess) return APH_E_DEVICE;
    c->nmat = 1;
    c->matsHost.assign(1, dm);
    if (uploadMaterialTextures(c.get()) != APH_OK) return APH_E_DEVICE;
    // the allocation-free render's memory on this device (a failure only disables that entry point)
    c->wfStaticStatus = (int)wavefrontScratchState(c->wfStatic);
    *out = c.release();
    return APH_OK;
}

void aph_destroy(aph_ctx_t c)
{
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    resolveTimings(c);
    c->meshes.clear();
    c->tlases.clear();
    c->targets.clear();
    wavefrontRelease(c->wf, c->stream);
    hipStreamDestroy(c->stream);
    delete c;
}

const char* aph_last_error(aph_ctx_t c) { return c ? c->err.c_str() : "null context"; }

int aph_get_stream(aph_ctx_t c, void** stream)
{
    if (!c || !stream) return APH_E_INVALID;
    *stream = (void*)c->stream;
    return APH_OK;
}

int aph_inverse3x4(const float mf[12], float out[12])
{
    if (!mf || !out) return APH_E_INVALID;
    double m[12];
    for (int i = 0; i < 12; i++) m[i] = (double)mf[i];
    double a = m[0], b = m[1], cc = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    double A = e * k - f * h, B = f * g - d * k, C = d * h - e * g;
    double det = a * A + b * B + cc * C;
    if (det == 0.0) {
        std::memset(out, 0, 12 * sizeof(float));
        return APH_E_INVALID;
    }
    double r[9] = {A / det,           (cc * h - b * k) / det, (b * f - cc * e) / det,
                   B / det,           (a * k - cc * g) / det, (cc * d - a * f) / det,
                   C / det,           (b * g - a * h) / det,  (a * e - b * d) / det};
    double tx = m[3], ty = m[7], tz = m[11];
    for (int i = 0; i < 3; i++) {
        out[4 * i + 0] = (float)r[3 * i + 0];
        out[4 * i + 1] = (float)r[3 * i + 1];
        out[4 * i + 2] = (float)r[3 * i + 2];
        out[4 * i + 3] = (float)(-((r[3 * i + 0] * tx + r[3 * i + 1] * ty) + r[3 * i + 2] * tz));
    }
    return APH_OK;
}

int aph_mesh_upload(aph_ctx_t c, const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* nrm,
                    const float* uv, uint64_t* mesh)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!mesh) return fail(c, APH_E_INVALID, "mesh output pointer is null");
    *mesh = 0;
    if (!pos || nv == 0) return fail(c, APH_E_INVALID, "positions are required (GeometryBase.cpp:84-87)");
    if (nt > (kLeafFirstMask + 1u)) return fail(c, APH_E_INVALID, "too many triangles (%u)", nt);
    std::vector<uint32_t> seq;
    if (!idx) {
        // No indices: sequential triangles (HGIGeometry.cpp:37-47).
        if ((uint64_t)nt * 3u > nv) return fail(c, APH_E_INVALID, "non-indexed mesh needs 3*nt <= nv");
        seq.resize((size_t)nt * 3);
        for (size_t i = 0; i < seq.size(); i++) seq[i] = (uint32_t)i;
        idx = seq.data();
    }
    for (size_t i = 0; i < (size_t)nt * 3; i++)
        if (idx[i] >= nv) return fail(c, APH_E_INVALID, "index %zu = %u out of range (nv = %u)", i, idx[i], nv);
    std::unique_ptr<Mesh> m(new Mesh());
    m->nv = nv;
    m->nt = nt;
    m->hasNrm = nrm != nullptr;
    HIPCALL(c, m->pos.alloc((size_t)nv * 12), "alloc positions");
    HIPCALL(c, hipMemcpyAsync(m->pos.p, pos, (size_t)nv * 12, hipMemcpyHostToDevice, c->stream), "copy positions");
    HIPCALL(c, m->idx.alloc((size_t)nt * 12), "alloc indices");
    if (nt) HIPCALL(c, hipMemcpyAsync(m->idx.p, idx, (size_t)nt * 12, hipMemcpyHostToDevice, c->stream), "copy indices");
    if (nrm) {
        HIPCALL(c, m->nrm.alloc((size_t)nv * 12), "alloc normals");
        HIPCALL(c, hipMemcpyAsync(m->nrm.p, nrm, (size_t)nv * 12, hipMemcpyHostToDevice, c->stream), "copy normals");
    }
    // Texture coordinates; without them u = v = i / vertex count (GeometryBase.cpp:114-124).
    std::vector<float> genUv;
    if (!uv) {
        genUv.resize((size_t)nv * 2);
        for (uint32_t i = 0; i < nv; i++) genUv[2 * (size_t)i] = genUv[2 * (size_t)i + 1] = (float)i / (float)nv;
        uv = genUv.data();
    }
    HIPCALL(c, m->uv.alloc((size_t)nv * 8), "alloc uv");
    HIPCALL(c, hipMemcpyAsync(m->uv.p, uv, (size_t)nv * 8, hipMemcpyHostToDevice, c->stream), "copy uv");
    HIPCALL(c, hipStreamSynchronize(c->stream), "mesh upload");
    uint64_t h = c->nextHandle++;
    c->meshes[h] = std::move(m);
    *mesh = h;
    return APH_OK;
}

int aph_blas_build(aph_ctx_t c, uint64_t mesh, int builder)
{
    int rc = enter(c);
    if (rc) return rc;
    Mesh* m = lookup(c->meshes, mesh);
    if (!m) return fail(c, APH_E_INVALID, "unknown mesh handle %llu", (unsigned long long)mesh);
    if (builder != APH_BUILDER_LBVH && builder != APH_BUILDER_SAH)
        return fail(c, APH_E_INVALID, "unknown builder %d", builder);
    uint32_t n = m->nt;
    TimedScope ts(c, "blas_build");
    HIPCALL(c, m->nodes.alloc((size_t)(n > 1 ? n - 1 : 1) * sizeof(BvhNode)), "alloc BLAS nodes");
    HIPCALL(c, m->nodes4.alloc((size_t)(n > 1 ? n - 1 : 1) * sizeof(Bvh4Node)), "alloc BLAS nodes");
    HIPCALL(c, m->tris.alloc((size_t)(n ? n : 1) * sizeof(Tri)), "alloc BLAS triangles");
    if (n == 0) {
        m->built = true;
        m->info = BuildInfo{};
        return APH_OK;
    }
    DevBuf lo, hi, sorted;
    HIPCALL(c, lo.alloc((size_t)n * 16), "alloc boxes");
    HIPCALL(c, hi.alloc((size_t)n * 16), "alloc boxes");
    HIPCALL(c, sorted.alloc((size_t)n * 4), "alloc order");
    HIPCALL(c, triBoxes(m->pos.as<float>(), m->idx.as<uint32_t>(), n, c->stream, lo.as<float4>(), hi.as<float4>()),
            "triangle boxes");
    BuildInfo info{};
    if (builder == APH_BUILDER_SAH)
        HIPCALL(c, sahBuild(lo.as<float4>(), hi.as<float4>(), n, kSahMaxLeaf, false, nullptr, c->stream,
                            m->nodes.as<BvhNode>(), sorted.as<uint32_t>(), &info),
                "BLAS build (SAH)");
    else
        HIPCALL(c, lbvhBuild(lo.as<float4>(), hi.as<float4>(), n, false, nullptr, c->stream, m->nodes.as<BvhNode>(),
                             sorted.as<uint32_t>(), &info),
                "BLAS build (LBVH)");
    HIPCALL(c, reorderTris(m->pos.as<float>(), m->idx.as<uint32_t>(), sorted.as<uint32_t>(), n, c->stream,
                           m->tris.as<Tri>()),
            "reorder triangles");
    HIPCALL(c, collapse4(m->nodes.as<BvhNode>(), info.nodes, c->stream, m->nodes4.as<Bvh4Node>()), "BLAS nodes");
    HIPCALL(c, hipStreamSynchronize(c->stream), "BLAS build");
    m->info = info;
    m->built = true;
    return APH_OK;
}

int aph_blas_export(aph_ctx_t c, uint64_t mesh, void* nodes, uint64_t nodesBytes, void* tris, uint64_t trisBytes,
                    uint32_t* nodeCount, uint32_t* depth)
{
    int rc = enter(c);
    if (rc) return rc;
    Mesh* m = lookup(c->meshes, mesh);
    if (!m) return fail(c, APH_E_INVALID, "unknown mesh handle %llu", (unsigned long long)mesh);
    if (!m->built) return fail(c, APH_E_STATE, "BLAS of mesh %llu not built", (unsigned long long)mesh);
    if (nodeCount) *nodeCount = m->info.nodes;
    if (depth) *depth = m->info.depth;
    size_t nb = (size_t)m->info.nodes * sizeof(BvhNode), tb = (size_t)m->nt * sizeof(Tri);
    if (nodes) {
        if (nodesBytes < nb) return fail(c, APH_E_INVALID, "node buffer holds %llu of %zu bytes",
                                         (unsigned long long)nodesBytes, nb);
        HIPCALL(c, hipMemcpy(nodes, m->nodes.p, nb, hipMemcpyDeviceToHost), "export nodes");
    }
    if (tris) {
        if (trisBytes < tb) return fail(c, APH_E_INVALID, "triangle buffer holds %llu of %zu bytes",
                                        (unsigned long long)trisBytes, tb);
        HIPCALL(c, hipMemcpy(tris, m->tris.p, tb, hipMemcpyDeviceToHost), "export triangles");
    }
    return APH_OK;
}

// Fills `t` (a fresh Tlas) from the instances: the traversal and shading records and
// the tree over the instance world boxes - a new binned-SAH build, or, with refitFrom,
// the tree of that TLAS with every box recomputed bottom-up for the new instances
// (its topology kept: boxes only cull, so any grouping gives the same hits).
static int buildTlas(aph_ctx_t c, const aph_instance* inst, uint32_t n, Tlas* t, const Tlas* refitFrom)
{
    if (n && !inst) return fail(c, APH_E_INVALID, "instances pointer is null");
    t->n = n;
    std::vector<InstTrav> trav(n ? n : 1);
    std::vector<InstShade> shade(n ? n : 1);
    std::vector<float4> blo, bhi;
    std::vector<uint32_t> remap;
    double extent = 0.0, kappa = 1.0;
    uint32_t maxBlasDepth = 0;
    std::set<const Mesh*> usedMeshes;
    for (uint32_t i = 0; i < n; i++) {
        const aph_instance& in = inst[i];
        Mesh* m = lookup(c->meshes, (uint64_t)in.mesh);
        if (m) usedMeshes.insert(m);
        if (!m) return fail(c, APH_E_INVALID, "instance %u: unknown mesh handle %u", i, in.mesh);
        if (!m->built) return fail(c, APH_E_STATE, "instance %u: BLAS of mesh %u not built", i, in.mesh);
        float w2o[12];
        bool zero = true;
        for (int k = 0; k < 12; k++) zero = zero && in.w2o[k] == 0.0f;
        if (zero) {
            if (aph_inverse3x4(in.o2w, w2o) != APH_OK) return fail(c, APH_E_INVALID, "instance %u: singular transform", i);
        } else {
            std::memcpy(w2o, in.w2o, sizeof(w2o));
        }
        InstTrav& tv = trav[i];
        std::memcpy(tv.w2o, w2o, sizeof(w2o));
        tv.nodes = m->nodes4.as<Bvh4Node>();
        tv.tris = m->tris.as<Tri>();
        InstShade& sh = shade[i];
        std::memcpy(sh.o2w, in.o2w, sizeof(sh.o2w));
        sh.pos = m->pos.as<float>();
        sh.nrm = m->hasNrm ? m->nrm.as<float>() : nullptr;
        sh.uv = m->uv.as<float>();
        sh.idx = m->idx.as<uint32_t>();
        sh.material = in.material;
        sh.flags = in.flags;
        sh.layerFirst = 0;
        sh.layerCount = 0;
        if (m->nt == 0) continue;
        // condition number bound (infinity norm) of the linear part
        double na = 0.0, nb = 0.0;
        for (int r = 0; r < 3; r++) {
            na = std::fmax(na, std::fabs(in.o2w[4 * r]) + std::fabs(in.o2w[4 * r + 1]) + std::fabs(in.o2w[4 * r + 2]));
            nb = std::fmax(nb, std::fabs(w2o[4 * r]) + std::fabs(w2o[4 * r + 1]) + std::fabs(w2o[4 * r + 2]));
        }
        kappa = std::fmax(kappa, na * nb);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int cn = 0; cn < 8; cn++) {
            double p[3] = {(cn & 1) ? m->info.rootHi[0] : m->info.rootLo[0], (cn & 2) ? m->info.rootHi[1] : m->info.rootLo[1],
                           (cn & 4) ? m->info.rootHi[2] : m->info.rootLo[2]};
            for (int k = 0; k < 3; k++) {
                const float* r = in.o2w + 4 * k;
                double w = (double)r[0] * p[0] + (double)r[1] * p[1] + (double)r[2] * p[2] + (double)r[3];
                lo[k] = std::fmin(lo[k], std::nextafter((float)w, -INFINITY));
                hi[k] = std::fmax(hi[k], std::nextafter((float)w, INFINITY));
            }
        }
        for (int k = 0; k < 3; k++) extent = std::fmax(extent, std::fmax(std::fabs(lo[k]), std::fabs(hi[k])));
        blo.push_back(make_float4(lo[0], lo[1], lo[2], 0.0f));
        bhi.push_back(make_float4(hi[0], hi[1], hi[2], 0.0f));
        remap.push_back(i);
        maxBlasDepth = std::max(maxBlasDepth, m->info.depth);
    }
    uint32_t nv = (uint32_t)blo.size();
    if (refitFrom && (refitFrom->n != n || refitFrom->remap != remap))
        return fail(c, APH_E_INVALID, "refit needs the same instances with geometry as the TLAS was built with "
                    "(%u of %u before, %u of %u now); rebuild instead", refitFrom->nvalid, refitFrom->n, nv, n);
    t->nvalid = nv;
    t->remap = remap;
    HIPCALL(c, t->trav.alloc(sizeof(InstTrav) * (n ? n : 1)), "alloc instances");
    t->shadeHost = shade;
    HIPCALL(c, t->layerMats.alloc(16), "alloc layers");
    HIPCALL(c, t->shade.alloc(sizeof(InstShade) * (n ? n : 1)), "alloc instances");
    HIPCALL(c, hipMemcpyAsync(t->trav.p, trav.data(), sizeof(InstTrav) * (n ? n : 1), hipMemcpyHostToDevice, c->stream),
            "copy instances");
    HIPCALL(c, hipMemcpyAsync(t->shade.p, shade.data(), sizeof(InstShade) * (n ? n : 1), hipMemcpyHostToDevice,
                              c->stream),
            "copy instances");
    HIPCALL(c, t->nodes.alloc(sizeof(BvhNode) * (nv > 1 ? nv - 1 : 1)), "alloc TLAS nodes");
    HIPCALL(c, t->nodes4.alloc(sizeof(Bvh4Node) * (nv > 1 ? nv - 1 : 1)), "alloc TLAS nodes");
    if (nv) {
        TimedScope ts(c, refitFrom ? "tlas_refit" : "tlas_build");
        BuildInfo info{};
        if (refitFrom) {
            // the old tree, every child box recomputed from its subtree's new instance boxes
            const uint32_t nn = refitFrom->nodeCount;
            std::vector<BvhNode> nodes(nn);
            HIPCALL(c, hipMemcpyAsync(nodes.data(), refitFrom->nodes.p, sizeof(BvhNode) * nn, hipMemcpyDeviceToHost,
                                      c->stream),
                    "refit download");
            HIPCALL(c, hipStreamSynchronize(c->stream), "refit download");
            std::vector<float4> boxLo(n), boxHi(n);
            for (uint32_t p = 0; p < nv; p++) {
                boxLo[remap[p]] = blo[p];
                boxHi[remap[p]] = bhi[p];
            }
            refitTree(nodes, 0, boxLo, boxHi);
            HIPCALL(c, hipMemcpyAsync(t->nodes.p, nodes.data(), sizeof(BvhNode) * nn, hipMemcpyHostToDevice, c->stream),
                    "refit upload");
            HIPCALL(c, hipStreamSynchronize(c->stream), "refit upload");
            info.nodes = nn;
            info.depth = refitFrom->depth;
        } else {
            DevBuf lo, hi, sorted, rm;
            HIPCALL(c, lo.alloc((size_t)nv * 16), "alloc");
            HIPCALL(c, hi.alloc((size_t)nv * 16), "alloc");
            HIPCALL(c, sorted.alloc((size_t)nv * 4), "alloc");
            HIPCALL(c, rm.alloc((size_t)nv * 4), "alloc");
            HIPCALL(c, hipMemcpyAsync(lo.p, blo.data(), (size_t)nv * 16, hipMemcpyHostToDevice, c->stream), "copy");
            HIPCALL(c, hipMemcpyAsync(hi.p, bhi.data(), (size_t)nv * 16, hipMemcpyHostToDevice, c->stream), "copy");
            HIPCALL(c, hipMemcpyAsync(rm.p, remap.data(), (size_t)nv * 4, hipMemcpyHostToDevice, c->stream), "copy");
            // binned SAH over the instance world boxes, one instance per leaf
            HIPCALL(c, sahBuild(lo.as<float4>(), hi.as<float4>(), nv, 1, true, rm.as<uint32_t>(), c->stream,
                                t->nodes.as<BvhNode>(), sorted.as<uint32_t>(), &info),
                    "TLAS build");
            HIPCALL(c, hipStreamSynchronize(c->stream), "TLAS build");  // the scratch buffers are freed on return
        }
        HIPCALL(c, collapse4(t->nodes.as<BvhNode>(), info.nodes, c->stream, t->nodes4.as<Bvh4Node>()), "TLAS nodes");
        HIPCALL(c, hipStreamSynchronize(c->stream), "TLAS build");
        t->depth = info.depth;
        t->nodeCount = info.nodes;
        // Stack bound: a four-child level (two levels of the two-child tree) leaves
        // at most 3 siblings pending; entering an instance adds the pending TLAS
        // entry and the exit marker.
        const uint32_t need = 3 * ((info.depth + 1) / 2) + 2 + 3 * ((maxBlasDepth + 1) / 2);
        if (need > (uint32_t)kStackCap)
            return fail(c, APH_E_OVERFLOW,
                        "BVH too deep for the traversal stack (TLAS depth %u, BLAS depth %u: %u > %d entries)",
                        info.depth, maxBlasDepth, need, kStackCap);
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "TLAS upload");
    t->sc.tlas = t->nodes4.as<Bvh4Node>();
    t->sc.inst = t->trav.as<InstTrav>();
    t->sc.extent = (float)(3.0 * extent);
    t->sc.padScale = (float)(1e-5 * kappa);
    t->sc.ninst = nv;
    t->bytes = t->nodes.bytes + t->nodes4.bytes + t->trav.bytes + t->shade.bytes;
    for (const Mesh* m : usedMeshes)
        t->bytes += m->pos.bytes + m->nrm.bytes + m->uv.bytes + m->idx.bytes + m->nodes.bytes + m->nodes4.bytes +
                    m->tris.bytes;
    return APH_OK;
}

int aph_tlas_build(aph_ctx_t c, const aph_instance* inst, uint32_t n, uint64_t* tlas)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!tlas) return fail(c, APH_E_INVALID, "tlas output pointer is null");
    *tlas = 0;
    std::unique_ptr<Tlas> t(new Tlas());
    if ((rc = buildTlas(c, inst, n, t.get(), nullptr))) return rc;
    uint64_t h = c->nextHandle++;
    c->tlases[h] = std::move(t);
    *tlas = h;
    return APH_OK;
}

int aph_tlas_update(aph_ctx_t c, uint64_t tlas, const aph_instance* inst, uint32_t n, int mode)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* old = lookup(c->tlases, tlas);
    if (!old) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    if (mode != APH_TLAS_REBUILD && mode != APH_TLAS_REFIT) return fail(c, APH_E_INVALID, "unknown update mode %d", mode);
    Tlas fresh;
    if ((rc = buildTlas(c, inst, n, &fresh, mode == APH_TLAS_REFIT ? old : nullptr))) return rc;
    // The handle keeps its environment and ground plane, and its layers while the instance
    // count is unchanged.
    fresh.environment = old->environment;
    fresh.ground = old->ground;
    if (n == old->n) {
        fresh.layerMats = std::move(old->layerMats);
        fresh.layerUV = std::move(old->layerUV);
        for (uint32_t k = 0; k < n && k < old->shadeHost.size(); k++) {
            fresh.shadeHost[k].layerFirst = old->shadeHost[k].layerFirst;
            fresh.shadeHost[k].layerCount = old->shadeHost[k].layerCount;
        }
        HIPCALL(c, hipMemcpy(fresh.shade.p, fresh.shadeHost.data(), sizeof(InstShade) * fresh.shadeHost.size(),
                             hipMemcpyHostToDevice),
                "instance layers");
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "TLAS update");  // renders in flight read the old records
    *old = std::move(fresh);  // same object: servers and pending jobs keep their pointer
    return APH_OK;
}

int aph_tlas_set_layers(aph_ctx_t c, uint64_t tlas, const uint32_t* offsets, const uint32_t* materials)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* t = lookup(c->tlases, tlas);
    if (!t) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    if (!offsets) return fail(c, APH_E_INVALID, "null layer offsets");
    uint32_t n = t->n;
    if (offsets[0] != 0) return fail(c, APH_E_INVALID, "layer offsets must start at 0");
    for (uint32_t i = 0; i < n; i++)
        if (offsets[i + 1] < offsets[i]) return fail(c, APH_E_INVALID, "layer offsets decrease at instance %u", i);
    uint32_t total = offsets[n];
    if (total && !materials) return fail(c, APH_E_INVALID, "null layer materials");
    if (total > (1u << 24)) return fail(c, APH_E_INVALID, "too many layers");
    // Indices are resolved at render time against the material table then current;
    // one beyond it reads material 0, like an instance's own material index.
    HIPCALL(c, hipStreamSynchronize(c->stream), "layers");
    t->layerUV.reset();  // new layers: every layer reads the instance's texture coordinates again
    HIPCALL(c, t->layerMats.alloc((size_t)(total ? total : 1) * 4), "alloc layers");
    if (total) HIPCALL(c, hipMemcpy(t->layerMats.p, materials, (size_t)total * 4, hipMemcpyHostToDevice), "layers");
    // Instance k of the TLAS is the k-th instance given to aph_tlas_build.
    for (size_t k = 0; k < t->shadeHost.size() && k < n; k++) {
        t->shadeHost[k].layerFirst = offsets[k];
        t->shadeHost[k].layerCount = offsets[k + 1] - offsets[k];
    }
    HIPCALL(c, hipMemcpy(t->shade.p, t->shadeHost.data(), sizeof(InstShade) * t->shadeHost.size(),
                         hipMemcpyHostToDevice),
            "layers");
    return APH_OK;
}

int aph_tlas_set_layer_geometry(aph_ctx_t c, uint64_t tlas, const uint64_t* meshes)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* t = lookup(c->tlases, tlas);
    if (!t) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    uint32_t total = 0;
    for (const InstShade& sh : t->shadeHost) total = std::max(total, sh.layerFirst + sh.layerCount);
    if (total == 0) return fail(c, APH_E_STATE, "TLAS %llu has no material layers", (unsigned long long)tlas);
    if (!meshes) return fail(c, APH_E_INVALID, "null layer geometry");
    // vertex count of each instance's mesh, found by its texture coordinate array
    std::map<const float*, uint32_t> nvOf;
    for (const auto& kv : c->meshes) nvOf[kv.second->uv.as<float>()] = kv.second->nv;
    std::vector<const float*> uv(total, nullptr);
    for (size_t k = 0; k < t->shadeHost.size() && k < t->n; k++) {
        const InstShade& sh = t->shadeHost[k];
        for (uint32_t j = sh.layerFirst; j < sh.layerFirst + sh.layerCount; j++) {
            if (!meshes[j]) continue;  // the instance's own texture coordinates
            Mesh* m = lookup(c->meshes, meshes[j]);
            if (!m) return fail(c, APH_E_INVALID, "layer %u: unknown mesh handle %llu", j, (unsigned long long)meshes[j]);
            // PTScene.cpp:1102-1103: a layer's geometry matches its instance's vertex count
            auto it = nvOf.find(sh.uv);
            if (it == nvOf.end() || it->second != m->nv)
                return fail(c, APH_E_INVALID, "layer %u: mesh %llu has %u vertices, its instance's mesh %u", j,
                            (unsigned long long)meshes[j], m->nv, it == nvOf.end() ? 0u : it->second);
            uv[j] = m->uv.as<float>();
        }
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "layer geometry");  // renders in flight read the old table
    HIPCALL(c, t->layerUV.alloc(sizeof(const float*) * total), "alloc layer geometry");
    HIPCALL(c, hipMemcpy(t->layerUV.p, uv.data(), sizeof(const float*) * total, hipMemcpyHostToDevice), "layer geometry");
    return APH_OK;
}

void aph_material_defaults(aph_material* m)
{
    // Standard Surface defaults, MaterialBase.cpp:109-167.
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->base = 0.8f;
    for (int i = 0; i < 3; i++) {
        m->base_color[i] = 1.0f;
        m->specular_color[i] = 1.0f;
        m->transmission_color[i] = 1.0f;
        m->subsurface_color[i] = 1.0f;
        m->subsurface_radius[i] = 1.0f;
        m->sheen_color[i] = 1.0f;
        m->coat_color[i] = 1.0f;
        m->emission_color[i] = 1.0f;
        m->opacity[i] = 1.0f;
    }
    m->specular = 1.0f;
    m->specular_roughness = 0.2f;
    m->specular_ior = 1.5f;
    m->subsurface_scale = 1.0f;
    m->sheen_roughness = 0.3f;
    m->coat_roughness = 0.1f;
    m->coat_ior = 1.5f;
}

namespace {

// Word offsets (from offset 224) of each texture slot's flag, offset, scale,
// pivot and rotation in MaterialConstants (DefaultMaterialUniformBuffer.slang:43-71).
const int kTexWord[kTexSlots][5] = {
    {0, 1, 4, 6, 8}, {9, 10, 12, 14, 16}, {17, 18, 20, 22, 24}, {25, 26, 28, 30, 32}, {33, 34, 36, 38, 40}};

float blockFloat(const aph_material& m, int w)
{
    float f;
    std::memcpy(&f, &m.texture_block[w], 4);
    return f;
}

// Device texture table and per-material slots from the current materials and bindings.
int uploadMaterialTextures(aph_ctx* c)
{
    const uint32_t n = c->nmat;
    std::vector<DevMatTex> mt(n ? n : 1);
    std::vector<DevTexture> tex;
    c->texturesMissing.clear();
    for (uint32_t i = 0; i < n; i++) {
        const aph_material& m = c->matsHost[i];
        DevMatTex& d = mt[i];
        std::memset(&d, 0, sizeof(d));
        for (int k = 0; k < kTexSlots; k++) {
            const int* w = kTexWord[k];
            d.tex[k] = -1;
            d.xf[k][0] = blockFloat(m, w[1]);
            d.xf[k][1] = blockFloat(m, w[1] + 1);
            d.xf[k][2] = blockFloat(m, w[2]);
            d.xf[k][3] = blockFloat(m, w[2] + 1);
            d.xf[k][4] = blockFloat(m, w[3]);
            d.xf[k][5] = blockFloat(m, w[3] + 1);
            d.xf[k][6] = blockFloat(m, w[4]);
            if (!m.texture_block[w[0]]) continue;
            const aph_texture_binding* b = c->texBindings.empty() ? nullptr : &c->texBindings[(size_t)i * kTexSlots + k];
            Image* img = b && b->image ? lookup(c->images, b->image) : nullptr;
            if (!img) {
                if (k != kTexEmissionColor) c->texturesMissing.push_back(i * kTexSlots + k);
                continue;
            }
            d.tex[k] = (int32_t)tex.size();
            tex.push_back(DevTexture{img->texels.as<float4>(), img->w, img->h, b->address_u, b->address_v});
        }
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "material textures");
    HIPCALL(c, c->matTex.alloc(sizeof(DevMatTex) * mt.size()), "alloc material textures");
    HIPCALL(c, hipMemcpy(c->matTex.p, mt.data(), sizeof(DevMatTex) * mt.size(), hipMemcpyHostToDevice),
            "copy material textures");
    HIPCALL(c, c->textures.alloc(sizeof(DevTexture) * (tex.empty() ? 1 : tex.size())), "alloc texture table");
    if (!tex.empty())
        HIPCALL(c, hipMemcpy(c->textures.p, tex.data(), sizeof(DevTexture) * tex.size(), hipMemcpyHostToDevice),
                "copy texture table");
    return APH_OK;
}

}  // namespace

int aph_materials_upload(aph_ctx_t c, const aph_material* mats, uint32_t n)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!mats || n == 0) return fail(c, APH_E_INVALID, "need at least one material");
    std::vector<DevMaterial> d(n);
    for (uint32_t i = 0; i < n; i++) {
        d[i] = toDev(mats[i]);
        for (int k = 0; k < kTexSlots; k++)
            if (mats[i].texture_block[kTexWord[k][0]]) d[i].texMask |= 1 << k;
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "materials");
    HIPCALL(c, c->mats.alloc(sizeof(DevMaterial) * n), "alloc materials");
    HIPCALL(c, hipMemcpy(c->mats.p, d.data(), sizeof(DevMaterial) * n, hipMemcpyHostToDevice), "copy materials");
    c->nmat = n;
    c->matsHost.assign(mats, mats + n);
    c->matVersion++;
    c->texBindings.clear();
    return uploadMaterialTextures(c);
}

int aph_materials_update(aph_ctx_t c, uint32_t first, const aph_material* mats, uint32_t n)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!mats || n == 0) return fail(c, APH_E_INVALID, "need at least one material");
    if ((uint64_t)first + n > c->nmat)
        return fail(c, APH_E_INVALID, "materials %u..%u beyond the table of %u", first, first + n - 1, c->nmat);
    std::vector<DevMaterial> d(n);
    for (uint32_t i = 0; i < n; i++) {
        d[i] = toDev(mats[i]);
        for (int k = 0; k < kTexSlots; k++)
            if (mats[i].texture_block[kTexWord[k][0]]) d[i].texMask |= 1 << k;
    }
    HIPCALL(c, hipStreamSynchronize(c->stream), "materials");  // renders in flight read the table
    HIPCALL(c, hipMemcpy(c->mats.as<DevMaterial>() + first, d.data(), sizeof(DevMaterial) * n, hipMemcpyHostToDevice),
            "patch materials");
    std::copy(mats, mats + n, c->matsHost.begin() + first);
    c->matVersion++;
    // texture flags and transforms may have changed; the bindings stay
    return uploadMaterialTextures(c);
}

int aph_materials_bind_textures(aph_ctx_t c, const aph_texture_binding* bindings, uint32_t n)
{
    int rc = enter(c);
    if (rc) return rc;
    if (n != c->nmat) return fail(c, APH_E_INVALID, "%u bindings' materials for a table of %u", n, c->nmat);
    if (!bindings && n) return fail(c, APH_E_INVALID, "null bindings");
    for (uint32_t i = 0; i < n * (uint32_t)kTexSlots; i++) {
        const aph_texture_binding& b = bindings[i];
        if (b.image && !lookup(c->images, b.image))
            return fail(c, APH_E_INVALID, "material %u slot %u: unknown image %llu", i / kTexSlots, i % kTexSlots,
                        (unsigned long long)b.image);
        if (b.address_u < APH_ADDRESS_WRAP || b.address_u > APH_ADDRESS_BORDER || b.address_v < APH_ADDRESS_WRAP ||
            b.address_v > APH_ADDRESS_BORDER)
            return fail(c, APH_E_INVALID, "material %u slot %u: unknown address mode", i / kTexSlots,
                        i % kTexSlots);
    }
    c->texBindings.assign(bindings, bindings + (size_t)n * kTexSlots);
    return uploadMaterialTextures(c);
}

int aph_image_create(aph_ctx_t c, const aph_image_desc* desc, uint64_t* image)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!desc || !image || !desc->pixels) return fail(c, APH_E_INVALID, "null image description or output");
    *image = 0;
    const uint32_t w = desc->width, h = desc->height;
    if (w == 0 || h == 0 || w > 32768 || h > 32768) return fail(c, APH_E_INVALID, "bad image size %ux%u", w, h);
    if (desc->format < APH_IMAGE_RGBA8 || desc->format > APH_IMAGE_RGB32F)
        return fail(c, APH_E_INVALID, "unknown image format %d", desc->format);
    const size_t n = (size_t)w * h;
    std::vector<float4> px(n);
    if (desc->format == APH_IMAGE_RGBA8) {
        // HgiFormatUNorm8Vec4(srgb): unorm, and with linearize the sRGB transfer
        // function undone per channel (alpha stays linear) before filtering.
        const uint8_t* p = (const uint8_t*)desc->pixels;
        float lut[256], lin[256];
        for (int v = 0; v < 256; v++) {
            lut[v] = (float)v / 255.0f;
            lin[v] = srgbToLinear(lut[v]);
        }
        const float* rgb = desc->linearize ? lin : lut;
        for (size_t i = 0; i < n; i++)
            px[i] = make_float4(rgb[p[4 * i]], rgb[p[4 * i + 1]], rgb[p[4 * i + 2]], lut[p[4 * i + 3]]);
    } else if (desc->format == APH_IMAGE_RGBA32F) {
        std::memcpy(px.data(), desc->pixels, n * 16);
    } else {
        const float* p = (const float*)desc->pixels;
        for (size_t i = 0; i < n; i++) px[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], 1.0f);
    }
    std::unique_ptr<Image> img(new Image());
    img->w = w;
    img->h = h;
    HIPCALL(c, img->texels.alloc(n * 16), "alloc image");
    HIPCALL(c, hipMemcpy(img->texels.p, px.data(), n * 16, hipMemcpyHostToDevice), "copy image");
    uint64_t hd = c->nextHandle++;
    c->images[hd] = std::move(img);
    *image = hd;
    return APH_OK;
}

int aph_image_destroy(aph_ctx_t c, uint64_t image)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!lookup(c->images, image)) return fail(c, APH_E_INVALID, "unknown image %llu", (unsigned long long)image);
    for (const aph_texture_binding& b : c->texBindings)
        if (b.image == image) return fail(c, APH_E_STATE, "image %llu is bound to the material table",
                                          (unsigned long long)image);
    HIPCALL(c, hipStreamSynchronize(c->stream), "image destroy");
    c->images.erase(image);
    return APH_OK;
}

int aph_alias_build(const float* rgb, uint32_t w, uint32_t h, void* entries, float* integral)
{
    if (!rgb || !entries || !integral || w == 0 || h == 0) return APH_E_INVALID;
    aliasBuild(rgb, w, h, (AliasEntry*)entries, integral);
    return APH_OK;
}

// Builds the device copy of an environment (EnvironmentBase::updateGPUStruct,
// EnvironmentBase.cpp:52-71, and the alias map of AliasMap.cpp:31-166) into `out`.
static int buildEnvironment(aph_ctx_t c, const aph_environment* e, Environment& out)
{
    if (!e) return fail(c, APH_E_INVALID, "environment is null");
    if (e->filter != 0 && e->filter != 1) return fail(c, APH_E_INVALID, "filter must be 0 or 1");
    DevEnv& d = out.env;
    std::memset(&d, 0, sizeof(d));
    d.lightTop = f3{e->light_top[0], e->light_top[1], e->light_top[2]};
    d.lightBottom = f3{e->light_bottom[0], e->light_bottom[1], e->light_bottom[2]};
    d.bgTop = f3{e->background_top[0], e->background_top[1], e->background_top[2]};
    d.bgBottom = f3{e->background_bottom[0], e->background_bottom[1], e->background_bottom[2]};
    std::memcpy(d.lightXf, e->light_transform, sizeof(d.lightXf));
    std::memcpy(d.bgXf, e->background_transform, sizeof(d.bgXf));
    if (!inverse4(e->light_transform, d.lightXfInv)) return fail(c, APH_E_INVALID, "light transform is singular");
    d.bgUseScreen = e->background_use_screen;
    d.filter = e->filter;
    if (e->light_rgb) {
        uint32_t w = e->light_width, h = e->light_height;
        if (w == 0 || h == 0) return fail(c, APH_E_INVALID, "light image has zero size");
        size_t n = (size_t)w * h;
        std::vector<float4> tex(n);
        for (size_t i = 0; i < n; i++)
            tex[i] = make_float4(e->light_rgb[3 * i], e->light_rgb[3 * i + 1], e->light_rgb[3 * i + 2], 1.0f);
        std::vector<AliasEntry> alias(n);
        float integral = 0.0f;
        aliasBuild(e->light_rgb, w, h, alias.data(), &integral);
        // the light sampling records: each entry with its alias's pdf and both radiances
        std::vector<EnvSample> samples(n);
        for (size_t i = 0; i < n; i++) {
            const AliasEntry& ae = alias[i];
            const size_t j = ae.alias < n ? ae.alias : i;
            EnvSample& sr = samples[i];
            sr.alias = ae.alias;
            sr.prob = ae.prob;
            sr.pdf = ae.pdf;
            sr.pdfAlias = alias[j].pdf;
            sr.r = tex[i].x; sr.g = tex[i].y; sr.b = tex[i].z;
            sr.ra = tex[j].x; sr.ga = tex[j].y; sr.ba = tex[j].z;
            sr.pad[0] = sr.pad[1] = 0.0f;
        }
        HIPCALL(c, out.light.alloc(n * 16), "alloc light image");
        HIPCALL(c, out.samples.alloc(n * sizeof(EnvSample)), "alloc light sampling records");
        HIPCALL(c, hipMemcpy(out.light.p, tex.data(), n * 16, hipMemcpyHostToDevice), "copy light image");
        HIPCALL(c, hipMemcpy(out.samples.p, samples.data(), n * sizeof(EnvSample), hipMemcpyHostToDevice),
                "copy light sampling records");
        d.hasLightTex = 1;
        d.lightTex = out.light.as<float4>();
        d.samples = out.samples.as<EnvSample>();
        d.lw = w;
        d.lh = h;
        d.lightIntegral = integral;
    }
    if (e->background_rgb) {
        uint32_t w = e->background_width, h = e->background_height;
        if (w == 0 || h == 0) return fail(c, APH_E_INVALID, "background image has zero size");
        size_t n = (size_t)w * h;
        std::vector<float4> tex(n);
        for (size_t i = 0; i < n; i++)
            tex[i] = make_float4(e->background_rgb[3 * i], e->background_rgb[3 * i + 1], e->background_rgb[3 * i + 2],
                                 1.0f);
        HIPCALL(c, out.bg.alloc(n * 16), "alloc background image");
        HIPCALL(c, hipMemcpy(out.bg.p, tex.data(), n * 16, hipMemcpyHostToDevice), "copy background image");
        d.hasBgTex = 1;
        d.bgTex = out.bg.as<float4>();
        d.bw = w;
        d.bh = h;
    }
    return APH_OK;
}

int aph_environment_upload(aph_ctx_t c, const aph_environment* e)
{
    int rc = enter(c);
    if (rc) return rc;
    std::unique_ptr<Environment> env(new Environment());
    if ((rc = buildEnvironment(c, e, *env))) return rc;
    HIPCALL(c, hipStreamSynchronize(c->stream), "environment");  // renders in flight read the old one
    c->env = std::move(env);
    return APH_OK;
}

int aph_environment_create(aph_ctx_t c, const aph_environment* e, uint64_t* env)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!env) return fail(c, APH_E_INVALID, "environment output pointer is null");
    *env = 0;
    std::unique_ptr<Environment> en(new Environment());
    if ((rc = buildEnvironment(c, e, *en))) return rc;
    uint64_t h = c->nextHandle++;
    c->environments[h] = std::move(en);
    *env = h;
    return APH_OK;
}

int aph_environment_destroy(aph_ctx_t c, uint64_t env)
{
    int rc = enter(c);
    if (rc) return rc;
    Environment* en = lookup(c->environments, env);
    if (!en) return fail(c, APH_E_INVALID, "unknown environment handle %llu", (unsigned long long)env);
    for (const auto& kv : c->tlases)
        if (kv.second->environment == en)
            return fail(c, APH_E_STATE, "environment %llu is set on TLAS %llu", (unsigned long long)env,
                        (unsigned long long)kv.first);
    HIPCALL(c, hipStreamSynchronize(c->stream), "environment");
    c->environments.erase(env);
    return APH_OK;
}

int aph_tlas_set_environment(aph_ctx_t c, uint64_t tlas, uint64_t env)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* tl = lookup(c->tlases, tlas);
    if (!tl) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    const Environment* en = nullptr;
    if (env) {
        en = lookup(c->environments, env);
        if (!en) return fail(c, APH_E_INVALID, "unknown environment handle %llu", (unsigned long long)env);
    }
    tl->environment = en;
    return APH_OK;
}

void aph_ground_plane_defaults(aph_ground_plane* gp)
{
    if (!gp) return;
    *gp = aph_ground_plane{};
    gp->enabled = 1;
    gp->normal[1] = 1.0f;
    gp->shadow_opacity = 1.0f;
    gp->reflection_opacity = 0.5f;
    for (int k = 0; k < 3; k++) gp->reflection_color[k] = 0.5f;
    gp->reflection_roughness = 0.1f;
}

int aph_tlas_set_ground_plane(aph_ctx_t c, uint64_t tlas, const aph_ground_plane* gp)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* tl = lookup(c->tlases, tlas);
    if (!tl) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    DevGround g{};
    if (gp && gp->enabled) {
        g.enabled = 1;
        g.position = f3{gp->position[0], gp->position[1], gp->position[2]};
        g.normal = f3{gp->normal[0], gp->normal[1], gp->normal[2]};
        // the plane's frame (PTGroundPlane.cpp buildBasis): from +Y, or +X for a normal near +-Y
        f3 b = std::fabs(g.normal.y) < 0.999f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
        g.tangent = normalize(cross(b, g.normal));
        g.bitangent = cross(g.normal, g.tangent);
        g.shadowOpacity = gp->shadow_opacity;
        g.shadowColor = f3{gp->shadow_color[0], gp->shadow_color[1], gp->shadow_color[2]};
        g.reflectionOpacity = gp->reflection_opacity;
        g.reflectionColor = f3{gp->reflection_color[0], gp->reflection_color[1], gp->reflection_color[2]};
        g.reflectionRoughness = gp->reflection_roughness;
    }
    // renders already enqueued read the TLAS's plane from their own call tables
    tl->ground = g;
    return APH_OK;
}

static int traceCommon(aph_ctx_t c, uint64_t tlas, const aph_ray* rays, uint32_t n, aph_hit* hits, uint8_t* occ,
                       bool any, bool device)
{
    int rc = enter(c);
    if (rc) return rc;
    Tlas* t = lookup(c->tlases, tlas);
    if (!t) return fail(c, APH_E_INVALID, "unknown tlas handle %llu", (unsigned long long)tlas);
    if (n == 0) return APH_OK;
    if (!rays || (any ? !occ : !hits)) return fail(c, APH_E_INVALID, "null ray or result pointer");
    const aph_ray* dr = rays;
    aph_hit* dh = hits;
    uint8_t* doc = occ;
    if (!device) {
        HIPCALL(c, c->scratchA.reserve((size_t)n * sizeof(aph_ray)), "alloc rays");
        HIPCALL(c, c->scratchB.reserve((size_t)n * (any ? 1 : sizeof(aph_hit))), "alloc hits");
        HIPCALL(c, hipMemcpyAsync(c->scratchA.p, rays, (size_t)n * sizeof(aph_ray), hipMemcpyHostToDevice, c->stream),
                "copy rays");
        dr = c->scratchA.as<aph_ray>();
        dh = any ? nullptr : c->scratchB.as<aph_hit>();
        doc = any ? c->scratchB.as<uint8_t>() : nullptr;
    }
    {
        TimedScope ts(c, any ? "trace_any" : "trace_closest");
        HIPCALL(c, launchTrace(t->sc, dr, n, dh, doc, c->counters.as<DevCounters>(), any, c->statsLevel > 0, c->stream),
                "trace launch");
    }
    if (!device) {
        if (any) HIPCALL(c, hipMemcpyAsync(occ, doc, n, hipMemcpyDeviceToHost, c->stream), "copy occlusion");
        else HIPCALL(c, hipMemcpyAsync(hits, dh, (size_t)n * sizeof(aph_hit), hipMemcpyDeviceToHost, c->stream),
                     "copy hits");
        HIPCALL(c, hipStreamSynchronize(c->stream), "trace");
        DevCounters k;
        HIPCALL(c, hipMemcpy(&k, c->counters.p, sizeof(k), hipMemcpyDeviceToHost), "counters");
        if (k.overflow) return fail(c, APH_E_OVERFLOW, "traversal stack overflow on %llu rays", k.overflow);
    }
    return APH_OK;
}

int aph_trace_closest(aph_ctx_t c, uint64_t tlas, const aph_ray* rays, uint32_t n, aph_hit* hits)
{
    return traceCommon(c, tlas, rays, n, hits, nullptr, false, false);
}
int aph_trace_any(aph_ctx_t c, uint64_t tlas, const aph_ray* rays, uint32_t n, uint8_t* occluded)
{
    return traceCommon(c, tlas, rays, n, nullptr, occluded, true, false);
}
int aph_trace_closest_device(aph_ctx_t c, uint64_t tlas, const aph_ray* d_rays, uint32_t n, aph_hit* d_hits)
{
    return traceCommon(c, tlas, d_rays, n, d_hits, nullptr, false, true);
}
int aph_trace_any_device(aph_ctx_t c, uint64_t tlas, const aph_ray* d_rays, uint32_t n, uint8_t* d_occluded)
{
    return traceCommon(c, tlas, d_rays, n, nullptr, d_occluded, true, true);
}

int aph_target_create(aph_ctx_t c, uint32_t width, uint32_t height, uint64_t* target)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!target) return fail(c, APH_E_INVALID, "target output pointer is null");
    *target = 0;
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(c, APH_E_INVALID, "bad target size %ux%u", width, height);
    std::unique_ptr<Target> t(new Target());
    t->w = width;
    t->h = height;
    HIPCALL(c, t->accum.alloc((size_t)width * height * 16), "alloc target");
    HIPCALL(c, hipMemsetAsync(t->accum.p, 0, (size_t)width * height * 16, c->stream), "clear target");
    uint64_t h = c->nextHandle++;
    c->targets[h] = std::move(t);
    *target = h;
    return APH_OK;
}

int aph_target_destroy(aph_ctx_t c, uint64_t target)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!lookup(c->targets, target)) return fail(c, APH_E_INVALID, "unknown target");
    HIPCALL(c, hipStreamSynchronize(c->stream), "target destroy");
    c->targets.erase(target);
    return APH_OK;
}

int aph_target_device_ptr(aph_ctx_t c, uint64_t target, void** ptr)
{
    int rc = enter(c);
    if (rc) return rc;
    Target* t = lookup(c->targets, target);
    if (!t || !ptr) return fail(c, APH_E_INVALID, "unknown target or null pointer");
    *ptr = t->accum.p;
    return APH_OK;
}

int aph_render(aph_ctx_t c, const aph_request* reqs, uint32_t nreq, int mode)
{
    int rc = enter(c);
    if (rc) return rc;
    if (nreq == 0) return APH_OK;
    if ((rc = checkTextures(c))) return rc;
    if (!reqs) return fail(c, APH_E_INVALID, "requests pointer is null");
    if (mode != APH_MODE_MEGAKERNEL && mode != APH_MODE_WAVEFRONT) return fail(c, APH_E_INVALID, "unknown mode %d", mode);
    // Every request is checked before any is launched, so a bad request renders nothing.
    std::vector<WavefrontRequest> wreq;
    if ((rc = buildRequests(c, reqs, nreq, wreq))) return rc;
    if (mode == APH_MODE_MEGAKERNEL) {
        for (const WavefrontRequest& w : wreq) {
            TimedScope ts(c, "render_megakernel");
            HIPCALL(c, launchRenderMega(w.rs, w.frame, w.width, w.height, w.sampleStart, w.sampleCount, w.seedOffset,
                                        w.accum, c->counters.as<DevCounters>(), c->statsLevel > 0, c->stream),
                    "megakernel launch");
        }
        return APH_OK;
    }
    // The wavefront folds each chunk's samples into every target in one pass, so
    // two requests on one target would race: one target per request of a call.
    for (uint32_t i = 0; i < nreq; i++)
        for (uint32_t k = 0; k < i; k++)
            if (reqs[k].target == reqs[i].target)
                return fail(c, APH_E_INVALID, "requests %u and %u share target %llu (one request per target per call)",
                            k, i, (unsigned long long)reqs[i].target);
    hipError_t e = wavefrontRender(c->wf, wreq.data(), (uint32_t)wreq.size(), c->counters.as<DevCounters>(),
                                   c->statsLevel > 0, c->stream, c->timing ? &c->pending : nullptr);
    if (e != hipSuccess) return hipFail(c, e, "wavefront render");
    return APH_OK;
}

int aph_render_noalloc(aph_ctx_t c, const aph_request* reqs, uint32_t nreq)
{
    int rc = enter(c);
    if (rc) return rc;
    if (nreq == 0) return APH_OK;
    if ((rc = checkTextures(c))) return rc;
    if (!reqs) return fail(c, APH_E_INVALID, "requests pointer is null");
    if (c->wfStaticStatus != (int)hipSuccess)
        return fail(c, APH_E_STATE, "allocation-free render memory unavailable on device %d: %s", c->device,
                    hipGetErrorString((hipError_t)c->wfStaticStatus));
    std::vector<WavefrontRequest> wreq;
    if ((rc = buildRequests(c, reqs, nreq, wreq))) return rc;
    for (uint32_t i = 0; i < nreq; i++) {
        for (uint32_t k = 0; k < i; k++)
            if (reqs[k].target == reqs[i].target)
                return fail(c, APH_E_INVALID, "requests %u and %u share target %llu (one request per target per call)",
                            k, i, (unsigned long long)reqs[i].target);
        if ((uint64_t)wreq[i].width * wreq[i].height > kNoAllocPaths)
            return fail(c, APH_E_INVALID, "request %u: %ux%u pixels exceed the allocation-free render's %u paths",
                        i, wreq[i].width, wreq[i].height, kNoAllocPaths);
    }
    if (wavefrontCallTableBytes(wreq.data(), nreq, kNoAllocPaths) > kNoAllocTableBytes)
        return fail(c, APH_E_INVALID, "%u requests and their sample segments exceed the allocation-free render's "
                    "%zu-byte call table", nreq, kNoAllocTableBytes);
    c->wfStatic.tuning = c->wf.tuning;
    hipError_t e = wavefrontRender(c->wfStatic, wreq.data(), nreq, c->counters.as<DevCounters>(), c->statsLevel > 0,
                                   c->stream, nullptr, kNoAllocPaths);
    if (e == hipErrorOutOfMemory) return fail(c, APH_E_INVALID, "the call does not fit the allocation-free render");
    if (e != hipSuccess) return hipFail(c, e, "allocation-free render");
    return APH_OK;
}

int aph_set_launch_mode(aph_ctx_t c, int mode)
{
    int rc = enter(c);
    if (rc) return rc;
    if (mode != APH_LAUNCH_GRAPH && mode != APH_LAUNCH_EAGER) return fail(c, APH_E_INVALID, "unknown launch mode");
    c->wf.useGraph = mode == APH_LAUNCH_GRAPH;
    return APH_OK;
}

int aph_get_tuning(aph_ctx_t c, aph_tuning* t)
{
    if (!c || !t) return APH_E_INVALID;
    t->chunk_paths = c->wf.tuning.chunkPaths;
    t->trace_blocks_per_cu = c->wf.tuning.tracePerCU;
    t->shade_blocks_per_cu = c->wf.tuning.shadePerCU;
    t->trace_refill_lanes = c->wf.tuning.refillLanes;
    t->trace_leaf_wait_lanes = c->wf.tuning.leafWaitLanes;
    t->ms = 0.0f;
    return APH_OK;
}

int aph_set_tuning(aph_ctx_t c, const aph_tuning* t)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!t || t->chunk_paths < (1u << 16) || t->chunk_paths > (1u << 26) || t->trace_blocks_per_cu == 0 ||
        t->trace_blocks_per_cu > 64 || t->shade_blocks_per_cu == 0 || t->shade_blocks_per_cu > 64 ||
        t->trace_refill_lanes == 0 || t->trace_refill_lanes > 64 || t->trace_leaf_wait_lanes > 63)
        return fail(c, APH_E_INVALID, "tuning out of range");
    HIPCALL(c, hipStreamSynchronize(c->stream), "tuning");
    c->wf.tuning.chunkPaths = t->chunk_paths;
    c->wf.tuning.tracePerCU = t->trace_blocks_per_cu;
    c->wf.tuning.shadePerCU = t->shade_blocks_per_cu;
    c->wf.tuning.refillLanes = t->trace_refill_lanes;
    c->wf.tuning.leafWaitLanes = t->trace_leaf_wait_lanes;
    return APH_OK;
}

int aph_wavefront_autotune(aph_ctx_t c, const aph_request* reqs, uint32_t nreq, aph_tuning* best)
{
    int rc = enter(c);
    if (rc) return rc;
    if ((rc = checkTextures(c))) return rc;
    if (!reqs || nreq == 0) return fail(c, APH_E_INVALID, "autotune needs at least one request");
    std::vector<WavefrontRequest> wreq;
    if ((rc = buildRequests(c, reqs, nreq, wreq))) return rc;
    // Trials render into scratch targets of the requests' sizes, so the caller's
    // targets are untouched (a request with sample_start > 0 would blend into them).
    std::vector<DevBuf> scratch(nreq);
    for (uint32_t i = 0; i < nreq; i++) {
        const size_t bytes = (size_t)wreq[i].width * wreq[i].height * sizeof(float4);
        HIPCALL(c, scratch[i].alloc(bytes), "autotune scratch target");
        HIPCALL(c, hipMemsetAsync(scratch[i].p, 0, bytes, c->stream), "autotune scratch target");
        wreq[i].accum = scratch[i].as<float4>();
    }
    const uint32_t chunks[] = {4u << 20, 8u << 20, 16u << 20, 32u << 20, 64u << 20};
    const uint32_t traces[] = {4, 8, 16};
    const uint32_t shades[] = {4, 8, 16};
    const uint32_t refills[] = {16, 24, 32, 40, 48, 56, 64};
    const uint32_t leafWaits[] = {0, 8, 16, 24, 32};
    const WavefrontTuning keep = c->wf.tuning;
    aph_tuning win{keep.chunkPaths, keep.tracePerCU, keep.shadePerCU, keep.refillLanes, keep.leafWaitLanes, INFINITY};
    hipEvent_t a, b;
    HIPCALL(c, hipEventCreate(&a), "autotune");
    HIPCALL(c, hipEventCreate(&b), "autotune");
    // Renders the requests twice under t (the first run captures the graph for
    // this launch shape) and times the second; keeps t if it is the fastest.  A
    // candidate whose buffers do not fit in device memory is skipped.
    auto trial = [&](aph_tuning t) -> int {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return APH_E_DEVICE;
        c->wf.tuning = WavefrontTuning{t.chunk_paths, t.trace_blocks_per_cu, t.shade_blocks_per_cu,
                                       t.trace_refill_lanes, t.trace_leaf_wait_lanes};
        DevCounters* ctr = c->counters.as<DevCounters>();
        hipError_t e = wavefrontRender(c->wf, wreq.data(), nreq, ctr, c->statsLevel > 0, c->stream, nullptr);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            return APH_OK;
        }
        if (e != hipSuccess) return hipFail(c, e, "autotune render");
        if (hipEventRecord(a, c->stream) != hipSuccess) return APH_E_DEVICE;
        e = wavefrontRender(c->wf, wreq.data(), nreq, ctr, c->statsLevel > 0, c->stream, nullptr);
        if (e != hipSuccess) return hipFail(c, e, "autotune render");
        if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) return APH_E_DEVICE;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, a, b);
        if (ms < win.ms) {
            win = t;
            win.ms = ms;
        }
        return APH_OK;
    };
    int status = APH_OK;
    // Stage 1: chunk size x traversal grid x shade grid at the current refill
    // threshold and leaf wait; stage 2: the refill threshold at the stage-1
    // winner; stage 3: the leaf wait at the stage-2 winner.
    for (uint32_t ch : chunks)
        for (uint32_t tr : traces)
            for (uint32_t sh : shades)
                if (!status) status = trial(aph_tuning{ch, tr, sh, keep.refillLanes, keep.leafWaitLanes, 0.0f});
    const aph_tuning grid = win;
    for (uint32_t rf : refills)
        if (!status && rf != grid.trace_refill_lanes)
            status = trial(aph_tuning{grid.chunk_paths, grid.trace_blocks_per_cu, grid.shade_blocks_per_cu, rf,
                                      grid.trace_leaf_wait_lanes, 0.0f});
    const aph_tuning withRefill = win;
    for (uint32_t lw : leafWaits)
        if (!status && lw != withRefill.trace_leaf_wait_lanes)
            status = trial(aph_tuning{withRefill.chunk_paths, withRefill.trace_blocks_per_cu,
                                      withRefill.shade_blocks_per_cu, withRefill.trace_refill_lanes, lw, 0.0f});
    hipEventDestroy(a);
    hipEventDestroy(b);
    hipStreamSynchronize(c->stream);  // the scratch targets are freed on return
    if (!status && !std::isfinite(win.ms)) status = fail(c, APH_E_MEMORY, "no candidate fits in device memory");
    if (status) {
        c->wf.tuning = keep;
        return status;
    }
    c->wf.tuning = WavefrontTuning{win.chunk_paths, win.trace_blocks_per_cu, win.shade_blocks_per_cu,
                                   win.trace_refill_lanes, win.trace_leaf_wait_lanes};
    if (best) *best = win;
    return APH_OK;
}

int aph_wait(aph_ctx_t c)
{
    int rc = enter(c);
    if (rc) return rc;
    HIPCALL(c, hipStreamSynchronize(c->stream), "wait");
    DevCounters k;
    HIPCALL(c, hipMemcpy(&k, c->counters.p, sizeof(k), hipMemcpyDeviceToHost), "counters");
    if (k.overflow) return fail(c, APH_E_OVERFLOW, "traversal stack overflow on %llu rays", k.overflow);
    return APH_OK;
}

int aph_readback(aph_ctx_t c, uint64_t target, void* dst, size_t bytes)
{
    int rc = enter(c);
    if (rc) return rc;
    Target* t = lookup(c->targets, target);
    if (!t || !dst) return fail(c, APH_E_INVALID, "unknown target or null destination");
    size_t need = (size_t)t->w * t->h * 16;
    if (bytes < need) return fail(c, APH_E_INVALID, "destination too small (%zu < %zu)", bytes, need);
    HIPCALL(c, hipMemcpyAsync(dst, t->accum.p, need, hipMemcpyDeviceToHost, c->stream), "readback");
    HIPCALL(c, hipStreamSynchronize(c->stream), "readback");
    return APH_OK;
}

int aph_post(aph_ctx_t c, uint64_t target, const float brightness[3], int tone_map, int gamma, float* dst, size_t bytes)
{
    int rc = enter(c);
    if (rc) return rc;
    Target* t = lookup(c->targets, target);
    if (!t || !dst || !brightness) return fail(c, APH_E_INVALID, "unknown target or null pointer");
    size_t n = (size_t)t->w * t->h;
    if (bytes < n * 16) return fail(c, APH_E_INVALID, "destination too small");
    HIPCALL(c, c->scratchB.reserve(n * 16), "alloc post");
    {
        TimedScope ts(c, "post");
        HIPCALL(c, launchPost(t->accum.as<float4>(), (uint32_t)n, brightness, tone_map, gamma, c->scratchB.as<float4>(),
                              c->stream),
                "post launch");
    }
    HIPCALL(c, hipMemcpyAsync(dst, c->scratchB.p, n * 16, hipMemcpyDeviceToHost, c->stream), "post readback");
    HIPCALL(c, hipStreamSynchronize(c->stream), "post");
    return APH_OK;
}

int aph_set_stats_level(aph_ctx_t c, int level)
{
    if (!c) return APH_E_INVALID;
    if (level < 0 || level > 1) return fail(c, APH_E_INVALID, "stats level must be 0 or 1");
    c->statsLevel = level;
    return APH_OK;
}

int aph_stats_get(aph_ctx_t c, aph_stats* out)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!out) return fail(c, APH_E_INVALID, "null stats pointer");
    HIPCALL(c, hipStreamSynchronize(c->stream), "stats");
    DevCounters k;
    HIPCALL(c, hipMemcpy(&k, c->counters.p, sizeof(k), hipMemcpyDeviceToHost), "stats");
    out->extension_rays = k.ext;
    out->shadow_rays = k.shadow;
    out->node_visits = k.nodes;
    out->tri_tests = k.tris;
    out->inst_visits = k.insts;
    out->paths = k.paths;
    out->shadow_node_visits = k.snodes;
    out->shadow_tri_tests = k.stris;
    out->shadow_inst_visits = k.sinsts;
    out->hits = k.hits;
    return APH_OK;
}

int aph_stats_reset(aph_ctx_t c)
{
    int rc = enter(c);
    if (rc) return rc;
    HIPCALL(c, hipStreamSynchronize(c->stream), "stats reset");
    HIPCALL(c, hipMemset(c->counters.p, 0, sizeof(DevCounters)), "stats reset");
    return APH_OK;
}

int aph_timing_enable(aph_ctx_t c, int enable)
{
    if (!c) return APH_E_INVALID;
    c->timing = enable != 0;
    return APH_OK;
}

int aph_timing_get(aph_ctx_t c, const char* kernel, double* total_ms, uint64_t* launches)
{
    int rc = enter(c);
    if (rc) return rc;
    if (!kernel || !total_ms || !launches) return fail(c, APH_E_INVALID, "null pointer");
    resolveTimings(c);
    auto it = c->timings.find(kernel);
    *total_ms = it == c->timings.end() ? 0.0 : it->second.ms;
    *launches = it == c->timings.end() ? 0 : it->second.launches;
    return APH_OK;
}

int aph_timing_reset(aph_ctx_t c)
{
    int rc = enter(c);
    if (rc) return rc;
    resolveTimings(c);
    c->timings.clear();
    return APH_OK;
}

}  // extern "C"
