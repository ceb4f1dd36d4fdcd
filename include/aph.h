/*
 * aph.h - C-ABI of the MI355X (gfx950) path-tracing backend for Aurora.
 *
 * This is the thin boundary between Aurora-side C++ host code (a HIP renderer
 * type behind IRenderer/IScene) and the hand-written HIP kernels in
 * libaph.so.  Plain C types only: pointers, sizes, POD structs, int status.
 * Reference citations are relative to the Aurora source tree.
 *
 * Conventions
 *   - Every call returns APH_OK (0) or a negative APH_E_* code; the message is
 *     available from aph_last_error(ctx).  No C++ exception crosses the ABI.
 *     (Aurora reports the same conditions through AU_ERROR / AU_FAIL,
 *     Libraries/Foundation/API/Aurora/Foundation/Log.h:79-113.)
 *   - Host arrays passed in are caller-owned and copied before the call
 *     returns (GeometryBase.cpp:94-112 copies client geometry the same way).
 *   - Device objects (meshes, BVHs, targets) are owned by the context and named
 *     by 64-bit handles; 0 is never a valid handle.
 *   - Calls on one context are externally serialized (Aurora has no internal
 *     threading contract, Aurora.h:1068-1227); separate contexts may be used
 *     from separate threads.
 *   - Work is enqueued on the context's stream; aph_wait() blocks until it
 *     drains.  Entry points taking *host* output pointers are synchronous.
 */
#ifndef APH_H
#define APH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APH_VERSION 1

enum {
    APH_OK = 0,
    APH_E_INVALID = -1,   /* bad argument or handle */
    APH_E_DEVICE = -2,    /* HIP runtime error */
    APH_E_MEMORY = -3,    /* allocation failed / over budget */
    APH_E_STATE = -4,     /* call order violated (e.g. BLAS not built) */
    APH_E_UNSUPPORTED = -5,
    APH_E_OVERFLOW = -6   /* traversal stack or queue capacity exceeded */
};

enum { APH_BUILDER_LBVH = 0, APH_BUILDER_SAH = 1 };

typedef struct aph_ctx* aph_ctx_t;

/* Ray and hit records, 32 B / 20 B.  A miss has inst == prim == 0xFFFFFFFF and
 * t == +inf.  (u, v) are the barycentric weights of vertices 1 and 2, the same
 * convention as BuiltInTriangleIntersectionAttributes (Geometry.slang:115-118). */
typedef struct { float org[3]; float tmin; float dir[3]; float tmax; } aph_ray;
typedef struct { float t; uint32_t inst; uint32_t prim; float u; float v; } aph_hit;

/* Instance record, 112 B.  o2w is the row-major 3x4 object-to-world transform,
 * the layout of the HGI instance transform (HGIScene.h:44-48).  w2o may be left
 * all zero: the library then fills it with aph_inverse3x4(o2w). */
typedef struct {
    float o2w[12];
    float w2o[12];
    uint32_t mesh;      /* 1-based mesh handle (low 32 bits) */
    uint32_t material;  /* index into the table given to aph_materials_upload */
    uint32_t mask;      /* reserved, 0xFF */
    uint32_t flags;     /* reserved, 0 */
} aph_instance;

/* Standard Surface constants, 400 B, byte-for-byte the MaterialConstants layout
 * of Shaders/DefaultMaterialUniformBuffer.slang:3-72 (offsets in comments). */
typedef struct {
    float base;                 /*   0 */
    float base_color[3];        /*   4 */
    float diffuse_roughness;    /*  16 */
    float metalness;            /*  20 */
    float specular;             /*  24 */
    int32_t _pad0;              /*  28 */
    float specular_color[3];    /*  32 */
    float specular_roughness;   /*  44 */
    float specular_ior;         /*  48 */
    float specular_anisotropy;  /*  52 */
    float specular_rotation;    /*  56 */
    float transmission;         /*  60 */
    float transmission_color[3];/*  64 */
    float subsurface;           /*  76 */
    float subsurface_color[3];  /*  80 */
    int32_t _pad1;              /*  92 */
    float subsurface_radius[3]; /*  96 */
    float subsurface_scale;     /* 108 */
    float subsurface_anisotropy;/* 112 */
    float sheen;                /* 116 */
    int32_t _pad2, _pad3;       /* 120 */
    float sheen_color[3];       /* 128 */
    float sheen_roughness;      /* 140 */
    float coat;                 /* 144 */
    float coat_color[3];        /* 148 */
    float coat_roughness;       /* 160 */
    float coat_anisotropy;      /* 164 */
    float coat_rotation;        /* 168 */
    float coat_ior;             /* 172 */
    float coat_affect_color;    /* 176 */
    float coat_affect_roughness;/* 180 */
    float emission;             /* 184 */
    int32_t _pad4;              /* 188 */
    float emission_color[3];    /* 192 */
    int32_t _pad5;              /* 204 */
    float opacity[3];           /* 208 */
    int32_t thin_walled;        /* 220 */
    int32_t texture_block[44];  /* 224..399: has_*_image flags and UV transforms */
} aph_material;

/* Environment description, the EnvironmentData of EnvironmentBase.h:33-55 plus
 * image pointers.  Transforms are row-major 4x4 applied as M*(v,0). */
typedef struct {
    float light_top[3];
    float light_bottom[3];
    float background_top[3];
    float background_bottom[3];
    float light_transform[16];
    float background_transform[16];
    int32_t background_use_screen;
    int32_t filter;             /* 0 nearest (HGI default sampler), 1 bilinear */
    const float* light_rgb;     /* RGB32F lat-long image or NULL (gradient) */
    uint32_t light_width, light_height;
    const float* background_rgb;/* RGB32F or NULL (gradient) */
    uint32_t background_width, background_height;
} aph_environment;

typedef struct {
    float color_intensity[4];   /* rgb, intensity */
    float direction[3];         /* toward the light */
    float cos_radius;
} aph_distant_light;            /* DistantLight, Frame.slang:33-43 */

/* Per-request camera and options: the fields of FrameData (Frame.slang:61-157,
 * RendererBase.cpp:225-295) that the path reads. */
typedef struct {
    float inv_view[16];         /* rows: right, up, back, eye = transpose(inverse(view)) */
    float view_size[2];         /* (2/P[0][0], 2/P[1][1]) */
    float jitter[2];
    int32_t is_ortho;
    float focal_distance;
    float lens_radius;
    int32_t trace_depth;        /* clamped to [1,10] like RendererBase.cpp:265 */
    float max_luminance;
    int32_t rr_enabled;
    int32_t rr_start_depth;
    int32_t force_opaque_shadows;
    int32_t display_errors;
    int32_t temporal_jitter;
    int32_t light_count;
    aph_distant_light lights[4];
} aph_frame;

/* One render request: accumulate samples [sample_start, sample_start+sample_count)
 * of one camera into one target (IRenderer::render(sampleStart, sampleCount),
 * Aurora.h:1196). */
typedef struct {
    uint64_t target;
    uint64_t tlas;
    aph_frame frame;
    uint32_t sample_start;
    uint32_t sample_count;
    uint32_t seed_offset;       /* SampleData.seedOffset (Frame.slang:18-25), 0 normally */
    uint32_t _pad;
} aph_request;

enum { APH_MODE_MEGAKERNEL = 0, APH_MODE_WAVEFRONT = 1 };

typedef struct {
    uint64_t extension_rays;    /* closest-hit queries */
    uint64_t shadow_rays;       /* any-hit queries */
    uint64_t node_visits;       /* BVH nodes fetched (stats mode only) */
    uint64_t tri_tests;         /* triangles tested (stats mode only) */
    uint64_t inst_visits;       /* instance transforms applied (stats mode only) */
    uint64_t paths;             /* pixel-samples started */
    uint64_t shadow_node_visits; /* the any-hit queries' share of node_visits */
    uint64_t shadow_tri_tests;   /* ... of tri_tests */
    uint64_t shadow_inst_visits; /* ... of inst_visits */
    uint64_t hits;               /* extension rays that hit a triangle (wavefront, stats mode only) */
} aph_stats;

/* ---- context: createRenderer(Backend, taskCount), Aurora.h:1233-1234;
 *      Aurora.cpp:31-69 returns nullptr when the backend is not valid ---- */
int aph_create(int device, aph_ctx_t* out);
void aph_destroy(aph_ctx_t ctx);
const char* aph_last_error(aph_ctx_t ctx);
int aph_version(void);
/* The HIP stream (hipStream_t) all work of the context is enqueued on. */
int aph_get_stream(aph_ctx_t ctx, void** stream);

/* ---- geometry: IRenderer::createGeometryPointer + GeometryBase copy
 *      (GeometryBase.cpp:22-125); positions required, normals/uv optional ---- */
int aph_mesh_upload(aph_ctx_t ctx, const float* pos, uint32_t nv, const uint32_t* idx, uint32_t nt,
                    const float* nrm, const float* uv, uint64_t* mesh);
/* BLAS build, replaces the driver build of HGIGeometry.cpp:127-160 /
 * PTGeometry.cpp:155-223.  APH_BUILDER_LBVH: Karras LBVH over 63-bit Morton
 * codes, one triangle per leaf (fastest build); APH_BUILDER_SAH: top-down
 * binned SAH (16 bins x 3 axes), up to 4 triangles per leaf (fastest trace).
 * The TLAS is always a binned-SAH tree over the instance world boxes. */
int aph_blas_build(aph_ctx_t ctx, uint64_t mesh, int builder);
/* Material layers of each instance of a TLAS: the DirectX backend's layer loop
 * (HitShaderEntryPoints.slang:84-113) over IScene::addInstance's
 * LayerDefinitions.  Instance i's layers are materials[offsets[i] ..
 * offsets[i+1]) (indices into the material table), ordered from just above the
 * base material to the top; a hit tries the top layer first, each layer's
 * opacity coin passing the ray on (throughput *= transparency / P) or stopping
 * at it.  Layers read the instance's texture coordinates until
 * aph_tlas_set_layer_geometry gives them their own; setting layers clears that. */
int aph_tlas_set_layers(aph_ctx_t ctx, uint64_t tlas, const uint32_t* offsets, const uint32_t* materials);
/* The geometry half of each LayerDefinition (IScene::addInstancePointer, Aurora.h:1027-1029;
 * PTScene.cpp:1051-1066, 1095-1107): meshes[j] supplies the texture coordinates of layer j
 * (the same indexing as aph_tlas_set_layers' materials), read at the hit's vertices
 * (computeLayerUV, GlobalBufferAccessors.slang:116-127) by that layer's opacity and material
 * textures; 0 keeps the instance's own.  A layer mesh has its instance's vertex count
 * (PTScene.cpp:1102-1103, APH_E_INVALID otherwise); APH_E_STATE without layers. */
int aph_tlas_set_layer_geometry(aph_ctx_t ctx, uint64_t tlas, const uint64_t* meshes);
/* Copies a built BLAS to host memory: *node_count 64-byte nodes (two child
 * boxes; child references in a.w / b.w, leaf = bit 31 | (count-1) << 27 |
 * first) and the triangles in leaf order (48 bytes, original primitive index in
 * v0.w).  Null buffers query node_count and depth only.  No counterpart in the
 * reference, whose BVH lives inside the driver: BVH inspection for tests and
 * tools. */
int aph_blas_export(aph_ctx_t ctx, uint64_t mesh, void* nodes, uint64_t nodes_bytes, void* tris,
                    uint64_t tris_bytes, uint32_t* node_count, uint32_t* depth);
/* TLAS build over instances, replaces HGIScene::rebuildAccelerationStructure
 * (HGIScene.cpp:377-449) / PTScene::buildTLAS (PTScene.cpp:1460-1552).
 * Instance IDs are the indices in inst[] (active order, HGIScene.cpp:392). */
int aph_tlas_build(aph_ctx_t ctx, const aph_instance* inst, uint32_t n, uint64_t* tlas);
/* Incremental scene update (IScene::setInstanceProperties, Aurora.h:1003-1009, addInstance /
 * removeInstance, Aurora.h:892-1042), in place of the reference's rebuild of the instance list,
 * acceleration structure and pipeline on any change (HGIScene.cpp:158-171, driven by
 * resource activation, ResourceStub.cpp:386-470).  The TLAS handle takes the new
 * instance list; every BLAS is kept.  APH_TLAS_REBUILD builds a new tree over the new
 * instances (any count); APH_TLAS_REFIT keeps the tree's topology and recomputes its
 * boxes bottom-up (the same instances with geometry as before; new transforms,
 * materials or meshes) - faster for small motions, and the hits are the same either
 * way because boxes only cull.  The handle keeps its environment, and its layers while
 * the instance count is unchanged.  Synchronizes with renders in flight. */
enum { APH_TLAS_REBUILD = 0, APH_TLAS_REFIT = 1 };
int aph_tlas_update(aph_ctx_t ctx, uint64_t tlas, const aph_instance* inst, uint32_t n, int mode);
/* Row-major 3x4 affine inverse in double, rounded to float (the w2o convention). */
int aph_inverse3x4(const float m[12], float out[12]);

/* ---- materials: MaterialBase::updateBuiltInMaterial (MaterialBase.cpp:193-214),
 *      uploaded per change like HGIMaterial.cpp:39-59 ---- */
int aph_materials_upload(aph_ctx_t ctx, const aph_material* mats, uint32_t n);
/* Patches materials [first, first + n) of the uploaded table in place
 * (IScene::setMaterialProperties, Aurora.h:998; MaterialBase::updateBuiltInMaterial,
 * MaterialBase.cpp:193-214, for the changed materials only): unlike a new upload,
 * texture bindings are kept. */
int aph_materials_update(aph_ctx_t ctx, uint32_t first, const aph_material* mats, uint32_t n);
void aph_material_defaults(aph_material* m);

/* ---- material textures: evaluateDefaultMaterial (DefaultMaterial.slang:171-233)
 * samples up to five images per material, each through its own UV transform
 * (applyUVTransform, :97-104: scale, then rotation in degrees, then offset,
 * about the pivot) read from the material's texture block (offsets 224-399),
 * only where its has_*_image flag is set.  Images are sampled as HGI's samplers
 * do (HGISampler.cpp:22-60): bilinear, one mip level (HGIImage.cpp:35), address
 * mode per binding, opaque-white border; an RGBA8 image with linearize set is
 * sRGB-decoded texel by texel before filtering (HgiFormatUNorm8Vec4srgb,
 * HGIImage.cpp:98-100).  Texel (x, y) of a w x h image is pixels[y*w + x], at
 * uv ((x + 0.5)/w, (y + 0.5)/h).  The normal image is a tangent-space normal map
 * (calculateNormalFromMap, :107-127) whose result replaces the shading normal
 * and frame of the hit (HitShaderEntryPoints.slang:251-255). ---- */
enum { APH_IMAGE_RGBA8 = 0, APH_IMAGE_RGBA32F = 1, APH_IMAGE_RGB32F = 2 };
typedef struct {
    uint32_t width, height;
    int32_t format;     /* APH_IMAGE_* */
    int32_t linearize;  /* RGBA8 only: decode sRGB (IImage::InitData::linearize) */
    const void* pixels; /* row-major, rows of width pixels */
} aph_image_desc;
/* IRenderer::createImagePointer (Aurora.h:1116; HGIRenderer.cpp:434, HGIImage.cpp:23-112). */
int aph_image_create(aph_ctx_t ctx, const aph_image_desc* desc, uint64_t* image);
/* Fails with APH_E_STATE while the image is bound to the material table. */
int aph_image_destroy(aph_ctx_t ctx, uint64_t image);
/* Sampler address modes (Names::AddressModes, HGISampler.cpp:45-60). */
enum { APH_ADDRESS_WRAP = 0, APH_ADDRESS_MIRROR = 1, APH_ADDRESS_CLAMP = 2, APH_ADDRESS_MIRROR_ONCE = 3,
       APH_ADDRESS_BORDER = 4 };
/* Texture slots, in the order of the material's has_*_image flags. */
enum { APH_TEX_BASE_COLOR = 0, APH_TEX_SPECULAR_ROUGHNESS = 1, APH_TEX_EMISSION_COLOR = 2, APH_TEX_OPACITY = 3,
       APH_TEX_NORMAL = 4, APH_TEX_SLOTS = 5 };
typedef struct {
    uint64_t image;              /* 0: unbound */
    int32_t address_u, address_v; /* APH_ADDRESS_* (default wrap) */
} aph_texture_binding;
/* Binds images to the material table: bindings[APH_TEX_SLOTS * i + slot] for
 * material i, n = the table's material count (HGIScene.cpp:226-380 gathers the
 * same per-instance texture and sampler indices).  A flagged emission slot
 * without an image reads as 0 (InstanceData.glsl:147-153); any other flagged
 * slot without an image makes renders fail with APH_E_STATE.  Uploading
 * materials clears the bindings. */
int aph_materials_bind_textures(aph_ctx_t ctx, const aph_texture_binding* bindings, uint32_t n);

/* ---- environment: EnvironmentBase::updateGPUStruct (EnvironmentBase.cpp:52-71)
 *      + AliasMap::build (AliasMap.cpp:31-166) ---- */
int aph_environment_upload(aph_ctx_t ctx, const aph_environment* env);
/* The environment as a resident object of its own, for scenes that differ in it
 * (IScene::setEnvironment, Aurora.h:974, SceneBase.h:81: each Aurora scene
 * holds one).
 * aph_tlas_set_environment(ctx, tlas, env) makes renders of that TLAS use it;
 * env 0 goes back to the context's (aph_environment_upload).  A request batch
 * may mix TLASes with different environments.  An environment set on a TLAS
 * cannot be destroyed (APH_E_STATE). */
int aph_environment_create(aph_ctx_t ctx, const aph_environment* env, uint64_t* handle);
int aph_environment_destroy(aph_ctx_t ctx, uint64_t handle);
int aph_tlas_set_environment(aph_ctx_t ctx, uint64_t tlas, uint64_t env);
/* ---- ground plane: IGroundPlane (Aurora.h), the DirectX backend's PTGroundPlane
 *      (DirectX/PTGroundPlane.cpp) evaluated by Shaders/GroundPlane.slang:24-290 ----
 * A matte plane seen by primary rays in front of their hit: with probability
 * reflection_opacity a GGX reflection of reflection_color / reflection_roughness continues
 * the path from the plane; otherwise the ray goes on, darkened towards shadow_color by
 * shadow_opacity where one environment sample from the plane is blocked (opaque shadow
 * ray).  The normal is used as given (not normalized). */
typedef struct {
    int32_t enabled;
    float position[3];
    float normal[3];
    float shadow_opacity;
    float shadow_color[3];
    float reflection_opacity;
    float reflection_color[3];
    float reflection_roughness;
} aph_ground_plane;
/* The property defaults of PTGroundPlane.cpp: enabled, at the origin facing +Y, shadow
 * opacity 1 in black, reflection opacity 0.5 of grey 0.5 at roughness 0.1. */
void aph_ground_plane_defaults(aph_ground_plane* gp);
/* IScene::setGroundPlanePointer (PTScene.cpp:660-667): renders of this TLAS use the plane;
 * NULL restores the default, a disabled plane (PTRenderer.cpp:713-727). */
int aph_tlas_set_ground_plane(aph_ctx_t ctx, uint64_t tlas, const aph_ground_plane* gp);

/* Alias table + luminance integral for an RGB32F lat-long image (host build,
 * the AliasMap::build algorithm); entries are 16 B {alias, prob, pdf, pad}. */
int aph_alias_build(const float* rgb, uint32_t w, uint32_t h, void* entries, float* integral);

/* ---- ray queries: TraceRay in traceInstanceRay (RayTrace.slang:185-210) and
 *      traceShadowRay (RayTrace.slang:221-270, forced opaque) ---- */
int aph_trace_closest(aph_ctx_t ctx, uint64_t tlas, const aph_ray* rays, uint32_t n, aph_hit* hits);
int aph_trace_any(aph_ctx_t ctx, uint64_t tlas, const aph_ray* rays, uint32_t n, uint8_t* occluded);
/* Same queries on device-resident buffers, asynchronous on the context stream. */
int aph_trace_closest_device(aph_ctx_t ctx, uint64_t tlas, const aph_ray* d_rays, uint32_t n, aph_hit* d_hits);
int aph_trace_any_device(aph_ctx_t ctx, uint64_t tlas, const aph_ray* d_rays, uint32_t n, uint8_t* d_occluded);

/* ---- render targets: IRenderer::createRenderBuffer (Aurora.h:1068-1227); the
 *      accumulation texture of HGIRenderer.cpp:155-165 (RGBA32F) ---- */
int aph_target_create(aph_ctx_t ctx, uint32_t width, uint32_t height, uint64_t* target);
int aph_target_destroy(aph_ctx_t ctx, uint64_t target);
/* Device pointer of a target's RGBA32F accumulation buffer. */
int aph_target_device_ptr(aph_ctx_t ctx, uint64_t target, void** ptr);

/* ---- render: IRenderer::render(sampleStart, sampleCount) (Aurora.h:1196;
 *      HGIRenderer::render, HGIRenderer.cpp:477-589).  Asynchronous; no host
 *      memory is referenced after the call returns.  Every request is checked
 *      before any is launched.
 *      mode APH_MODE_MEGAKERNEL: one megakernel launch per request, one lane per
 *      pixel looping over all of the request's samples in order (the reference
 *      dispatches once per sample, HGIRenderer.cpp:526-566);
 *      APH_MODE_WAVEFRONT: all requests and samples share one wavefront
 *      pipeline (raygen, extend, shade, shadow, accumulate); the requests of
 *      one call must name distinct targets (APH_E_INVALID otherwise). ---- */
int aph_render(aph_ctx_t ctx, const aph_request* reqs, uint32_t nreq, int mode);
/* Allocation-free render.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc).  The same wavefront pipeline as aph_render(..., APH_MODE_WAVEFRONT)
 * and the same images, but its path state, queues and tables live in one fixed array of
 * about 74 MB per GPU (held from the library's first use on that GPU until the process
 * ends), and the call makes no hipMalloc, hipFree or other allocating call; its chunks
 * launch eagerly (no graph).  Limits: each request at most 131072 pixels (the paths in
 * flight), and the call's request and sample-segment tables at most 1 MiB (about 1,000
 * requests, or some 60,000 requested samples); beyond them APH_E_INVALID.  Every call on
 * one GPU shares that GPU's array: calls on one context are serialized by its stream, but
 * allocation-free calls of two contexts on the same GPU that overlap in time race. */
int aph_render_noalloc(aph_ctx_t ctx, const aph_request* reqs, uint32_t nreq);
/* IRenderer::waitForTask (HGIRenderer.cpp:591-594 is a no-op because HGI is synchronous). */
int aph_wait(aph_ctx_t ctx);
/* IRenderBuffer::data (HGIRenderBuffer.cpp:87-136): blocking copy of the
 * RGBA32F accumulation buffer. */
int aph_readback(aph_ctx_t ctx, uint64_t target, void* dst, size_t bytes);
/* PostProcessing.glsl:37-64 on a target into RGBA32F host memory. */
int aph_post(aph_ctx_t ctx, uint64_t target, const float brightness[3], int tone_map, int gamma,
             float* dst, size_t bytes);

/* ---- instrumentation ---- */
/* stats_level 0: ray counts only; 1: also node/triangle/instance visit counts. */
int aph_set_stats_level(aph_ctx_t ctx, int level);
int aph_stats_get(aph_ctx_t ctx, aph_stats* out);   /* synchronizes */
int aph_stats_reset(aph_ctx_t ctx);
/* Per-kernel timing with HIP events on the context stream. */
int aph_timing_enable(aph_ctx_t ctx, int enable);
int aph_timing_get(aph_ctx_t ctx, const char* kernel, double* total_ms, uint64_t* launches);
int aph_timing_reset(aph_ctx_t ctx);

/* ---- launch control of the wavefront pipeline.  No reference counterpart: the
 * reference's driver schedules its own ray tracing. ---- */
enum { APH_LAUNCH_GRAPH = 0, APH_LAUNCH_EAGER = 1 };
/* APH_LAUNCH_GRAPH (default): each chunk's stage sequence (generation, per
 * bounce reset/extend/shade/shadow/resolve, accumulation) is one captured HIP
 * graph, replayed for every chunk; APH_LAUNCH_EAGER launches the kernels one by
 * one.  With timing enabled the pipeline always launches eagerly. */
int aph_set_launch_mode(aph_ctx_t ctx, int mode);
typedef struct {
    uint32_t chunk_paths;          /* paths in flight per chunk */
    uint32_t trace_blocks_per_cu;  /* persistent traversal kernels: 128-thread blocks per CU */
    uint32_t shade_blocks_per_cu;  /* shade kernel */
    uint32_t trace_refill_lanes;   /* idle lanes (1..64) of a traversal wave that trigger a refill; default 32 */
    uint32_t trace_leaf_wait_lanes;/* lanes (0..63) still looking for a leaf when a traversal wave processes
                                      leaves; default 16 */
    float ms;                      /* measured time of the winning configuration (autotune) */
} aph_tuning;
int aph_get_tuning(aph_ctx_t ctx, aph_tuning* t);
int aph_set_tuning(aph_ctx_t ctx, const aph_tuning* t);
/* Renders the given requests - into scratch targets of their sizes, so the
 * requests' own targets are left as they were - under each candidate launch
 * configuration - chunk of 4M/8M/16M/32M/64M paths x 4/8/16 traversal blocks per
 * CU x 4/8/16 shade blocks per CU, then refill thresholds 16..64 (step 8) at
 * the best of those, then leaf-wait lanes 0/8/16/24/32 at the best of those -
 * and keeps the fastest for this context.  A candidate whose buffers do not
 * fit in device memory is skipped (APH_E_MEMORY if none fits).
 * A chunk holds about 0.55 KB of device memory per path (35 GB at 64M paths),
 * allocated at the tuned size even for smaller frames. */
int aph_wavefront_autotune(aph_ctx_t ctx, const aph_request* reqs, uint32_t nreq, aph_tuning* best);

/* ---- guide images (AOVs) and temporal resolve: the DirectX backend's denoising
 * guides gDepthView, gNormalRoughness, gBaseColorMetalness, gMotionVectors
 * (GlobalPipelineState.slang:110-172, written by the primary hit/miss,
 * HitShaderEntryPoints.slang:150-215, MainEntryPoints.slang:385-405) and
 * TemporalResolve.hlsl:156-249.  An AOV set holds four planes per pixel:
 * view depth (float), normal+roughness, base colour+metalness and motion
 * (float4 each). ---- */
int aph_aov_create(aph_ctx_t ctx, uint32_t width, uint32_t height, uint64_t* aov);
int aph_aov_destroy(aph_ctx_t ctx, uint64_t aov);
/* Copies planes to host memory; any pointer may be NULL. */
int aph_aov_readback(aph_ctx_t ctx, uint64_t aov, float* depth, float* normal_roughness,
                     float* base_color_metalness, float* motion);

typedef struct {
    uint64_t aov;                 /* output set; its size is the image size */
    uint64_t tlas;
    aph_frame frame;              /* this frame's camera (frame.jitter with temporal_jitter) */
    float view_proj[16];          /* FrameData.cameraViewProj, row-major: clip = M * (p, 1) */
    float view_proj_prev[16];     /* cameraViewProjPrev */
    float inv_view_prev[16];      /* cameraInvViewPrev, the layout of frame.inv_view */
    float sky_view_z;             /* FrameData.skyViewZ (RendererBase.cpp:258) */
    uint32_t sample_index;        /* the camera ray of this sample index (plus seed_offset) */
    uint32_t seed_offset;
    uint32_t _pad;
} aph_aov_request;
/* Traces the camera ray of one sample per pixel and writes the guides of its
 * first segment (hit: view depth of origin + t*dir, normalized shading normal,
 * max(specular roughness, 0.05), base colour, metalness, 2.5D motion; miss:
 * sky_view_z, (0,0,0,0.1) twice, direction-only motion). */
int aph_render_aovs(aph_ctx_t ctx, const aph_aov_request* req);

typedef struct {
    uint64_t current;             /* target: this frame's accumulation (gAccumulation) */
    uint64_t aov;                 /* this frame's motion vectors and view depth */
    uint64_t history_in;          /* target: last resolved frame */
    uint64_t history_out;         /* target written (gTAAHistoryA/B ping-pong) */
    int32_t is_reset;             /* discard history this frame */
    float blend_alpha_min;        /* PTRenderer.cpp:2043: 1/16 */
    float blend_alpha_max;        /* :2044: 0.6 */
    float speed_reference;        /* :2046: 6 pixels per frame */
} aph_taa_request;
int aph_temporal_resolve(aph_ctx_t ctx, const aph_taa_request* req);
/* Sub-pixel camera jitter of a frame, Halton(2,3) centred on the pixel:
 * getTemporalJitter, PTRenderer.cpp:159-174 (phase_count 0: no wrap). */
void aph_halton_jitter(uint32_t frame_index, uint32_t phase_count, float out[2]);

/* ---- serving loop: continuous batching of render jobs against device-resident
 * TLASes.  No reference counterpart: Aurora renders one request at a time
 * (HGIRenderer::render, HGIRenderer.cpp:477-589, progressive
 * render(sampleStart, sampleCount) driven by SampleCounter, Timer.h:97-286).
 * Each aph_server_step admits waiting jobs in arrival order while a slot is
 * free and the HBM budget allows, then renders the next run of samples of every
 * admitted job in one shared wavefront.  The budget covers the resident scenes
 * (BVHs, meshes, instance records, their environments), the images bound to
 * the material table, the admitted jobs' RGBA32F accumulation buffers and the
 * wavefront working set; a step's chunk of paths in flight is the context's
 * tuned one, halved until the working set fits (never below the largest
 * active frame).  A job's image is the same alone and in any mix.  Errors are
 * reported through aph_last_error of the context. ---- */
typedef struct aph_server* aph_server_t;

typedef struct {
    uint64_t hbm_budget;    /* bytes: resident scenes + images + job buffers + working set; 0 = unlimited */
    uint32_t max_active;    /* jobs sharing the wavefront at once; 0 = unlimited */
    uint32_t step_samples;  /* pixel-samples per step over all active jobs; 0 = 8M */
    int mode;               /* APH_MODE_WAVEFRONT: batched; APH_MODE_MEGAKERNEL: one job at a time */
    int keep_results;       /* 1: a finished job keeps its image until aph_server_release */
} aph_server_config;

typedef struct {
    uint64_t tlas;          /* scene (built TLAS of the context) */
    aph_frame frame;        /* camera and options, as in aph_request */
    uint32_t width, height; /* image size */
    uint32_t spp;           /* samples 0 .. spp-1 */
    uint32_t seed_offset;
} aph_job;

enum { APH_JOB_PENDING = 0, APH_JOB_ACTIVE = 1, APH_JOB_DONE = 2, APH_JOB_FAILED = 3 };

typedef struct {
    int state;
    uint32_t samples_done;
    double t_arrival, t_admit, t_done; /* seconds on the server clock */
} aph_job_status;

int aph_server_create(aph_ctx_t ctx, const aph_server_config* cfg, aph_server_t* out);
void aph_server_destroy(aph_server_t s);
/* Seconds since aph_server_create (steady clock). */
double aph_server_clock(aph_server_t s);
/* Queues a job; t_arrival < 0 stamps the current clock. */
int aph_server_submit(aph_server_t s, const aph_job* job, double t_arrival, uint64_t* ticket);
/* Admission + one batched step (synchronous); reports jobs running and waiting. */
int aph_server_step(aph_server_t s, uint32_t* active, uint32_t* pending);
int aph_server_status(aph_server_t s, uint64_t ticket, aph_job_status* st);
/* RGBA32F image of a finished job kept with keep_results. */
int aph_server_result(aph_server_t s, uint64_t ticket, float* dst, size_t bytes);
/* Device memory charged to the budget (resident scenes and material images,
 * admitted job buffers) and the jobs running / waiting. */
int aph_server_usage(aph_server_t s, uint64_t* resident_bytes, uint64_t* job_bytes, uint32_t* active,
                     uint32_t* pending);
/* The wavefront working set of the last batched step (path state, queues and
 * tables, charged to the budget) and its paths in flight. */
int aph_server_working_set(aph_server_t s, uint64_t* wavefront_bytes, uint32_t* paths_in_flight);
/* Forgets a finished or waiting job and frees its buffer. */
int aph_server_release(aph_server_t s, uint64_t ticket);
/* Replays an arrival trace in real time: job i is submitted arrival[i] seconds
 * (non-decreasing) after the call; returns when every job is done. */
int aph_server_replay(aph_server_t s, const aph_job* jobs, const double* arrival, uint32_t n, uint64_t* tickets);


/* ---- on-disk inputs (host only, no context): the Wavefront OBJ + MTL translation of
 * Plasma's loader (Applications/Plasma/OBJLoader.cpp:77-470, tinyobjloader's parsing), the
 * glTF 2.0 translation of its glTF loader (glTFLoader.cpp:26-386, tinygltf's parsing) and
 * the Radiance HDR lat-long environments Plasma loads (Plasma.cpp:69, :1341).  Texture
 * maps are reported by path, not decoded; MaterialX documents are reported, not read. ---- */
typedef struct {
    uint32_t nv, nt;
    const float* positions;     /* 3 nv */
    const float* normals;       /* 3 nv: the file's, normalized, or angle-weighted ones
                                   (Foundation calculateNormals, Geometry.cpp:29-101) */
    const float* uvs;           /* 2 nv with V flipped, or NULL */
    const uint32_t* indices;    /* 3 nt */
    int32_t material;           /* the shape's first face's material, -1: the default */
} aph_obj_shape;
typedef struct {
    uint32_t nshape, nmat;
    const aph_obj_shape* shapes;        /* one per o/g group with faces */
    const aph_material* materials;      /* Standard Surface, OBJLoader.cpp:162-229 */
    const char* const* material_names;  /* nmat */
    const char* const* textures;        /* 5 per material (map_Kd, map_Pr, map_Ke, map_d, norm/bump,
                                           the APH_TEX_* slot order): file path or NULL */
} aph_obj_file;
/* Parses an OBJ file and the MTL libraries it names; *out stays valid until aph_obj_free. */
int aph_obj_load(const char* path, aph_obj_file** out);
void aph_obj_free(aph_obj_file* f);
/* glTF: one mesh per (mesh, TRIANGLES primitive) with a float POSITION accessor, one
 * instance per node that uses it (every scene's node tree, world = parent x local). */
typedef struct {
    uint32_t nv, nt;
    const float* positions;     /* 3 nv */
    const float* normals;       /* 3 nv: NORMAL as stored, or Foundation calculateNormals' */
    const float* uvs;           /* 2 nv: TEXCOORD_0 as stored, or NULL */
    const uint32_t* indices;    /* 3 nt: the primitive's, or sequential */
} aph_gltf_mesh;
typedef struct {
    uint32_t mesh;              /* index into meshes */
    int32_t material;           /* index into materials; -1: the default material */
    float o2w[12];              /* row-major 3x4 world transform of the node */
    int32_t node;               /* glTF node index */
    const char* materialx;      /* the node's AURORA_materials_materialX file, or NULL */
} aph_gltf_instance;
typedef struct {
    int32_t type;               /* 0 perspective, 1 orthographic */
    int32_t node;               /* glTF node index */
    float view[16];             /* row-major world-to-view: inverse of the node's own transform */
    float yfov, aspect, znear, zfar, xmag, ymag;
} aph_gltf_camera;
typedef struct {
    uint32_t nmesh, ninst, nmat, ncam;
    const aph_gltf_mesh* meshes;
    const aph_gltf_instance* instances;
    const aph_material* materials;   /* base_color, metalness, specular_roughness from
                                        pbrMetallicRoughness (glTFLoader.cpp:400-414) */
    const aph_gltf_camera* cameras;
} aph_gltf_file;
/* Parses a .gltf (data: or file buffers) or .glb file; *out stays valid until aph_gltf_free. */
int aph_gltf_load(const char* path, aph_gltf_file** out);
void aph_gltf_free(aph_gltf_file* f);
/* Reads a Radiance RGBE (.hdr) image: *rgb is a width x height x 3 float array, top row
 * first, freed with aph_host_free.  APH_E_UNSUPPORTED for other orientations. */
int aph_hdr_load(const char* path, float** rgb, uint32_t* width, uint32_t* height);
void aph_host_free(void* p);

/* ---- Learned denoiser ----
 * A U-Net over the noisy radiance of a target and the guides of an AOV set
 * (aph_render_aovs): log radiance, base colour + metalness, normal +
 * roughness, view depth.  Convolutions run on bf16 MFMA with fp32
 * accumulation; training (Adam, L1 loss on log radiance) runs on the GPU.
 * No reference counterpart: the reference's DirectX backend hands the same
 * guides to NVIDIA NRD (DirectX/Denoiser.h:24-35, Denoiser::denoise at :70)
 * from PTRenderer.cpp:1052 (submitDenoising); aph_denoise takes that call's
 * place.  Images are padded to multiples of 16 internally. */
typedef struct { uint32_t level, channels; } aph_dn_buffer; /* NHWC bf16 at 1/2^level resolution */
enum { APH_DN_OP_CONV = 0, APH_DN_OP_POOL = 1, APH_DN_OP_UPSAMPLE = 2 };
/* conv: layer; pool (2x2 mean) and upsample (nearest, 2x): channels
 * [src_offset, +channels) of buffer src to [dst_offset, +channels) of dst */
typedef struct { int32_t kind, layer, src, src_offset, channels, dst, dst_offset; } aph_dn_op;
typedef struct {
    char name[16];
    uint32_t cin, cin_padded, cout, cout_padded;
    uint32_t in_buffer, out_buffer, out_offset, relu;
    /* offsets into the parameter vector: weights [9][cout_padded][cin_padded]
     * (tap 3*ky+kx, zero beyond cin/cout), bias [cout_padded] */
    uint64_t weight_offset, bias_offset;
} aph_dn_layer;
/* The network's layers, activation buffers and forward schedule (pass a null
 * array with a count pointer to query a count). */
int aph_denoiser_arch(aph_dn_layer* layers, uint32_t* n_layers, aph_dn_buffer* buffers, uint32_t* n_buffers,
                      aph_dn_op* ops, uint32_t* n_ops);
uint64_t aph_denoiser_param_count(void);
/* He-initialised weights from seed. */
int aph_denoiser_create(aph_ctx_t ctx, uint64_t seed, uint64_t* denoiser);
int aph_denoiser_destroy(aph_ctx_t ctx, uint64_t denoiser);
/* Loads a parameter vector (count = aph_denoiser_param_count()); resets Adam. */
int aph_denoiser_set_params(aph_ctx_t ctx, uint64_t denoiser, const float* params, uint64_t count);
enum {
    APH_DN_PARAMS = 0,           /* fp32 parameter vector */
    APH_DN_PARAM_GRADS = 1,      /* its gradient from the last training step */
    APH_DN_ACTIVATIONS = 2,      /* buffer `index` of the last forward pass, NHWC; after
                                  * aph_denoise the channels an upsampling op writes are
                                  * not stored (the decoder convolutions read the lower
                                  * level directly), after a training step they are */
    APH_DN_ACTIVATION_GRADS = 3, /* its gradient from the last training step */
    APH_DN_PREDICTION = 4        /* fp32 [n][Hp][Wp][4] predicted log1p radiance */
};
/* Inspection of the last pass (synchronizes); count must equal the size. */
int aph_denoiser_tensor_size(aph_ctx_t ctx, uint64_t denoiser, int kind, uint32_t index, uint64_t* count);
int aph_denoiser_tensor(aph_ctx_t ctx, uint64_t denoiser, int kind, uint32_t index, float* out, uint64_t count);
/* targets and an AOV set of one size; transform (augmentation, applied to all
 * three): destination pixel (x, y) reads source (sx, sy) with (u, v) = (y, x)
 * if bit 2 (transpose, square images only) else (x, y), sx = W-1-u if bit 0
 * else u, sy = H-1-v if bit 1 else v */
typedef struct { uint64_t noisy, aov, reference; uint32_t transform, _pad; } aph_denoise_example;
enum {
    APH_DN_LOSS_L1_LOG = 0,        /* mean |p - log1p(ref)|, p the predicted log1p radiance (default) */
    APH_DN_LOSS_L2_TONEMAPPED = 1  /* mean (x/(1+x) - ref/(1+ref))^2 = (exp(-log1p(ref)) - exp(-p))^2 */
};
int aph_denoiser_set_loss(aph_ctx_t ctx, uint64_t denoiser, int kind);
/* One Adam step over a batch (all examples one size); returns the mean loss
 * before the step (synchronizes). */
int aph_denoiser_train_step(aph_ctx_t ctx, uint64_t denoiser, const aph_denoise_example* ex, uint32_t n, float lr,
                            float* loss);
/* output may be the noisy target itself; flags: APH_DN_ENSEMBLE averages the
 * outputs of the network over the mirrored (and, for square images,
 * transposed) inputs, 4 or 8 passes in one batch */
typedef struct { uint64_t noisy, aov, output; uint32_t flags, _pad; } aph_denoise_item;
enum { APH_DN_ENSEMBLE = 1 };
/* Denoises n images of one size in one batch (asynchronous, on the context's stream). */
int aph_denoise(aph_ctx_t ctx, uint64_t denoiser, const aph_denoise_item* items, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
