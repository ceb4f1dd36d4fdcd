/*
 * CPU oracle for the Aurora path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke entry point and bench.py's CPU-baseline leg may load it.  The product
 * (aurora_amd/libaph.so) never links, loads or calls it.
 *
 * It is a plain-C restatement of the reference path tracer (Autodesk Aurora,
 * HGI backend semantics).  Each function cites the reference file:line it
 * follows; paths are relative to the reference checkout root, shader files live
 * under Libraries/Aurora/Source/Shaders/.
 *
 * Precision contract:
 *   - RNG, alias-map build, instance inverse and the FP64 watertight
 *     ray/triangle test are specified to the bit (IEEE single/double, no FMA
 *     contraction: the Makefile passes -ffp-contract=off).
 *   - Shading is FP32 like the reference shaders; results are compared with the
 *     GPU statistically (per-pixel tolerance + mean-luminance bound).
 *
 * What pins each part (see tests/test_oracle_*.py):
 *   RNG          - independent numpy restatement of Jarzynski & Olano 2020 pcg2d,
 *                  uniformity chi^2, the [0,1] inclusive quirk.
 *   intersection - brute-force vs BVH, closed-form hits, barycentric
 *                  reconstruction, watertightness on shared edges/vertices.
 *   alias map    - closed-form integral, sum(prob)=N invariant, pdf closed form,
 *                  chi^2 of the sampled distribution.
 *   BSDF         - per-lobe closed forms (Lambert/Oren-Nayar, GGX D, Smith G,
 *                  Schlick), the reference's sheen LUT values, reciprocity,
 *                  sampled-vs-evaluated Monte Carlo consistency.
 *   camera/env   - closed forms (lat-long round trip, pinhole geometry).
 *   path tracer  - analytic scenes (empty scene = background, emissive quad,
 *                  white furnace bound) and the reference's teapot fixture.
 */
#ifndef ORA_H
#define ORA_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lobe IDs and ray types, BSDFCommon.slang:28-39. */
#define ORA_TRANSPARENCY_LOBE 1
#define ORA_COAT_LOBE 2
#define ORA_METAL_LOBE 3
#define ORA_GLOSSY_LOBE 4
#define ORA_TRANSMISSION_LOBE 5
#define ORA_SHEEN_LOBE 6
#define ORA_DIFFUSE_LOBE 7
#define ORA_NO_SECONDARY_RAY 8
#define ORA_CAMERA_RAY 9
#define ORA_GROUND_PLANE_REFLECTION_RAY 10

/* ---- RNG (Random.slang:96-137) ---- */
void ora_pcg2d_init(uint32_t sample_index, uint32_t width, uint32_t x, uint32_t y, uint32_t state[2]);
void ora_pcg2d(uint32_t state[2]);                      /* advances state in place  */
void ora_random2d(uint32_t state[2], float out[2]);     /* pcg2DNext2D              */
void ora_random_stream(uint32_t sx, uint32_t sy, uint32_t n, float* out /* 2n */);

/* ---- transforms ---- */
/* Inverse of a row-major 3x4 affine matrix, computed in double by cofactors and
 * rounded to float.  Both sides of the boundary use this exact definition. */
int ora_inverse3x4(const float m[12], float out[12]);

/* ---- scene ---- */
typedef struct {
    const float* pos;     /* 3*nv */
    const float* nrm;     /* 3*nv or NULL */
    const float* uv;      /* 2*nv or NULL */
    const uint32_t* idx;  /* 3*nt */
    uint32_t nv, nt;
} ora_mesh;

typedef struct {
    float o2w[12];        /* object->world, row-major 3x4 */
    float w2o[12];        /* world->object, ora_inverse3x4(o2w) */
    uint32_t mesh, material, mask, flags;
} ora_instance;           /* 112 bytes, same layout as aph_instance */

typedef struct {
    float org[3]; float tmin;
    float dir[3]; float tmax;
} ora_ray;                /* 32 bytes, same layout as aph_ray */

typedef struct {
    float t; uint32_t inst; uint32_t prim; float u; float v;
} ora_hit;                /* 20 bytes; inst = 0xFFFFFFFF on miss */

typedef struct ora_scene ora_scene;

ora_scene* ora_scene_create(const ora_mesh* meshes, uint32_t nmesh,
                            const ora_instance* inst, uint32_t ninst);
void ora_scene_destroy(ora_scene* s);
/* Material layers per instance (as aph_tlas_set_layers): bottom to top. */
void ora_scene_set_layers(ora_scene* s, const uint32_t* offsets, const uint32_t* materials);
/* Per layer (indexed as the materials above): the texture coordinates of the layer's own
 * geometry, 2 per vertex of its instance's mesh (copied), or NULL for the instance's. */
void ora_scene_set_layer_uvs(ora_scene* s, const float* const* uvs);
void ora_trace_closest(const ora_scene* s, const ora_ray* rays, uint32_t n, ora_hit* hits, int nthreads);
void ora_trace_closest_brute(const ora_scene* s, const ora_ray* rays, uint32_t n, ora_hit* hits);
/* SAH cost of a full-sweep SAH tree over n boxes (lo/hi float[3n]); checker for the BVH builders. */
double ora_sah_sweep_cost(const float* lo, const float* hi, uint32_t n, uint32_t maxLeaf, float costNode,
                          float costTri);
void ora_trace_any(const ora_scene* s, const ora_ray* rays, uint32_t n, uint8_t* occluded, int nthreads);
void ora_trace_any_brute(const ora_scene* s, const ora_ray* rays, uint32_t n, uint8_t* occluded);
/* single triangle, FP64 watertight test; returns 1 on hit */
/* Accepted and owned by this triangle (shared edges and vertices count once: transmitting shadow rays). */
int ora_triangle_hit_owned(const float org[3], const float dir[3], const float v0[3], const float v1[3],
                           const float v2[3]);
int ora_intersect_triangle(const float org[3], const float dir[3], float tmin, float tmax,
                           const float v0[3], const float v1[3], const float v2[3],
                           double* t, double* u, double* v);

/* ---- materials: 400-byte MaterialConstants (DefaultMaterialUniformBuffer.slang:3-72) ---- */
typedef struct {
    float base; float baseColor[3]; float diffuseRoughness; float metalness; float specular; int32_t _p0;
    float specularColor[3]; float specularRoughness; float specularIOR; float specularAnisotropy;
    float specularRotation; float transmission; float transmissionColor[3]; float subsurface;
    float subsurfaceColor[3]; int32_t _p1; float subsurfaceRadius[3]; float subsurfaceScale;
    float subsurfaceAnisotropy; float sheen; int32_t _p2, _p3; float sheenColor[3]; float sheenRoughness;
    float coat; float coatColor[3]; float coatRoughness; float coatAnisotropy; float coatRotation;
    float coatIOR; float coatAffectColor; float coatAffectRoughness; float emission; int32_t _p4;
    float emissionColor[3]; int32_t _p5; float opacity[3]; int32_t thinWalled;
    int32_t tex_block[44]; /* texture flags/transforms (offsets 224..399), unused by the oracle */
} ora_material_constants;
void ora_material_defaults(ora_material_constants* m); /* MaterialBase.cpp:109-167 */

/* ---- ground plane (DirectX: Shaders/GroundPlane.slang:24-290, PTGroundPlane.cpp) ---- */
typedef struct {
    int32_t enabled;
    float position[3];
    float normal[3];
    float shadow_opacity;
    float shadow_color[3];
    float reflection_opacity;
    float reflection_color[3];
    float reflection_roughness;
} ora_ground_plane;  /* same layout as aph_ground_plane */
/* The scene's ground plane (NULL or enabled = 0: none), evaluated on primary rays. */
void ora_scene_set_ground_plane(ora_scene* s, const ora_ground_plane* gp);
/* intersectGroundPlane (GroundPlane.slang:104-128): 1 and the hit when the ray meets the plane's front. */
int ora_ground_plane_hit(const ora_ground_plane* gp, const float org[3], const float dir[3], float pos[3],
                         float* dist);
/* sampleGGXSmithBRDF (BSDFCommon.slang:331-367) with anisotropy 0, in tangent space. */
void ora_ggx_smith_sample(const float V[3], const float f0[3], float roughness, const float rnd[2], float L[3],
                          float* pdf, float out[3]);

/* ---- material textures (DefaultMaterial.slang:84-233; HGISampler.cpp:22-60) ---- */
#define ORA_TEX_SLOTS 5 /* base colour, specular roughness, emission colour, opacity, normal */
typedef struct {
    const float* rgba;       /* linear RGBA float, w*h*4, row-major */
    uint32_t w, h;
    int32_t addr_u, addr_v;  /* 0 wrap, 1 mirror, 2 clamp, 3 mirror once, 4 border (opaque white) */
} ora_texture;
float ora_srgb_to_linear(float c);
void ora_uv_transform(const float uv[2], const float xf[7], float out[2]);
void ora_texture_sample(const ora_texture* t, const float uv[2], float out[4]);
void ora_normal_from_map(const float texel[3], const float N[3], const float T[3], float out[3]);
/* Interpolated texture coordinate of a hit (tests). */
void ora_hit_uv(const ora_scene* s, uint32_t inst, uint32_t prim, float bu, float bv, float out[2]);
/* slot_tex[5*m + slot]: texture index of material m's slot or -1; textures are copied. */
void ora_scene_set_textures(ora_scene* s, const ora_texture* tex, uint32_t ntex, const int32_t* slot_tex,
                            uint32_t nmat);

/* ---- environment (EnvironmentBase.cpp:52-71, Environment.slang) ---- */
typedef struct { uint32_t alias; float prob; float pdf; float pad; } ora_alias_entry;
void ora_alias_build(const float* rgb, uint32_t w, uint32_t h, ora_alias_entry* out, float* integral);

typedef struct {
    float light_top[3], light_bottom[3];
    float bg_top[3], bg_bottom[3];
    float light_xform[16];       /* row-major 4x4 applied as M*v */
    float light_xform_inv[16];
    float bg_xform[16];
    int32_t bg_use_screen;
    int32_t filter;              /* 0 = nearest (HGI default sampler), 1 = bilinear */
    const float* light_rgb;      /* RGB float, w*h*3, or NULL */
    uint32_t light_w, light_h;
    const ora_alias_entry* alias;
    float light_integral;
    const float* bg_rgb;         /* RGB float or NULL */
    uint32_t bg_w, bg_h;
} ora_env;

void ora_env_evaluate(const ora_env* env, const float dir[3], int as_background,
                      uint32_t px, uint32_t py, uint32_t W, uint32_t H, float out[3]);
void ora_env_sample(const ora_env* env, const float rnd[2], float L[3], float* pdf, float out[3]);
void ora_dir_to_latlong(const float dir[3], float uv[2]);
void ora_latlong_to_dir(const float uv[2], float dir[3]);

/* ---- BSDF (StandardSurfaceBSDF.slang) on an explicit shading frame ---- */
typedef struct { float N[3], X[3], Y[3]; } ora_frame;
void ora_bsdf_evaluate(const ora_material_constants* mc, const ora_frame* f,
                       const float V[3], const float L[3], float out[3]);
void ora_bsdf_evaluate_pdf(const ora_material_constants* mc, const ora_frame* f,
                           const float V[3], const float L[3], const float rnd[2],
                           float out[3], float* pdf, int* lobe);
void ora_bsdf_sample(const ora_material_constants* mc, const ora_frame* f,
                     const float V[3], const float rnd[2],
                     float out[3], float L[3], float* pdf, int* lobe);
float ora_sheen_albedo(float cos_theta, float roughness);

/* ---- renderer frame parameters (RendererBase.cpp:225-295, Frame.slang:61-157) ---- */
typedef struct {
    float color_intensity[4];
    float direction[3];
    float cos_radius;
} ora_distant_light;

typedef struct {
    float inv_view[16];          /* rows: right, up, back, eye (transposed inverse view) */
    float view_size[2];
    float jitter[2];
    int32_t is_ortho;
    float focal_distance;
    float lens_radius;
    int32_t trace_depth;
    float max_luminance;
    int32_t rr_enabled;
    int32_t rr_start_depth;
    int32_t force_opaque_shadows;  /* HGI forces 1 (HGIRenderer.cpp:490-491) */
    int32_t display_errors;
    int32_t temporal_jitter;
    int32_t light_count;
    ora_distant_light lights[4];
} ora_frame_params;

typedef struct {
    uint64_t extension_rays;
    uint64_t shadow_rays;
} ora_counters;

/* computeCameraRay (RayTrace.slang:130-182) of pixel (x, y) at sample_index. */
void ora_camera_ray(const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                    uint32_t sample_index, float org[3], float dir[3]);

/* Renders sample_count samples into accum (RGBA float, W*H*4) starting at
 * sample_start, with HGI running-mean accumulation (Accumulation.glsl:13-34).
 * pixels: optional list of pixel indices (y*W+x) to render (others untouched),
 * npix = 0 means all pixels. */
void ora_render(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat,
                const ora_env* env, const ora_frame_params* fp, uint32_t W, uint32_t H,
                uint32_t sample_start, uint32_t sample_count,
                const uint32_t* pixels, uint32_t npix, int nthreads,
                float* accum, ora_counters* counters);

/* One pixel-sample: the RayGenShader body (MainEntryPoints.slang:70-334). */
void ora_pixel_sample(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat,
                      const ora_env* env, const ora_frame_params* fp, uint32_t W, uint32_t H,
                      uint32_t x, uint32_t y, uint32_t sample_index, float out[4],
                      ora_counters* counters);

/* Post-processing, PostProcessing.glsl:37-64. */
/* Temporal resolve (TemporalResolve.hlsl:156-249) over RGBA images. */
void ora_temporal_resolve(const float* accum, const float* depth, const float* motion, const float* hist_in,
                          float* hist_out, uint32_t W, uint32_t H, int is_reset, float alpha_min, float alpha_max,
                          float speed_reference);
float ora_view_z(const float inv_view[16], const float p[3]);
void ora_screen_motion(const float view_proj[16], const float view_proj_prev[16], const float p[4], uint32_t W,
                       uint32_t H, float out[2]);
void ora_halton_jitter(uint32_t frame_index, uint32_t phase_count, float out[2]);
/* Guides of the first segment of one camera ray (the AOV pass): out = depth,
 * normal.xyz, roughness, base colour.xyz, metalness, motion.xyzw (13 floats). */
void ora_aov_pixel(const ora_scene* s, const ora_material_constants* mats, uint32_t nmat,
                   const ora_frame_params* fp, uint32_t W, uint32_t H, uint32_t x, uint32_t y,
                   uint32_t sample_index, const float view_proj[16], const float view_proj_prev[16],
                   const float inv_view_prev[16], float sky_view_z, float out[13]);

void ora_post(const float* accum, uint32_t n, const float brightness[3], int tone_map, int gamma,
              float* out);

#ifdef __cplusplus
}
#endif
#endif
