/* Oracle-internal scene and ray types (test infrastructure only). */
#ifndef ORA_GEOM_H
#define ORA_GEOM_H

#include "ora.h"

typedef struct {
    double o[3], d[3], inv[3];
    int dz[3];
    int kx, ky, kz;
    double Sx, Sy, Sz;
    double tmin, tmax;
} ora_wray;

typedef struct {
    float lo[3], hi[3];
    int leaf;
    uint32_t left, right, first, count;
} ora_bvh_node;

typedef struct {
    ora_bvh_node* nodes;
    uint32_t nnodes;
    uint32_t* prims;
} ora_bvh;

struct ora_scene {
    ora_mesh* meshes;
    uint32_t nmesh;
    ora_instance* inst;
    uint32_t ninst;
    ora_bvh* blas;
    ora_bvh tlas;
    double extent;   /* 3 x max |coordinate| of the instance world boxes */
    uint32_t* layer_off; /* material layers: instance i has layer_mat[layer_off[i] .. layer_off[i+1]) */
    uint32_t* layer_mat;
    float** layer_uv;    /* per layer: its geometry's texture coordinates (2 per vertex of the instance's
                            mesh), or NULL for the instance's; the array itself NULL: none set */
    float** uv_gen;      /* per mesh: generated texture coordinates (u = v = i / nv) or NULL */
    ora_texture* tex;    /* material textures (ora_scene_set_textures) */
    uint32_t ntex;
    int32_t* slot_tex;   /* ORA_TEX_SLOTS per material: texture index or -1 */
    uint32_t nmat_tex;
    ora_ground_plane ground;         /* enabled = 0: none */
    float ground_t[3], ground_b[3];  /* tangent and bitangent of the plane's normal */
};

/* Material texture slot at a texture coordinate (0 and return 0 when unbound). */
int ora_tex_flag(const ora_material_constants* mc, int slot);
int ora_tex_slot(const ora_scene* s, const ora_material_constants* mc, uint32_t mi, int slot, const float uv[2],
                 float out[3]);

void ora_wray_setup(ora_wray* r, const double o[3], const double d[3], double tmin, double tmax);
int ora_wray_triangle(const ora_wray* r, const float* p0, const float* p1, const float* p2,
                      double* t, double* u, double* v);
int ora_wray_hit_owned(const ora_wray* r, const float* p0, const float* p1, const float* p2);
void ora_object_ray(const float* w2o, const float org[3], const float dir[3], double o[3], double d[3]);
void ora_trace_closest_one(const ora_scene* s, const ora_ray* ray, ora_hit* hit);
int ora_trace_any_one(const ora_scene* s, const ora_ray* ray);
/* Shadow visibility through non-opaque instances: the product of the transmittances
 * (transmit: xyz per instance, w = 1 for an opaque instance) of every hit, or 0. */
void ora_trace_transmit_one(const ora_scene* s, const float* transmit, const ora_ray* ray, float vis[3]);

#endif
