/* Oracle: scene, FP64 watertight intersection, CPU BVH (test infrastructure only).
 *
 * The reference delegates traversal and intersection to the ray-tracing driver
 * (Shaders/RayTrace.slang:185-210,221-270 call TraceRay; BLAS/TLAS are built by
 * HGI/HGIGeometry.cpp:127-160 and HGI/HGIScene.cpp:377-420).  The exact hit
 * semantics are defined here instead and shared with the GPU path:
 *
 *   - world ray (float) -> object ray: o' = W2O*(o,1), d' = W2O*(d,0) evaluated
 *     in double (W2O = ora_inverse3x4(O2W) as float, products summed left to
 *     right) and rounded to float; every test of that instance uses this float
 *     object ray (exactly representable in double);
 *   - watertight test of Woop, Benthin & Wald (JCGT 2(1), 2013) evaluated in
 *     double with no FMA, edges accepted when U,V,W share a sign (zeros allowed),
 *     det == 0 rejected, t = T/det accepted when tmin <= t <= tmax;
 *   - closest hit = lexicographic minimum of (t, instance, primitive);
 *   - a ray with a non-finite origin or direction component misses;
 *   - barycentrics (u, v) = (V/det, W/det) are the weights of vertices 1 and 2.
 */
#include "ora.h"
#include "ora_geom.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- watertight ray / triangle (double) ---------------- */

void ora_wray_setup(ora_wray* r, const double o[3], const double d[3], double tmin, double tmax)
{
    for (int i = 0; i < 3; i++) { r->o[i] = o[i]; r->d[i] = d[i]; }
    r->tmin = tmin;
    r->tmax = tmax;
    double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
    int kz;
    if (ax >= ay && ax >= az) kz = 0;
    else if (ay >= az) kz = 1;
    else kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    /* one division: the shear constants are d[kx] and d[ky] scaled by the rounded 1 / d[kz] */
    r->Sz = 1.0 / d[kz];
    r->Sx = d[kx] * r->Sz;
    r->Sy = d[ky] * r->Sz;
    for (int i = 0; i < 3; i++) {
        r->dz[i] = d[i] == 0.0;
        r->inv[i] = r->dz[i] ? 0.0 : 1.0 / d[i];
    }
}

int ora_wray_triangle(const ora_wray* r, const float* p0, const float* p1, const float* p2,
                      double* t_out, double* u_out, double* v_out)
{
    int kx = r->kx, ky = r->ky, kz = r->kz;
    double A[3], B[3], C[3];
    for (int i = 0; i < 3; i++) {
        A[i] = (double)p0[i] - r->o[i];
        B[i] = (double)p1[i] - r->o[i];
        C[i] = (double)p2[i] - r->o[i];
    }
    double Ax = A[kx] - r->Sx * A[kz], Ay = A[ky] - r->Sy * A[kz];
    double Bx = B[kx] - r->Sx * B[kz], By = B[ky] - r->Sy * B[kz];
    double Cx = C[kx] - r->Sx * C[kz], Cy = C[ky] - r->Sy * C[kz];
    double U = Cx * By - Cy * Bx;
    double V = Ax * Cy - Ay * Cx;
    double W = Bx * Ay - By * Ax;
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return 0;
    double det = U + V + W;
    if (det == 0.0) return 0;
    double Az = r->Sz * A[kz], Bz = r->Sz * B[kz], Cz = r->Sz * C[kz];
    double T = U * Az + V * Bz + W * Cz;
    double t = T / det;
    if (!(t >= r->tmin && t <= r->tmax)) return 0;
    *t_out = t;
    *u_out = V / det;
    *v_out = W / det;
    return 1;
}

/* Whether an accepted hit belongs to this triangle when it lies exactly on an edge or
 * vertex (an edge function of zero): of the triangles sharing that edge only one counts
 * it, as ray-tracing hardware reports shared edges once.  The edge's 2D direction in the
 * sheared ray frame, in the triangle's winding (times the sign of det), must point to -y,
 * or to +x when horizontal (the top-left rule). */
static int edge_owns(double ex, double ey) { return ey < 0.0 || (ey == 0.0 && ex > 0.0); }

int ora_wray_hit_owned(const ora_wray* r, const float* p0, const float* p1, const float* p2)
{
    int kx = r->kx, ky = r->ky, kz = r->kz;
    double A[3], B[3], C[3];
    for (int i = 0; i < 3; i++) {
        A[i] = (double)p0[i] - r->o[i];
        B[i] = (double)p1[i] - r->o[i];
        C[i] = (double)p2[i] - r->o[i];
    }
    double Ax = A[kx] - r->Sx * A[kz], Ay = A[ky] - r->Sy * A[kz];
    double Bx = B[kx] - r->Sx * B[kz], By = B[ky] - r->Sy * B[kz];
    double Cx = C[kx] - r->Sx * C[kz], Cy = C[ky] - r->Sy * C[kz];
    double U = Cx * By - Cy * Bx;
    double V = Ax * Cy - Ay * Cx;
    double W = Bx * Ay - By * Ax;
    double s = (U + V + W) > 0.0 ? 1.0 : -1.0;
    if (U == 0.0 && !edge_owns(s * (Cx - Bx), s * (Cy - By))) return 0;
    if (V == 0.0 && !edge_owns(s * (Ax - Cx), s * (Ay - Cy))) return 0;
    if (W == 0.0 && !edge_owns(s * (Bx - Ax), s * (By - Ay))) return 0;
    return 1;
}

int ora_intersect_triangle(const float org[3], const float dir[3], float tmin, float tmax,
                           const float v0[3], const float v1[3], const float v2[3],
                           double* t, double* u, double* v)
{
    double o[3] = { org[0], org[1], org[2] }, d[3] = { dir[0], dir[1], dir[2] };
    ora_wray r;
    ora_wray_setup(&r, o, d, (double)tmin, (double)tmax);
    return ora_wray_triangle(&r, v0, v1, v2, t, u, v);
}

/* The hit of one triangle counted by a transmitting shadow ray: accepted and owned. */
int ora_triangle_hit_owned(const float org[3], const float dir[3], const float v0[3], const float v1[3],
                           const float v2[3])
{
    double o[3] = { org[0], org[1], org[2] }, d[3] = { dir[0], dir[1], dir[2] };
    ora_wray r;
    ora_wray_setup(&r, o, d, 0.0, INFINITY);
    double t, u, v;
    return ora_wray_triangle(&r, v0, v1, v2, &t, &u, &v) && ora_wray_hit_owned(&r, v0, v1, v2);
}

/* world -> object ray: double evaluation, fixed summation order, rounded to float */
void ora_object_ray(const float* w2o, const float org[3], const float dir[3], double o[3], double d[3])
{
    double ox = org[0], oy = org[1], oz = org[2];
    double dx = dir[0], dy = dir[1], dz = dir[2];
    for (int i = 0; i < 3; i++) {
        const float* m = w2o + 4 * i;
        o[i] = (double)(float)((((double)m[0] * ox + (double)m[1] * oy) + (double)m[2] * oz) + (double)m[3]);
        d[i] = (double)(float)(((double)m[0] * dx + (double)m[1] * dy) + (double)m[2] * dz);
    }
}

/* ---------------- CPU BVH (top-down median split) ---------------- */

typedef struct { float lo[3], hi[3]; float c[3]; uint32_t id; } prim_ref;

static int g_axis;
static int cmp_axis(const void* a, const void* b)
{
    float ca = ((const prim_ref*)a)->c[g_axis], cb = ((const prim_ref*)b)->c[g_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    uint32_t ia = ((const prim_ref*)a)->id, ib = ((const prim_ref*)b)->id;
    return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

static uint32_t build_rec(ora_bvh* b, prim_ref* refs, uint32_t first, uint32_t count)
{
    uint32_t ni = b->nnodes++;
    ora_bvh_node* n = &b->nodes[ni];
    float lo[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, hi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    float clo[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, chi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    for (uint32_t i = first; i < first + count; i++) {
        for (int k = 0; k < 3; k++) {
            if (refs[i].lo[k] < lo[k]) lo[k] = refs[i].lo[k];
            if (refs[i].hi[k] > hi[k]) hi[k] = refs[i].hi[k];
            if (refs[i].c[k] < clo[k]) clo[k] = refs[i].c[k];
            if (refs[i].c[k] > chi[k]) chi[k] = refs[i].c[k];
        }
    }
    for (int k = 0; k < 3; k++) {
        n->lo[k] = lo[k];
        n->hi[k] = hi[k];
    }
    if (count <= 4) {
        n->leaf = 1;
        n->first = first;
        n->count = count;
        return ni;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    if (chi[1] - clo[1] > ext) { axis = 1; ext = chi[1] - clo[1]; }
    if (chi[2] - clo[2] > ext) { axis = 2; }
    g_axis = axis;
    qsort(refs + first, count, sizeof(prim_ref), cmp_axis);
    uint32_t half = count / 2;
    n->leaf = 0;
    uint32_t l = build_rec(b, refs, first, half);
    uint32_t r = build_rec(b, refs, first + half, count - half);
    n = &b->nodes[ni];
    n->left = l;
    n->right = r;
    return ni;
}

static void bvh_build(ora_bvh* b, prim_ref* refs, uint32_t count)
{
    b->nodes = (ora_bvh_node*)calloc(2 * (size_t)count + 1, sizeof(ora_bvh_node));
    b->nnodes = 0;
    b->prims = (uint32_t*)malloc(sizeof(uint32_t) * (count ? count : 1));
    if (count == 0) return;
    build_rec(b, refs, 0, count);
    for (uint32_t i = 0; i < count; i++) b->prims[i] = refs[i].id;
}

ora_scene* ora_scene_create(const ora_mesh* meshes, uint32_t nmesh, const ora_instance* inst, uint32_t ninst)
{
    ora_scene* s = (ora_scene*)calloc(1, sizeof(ora_scene));
    s->nmesh = nmesh;
    s->ninst = ninst;
    s->meshes = (ora_mesh*)malloc(sizeof(ora_mesh) * (nmesh ? nmesh : 1));
    s->inst = (ora_instance*)malloc(sizeof(ora_instance) * (ninst ? ninst : 1));
    s->blas = (ora_bvh*)calloc(nmesh ? nmesh : 1, sizeof(ora_bvh));
    memcpy(s->meshes, meshes, sizeof(ora_mesh) * nmesh);
    memcpy(s->inst, inst, sizeof(ora_instance) * ninst);
    /* texture coordinates: without them u = v = i / vertex count (GeometryBase.cpp:114-124) */
    s->uv_gen = (float**)calloc(nmesh ? nmesh : 1, sizeof(float*));
    for (uint32_t m = 0; m < nmesh; m++) {
        if (s->meshes[m].uv) continue;
        uint32_t nv = s->meshes[m].nv;
        float* g = (float*)malloc(sizeof(float) * 2 * (nv ? nv : 1));
        for (uint32_t i = 0; i < nv; i++) g[2 * i] = g[2 * i + 1] = (float)i / (float)nv;
        s->uv_gen[m] = g;
        s->meshes[m].uv = g;
    }
    for (uint32_t m = 0; m < nmesh; m++) {
        const ora_mesh* me = &meshes[m];
        prim_ref* refs = (prim_ref*)malloc(sizeof(prim_ref) * (me->nt ? me->nt : 1));
        for (uint32_t t = 0; t < me->nt; t++) {
            prim_ref* r = &refs[t];
            r->id = t;
            for (int k = 0; k < 3; k++) { r->lo[k] = FLT_MAX; r->hi[k] = -FLT_MAX; }
            for (int v = 0; v < 3; v++) {
                const float* p = me->pos + 3 * (size_t)me->idx[3 * (size_t)t + v];
                for (int k = 0; k < 3; k++) {
                    if (p[k] < r->lo[k]) r->lo[k] = p[k];
                    if (p[k] > r->hi[k]) r->hi[k] = p[k];
                }
            }
            for (int k = 0; k < 3; k++) r->c[k] = 0.5f * (r->lo[k] + r->hi[k]);
        }
        bvh_build(&s->blas[m], refs, me->nt);
        free(refs);
    }
    /* instance-level BVH over world boxes of the transformed BLAS roots */
    prim_ref* refs = (prim_ref*)malloc(sizeof(prim_ref) * (ninst ? ninst : 1));
    uint32_t nref = 0;
    for (uint32_t i = 0; i < ninst; i++) {
        const ora_instance* in = &inst[i];
        const ora_bvh* b = &s->blas[in->mesh];
        if (b->nnodes == 0) continue;
        prim_ref* r = &refs[nref++];
        r->id = i;
        for (int k = 0; k < 3; k++) { r->lo[k] = FLT_MAX; r->hi[k] = -FLT_MAX; }
        const ora_bvh_node* root = &b->nodes[0];
        for (int c = 0; c < 8; c++) {
            double p[3] = { (c & 1) ? root->hi[0] : root->lo[0], (c & 2) ? root->hi[1] : root->lo[1],
                            (c & 4) ? root->hi[2] : root->lo[2] };
            for (int k = 0; k < 3; k++) {
                const float* m = in->o2w + 4 * k;
                double w = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
                float wl = (float)w, wh = (float)w;
                wl = nextafterf(wl, -FLT_MAX);
                wh = nextafterf(wh, FLT_MAX);
                if (wl < r->lo[k]) r->lo[k] = wl;
                if (wh > r->hi[k]) r->hi[k] = wh;
            }
        }
        for (int k = 0; k < 3; k++) r->c[k] = 0.5f * (r->lo[k] + r->hi[k]);
    }
    bvh_build(&s->tlas, refs, nref);
    s->extent = 0.0;
    for (uint32_t i = 0; i < nref; i++)
        for (int k = 0; k < 3; k++) {
            if (fabs(refs[i].lo[k]) > s->extent) s->extent = fabs(refs[i].lo[k]);
            if (fabs(refs[i].hi[k]) > s->extent) s->extent = fabs(refs[i].hi[k]);
        }
    s->extent *= 3.0;
    free(refs);
    return s;
}

static void free_layer_uvs(ora_scene* s)
{
    if (!s->layer_uv) return;
    uint32_t total = s->layer_off ? s->layer_off[s->ninst] : 0;
    for (uint32_t j = 0; j < total; j++) free(s->layer_uv[j]);
    free(s->layer_uv);
    s->layer_uv = NULL;
}

void ora_scene_set_layer_uvs(ora_scene* s, const float* const* uvs)
{
    free_layer_uvs(s);
    if (!s->layer_off || !uvs) return;
    uint32_t total = s->layer_off[s->ninst];
    s->layer_uv = (float**)calloc(total ? total : 1, sizeof(float*));
    for (uint32_t i = 0; i < s->ninst; i++) {
        uint32_t nv = s->meshes[s->inst[i].mesh].nv;
        for (uint32_t j = s->layer_off[i]; j < s->layer_off[i + 1]; j++) {
            if (!uvs[j]) continue;
            s->layer_uv[j] = (float*)malloc(sizeof(float) * 2 * (nv ? nv : 1));
            memcpy(s->layer_uv[j], uvs[j], sizeof(float) * 2 * nv);
        }
    }
}

void ora_scene_set_layers(ora_scene* s, const uint32_t* offsets, const uint32_t* materials)
{
    free_layer_uvs(s);
    free(s->layer_off);
    free(s->layer_mat);
    s->layer_off = (uint32_t*)malloc(sizeof(uint32_t) * (s->ninst + 1));
    memcpy(s->layer_off, offsets, sizeof(uint32_t) * (s->ninst + 1));
    uint32_t total = offsets[s->ninst];
    s->layer_mat = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
    if (total) memcpy(s->layer_mat, materials, sizeof(uint32_t) * total);
}

/* The plane's frame as PTGroundPlane.cpp builds it: start from +Y (or +X for a normal
 * near +-Y), tangent = normalize(cross(that, N)), bitangent = cross(N, tangent). */
void ora_scene_set_ground_plane(ora_scene* s, const ora_ground_plane* gp)
{
    memset(&s->ground, 0, sizeof(s->ground));
    if (!gp) return;
    s->ground = *gp;
    float n[3] = { gp->normal[0], gp->normal[1], gp->normal[2] };
    float b[3] = { 0.0f, 1.0f, 0.0f };
    if (!(fabsf(n[1]) < 0.999f)) { b[0] = 1.0f; b[1] = 0.0f; }
    float t[3] = { b[1] * n[2] - b[2] * n[1], b[2] * n[0] - b[0] * n[2], b[0] * n[1] - b[1] * n[0] };
    float l = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; k++) t[k] /= l;
    s->ground_t[0] = t[0]; s->ground_t[1] = t[1]; s->ground_t[2] = t[2];
    s->ground_b[0] = n[1] * t[2] - n[2] * t[1];
    s->ground_b[1] = n[2] * t[0] - n[0] * t[2];
    s->ground_b[2] = n[0] * t[1] - n[1] * t[0];
}

void ora_scene_destroy(ora_scene* s)
{
    if (!s) return;
    for (uint32_t m = 0; m < s->nmesh; m++) {
        free(s->blas[m].nodes);
        free(s->blas[m].prims);
    }
    free(s->tlas.nodes);
    free(s->tlas.prims);
    free_layer_uvs(s);
    free(s->layer_off);
    free(s->layer_mat);
    for (uint32_t m = 0; m < s->nmesh; m++) free(s->uv_gen[m]);
    free(s->uv_gen);
    for (uint32_t i = 0; i < s->ntex; i++) free((void*)s->tex[i].rgba);
    free(s->tex);
    free(s->slot_tex);
    free(s->blas);
    free(s->meshes);
    free(s->inst);
    free(s);
}

/* Slab test in double on the (float-valued) ray.  pad widens the box by an
 * absolute amount; the far distance is widened by a relative 1e-12, far above
 * the double rounding of the slab distances. */
static int box_hit(const ora_wray* r, const float lo[3], const float hi[3], double pad, double tmin, double tmax)
{
    for (int k = 0; k < 3; k++) {
        double l = (double)lo[k] - pad, h = (double)hi[k] + pad;
        if (r->dz[k]) {
            if (r->o[k] < l || r->o[k] > h) return 0;
            continue;
        }
        double t0 = (l - r->o[k]) * r->inv[k];
        double t1 = (h - r->o[k]) * r->inv[k];
        if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
    }
    return tmin <= tmax * (1.0 + 1e-12) + 1e-300;
}

static inline int better(double t, uint32_t inst, uint32_t prim, double bt, uint32_t binst, uint32_t bprim)
{
    if (t < bt) return 1;
    if (t > bt) return 0;
    if (inst != binst) return inst < binst;
    return prim < bprim;
}

typedef struct {
    double t, u, v;
    uint32_t inst, prim;
    int found;
    /* transparent any-hit (mode 2): per-instance transmittance (xyz, w = opaque) and the
     * visibility accumulated over every hit along the ray */
    const float* transmit;
    float vis[3];
} best_hit;

/* Closest (any_hit = 0), first found (any_hit = 1), or every hit with transmittance
 * (any_hit = 2: the shadow any-hit shader, HitShaderEntryPoints.slang:489-553) over one
 * instance's BLAS. */
static int blas_trace(const ora_scene* s, uint32_t ii, const ora_ray* ray, int any_hit, best_hit* best)
{
    const ora_instance* in = &s->inst[ii];
    const ora_mesh* me = &s->meshes[in->mesh];
    const ora_bvh* b = &s->blas[in->mesh];
    if (b->nnodes == 0) return 0;
    double o[3], d[3];
    ora_object_ray(in->w2o, ray->org, ray->dir, o, d);
    ora_wray wr;
    ora_wray_setup(&wr, o, d, (double)ray->tmin, (double)ray->tmax);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const ora_bvh_node* n = &b->nodes[stack[--sp]];
        double tlim = (best->found && any_hit != 2) ? best->t : wr.tmax;
        if (!box_hit(&wr, n->lo, n->hi, 0.0, wr.tmin, tlim)) continue;
        if (n->leaf) {
            for (uint32_t i = n->first; i < n->first + n->count; i++) {
                uint32_t prim = b->prims[i];
                const uint32_t* tri = me->idx + 3 * (size_t)prim;
                double t, u, v;
                if (!ora_wray_triangle(&wr, me->pos + 3 * (size_t)tri[0], me->pos + 3 * (size_t)tri[1],
                                       me->pos + 3 * (size_t)tri[2], &t, &u, &v))
                    continue;
                if (any_hit == 2) {
                    if (!ora_wray_hit_owned(&wr, me->pos + 3 * (size_t)tri[0], me->pos + 3 * (size_t)tri[1],
                                            me->pos + 3 * (size_t)tri[2]))
                        continue;
                    const float* tm = best->transmit + 4 * (size_t)ii;
                    if (tm[3] != 0.0f) {
                        best->vis[0] = best->vis[1] = best->vis[2] = 0.0f;
                        best->found = 1;
                        return 1;
                    }
                    for (int k = 0; k < 3; k++) best->vis[k] *= tm[k];
                    if (best->vis[0] == 0.0f && best->vis[1] == 0.0f && best->vis[2] == 0.0f) {
                        best->found = 1;
                        return 1;
                    }
                    continue;
                }
                if (!best->found || better(t, ii, prim, best->t, best->inst, best->prim)) {
                    best->found = 1; best->t = t; best->u = u; best->v = v; best->inst = ii; best->prim = prim;
                    if (any_hit) return 1;
                }
            }
        } else {
            stack[sp++] = n->left;
            stack[sp++] = n->right;
        }
    }
    return best->found;
}

static void trace_one(const ora_scene* s, const ora_ray* ray, int any_hit, best_hit* best)
{
    best->found = 0;
    if (s->tlas.nnodes == 0) return;
    /* non-finite rays miss (every triangle test would reject them) */
    for (int k = 0; k < 3; k++)
        if (!isfinite(ray->org[k]) || !isfinite(ray->dir[k])) return;
    double o[3] = { ray->org[0], ray->org[1], ray->org[2] }, d[3] = { ray->dir[0], ray->dir[1], ray->dir[2] };
    ora_wray wr;
    ora_wray_setup(&wr, o, d, (double)ray->tmin, (double)ray->tmax);
    /* Instance boxes are world boxes of the transformed BLAS roots, while the
     * triangle tests run on the rounded object-space ray: widen them by a pad
     * that bounds that rounding for any hit inside the scene. */
    double on = fabs(o[0]) + fabs(o[1]) + fabs(o[2]);
    double pad = 1e-4 * (on + s->extent);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const ora_bvh_node* n = &s->tlas.nodes[stack[--sp]];
        double tlim = (best->found && any_hit != 2) ? best->t : wr.tmax;
        if (!box_hit(&wr, n->lo, n->hi, pad, wr.tmin, tlim * (1.0 + 1e-6))) continue;
        if (n->leaf) {
            for (uint32_t i = n->first; i < n->first + n->count; i++) {
                if (blas_trace(s, s->tlas.prims[i], ray, any_hit, best) && any_hit) return;
            }
        } else {
            stack[sp++] = n->left;
            stack[sp++] = n->right;
        }
    }
}

void ora_trace_closest_one(const ora_scene* s, const ora_ray* ray, ora_hit* hit)
{
    best_hit b;
    trace_one(s, ray, 0, &b);
    if (b.found) {
        hit->t = (float)b.t; hit->inst = b.inst; hit->prim = b.prim; hit->u = (float)b.u; hit->v = (float)b.v;
    } else {
        hit->t = INFINITY; hit->inst = 0xFFFFFFFFu; hit->prim = 0xFFFFFFFFu; hit->u = 0.0f; hit->v = 0.0f;
    }
}

int ora_trace_any_one(const ora_scene* s, const ora_ray* ray)
{
    best_hit b;
    trace_one(s, ray, 1, &b);
    return b.found;
}

void ora_trace_transmit_one(const ora_scene* s, const float* transmit, const ora_ray* ray, float vis[3])
{
    best_hit b;
    b.transmit = transmit;
    b.vis[0] = b.vis[1] = b.vis[2] = 1.0f;
    trace_one(s, ray, 2, &b);
    for (int k = 0; k < 3; k++) vis[k] = b.found ? 0.0f : b.vis[k];
}

/* ---------------- threaded batch entry points ---------------- */

typedef struct {
    const ora_scene* s; const ora_ray* rays; uint32_t n; ora_hit* hits; uint8_t* occ;
    int tid, nthreads;
} trace_job;

static void* trace_worker(void* p)
{
    trace_job* j = (trace_job*)p;
    for (uint32_t i = (uint32_t)j->tid; i < j->n; i += (uint32_t)j->nthreads) {
        if (j->hits) ora_trace_closest_one(j->s, &j->rays[i], &j->hits[i]);
        else j->occ[i] = (uint8_t)ora_trace_any_one(j->s, &j->rays[i]);
    }
    return NULL;
}

static void run_trace(const ora_scene* s, const ora_ray* rays, uint32_t n, ora_hit* hits, uint8_t* occ, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    trace_job jobs[256];
    int started[256];
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setstacksize(&attr, (size_t)16 << 20);
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (trace_job){ s, rays, n, hits, occ, t, nthreads };
        started[t] = pthread_create(&th[t], &attr, trace_worker, &jobs[t]) == 0;
        if (!started[t]) trace_worker(&jobs[t]); /* no thread available: run this share inline */
    }
    pthread_attr_destroy(&attr);
    for (int t = 0; t < nthreads; t++)
        if (started[t]) pthread_join(th[t], NULL);
}

void ora_trace_closest(const ora_scene* s, const ora_ray* rays, uint32_t n, ora_hit* hits, int nthreads)
{
    run_trace(s, rays, n, hits, NULL, nthreads);
}

void ora_trace_any(const ora_scene* s, const ora_ray* rays, uint32_t n, uint8_t* occ, int nthreads)
{
    run_trace(s, rays, n, NULL, occ, nthreads);
}

/* Brute force: every instance x every triangle, no BVH. */
void ora_trace_closest_brute(const ora_scene* s, const ora_ray* rays, uint32_t n, ora_hit* hits)
{
    for (uint32_t r = 0; r < n; r++) {
        best_hit b;
        b.found = 0;
        for (uint32_t ii = 0; ii < s->ninst; ii++) {
            const ora_instance* in = &s->inst[ii];
            const ora_mesh* me = &s->meshes[in->mesh];
            double o[3], d[3];
            ora_object_ray(in->w2o, rays[r].org, rays[r].dir, o, d);
            ora_wray wr;
            ora_wray_setup(&wr, o, d, (double)rays[r].tmin, (double)rays[r].tmax);
            for (uint32_t t = 0; t < me->nt; t++) {
                const uint32_t* tri = me->idx + 3 * (size_t)t;
                double tt, u, v;
                if (!ora_wray_triangle(&wr, me->pos + 3 * (size_t)tri[0], me->pos + 3 * (size_t)tri[1],
                                       me->pos + 3 * (size_t)tri[2], &tt, &u, &v))
                    continue;
                if (!b.found || better(tt, ii, t, b.t, b.inst, b.prim)) {
                    b.found = 1; b.t = tt; b.u = u; b.v = v; b.inst = ii; b.prim = t;
                }
            }
        }
        if (b.found) {
            hits[r].t = (float)b.t; hits[r].inst = b.inst; hits[r].prim = b.prim;
            hits[r].u = (float)b.u; hits[r].v = (float)b.v;
        } else {
            hits[r].t = INFINITY; hits[r].inst = 0xFFFFFFFFu; hits[r].prim = 0xFFFFFFFFu;
            hits[r].u = 0.0f; hits[r].v = 0.0f;
        }
    }
}

void ora_trace_any_brute(const ora_scene* s, const ora_ray* rays, uint32_t n, uint8_t* occ)
{
    for (uint32_t r = 0; r < n; r++) {
        occ[r] = 0;
        for (uint32_t ii = 0; ii < s->ninst && !occ[r]; ii++) {
            const ora_instance* in = &s->inst[ii];
            const ora_mesh* me = &s->meshes[in->mesh];
            double o[3], d[3];
            ora_object_ray(in->w2o, rays[r].org, rays[r].dir, o, d);
            ora_wray wr;
            ora_wray_setup(&wr, o, d, (double)rays[r].tmin, (double)rays[r].tmax);
            for (uint32_t t = 0; t < me->nt; t++) {
                const uint32_t* tri = me->idx + 3 * (size_t)t;
                double tt, u, v;
                if (ora_wray_triangle(&wr, me->pos + 3 * (size_t)tri[0], me->pos + 3 * (size_t)tri[1],
                                      me->pos + 3 * (size_t)tri[2], &tt, &u, &v)) {
                    occ[r] = 1;
                    break;
                }
            }
        }
    }
}

/* ---------------- SAH reference cost (checker for the GPU builders) ----------------
 * No reference function: the reference's BVH is built inside the driver.  A
 * top-down full-sweep SAH build (MacDonald & Booth 1990; Wald 2007 section 3:
 * every centroid-sorted split position on all three axes) returning the tree's
 * SAH cost  sum_internal A(n)/A(root) * costNode + sum_leaves A(l)/A(root) *
 * costTri * count.  Leaf rule as in the GPU builder: a non-root node of at most
 * maxLeaf primitives becomes a leaf when costTri * count <= costNode + costTri *
 * (A_L N_L + A_R N_R) / A. */
typedef struct { float lo[3], hi[3], c[3]; } sah_prim;
static const sah_prim* g_sah_prims;
static int g_sah_axis_sort;

static int sah_cmp(const void* a, const void* b)
{
    uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
    float ci = g_sah_prims[i].c[g_sah_axis_sort], cj = g_sah_prims[j].c[g_sah_axis_sort];
    if (ci < cj) return -1;
    if (ci > cj) return 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

static double sah_area(const float lo[3], const float hi[3])
{
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static void sah_grow(float lo[3], float hi[3], const sah_prim* p)
{
    for (int k = 0; k < 3; k++) {
        if (p->lo[k] < lo[k]) lo[k] = p->lo[k];
        if (p->hi[k] > hi[k]) hi[k] = p->hi[k];
    }
}

static double sah_rec(uint32_t* idx, uint32_t n, uint32_t maxLeaf, double cn, double ct, double rootA, int root,
                      double* rightA)
{
    float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (uint32_t i = 0; i < n; i++) sah_grow(lo, hi, &g_sah_prims[idx[i]]);
    double A = sah_area(lo, hi);
    double leafCost = ct * (double)n;
    if (n == 1) return A / rootA * leafCost;
    double best = INFINITY;
    int bestAxis = 0;
    uint32_t bestK = n / 2;
    for (int a = 0; a < 3; a++) {
        g_sah_axis_sort = a;
        qsort(idx, n, sizeof(uint32_t), sah_cmp);
        float rl[3] = { INFINITY, INFINITY, INFINITY }, rh[3] = { -INFINITY, -INFINITY, -INFINITY };
        for (uint32_t i = n; i-- > 1;) {
            sah_grow(rl, rh, &g_sah_prims[idx[i]]);
            rightA[i] = sah_area(rl, rh);
        }
        float ll[3] = { INFINITY, INFINITY, INFINITY }, lh[3] = { -INFINITY, -INFINITY, -INFINITY };
        for (uint32_t k = 1; k < n; k++) {
            sah_grow(ll, lh, &g_sah_prims[idx[k - 1]]);
            double c = sah_area(ll, lh) * (double)k + rightA[k] * (double)(n - k);
            if (c < best) { best = c; bestAxis = a; bestK = k; }
        }
    }
    double splitCost = A > 0.0 ? cn + ct * best / A : cn + leafCost;
    if (!root && n <= maxLeaf && leafCost <= splitCost) return A / rootA * leafCost;
    g_sah_axis_sort = bestAxis;
    qsort(idx, n, sizeof(uint32_t), sah_cmp);
    return A / rootA * cn + sah_rec(idx, bestK, maxLeaf, cn, ct, rootA, 0, rightA) +
           sah_rec(idx + bestK, n - bestK, maxLeaf, cn, ct, rootA, 0, rightA);
}

double ora_sah_sweep_cost(const float* lo, const float* hi, uint32_t n, uint32_t maxLeaf, float costNode,
                          float costTri)
{
    if (n == 0) return 0.0;
    sah_prim* p = (sah_prim*)malloc(sizeof(sah_prim) * n);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    double* rightA = (double*)malloc(sizeof(double) * (n + 1));
    float rlo[3] = { INFINITY, INFINITY, INFINITY }, rhi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            p[i].lo[k] = lo[3 * i + k];
            p[i].hi[k] = hi[3 * i + k];
            p[i].c[k] = 0.5f * (p[i].lo[k] + p[i].hi[k]);
        }
        idx[i] = i;
        sah_grow(rlo, rhi, &p[i]);
    }
    g_sah_prims = p;
    double rootA = sah_area(rlo, rhi);
    double cost = rootA > 0.0 ? sah_rec(idx, n, maxLeaf, costNode, costTri, rootA, 1, rightA) : 0.0;
    free(p);
    free(idx);
    free(rightA);
    return cost;
}
