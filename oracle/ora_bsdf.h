/* Oracle-internal material and shading types (test infrastructure only). */
#ifndef ORA_BSDF_H
#define ORA_BSDF_H

#include "ora.h"
#include "ora_math.h"

/* Material, Shaders/Material.slang:20-57 (fields used by the default material). */
typedef struct {
    float base; v3 baseColor; float diffuseRoughness; float metalness; v3 metalColor;
    float specular; v3 specularColor; float specularRoughness; float specularIOR;
    float specularAnisotropy; float specularRotation; float transmission; v3 transmissionColor;
    float subsurface; v3 subsurfaceColor; float sheen; v3 sheenColor; float sheenRoughness;
    float coat; v3 coatColor; float coatRoughness; float coatAnisotropy; float coatRotation;
    float coatIOR; float coatAffectColor; float coatAffectRoughness; float emission;
    v3 emissionColor; v3 opacity; int thinWalled;
} ora_material;

/* ShadingData fields used by the path, Shaders/Geometry.slang:44-104. */
typedef struct {
    v3 normal, position, geomPosition, tangent, bitangent;
    float texCoord[2];
} ora_shading;

void ora_material_from_constants(const ora_material_constants* c, ora_material* m);
v3 ora_eval_material(const ora_material* m, const ora_shading* sh, v3 Vw, v3 Lw);
v3 ora_eval_material_pdf(const ora_material* m, const ora_shading* sh, v3 Vw, v3 Lw, const float r[2],
                         float* pdf, int* lobe);
/* sampleGGXSmithBRDF (BSDFCommon.slang:331-367), anisotropy 0, tangent space. */
v3 ora_ggx_smith_brdf_sample(v3 V, v3 f0, float roughness, const float r[2], v3* L, float* pdf);
v3 ora_sample_material(const ora_material* m, const ora_shading* sh, v3 Vw, const float r[2],
                       v3* Lw, float* pdf, int* lobe);

#endif
