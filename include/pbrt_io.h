/*
 * pbrt_io.h - scene front-end and image files for the frame path.
 *
 * The reference reads scenes through assimp (version 5.0/5.1, not vendored:
 * external/vsgXchange/src/assimp/build_vars.cmake:1-4) into a VSG scene graph
 * (external/vsgXchange/src/assimp/assimp.cpp:380-470, materials :540-747), which
 * RayTracingSceneDescriptorCreationVisitor packs for the tracer
 * (source/scene/RayTracingVisitor.cpp:29-560).  This header covers that path for
 * Wavefront OBJ/MTL and the image files its materials reference, producing the
 * input of pbrt_scene_create() (include/pbrt_c.h), whose packing applies the
 * visitor's per-mesh rules (area-weighted normals, emissive meshes as area
 * lights, opacity from diffuse alpha, the default sun).
 *
 * OBJ semantics are assimp's with the reference's import flags
 * (aiProcess_Triangulate | FlipUVs | OptimizeMeshes | SortByPType | GenUVCoords,
 * assimp.cpp:316):
 *   - every face corner becomes its own vertex (no vertex joining);
 *   - quads split at their reflex corner (else at corner 0), larger polygons
 *     as a fan from corner 0; points and lines are dropped;
 *   - v of every texture coordinate becomes 1 - v; corners without a texture
 *     coordinate get (0, 0), without a normal (0, 0, 0) (assimp.cpp:415-430);
 *   - one mesh per material, in order of first use (OptimizeMeshes merges the
 *     per-object meshes that share a material); indices are 16-bit when a mesh
 *     has fewer than 65535 of them (assimp.cpp:437-452);
 *   - faces before any usemtl, or naming an unknown material, use assimp's
 *     default material (Kd 0.6).
 * MTL keys: Ka Kd Ks Ke Tf Ns Ni d Tr illum, map_Kd map_Ks map_Ke (map_emissive),
 * map_Kn / norm (normal map); map_bump / bump are height maps, which the
 * reference does not bind.  Materials convert as vsgXchange's Phong path
 * (assimp.cpp:672-745) then RayTracingVisitor.cpp:279-298:
 *   Ns < 0.01 -> shininess 0 and specular (0,0,0,0); roughness sqrt(2/(Ns+2));
 *   Tf != (1,1,1) -> illum 7 (refraction); emission Ke; ior Ni; alpha mask 1,
 *   cutoff 0.5.
 *
 * Images decode to RGBA8, first row at the top: PNG (8/16-bit, gray, gray+alpha,
 * RGB, RGBA, palette with tRNS; not interlaced), TGA (true-colour and grey,
 * raw or RLE), PPM/PGM (P5/P6) and PAM (P7).  JPEG is not decoded: this build
 * ships no decoder for it (the reference uses stb_image through vsgXchange).
 */
#ifndef PBRT_IO_H
#define PBRT_IO_H

#include "pbrt_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* vsg::PhongMaterial (external/vsg/include/vsg/state/material.h:51-62), 96 bytes */
typedef struct pbrt_phong_material {
    float ambient[4];
    float diffuse[4];
    float specular[4];
    float emissive[4];
    float transmissive[3];
    float shininess;
    float alpha_mask;
    float alpha_mask_cutoff;
    float index_of_refraction;
    uint32_t category_id;
} pbrt_phong_material;

/* vsg::PbrMaterial (external/vsg/include/vsg/state/material.h:97-109), 100 bytes */
typedef struct pbrt_pbr_material {
    float base_color[4];
    float emissive[4];
    float diffuse[4];
    float specular[4];
    float transmission[3];
    float metallic;
    float roughness;
    float alpha_mask;
    float alpha_mask_cutoff;
    float index_of_refraction;
    uint32_t category_id;
} pbrt_pbr_material;

/* The packed tracer material of a Phong / PBR material and whether the mesh is
 * emissive (RayTracingVisitor.cpp:253-298).  `emissive` may be NULL. */
int32_t pbrt_pack_phong_material(const pbrt_phong_material* material, pbrt_material* out, uint32_t* emissive);
int32_t pbrt_pack_pbr_material(const pbrt_pbr_material* material, pbrt_material* out, uint32_t* emissive);

typedef struct pbrt_loaded_scene pbrt_loaded_scene;

typedef struct pbrt_load_options {
    uint32_t instance_count;          /* 0: the file once, at identity */
    uint32_t bvh_builder;             /* enum pbrt_bvh_builder written to the description */
    const float* instance_transforms; /* instance_count column-major 4x4 object-to-world */
} pbrt_load_options;

typedef struct pbrt_loaded_scene_stats {
    uint32_t mesh_count;
    uint32_t triangle_count;      /* per instance of the file */
    uint32_t vertex_count;
    uint32_t texture_count;
    uint32_t emissive_mesh_count;
    uint32_t instance_count;
    float bounds_min[3];          /* object-space bounds of the file's vertices */
    float bounds_max[3];
    double parse_ms;              /* host time of the OBJ, MTL and image reads */
    uint64_t file_bytes;          /* bytes of OBJ + MTL + image files read */
} pbrt_loaded_scene_stats;

/* Reads `path` (OBJ) with its MTL libraries and textures, relative paths against
 * the OBJ's folder.  `options` may be NULL. */
int32_t pbrt_load_obj(const char* path, const pbrt_load_options* options, pbrt_loaded_scene** out_scene);
/* The scene description, valid until the loaded scene is destroyed. */
int32_t pbrt_loaded_scene_desc(const pbrt_loaded_scene* scene, pbrt_scene_desc* out_desc);
int32_t pbrt_loaded_scene_stats_get(const pbrt_loaded_scene* scene, pbrt_loaded_scene_stats* out_stats);
int32_t pbrt_loaded_scene_destroy(pbrt_loaded_scene* scene);

/* An image file as RGBA8 (first row at the top); free the pixels with
 * pbrt_image_free. */
int32_t pbrt_image_load_rgba8(const char* path, uint32_t* width, uint32_t* height, uint8_t** pixels);
void pbrt_image_free(uint8_t* pixels);
/* 8-bit RGBA PNG of `pixels`; `bgra` != 0 when the pixels are B8G8R8A8 (the
 * tracer's final image format, BMFR.cpp:81). */
int32_t pbrt_image_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* pixels,
                             uint32_t bgra);

#ifdef __cplusplus
}
#endif

#endif /* PBRT_IO_H */
