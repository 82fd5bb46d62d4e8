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
 * pbrt_io_free. */
int32_t pbrt_image_load_rgba8(const char* path, uint32_t* width, uint32_t* height, uint8_t** pixels);
/* Releases memory returned by the readers of this header. */
void pbrt_io_free(void* memory);
/* 8-bit RGBA PNG of `pixels`; `bgra` != 0 when the pixels are B8G8R8A8 (the
 * tracer's final image format, BMFR.cpp:81). */
int32_t pbrt_image_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* pixels,
                             uint32_t bgra);

/* ------------------------------------------------------------------------- */
/* Offline frames (source/io/RenderIO.cpp): camera matrices, G-buffer and     */
/* illumination images.  The reference stores images as OpenEXR through      */
/* vsgXchange; this build uses PFM (portable float map) for the float images. */
/* ------------------------------------------------------------------------- */

/* CameraMatrices (source/io/RenderIO.hpp:27-33); column-major like vsg::mat4. */
typedef struct pbrt_camera_matrices {
    float view[16];     /* without a projection: the combined view-projection */
    float inv_view[16];
    float proj[16];     /* valid when has_proj */
    float inv_proj[16];
    uint32_t has_proj;  /* 1: "ModelView+Projection", 0: "ModelViewProjection" */
    uint32_t pad[3];
} pbrt_camera_matrices;

/* MatrixIO::import_matrices (RenderIO.cpp:593-666): a .json path holds
 * {"amtOfFrames": n, "matrices": [{"type", "storageType", "view", "invView",
 * "proj", "invProj"}]} with 16 numbers per matrix in column order; any other path
 * holds a plain list of 16-number matrices (commas and braces allowed), each a
 * combined matrix whose inverse is computed.  Free *out with pbrt_io_free. */
int32_t pbrt_matrices_read(const char* path, pbrt_camera_matrices** out, uint32_t* count);
/* MatrixIO::export_matrices (RenderIO.cpp:668-716): the JSON form. */
int32_t pbrt_matrices_write(const char* path, const pbrt_camera_matrices* matrices, uint32_t count);

/* PFM with 1 ("Pf") or 3 ("PF") channels, little-endian.  `pixels` holds
 * `stride` floats per pixel (>= channels), first row at the top; the file stores
 * the bottom row first. */
int32_t pbrt_pfm_write(const char* path, uint32_t width, uint32_t height, uint32_t channels, const float* pixels,
                       uint32_t stride);
/* Reads a PFM file into width * height * channels floats, first row at the top;
 * free with pbrt_io_free. */
int32_t pbrt_pfm_read(const char* path, uint32_t* width, uint32_t* height, uint32_t* channels, float** pixels);

/* GBufferIO conversions over n pixels (RenderIO.cpp), in the reference's FP32
 * arithmetic. */
/* convert_normal_to_spherical (:160-178): (x, y, z, w) -> (acos z, atan2(y, x)) */
int32_t pbrt_normals_to_spherical(const float* xyzw, uint64_t n, float* theta_phi);
/* spherical_to_cartesian (:312-329): -> (cos phi sin theta, sin phi sin theta, cos theta, 1) */
int32_t pbrt_spherical_to_normals(const float* theta_phi, uint64_t n, float* xyzw);
/* compress_albedo, float input (:183-189): 255 * value, truncated (clamped to 0..255) */
int32_t pbrt_albedo_from_float(const float* rgba, uint64_t n, uint8_t* rgba8);
/* compress_albedo, 16-bit float input (:197-209), with its half-to-float bit
 * conversion (no subnormal or infinity handling) */
int32_t pbrt_albedo_from_half(const uint16_t* rgba, uint64_t n, uint8_t* rgba8);
/* unorm_to_float (:331-346): value / 255 */
int32_t pbrt_unorm_to_float(const uint8_t* rgba8, uint64_t n, float* rgba);
/* depth_to_position (:348-382): the camera ray of each pixel centre (the tracer's
 * createRay) scaled by the depth, from the camera position; w = 1.  Needs has_proj. */
int32_t pbrt_depth_to_position(const float* depth, uint32_t width, uint32_t height,
                               const pbrt_camera_matrices* matrices, float* xyzw);
/* import_g_buffer_position (:96-110): depth = |camera - position|.  The camera is
 * the translation column of inv_view (the reference divides column 2, the view
 * axis, by its w, which is 0 for a view matrix: this build takes column 3). */
int32_t pbrt_position_to_depth(const float* xyzw, uint64_t n, const pbrt_camera_matrices* matrices, float* depth);

#ifdef __cplusplus
}
#endif

#endif /* PBRT_IO_H */
