/*
 * pbrt_c.h - C ABI of the MI355X (gfx950) frame path: BVH build, path tracing,
 * temporal accumulation, BMFR/BFR denoising, TAA and format conversion.
 *
 * Every entry point returns an int32_t status (PBRT_OK == 0, negative on error)
 * and never throws.  pbrt_last_error() returns the message of the last failure
 * on the calling thread.  Images are caller-owned device pointers, row-major,
 * tightly packed (pitch == width), pixel (x, y) at index y * width + x.
 * `stream` is a hipStream_t passed as an opaque pointer (NULL = null stream).
 *
 * The reference has no C ABI: its "FFI" is the Vulkan command buffer recorded by
 * the render modules in source/renderModules.  Each entry point below names the
 * reference pipeline/dispatch it replaces.
 */
#ifndef PBRT_C_H
#define PBRT_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBRT_ABI_VERSION 1

enum pbrt_status {
    PBRT_OK = 0,
    PBRT_ERR_INVALID_ARG = -1,
    PBRT_ERR_UNSUPPORTED = -2,
    PBRT_ERR_OUT_OF_MEMORY = -3,
    PBRT_ERR_HIP = -4,
    PBRT_ERR_INTERNAL = -5
};

/* ------------------------------------------------------------------------- */
/* Plain data contracts (std430 layouts of the reference's SSBOs)            */
/* ------------------------------------------------------------------------- */

/* RayTracingPushConstants (source/renderModules/PipelineStructs.hpp:6-13)
 * == layoutPTPushConstants.glsl:4-11.  Matrices are column-major (vsg/GLSL). */
typedef struct pbrt_push_constants {
    float view_inverse[16];
    float proj_inverse[16];
    float prev_view[16];
    uint32_t frame_number;
    uint32_t sample_number;
} pbrt_push_constants; /* 200 bytes */

/* vsg::Light::PackedLight (external/vsg/include/vsg/nodes/Light.h:23-37)
 * == struct Light in shaders/ptStructures.glsl:54-63.  v0_type[3] holds the
 * light type (1 directional, 2 point, 5 area), strengths[3] the inclusive
 * strength prefix sum written by scene packing. */
typedef struct pbrt_light {
    float v0_type[4];
    float v1_strength[4];
    float v2_angle[4];
    float dir_angle2[4];
    float col_ambient[4];
    float col_diffuse[4];
    float col_specular[4];
    float strengths[4];
} pbrt_light; /* 128 bytes */

/* WaveFrontMaterialPacked (source/scene/RayTracingVisitor.hpp:70-79)
 * == shaders/ptStructures.glsl:18-27. */
typedef struct pbrt_material {
    float ambient_roughness[4];
    float diffuse_ior[4];
    float specular_dissolve[4];
    float transmittance_illum[4];
    float emission_texture_id[4];
    uint32_t category_id;
    float pad[3];
} pbrt_material; /* 96 bytes */

/* RGBA8 UNORM texture, row-major, sampled LINEAR / REPEAT / mip 0
 * (external/vsg/include/vsg/state/Sampler.h:29-43). */
typedef struct pbrt_texture_desc {
    uint32_t width;
    uint32_t height;
    const uint8_t* rgba8;
} pbrt_texture_desc;

/* One mesh (a vsg::VertexIndexDraw plus its descriptor set).  A mesh's
 * material and texture set are indexed by the mesh id, as the reference's
 * closest-hit shader does (shaders/ptClosesthit.rchit:21-22,32,48,50,57,84). */
typedef struct pbrt_mesh_desc {
    uint32_t vertex_count;
    const float* positions; /* 3 * vertex_count */
    const float* normals;   /* 3 * vertex_count; a zero first normal requests area-weighted
                               normals (RayTracingVisitor.cpp:52-118); NULL == all zero */
    const float* uvs;       /* 2 * vertex_count, or NULL for the default per-triangle
                               uvs (RayTracingVisitor.cpp:122-159) */
    uint32_t index_count;   /* multiple of 3 */
    uint32_t index_stride;  /* 2 (uint16) or 4 (uint32) */
    const void* indices;
    pbrt_material material;
    int32_t diffuse_texture;  /* texture index or -1 for the 1x1 white default */
    int32_t normal_texture;
    int32_t specular_texture;
    int32_t emissive_texture;
} pbrt_mesh_desc;

/* One TLAS instance: object-to-world transform (column-major 4x4) and mesh. */
typedef struct pbrt_instance_desc {
    float transform[16];
    uint32_t mesh;
    uint32_t pad[3];
} pbrt_instance_desc;

enum pbrt_bvh_builder {
    PBRT_BVH_SAH_HOST = 0,  /* binned SAH on the host */
    PBRT_BVH_LBVH_GPU = 1,  /* Morton-code LBVH built by gfx950 kernels */
    PBRT_BVH_SBVH_HOST = 2, /* binned SAH with spatial splits on the host: fewer node visits per
                               ray, a few percent more triangle references, slower build */
    PBRT_BVH_PLOC_GPU = 3   /* parallel locally-ordered clustering (nearest-neighbour merges
                               along the Morton order) with a SAH leaf collapse, built by gfx950
                               kernels: SAH-quality trees in milliseconds */
};

typedef struct pbrt_scene_desc {
    uint32_t mesh_count;
    const pbrt_mesh_desc* meshes;
    uint32_t instance_count;
    const pbrt_instance_desc* instances;
    uint32_t texture_count;
    const pbrt_texture_desc* textures;
    uint32_t light_count;     /* explicit lights, appended after the emissive-triangle lights */
    const pbrt_light* lights;
    uint32_t bvh_builder;     /* enum pbrt_bvh_builder */
    uint32_t reserved;
} pbrt_scene_desc;

typedef struct pbrt_scene_info {
    uint32_t triangle_count;   /* world-space triangles over all instances */
    uint32_t instance_count;
    uint32_t mesh_count;
    uint32_t light_count;      /* Infos.lightCount (layoutPTUniform.glsl:4-9) */
    float light_strength_sum;  /* Infos.lightStrengthSum */
    uint32_t bvh_node_count;
    uint32_t bvh_depth;
    uint32_t light_sampling;   /* enum pbrt_light_sampling chosen at packing */
    uint64_t device_bytes;
    uint32_t bvh_builder;      /* enum pbrt_bvh_builder that built the tree (LBVH and PLOC need >= 2
                                  triangles; a single triangle gets the host SAH tree) */
    float bvh_build_ms;        /* host wall time (SAH, SBVH) or device time (LBVH, PLOC) of the build */
} pbrt_scene_info;

typedef struct pbrt_scene pbrt_scene;

/* ------------------------------------------------------------------------- */
/* Library                                                                    */
/* ------------------------------------------------------------------------- */

int32_t pbrt_abi_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* pbrt_last_error(void);

/* ------------------------------------------------------------------------- */
/* Scene packing + BVH build                                                  */
/* Replaces RayTracingSceneDescriptorCreationVisitor (source/scene/           */
/* RayTracingVisitor.cpp:3-560), PBRTPipeline::setup_pipeline                  */
/* (source/renderModules/PBRTPipeline.cpp:85-170) and the BLAS/TLAS build      */
/* (external/vsg/src/vsg/raytracing/BuildAccelerationStructureTraversal.cpp:   */
/* 63-90, vsg/vk/Context.cpp:214-222).                                         */
/* ------------------------------------------------------------------------- */

int32_t pbrt_scene_create(const pbrt_scene_desc* desc, pbrt_scene** out_scene);
int32_t pbrt_scene_destroy(pbrt_scene* scene);
int32_t pbrt_scene_get_info(const pbrt_scene* scene, pbrt_scene_info* out_info);
/* Copies the packed light list (PBRTPipeline.cpp:154-162 data) to host memory. */
int32_t pbrt_scene_get_lights(const pbrt_scene* scene, pbrt_light* out_lights, uint32_t capacity);

/* Stable LSD radix sort of n (key, value) pairs of device uint32 by key bits [0, bits), the
 * device sort of the GPU BVH builders (Morton codes), on `stream`.  keys_tmp / vals_tmp are
 * device arrays of n entries and `scratch` holds pbrt_radix_sort_scratch_size(n) bytes; the
 * sorted pairs end in keys / vals.  No counterpart in the reference (the Vulkan driver builds
 * its acceleration structures, AccelerationStructure.cpp:44). */
int32_t pbrt_radix_sort_scratch_size(uint32_t n, size_t* out_bytes);
int32_t pbrt_radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                              uint32_t bits, void* scratch, size_t scratch_bytes, void* stream);

/* ------------------------------------------------------------------------- */
/* Path tracing: replaces vkCmdTraceRaysKHR over ptRaygen.rgen /              */
/* ptClosesthit.rchit / ptAlphaHit.rahit / ptMiss.rmiss / shadow.rmiss        */
/* (PBRTPipeline::add_trace_rays_to_command_graph, PBRTPipeline.cpp:65-80).    */
/* ------------------------------------------------------------------------- */

enum pbrt_light_sampling {
    PBRT_LIGHT_SAMPLE_SURFACE_STRENGTH = 0, /* lighting.glsl:18-96  */
    PBRT_LIGHT_SAMPLE_LIGHT_STRENGTH = 1,   /* lighting.glsl:98-168 */
    PBRT_LIGHT_SAMPLE_UNIFORM = 2,          /* lighting.glsl:170-227 */
    PBRT_LIGHT_SAMPLE_AUTO = 0xFFFFFFFFu    /* >800 lights => uniform (PBRTPipeline.cpp:92-96) */
};

enum pbrt_illumination_output {
    PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT = 0, /* IlluminationBufferDemodulatedFloat, ptRaygen.rgen:83-89 */
    PBRT_OUTPUT_FINAL_IMAGE = 1               /* IlluminationBufferFinalFloat, ptRaygen.rgen:91-103 */
};

enum pbrt_tracer_kind {
    PBRT_TRACER_MEGAKERNEL = 0, /* one thread per path, whole path in one launch */
    PBRT_TRACER_WAVEFRONT = 1,  /* ray queues: extend / shade / shadow launches per bounce */
    PBRT_TRACER_PERSISTENT = 2  /* megakernel with path regeneration: resident waves draw pixels
                                   from a frame counter, a lane whose path ends starts the next
                                   pixel (same per-pixel results as the megakernel) */
};

typedef struct pbrt_trace_params {
    uint32_t width;
    uint32_t height;
    uint32_t max_recursion_depth; /* Infos.maxRecursionDepth, 2 in the reference (PBRTPipeline.cpp:27) */
    uint32_t min_recursion_depth; /* Infos.minRecursionDepth, 0 in the reference */
    uint32_t light_sampling;      /* enum pbrt_light_sampling */
    uint32_t output;              /* enum pbrt_illumination_output */
    uint32_t write_gbuffer;       /* GBUFFER define */
    uint32_t tracer;              /* enum pbrt_tracer_kind */
    /* Launch shape of the wavefront tracer (0 = defaults; pbrt_autotune_trace picks them):
     * threads per workgroup of the traversal kernels (64, 128 or 256), and paths in
     * flight (the frame is traced in slices of this many pixels, rounded up to 8x8
     * tiles, so the ray queues and path state stay in cache; 0 = the whole frame). */
    uint32_t block_threads;
    uint32_t queue_paths;
} pbrt_trace_params;

typedef struct pbrt_trace_images {
    float* depth;          /* R32F     (GBuffer.cpp, layoutPTImages.glsl:9) */
    float* normal;         /* RG32F    spherical (acos z, atan(y, x)) */
    uint8_t* material;     /* RGBA8    category / 255 */
    uint8_t* albedo;       /* RGBA8    diffuse + specular colour */
    float* illumination;   /* RGBA32F  demodulated radiance, or the progressive final image */
    void* workspace;       /* wavefront tracer: device scratch of pbrt_trace_workspace_size() bytes
                              (path state and ray queues); unused by the megakernel */
    size_t workspace_bytes;
} pbrt_trace_images;

/* Autotuner: traces the frame once per candidate launch configuration (megakernel;
 * persistent megakernel; wavefront with 64 / 128 / 256-thread traversal workgroups x whole-frame, half and
 * quarter queues), `repeats` timed runs each after a warm-up, on `stream`, and returns
 * the fastest as `best` (a copy of `params` with tracer, block_threads and
 * queue_paths set).  No counterpart in the reference (the Vulkan driver schedules
 * its ray tracing pipeline).  Synchronizes the stream.  Every candidate writes the
 * output images as pbrt_trace does: the G-buffer and illumination hold the last
 * candidate's frame afterwards, and with PBRT_OUTPUT_FINAL_IMAGE and
 * sample_number > 0 each run blends into the progressive image again, so pass
 * scratch images (or re-trace the frame) when their contents matter. */
#define PBRT_TUNING_MAX_CANDIDATES 16
typedef struct pbrt_trace_tuning {
    pbrt_trace_params best;
    uint32_t candidates;
    uint32_t tracer[PBRT_TUNING_MAX_CANDIDATES];
    uint32_t block_threads[PBRT_TUNING_MAX_CANDIDATES];
    uint32_t queue_paths[PBRT_TUNING_MAX_CANDIDATES];
    float ms[PBRT_TUNING_MAX_CANDIDATES];
} pbrt_trace_tuning;

int32_t pbrt_autotune_trace(const pbrt_scene* scene, const pbrt_trace_params* params,
                            const pbrt_push_constants* push_constants, const pbrt_trace_images* images,
                            uint32_t repeats, pbrt_trace_tuning* out, void* stream);

/* Device scratch the wavefront tracer needs for a width x height launch. */
int32_t pbrt_trace_workspace_size(uint32_t width, uint32_t height, size_t* out_bytes);

/* Optional per-launch counters (device pointer to 6 x uint64, accumulated with
 * atomics; NULL disables counting): [0] closest-hit rays, [1] shadow rays,
 * [2] BVH nodes visited and [3] triangles tested by closest-hit rays,
 * [4] BVH nodes visited and [5] triangles tested by shadow rays.  The counts
 * behind the Mrays/s figure of the reference's GUI stat line (Gui.hpp:39-44). */
int32_t pbrt_trace(const pbrt_scene* scene, const pbrt_trace_params* params,
                   const pbrt_push_constants* push_constants, const pbrt_trace_images* images,
                   unsigned long long* counters, void* stream);

/* Per-stage GPU time of pbrt_trace calls (no counterpart in the reference, whose
 * GUI shows only the frame time).  Between _begin and _end every pbrt_trace on
 * this scene records HIP events on its stream at each stage boundary: wavefront
 * extend / shade / shadow launches, the megakernel launch, and the rest
 * (generate, finalize, bookkeeping).  _end waits for the last recorded event and
 * returns the summed event times.  Not for use while the stream is captured. */
typedef struct pbrt_trace_timing {
    double extend_ms;     /* closest-hit traversal launches (wavefront) */
    double shade_ms;      /* shading launches (wavefront) */
    double shadow_ms;     /* any-hit shadow launches (wavefront) */
    double megakernel_ms; /* megakernel launches */
    double other_ms;      /* queue set-up, camera rays, finalize */
    uint64_t extend_launches;
    uint64_t shade_launches;
    uint64_t shadow_launches;
    uint64_t megakernel_launches;
    uint64_t frames;      /* pbrt_trace calls recorded */
} pbrt_trace_timing;

int32_t pbrt_trace_timing_begin(const pbrt_scene* scene);
int32_t pbrt_trace_timing_end(const pbrt_scene* scene, pbrt_trace_timing* out);

/* Closest-hit query over the scene BVH (the traversal traceRayEXT performs).
 * rays: n x 8 floats (origin xyz, tmin, direction xyz, tmax).
 * hits: n x 4 words (t as float bits, instance index, primitive index, 0xFFFFFFFF on miss)
 * and barycentrics: n x 2 floats (u, v) as in hitAttributeEXT. */
int32_t pbrt_intersect(const pbrt_scene* scene, const float* rays, uint32_t n, uint32_t* hits,
                       float* barycentrics, void* stream);
/* Any-hit shadow query (gl_RayFlagsTerminateOnFirstHit | SkipClosestHit with the
 * alpha test of ptAlphaHit.rahit on non-opaque instances); out: n bytes, 1 = occluded. */
int32_t pbrt_occluded(const pbrt_scene* scene, const float* rays, uint32_t n, uint8_t* occluded,
                      void* stream);

/* RNG streams of shaders/random.glsl (rEInit + randomFloat) for a width x height
 * launch: out[(y * width + x) * count + k] = bit pattern of the k-th randomFloat. */
int32_t pbrt_rng_sequence(uint32_t width, uint32_t height, uint32_t frame_index, uint32_t count,
                          uint32_t* out, void* stream);

/* ------------------------------------------------------------------------- */
/* Temporal accumulation: replaces accumulator.comp (SEPARATE_MATRICES form,   */
/* shaders/accumulator.comp:33-105) dispatched by Accumulator::                */
/* add_dispatch_to_command_graph (source/renderModules/Accumulator.cpp:72-83). */
/* ------------------------------------------------------------------------- */

/* Accumulator push constants as set by Accumulator::set_camera_matrices
 * (Accumulator.cpp:85-117) with separate matrices. */
typedef struct pbrt_accumulate_params {
    uint32_t width;
    uint32_t height;
    uint32_t frame_number;
    uint32_t reserved;
    float inv_proj[16];    /* current inverse projection ("view" push constant) */
    float inv_view[16];    /* current inverse view */
    float prev_view[16];   /* previous frame's view matrix */
    float prev_origin[4];  /* inverse(prev_view)[3] with w = 1 */
} pbrt_accumulate_params;

typedef struct pbrt_accumulate_images {
    const float* illumination;         /* RGBA32F  noisy demodulated radiance (tracer output) */
    const float* depth;                /* R32F     current G-buffer depth */
    const float* prev_depth;           /* R32F     previous depth (sampled bilinear) */
    const uint8_t* prev_spp;           /* R8       previous sample counts (sampled bilinear) */
    const uint16_t* prev_accumulated;  /* RGBA16F  previous accumulated illumination (sampled bilinear) */
    uint16_t* motion;                  /* RG16F    out: previous-frame uv, or (-1, -1) */
    uint8_t* spp;                      /* R8       out: sample count / 256 */
    uint16_t* accumulated;             /* RGBA16F  out: accumulated demodulated radiance */
} pbrt_accumulate_images;

int32_t pbrt_accumulate(const pbrt_accumulate_params* params, const pbrt_accumulate_images* images, void* stream);

/* AccumulationBuffer::copy_to_back_images (source/buffers/AccumulationBuffer.cpp:72-244):
 * current depth, spp and accumulated illumination copied to their previous-frame images.
 * Any destination may be NULL to skip it (a frame graph may ping-pong instead). */
int32_t pbrt_copy_back(uint32_t width, uint32_t height, const float* depth, const uint8_t* spp,
                       const uint16_t* accumulated, float* prev_depth, uint8_t* prev_spp,
                       uint16_t* prev_accumulated, void* stream);

/* ------------------------------------------------------------------------- */
/* TAA: replaces taa.comp (shaders/taa.comp:48-104) plus Taa::copy_final_image */
/* (source/renderModules/Taa.cpp:99-143).                                     */
/* ------------------------------------------------------------------------- */
typedef struct pbrt_taa_images {
    const uint16_t* motion;     /* RG16F    previous-frame uv (accumulator output) */
    const uint8_t* denoised;    /* B8G8R8A8 denoiser final image */
    uint8_t* history;           /* R8G8B8A8 accumulation image; receives a raw byte copy of `final` */
    uint8_t* final_image;       /* B8G8R8A8 out */
} pbrt_taa_images;

int32_t pbrt_taa(uint32_t width, uint32_t height, uint32_t frame_number, const pbrt_taa_images* images,
                 void* stream);

/* formatConverter.comp:9-14 with FORMAT rgba8 on a B8G8R8A8 image
 * (source/renderModules/FormatConverter.cpp:4-82): RGBA32F -> B8G8R8A8. */
int32_t pbrt_convert(uint32_t width, uint32_t height, const float* source, uint8_t* bgra8, void* stream);

/* Running mean of RGBA32F frames: dst += (src - dst) / (n + 1) with dst holding the
 * mean of n earlier frames (n = 0 copies).  No counterpart in the reference; it
 * builds the many-sample demodulated references the learned denoiser trains on. */
int32_t pbrt_running_mean(uint32_t width, uint32_t height, const float* src, float* dst, uint32_t n, void* stream);

/* ------------------------------------------------------------------------- */
/* BMFR denoiser: replaces bmfrPre/bmfrFit/bmfrPost.comp (shaders/bmfrPre.comp */
/* :5-98, bmfrFit.comp:5-92, bmfrPost.comp:5-124) dispatched by               */
/* BMFR::add_dispatch_to_command_graph (source/renderModules/denoisers/        */
/* BMFR.cpp:203-230).  Blocks are block_size^2 pixels on the padded grid       */
/* ((W / b + 2) * b) x ((H / b + 2) * b), shifted by pixelOffsets[frame % 16]. */
/* ------------------------------------------------------------------------- */

enum pbrt_bmfr_fit {
    /* the reference algorithm: FP32 Householder QR of the fp16 feature block
     * (bmfrFit.comp:27-81) with wave64 reductions */
    PBRT_BMFR_FIT_HOUSEHOLDER = 0,
    /* pre and fit fused (no feature buffer round trip): the columns are projected
     * against the noised constant column in FP64 and scaled by powers of two per
     * block, their Gram matrix is built on FP16 MFMA (hi/lo split, FP32
     * accumulation inside each 32-row chunk, FP64 across chunks) and solved by
     * FP64 elimination.  Accurate on well-conditioned blocks only: on planar or
     * sky blocks cond(A^T A) reaches 1e10 and the fit error grows to ~1e-1. */
    PBRT_BMFR_FIT_GRAM_F16 = 1,
    /* as PBRT_BMFR_FIT_GRAM_F16 but the Gram matrix is accumulated by FP64 MFMA
     * (v_mfma_f64_16x16x4f64) from the fp32 projected columns; fitted values
     * stay within ~1e-5 of the FP64 least-squares solution on every block. */
    PBRT_BMFR_FIT_GRAM_F64 = 2
};

typedef struct pbrt_bmfr_params {
    uint32_t width;
    uint32_t height;
    uint32_t block_size;     /* 8, 16 or 32 (DenoisingBlockSize X8 / X16 / X32) */
    uint32_t frame_number;   /* push constant frameNumber */
    uint32_t fit;            /* enum pbrt_bmfr_fit */
    uint32_t fit_threads;    /* Gram fits: workgroup size 256, 512 or 1024 (0 = 256), i.e. how a
                                block's rows are tiled over the waves' MFMA Gram accumulations;
                                pbrt_autotune_bmfr picks it */
    uint32_t reserved[2];
} pbrt_bmfr_params;

typedef struct pbrt_bmfr_images {
    const float* depth;         /* R32F G-buffer depth */
    const float* normal;        /* RG32F spherical normal */
    const uint8_t* albedo;      /* RGBA8 */
    const uint16_t* motion;     /* RG16F accumulator motion */
    const uint8_t* spp;         /* R8 accumulator sample count / 256 */
    const uint16_t* noisy;      /* RGBA16F accumulated demodulated illumination */
    uint16_t* denoised;         /* RGBA16F x 2 layers ([2][H][W][4]); reads layer frame & 1,
                                   writes layer (frame & 1) ^ 1 */
    uint8_t* final_image;       /* B8G8R8A8 out: remodulated, gamma 1/2.2 */
    void* workspace;            /* device scratch of pbrt_bmfr_workspace_size() bytes */
    size_t workspace_bytes;
} pbrt_bmfr_images;

/* Workspace layout (each part 256-byte aligned): fp16 feature buffer [13][Hp][Wp]
 * (Householder fit only), float weights [blocks_y][blocks_x][30] (feature-major,
 * rgb inner), the frame's noise table float [10][block_size^2] (Gram fits: the
 * addRandom values, computed once per frame), then float depth ranges
 * [blocks_y][blocks_x][2] (each block's depth min and max, written by the fit's
 * pass and read by the post pass). */
int32_t pbrt_bmfr_workspace_size(const pbrt_bmfr_params* params, size_t* out_bytes);
int32_t pbrt_bmfr(const pbrt_bmfr_params* params, const pbrt_bmfr_images* images, void* stream);

/* pbrt_bmfr without a workspace and without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * results as pbrt_bmfr; images->workspace and workspace_bytes are ignored.  The
 * scratch (the fit's weights, plus the fp16 feature array for the Householder fit)
 * lives in a device array of PBRT_BMFR_NOALLOC_SCRATCH_BYTES that the library holds
 * on each GPU it runs on until the process ends; a frame needing more is rejected
 * with PBRT_ERR_INVALID_ARG (Gram fits: 32 floats per block, i.e. up to 3840x2160 at
 * 16-pixel blocks; Householder: about 26 bytes per padded pixel, i.e. up to ~640x480).
 * Every call on one GPU shares that array: calls that overlap on one GPU (two streams,
 * two threads) race.  Nothing in the scratch needs clearing between calls. */
#define PBRT_BMFR_NOALLOC_SCRATCH_BYTES (8u << 20)
int32_t pbrt_bmfr_noalloc(const pbrt_bmfr_params* params, const pbrt_bmfr_images* images, void* stream);

/* Autotuner of the Gram fits' tile shape: runs pbrt_bmfr once per workgroup size
 * (256, 512, 1024 threads), `repeats` timed runs each after a warm-up, on `stream`,
 * and returns the fastest as `best` (a copy of `params` with fit_threads set).  The
 * fit must be PBRT_BMFR_FIT_GRAM_F16 or _F64.  No counterpart in the reference
 * (BMFR.cpp:155-170 fixes fitting_kernel = 256).  Writes the images pbrt_bmfr writes;
 * synchronizes the stream. */
typedef struct pbrt_bmfr_tuning {
    pbrt_bmfr_params best;
    uint32_t candidates;
    uint32_t fit_threads[4];
    float ms[4];
} pbrt_bmfr_tuning;
int32_t pbrt_autotune_bmfr(const pbrt_bmfr_params* params, const pbrt_bmfr_images* images, uint32_t repeats,
                           pbrt_bmfr_tuning* out, void* stream);

/* Batched BMFR: every block of many frames (any mix of resolutions) fitted in
 * one launch and post-processed in a second, instead of one BMFR dispatch chain
 * per frame (BMFR::add_dispatch_to_command_graph, BMFR.cpp:206-229).  All frames
 * share block_size and fit, and the fit is PBRT_BMFR_FIT_GRAM_F16 or _F64 (the
 * Householder path stages features per frame); the batch fits with 256-thread
 * workgroups whatever fit_threads says.  images.workspace of each frame
 * is ignored: the batch workspace holds the frame table and every frame's
 * weights.  The frame table is staged through a pinned host buffer owned by the
 * library; a call waits only for the previous call's table upload. */
typedef struct pbrt_bmfr_batch_frame {
    pbrt_bmfr_params params;
    pbrt_bmfr_images images;
} pbrt_bmfr_batch_frame;

int32_t pbrt_bmfr_batched_workspace_size(const pbrt_bmfr_batch_frame* frames, uint32_t count, size_t* out_bytes);
/* Workspace layout: frame table, then per frame (each 256-byte aligned) float
 * weights [blocks_y][blocks_x][30], the frame's noise table [10][block_size^2] and
 * its blocks' depth ranges [blocks_y][blocks_x][2], in frame order. */
int32_t pbrt_bmfr_batched(const pbrt_bmfr_batch_frame* frames, uint32_t count, void* workspace,
                          size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------- */
/* BFR (blockwise feature regression by 40 Adam steps): replaces bfr.comp      */
/* (shaders/bfr.comp:202-308, parallel_reduction_alpha :100-142) dispatched by */
/* BFR::add_dispatch_to_command_graph (source/renderModules/denoisers/        */
/* BFR.cpp:130-140) over (W/b + 2) x (H/b + 2) blocks of b x b pixels.         */
/* ------------------------------------------------------------------------- */

typedef struct pbrt_bfr_params {
    uint32_t width;
    uint32_t height;
    uint32_t block_size;   /* 8, 16 or 32 (DenoiserUtils.cpp:25-70) */
    uint32_t frame_number; /* push constant frameNumber: pixel offset and ping-pong layer */
} pbrt_bfr_params;

typedef struct pbrt_bfr_images {
    const float* depth;         /* R32F G-buffer depth */
    const float* normal;        /* RG32F spherical normal */
    const uint8_t* albedo;      /* RGBA8 G-buffer albedo */
    const uint16_t* motion;     /* RG16F accumulator motion (x < 0: not reprojected) */
    const uint8_t* spp;         /* R8 accumulator sample count / 256 */
    const uint16_t* noisy;      /* RGBA16F accumulated demodulated illumination */
    uint16_t* denoised;         /* RGBA16F x 2 layers: reads layer frame&1, writes (frame&1)^1 */
    uint8_t* final_image;       /* B8G8R8A8 remodulated, gamma 1/2.2 */
    float* weights;             /* optional (NULL: not written): fitted alpha per block,
                                   [blocks_y][blocks_x][7 features][rgb] */
} pbrt_bfr_images;

int32_t pbrt_bfr(const pbrt_bfr_params* params, const pbrt_bfr_images* images, void* stream);

/* BFR blender (shaders/bfrBlender.comp:22-69, BFRBlender.cpp:84-90): per-pixel
 * standard deviation of the mean illumination over a (2r+1)^2 window mixed with
 * the pixel's own mean and mean-square, then a three-way blend of the 32x32 (low
 * deviation), 16x16 and 8x8 (high deviation) denoiser outputs.  Window taps
 * outside the image read as zero (robust image access). */
typedef struct pbrt_bfr_blend_images {
    const uint16_t* average;         /* RGBA16F mean illumination */
    const uint16_t* average_squared; /* RGBA16F mean squared illumination */
    const uint8_t* denoised8;        /* B8G8R8A8 output of the 8x8 denoiser */
    const uint8_t* denoised16;       /* B8G8R8A8 output of the 16x16 denoiser */
    const uint8_t* denoised32;       /* B8G8R8A8 output of the 32x32 denoiser */
    uint8_t* final_image;            /* B8G8R8A8 */
} pbrt_bfr_blend_images;

int32_t pbrt_bfr_blend(uint32_t width, uint32_t height, uint32_t filter_radius,
                       const pbrt_bfr_blend_images* images, void* stream);

/* ------------------------------------------------------------------------- */
/* Learned kernel-predicting denoiser.  No counterpart in the reference: it    */
/* stands in for BMFR's per-block fit (bmfrFit.comp) with a network that reads */
/* the same accumulator / G-buffer images as bmfrPre.comp and writes the same  */
/* outputs as bmfrPost.comp (demodulated denoised radiance, remodulated BGRA8).*/
/*   features (12): log1p(noisy rgb), albedo rgb, normal xyz, depth/(1+depth), */
/*                  min(spp,16)/16, motion valid                               */
/*   6 x conv3x3 (12->32->32->32->32->32, ReLU, ->25 logits), BF16 weights and */
/*   activations, FP32 accumulation on the matrix cores; softmax = per-pixel   */
/*   5x5 kernel (taps clamped to the image) applied to the noisy illumination. */
/*   loss: relative squared error of the remodulated radiance, mean over pixels */
/*   and rgb of (a (out - target))^2 / ((a target)^2 + 0.01), a = albedo + 1e-6.*/
/* Parameters in canonical order, per layer W[cout][cin][ky][kx] then b[cout]. */
/* ------------------------------------------------------------------------- */

typedef struct pbrt_cnn pbrt_cnn;

/* A network of `layers` (2..6) convolutions: 12 -> 32 -> ... -> 32 -> 25 logits,
 * otherwise as above.  pbrt_cnn_create is the 6-layer network of the learned-
 * denoiser configuration.  A two-layer network runs inference as one fused kernel
 * per 16x16 tile with every intermediate in LDS (images.workspace is not used):
 * the equal-quality, lower-cost alternative to BMFR (no counterpart in the
 * reference; BMFR::add_dispatch_to_command_graph, BMFR.cpp:206-229, is what it
 * stands in for).  Parameter count and order follow the layer count. */
int32_t pbrt_cnn_create_layers(uint32_t layers, uint64_t seed, pbrt_cnn** out_net);
/* The compact network: two layers whose second convolution is pointwise (1x1; stored as
 * a 3x3 one whose eight outer taps stay zero: set_parameters zeroes them, training masks
 * their gradient), so inference needs no hidden halo and runs as one fused kernel that
 * keeps both layers in registers per 16-pixel row: the fastest learned stand-in for
 * BMFR (BMFR::add_dispatch_to_command_graph, BMFR.cpp:206-229; no counterpart in the
 * reference).  Parameters as for two layers; pbrt_cnn_layers reports 2. */
int32_t pbrt_cnn_create_compact(uint64_t seed, pbrt_cnn** out_net);
int32_t pbrt_cnn_is_compact(const pbrt_cnn* net, uint32_t* out_compact);
int32_t pbrt_cnn_layers(const pbrt_cnn* net, uint32_t* out_layers);
int32_t pbrt_cnn_layers_parameter_count(uint32_t layers, uint32_t* out_count);

typedef struct pbrt_cnn_images {
    uint32_t width;
    uint32_t height;
    uint32_t count;           /* images of width x height stored back to back */
    const uint16_t* noisy;    /* RGBA16F demodulated illumination (accumulator output) */
    const float* depth;       /* R32F */
    const float* normal;      /* RG32F spherical */
    const uint8_t* albedo;    /* RGBA8 */
    const uint8_t* spp;       /* R8 sample count / 256 */
    const uint16_t* motion;   /* RG16F */
    float* denoised;          /* RGBA32F out (denoise): demodulated denoised radiance, may be NULL */
    uint8_t* final_image;     /* B8G8R8A8 out (denoise): remodulated, gamma 1/2.2, may be NULL */
    const float* target;      /* RGBA32F reference radiance (training and loss) */
    void* workspace;          /* device scratch of pbrt_cnn_workspace_size() bytes */
    size_t workspace_bytes;
} pbrt_cnn_images;

int32_t pbrt_cnn_create(uint64_t seed, pbrt_cnn** out_net); /* He-normal init from the seed */
int32_t pbrt_cnn_destroy(pbrt_cnn* net);
int32_t pbrt_cnn_parameter_count(uint32_t* out_count);
int32_t pbrt_cnn_get_parameters(const pbrt_cnn* net, float* host_params);
int32_t pbrt_cnn_set_parameters(pbrt_cnn* net, const float* host_params); /* also resets the optimiser */
int32_t pbrt_cnn_workspace_size(uint32_t width, uint32_t height, uint32_t count, uint32_t training,
                                size_t* out_bytes);
/* Inference: denoised and/or final_image of every image. */
int32_t pbrt_cnn_denoise(const pbrt_cnn* net, const pbrt_cnn_images* images, void* stream);
/* Loss and its gradient (device pointers: 1 float, parameter_count floats in
 * canonical order) without changing the network. */
int32_t pbrt_cnn_gradients(const pbrt_cnn* net, const pbrt_cnn_images* images, float* loss, float* grad,
                           void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch; loss (device, 1 float,
 * may be NULL) receives the batch loss before the step. */
int32_t pbrt_cnn_train_step(pbrt_cnn* net, const pbrt_cnn_images* images, float learning_rate, float* loss,
                            void* stream);

/* ------------------------------------------------------------------------- */
/* Frame graph: replaces the per-frame recordAndSubmit of the recorded command */
/* graph (source/VulkanPBRT.cpp:543,588).  begin() starts capturing, on        */
/* `stream`, the launches the calling thread issues there (every entry point  */
/* of a frame: trace, accumulate, denoise, TAA, convert); end() closes the     */
/* capture, updates the executable graph in place when the launch topology is  */
/* unchanged (kernel arguments such as the camera change every frame) or       */
/* instantiates it anew, and launches it on `stream`.  `stream` must not be    */
/* the null stream.  The frame entry points issue no memset, host copy or      */
/* synchronisation, so a whole frame captures into one graph.                  */
/* ------------------------------------------------------------------------- */
typedef struct pbrt_frame_graph pbrt_frame_graph;

int32_t pbrt_frame_graph_create(pbrt_frame_graph** out_graph);
int32_t pbrt_frame_graph_destroy(pbrt_frame_graph* graph); /* ends an open capture unlaunched */
int32_t pbrt_frame_graph_begin(pbrt_frame_graph* graph, void* stream);
/* node_count (may be NULL): kernel nodes of the captured frame; instantiated (may be
 * NULL): 1 when the executable graph was rebuilt rather than updated. */
int32_t pbrt_frame_graph_end(pbrt_frame_graph* graph, void* stream, uint32_t* node_count, uint32_t* instantiated);
/* Ends an open capture without launching it (the frame's work is dropped) and keeps
 * the handle and its executable graph for the next begin(); a no-op when no capture
 * is open.  For callers whose frame failed between begin() and end(). */
int32_t pbrt_frame_graph_abort(pbrt_frame_graph* graph);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PBRT_C_H */
