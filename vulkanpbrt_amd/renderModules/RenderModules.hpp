// The reference's render-module API (source/renderModules, source/util/
// DenoiserUtils) on top of the gfx950 C ABI.  Class names, constructor
// arguments, method names and call order follow the reference so that the
// setup code of source/VulkanPBRT.cpp:356-505 keeps its shape:
//
//   auto pbrt = PBRTPipeline::create(scene, g_buffer, illumination, true, RayTracingRayOrigin::CAMERA);
//   pbrt->set_tlas(scene);
//   pbrt->compile(ctx); pbrt->add_trace_rays_to_command_graph(commands, push_constants);
//   auto acc = Accumulator::create(g_buffer, pbrt->get_illumination_buffer(), true);
//   acc->compile_images(ctx); acc->add_dispatch_to_command_graph(commands);
//   add_denoiser_to_commands(DenoisingType::BMFR, DenoisingBlockSize::X32, commands, ctx, w, h,
//                            push_constants, g_buffer, acc->accumulated_illumination,
//                            acc->accumulation_buffer, final_image);
//
// VSG objects are replaced by the types of Core.hpp; errors throw
// vkpbrt::Exception (the reference throws vsg::Exception or prints and returns).
#pragma once

#include <array>
#include <memory>
#include <optional>
#include <string>

#include "Buffers.hpp"
#include "Core.hpp"

namespace vkpbrt {

// PipelineStructs.hpp:6-30
struct RayTracingPushConstants {
    std::array<float, 16> view_inverse;
    std::array<float, 16> proj_inverse;
    std::array<float, 16> prev_view;
    uint32_t frame_number;
    uint32_t sample_number;
};
using PushConstantsPtr = std::shared_ptr<PushConstants<RayTracingPushConstants>>;

// LEARNED (the kernel-predicting network of pbrt_cnn_*) has no counterpart in the
// reference's DenoiserUtils.hpp; it reads and writes the same images as BMFR.
enum class DenoisingType { NONE, BMFR, BFR, SVG, LEARNED };
enum class DenoisingBlockSize { X8, X16, X32, X64, X8X16X32 };
enum class RayTracingRayOrigin { CAMERA, GBUFFER };

// RenderIO.hpp CameraMatrices (column-major float matrices)
struct CameraMatrices {
    std::array<float, 16> view{}, inv_view{};
    std::optional<std::array<float, 16>> proj, inv_proj;
};

class PBRTPipeline {
public:
    PBRTPipeline(pbrt_scene* scene, GBufferPtr g_buffer, IlluminationBufferPtr illumination_buffer,
                 bool write_g_buffer, RayTracingRayOrigin ray_tracing_ray_origin);
    static std::shared_ptr<PBRTPipeline> create(pbrt_scene* scene, GBufferPtr g, IlluminationBufferPtr i, bool w,
                                                RayTracingRayOrigin o) {
        return std::make_shared<PBRTPipeline>(scene, g, i, w, o);
    }
    // The BVH is built with the scene by pbrt_scene_create; set_tlas binds it.
    void set_tlas(pbrt_scene* as);
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    void add_trace_rays_to_command_graph(const CommandsPtr& command_graph, const PushConstantsPtr& push_constants);
    IlluminationBufferPtr get_illumination_buffer() const { return _illumination_buffer; }
    enum class LightSamplingMethod { SAMPLE_SURFACE_STRENGTH, SAMPLE_LIGHT_STRENGTH, SAMPLE_UNIFORM } light_sampling_method =
        LightSamplingMethod::SAMPLE_SURFACE_STRENGTH;
    uint32_t max_recursion_depth = 2;  // PBRTPipeline.cpp:27
    uint32_t tracer = PBRT_TRACER_WAVEFRONT;
    unsigned long long* counters = nullptr;  // optional device counters (pbrt_trace)

private:
    pbrt_scene* _scene;
    pbrt_scene* _tlas = nullptr;
    uint32_t _width, _height;
    bool _write_g_buffer;
    GBufferPtr _g_buffer;
    IlluminationBufferPtr _illumination_buffer;
    std::unique_ptr<DeviceBuffer> _workspace;
};

class Accumulator {
public:
    Accumulator(GBufferPtr g_buffer, IlluminationBufferPtr illumination_buffer, bool separate_matrices,
                int work_width = 16, int work_height = 16);
    static std::shared_ptr<Accumulator> create(GBufferPtr g, IlluminationBufferPtr i, bool separate, int ww = 16,
                                               int wh = 16) {
        return std::make_shared<Accumulator>(g, i, separate, ww, wh);
    }
    void compile_images(Context& context) const;
    void update_image_layouts(Context&) const {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph);
    void set_camera_matrices(int frame_index, const CameraMatrices& cur, const CameraMatrices& prev);
    IlluminationBufferPtr accumulated_illumination;
    AccumulationBufferPtr accumulation_buffer;

private:
    int _width, _height;
    GBufferPtr _g_buffer;
    IlluminationBufferPtr _original_illumination;
    pbrt_accumulate_params _params{};
};

class BMFR {
public:
    BMFR(uint32_t width, uint32_t height, uint32_t work_width, uint32_t work_height, GBufferPtr g_buffer,
         IlluminationBufferPtr illu_buffer, AccumulationBufferPtr acc_buffer, uint32_t fitting_kernel = 256);
    static std::shared_ptr<BMFR> create(uint32_t w, uint32_t h, uint32_t ww, uint32_t wh, GBufferPtr g,
                                        IlluminationBufferPtr i, AccumulationBufferPtr a, uint32_t fk = 256) {
        return std::make_shared<BMFR>(w, h, ww, wh, g, i, a, fk);
    }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph, const PushConstantsPtr& push_constants);
    DeviceImagePtr get_final_descriptor_image() const { return _final_illumination; }
    uint32_t fit = PBRT_BMFR_FIT_GRAM_F64;

private:
    uint32_t _width, _height, _work_width;
    GBufferPtr _g_buffer;
    IlluminationBufferPtr _illumination;
    AccumulationBufferPtr _acc;
    DeviceImagePtr _accumulated_illumination;  // RGBA16F x 2 layers
    DeviceImagePtr _final_illumination;        // BGRA8
    std::unique_ptr<DeviceBuffer> _workspace;
};

// BFR.hpp: blockwise feature regression by Adam (bfr.comp)
class BFR {
public:
    BFR(uint32_t width, uint32_t height, uint32_t work_width, uint32_t work_height, GBufferPtr g_buffer,
        IlluminationBufferPtr illu_buffer, AccumulationBufferPtr acc_buffer);
    static std::shared_ptr<BFR> create(uint32_t w, uint32_t h, uint32_t ww, uint32_t wh, GBufferPtr g,
                                       IlluminationBufferPtr i, AccumulationBufferPtr a) {
        return std::make_shared<BFR>(w, h, ww, wh, g, i, a);
    }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph, const PushConstantsPtr& push_constants);
    DeviceImagePtr get_final_descriptor_image() const { return _final_illumination; }

private:
    uint32_t _width, _height, _work_width;
    GBufferPtr _g_buffer;
    IlluminationBufferPtr _illumination;
    AccumulationBufferPtr _acc;
    DeviceImagePtr _accumulated_illumination;  // RGBA16F x 2 layers
    DeviceImagePtr _final_illumination;        // BGRA8
};

// BFRBlender.hpp: per-pixel blend of the 8x8 / 16x16 / 32x32 denoiser outputs
class BFRBlender {
public:
    BFRBlender(uint32_t width, uint32_t height, DeviceImagePtr average_image, DeviceImagePtr average_squared_image,
               DeviceImagePtr denoised0, DeviceImagePtr denoised1, DeviceImagePtr denoised2,
               uint32_t work_width = 16, uint32_t work_height = 16, uint32_t filter_radius = 2);
    static std::shared_ptr<BFRBlender> create(uint32_t w, uint32_t h, DeviceImagePtr avg, DeviceImagePtr avg_sq,
                                              DeviceImagePtr d0, DeviceImagePtr d1, DeviceImagePtr d2) {
        return std::make_shared<BFRBlender>(w, h, avg, avg_sq, d0, d1, d2);
    }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph);
    DeviceImagePtr get_final_descriptor_image() const { return _final; }

private:
    uint32_t _width, _height, _filter_radius;
    DeviceImagePtr _average, _average_squared, _den8, _den16, _den32, _final;
};

// Learned kernel-predicting denoiser (pbrt_cnn_denoise): trained parameters are
// loaded from a raw little-endian float32 file in the network's canonical order.
class LearnedDenoiser {
public:
    LearnedDenoiser(uint32_t width, uint32_t height, GBufferPtr g_buffer, IlluminationBufferPtr illu_buffer,
                    AccumulationBufferPtr acc_buffer, std::string weights_path = default_weights_path());
    ~LearnedDenoiser();
    LearnedDenoiser(const LearnedDenoiser&) = delete;
    LearnedDenoiser& operator=(const LearnedDenoiser&) = delete;
    static std::shared_ptr<LearnedDenoiser> create(uint32_t w, uint32_t h, GBufferPtr g, IlluminationBufferPtr i,
                                                   AccumulationBufferPtr a) {
        return std::make_shared<LearnedDenoiser>(w, h, g, i, a);
    }
    // weights/kpn6.bin beside libpbrt.so (6 layers); a file of 2-layer
    // parameters (weights/kpn2.bin) selects the fused two-layer network, and one
    // whose second layer holds only centre taps (weights/kpnc.bin) the compact one
    static std::string default_weights_path();
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph);
    DeviceImagePtr get_final_descriptor_image() const { return _final_illumination; }

private:
    uint32_t _width, _height;
    GBufferPtr _g_buffer;
    IlluminationBufferPtr _illumination;
    AccumulationBufferPtr _acc;
    std::string _weights_path;
    pbrt_cnn* _net = nullptr;
    DeviceImagePtr _final_illumination;  // BGRA8
    std::unique_ptr<DeviceBuffer> _workspace;
};

class Taa {
public:
    Taa(uint32_t width, uint32_t height, uint32_t work_width, uint32_t work_height, GBufferPtr g_buffer,
        AccumulationBufferPtr acc_buffer, DeviceImagePtr denoised);
    static std::shared_ptr<Taa> create(uint32_t w, uint32_t h, uint32_t ww, uint32_t wh, GBufferPtr g,
                                       AccumulationBufferPtr a, DeviceImagePtr d) {
        return std::make_shared<Taa>(w, h, ww, wh, g, a, d);
    }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    // taa.comp reads frameNumber from the push constants last pushed into the
    // command list, by the denoiser (Taa.cpp:99-107): the block the modules
    // recorded with Commands::push_constants before this call.
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph);
    DeviceImagePtr get_final_descriptor_image() const { return _final_image; }

private:
    uint32_t _width, _height;
    AccumulationBufferPtr _acc;
    DeviceImagePtr _denoised, _final_image, _accumulation_image;
};

class FormatConverter {
public:
    FormatConverter(DeviceImagePtr src_image, Format dst_format, int work_width = 16, int work_height = 16);
    void compile_images(Context& context) const;
    void update_image_layouts(Context&) const {}
    void add_dispatch_to_command_graph(const CommandsPtr& command_graph);
    DeviceImagePtr final_image;

private:
    DeviceImagePtr _src;
};

// DenoiserUtils.hpp:9-16
void add_denoiser_to_commands(DenoisingType denoising_type, DenoisingBlockSize denoising_size,
                              const CommandsPtr& commands, Context& compile, int width, int height,
                              const PushConstantsPtr& compute_constants, const GBufferPtr& g_buffer,
                              const IlluminationBufferPtr& illumination_buffer,
                              const AccumulationBufferPtr& accumulation_buffer, DeviceImagePtr& final_descriptor_image);

// vsg::perspective / vsg::lookAt (transform.h:148-183) in double, for callers
// building push constants on the host; column-major outputs.
std::array<double, 16> perspective(double fovy_radians, double aspect, double near_distance, double far_distance);
std::array<double, 16> look_at(const double* eye, const double* center, const double* up);
std::array<double, 16> inverse(const std::array<double, 16>& m);
std::array<float, 16> to_float(const std::array<double, 16>& m);

}  // namespace vkpbrt
