// Render modules and frame buffers on top of the C ABI (see RenderModules.hpp).
#include "RenderModules.hpp"

#include <dlfcn.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

namespace vkpbrt {

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------
GBuffer::GBuffer(uint32_t w, uint32_t h)
    : depth(std::make_shared<DeviceImage>(w, h, Format::R32F)),
      normal(std::make_shared<DeviceImage>(w, h, Format::RG32F)),
      material(std::make_shared<DeviceImage>(w, h, Format::RGBA8)),
      albedo(std::make_shared<DeviceImage>(w, h, Format::RGBA8)) {}

void GBuffer::compile(Context&) {
    depth->compile();
    normal->compile();
    material->compile();
    albedo->compile();
}

IlluminationBuffer::IlluminationBuffer(IlluminationKind k, uint32_t w, uint32_t h) : kind(k) {
    if (k == IlluminationKind::Demodulated) {
        illumination_images.push_back(std::make_shared<DeviceImage>(w, h, Format::RGBA16F));
        illumination_images.push_back(std::make_shared<DeviceImage>(w, h, Format::RGBA16F));
    } else {
        illumination_images.push_back(std::make_shared<DeviceImage>(w, h, Format::RGBA32F));
    }
}

void IlluminationBuffer::compile(Context&) {
    for (auto& i : illumination_images) i->compile();
}

AccumulationBuffer::AccumulationBuffer(uint32_t w, uint32_t h)
    : spp(std::make_shared<DeviceImage>(w, h, Format::R8)),
      prev_spp(std::make_shared<DeviceImage>(w, h, Format::R8)),
      prev_depth(std::make_shared<DeviceImage>(w, h, Format::R32F)),
      prev_normal(std::make_shared<DeviceImage>(w, h, Format::RG32F)),
      motion(std::make_shared<DeviceImage>(w, h, Format::RG16F)),
      prev_illu(std::make_shared<DeviceImage>(w, h, Format::RGBA16F)),
      prev_illu_squared(std::make_shared<DeviceImage>(w, h, Format::RGBA16F)),
      width(w),
      height(h) {}

void AccumulationBuffer::compile(Context&) {
    for (auto* i : {&spp, &prev_spp, &prev_depth, &prev_normal, &motion, &prev_illu, &prev_illu_squared})
        (*i)->compile();
}

void AccumulationBuffer::copy_to_back_images(const CommandsPtr& commands, const GBufferPtr& g_buffer,
                                             const IlluminationBufferPtr& illumination_buffer) {
    if (illumination_buffer->kind != IlluminationKind::Demodulated)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Error: AccumulationBuffer::copyToBackImages(...) Illumination buffer not supported.");
    DeviceImagePtr depth = g_buffer->depth, acc = illumination_buffer->illumination_images[0];
    commands->addChild([this, depth, acc](hipStream_t s) {
        check(pbrt_copy_back(width, height, depth->as<float>(), spp->as<uint8_t>(), acc->as<uint16_t>(),
                             prev_depth->as<float>(), prev_spp->as<uint8_t>(), prev_illu->as<uint16_t>(), s));
    });
}

// ---------------------------------------------------------------------------
// PBRTPipeline
// ---------------------------------------------------------------------------
PBRTPipeline::PBRTPipeline(pbrt_scene* scene, GBufferPtr g_buffer, IlluminationBufferPtr illumination_buffer,
                           bool write_g_buffer, RayTracingRayOrigin origin)
    : _scene(scene), _write_g_buffer(write_g_buffer), _g_buffer(g_buffer), _illumination_buffer(illumination_buffer) {
    if (!scene) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline: scene is NULL");
    if (!illumination_buffer) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline: illumination buffer is NULL");
    if (write_g_buffer && !g_buffer) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline: G-buffer requested but NULL");
    if (origin == RayTracingRayOrigin::GBUFFER)
        throw Exception(PBRT_ERR_UNSUPPORTED, "PBRTPipeline: G-buffer ray origin is not implemented");
    if (illumination_buffer->kind == IlluminationKind::Demodulated)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Error: PBRTPipeline::setupRaygenShader(...) Illumination buffer not supported.");
    _width = illumination_buffer->illumination_images[0]->width();
    _height = illumination_buffer->illumination_images[0]->height();
    pbrt_scene_info info{};
    check(pbrt_scene_get_info(scene, &info));
    if (info.light_count > 800) light_sampling_method = LightSamplingMethod::SAMPLE_UNIFORM;  // PBRTPipeline.cpp:92-96
}

void PBRTPipeline::set_tlas(pbrt_scene* as) {
    if (as != _scene) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline::set_tlas: BVH of another scene");
    _tlas = as;
}

void PBRTPipeline::compile(Context& context) {
    _illumination_buffer->compile(context);
    if (_g_buffer) _g_buffer->compile(context);
    size_t bytes = 0;
    check(pbrt_trace_workspace_size(_width, _height, &bytes));
    _workspace = std::make_unique<DeviceBuffer>(bytes);
    _workspace->compile();
}

void PBRTPipeline::add_trace_rays_to_command_graph(const CommandsPtr& command_graph,
                                                   const PushConstantsPtr& push_constants) {
    if (!_tlas) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline: set_tlas() was not called");
    if (!_workspace) throw Exception(PBRT_ERR_INVALID_ARG, "PBRTPipeline: compile() was not called");
    command_graph->push_constants(push_constants);
    command_graph->addChild([this, push_constants](hipStream_t s) {
        pbrt_trace_params p{};
        p.width = _width;
        p.height = _height;
        p.max_recursion_depth = max_recursion_depth;
        p.min_recursion_depth = 0;
        p.light_sampling = (uint32_t)light_sampling_method;
        p.output = _illumination_buffer->kind == IlluminationKind::FinalFloat ? PBRT_OUTPUT_FINAL_IMAGE
                                                                               : PBRT_OUTPUT_DEMOD_ILLUMINATION_FLOAT;
        p.write_gbuffer = _write_g_buffer ? 1 : 0;
        p.tracer = tracer;
        pbrt_push_constants pc{};
        const RayTracingPushConstants& v = push_constants->value();
        std::memcpy(pc.view_inverse, v.view_inverse.data(), 64);
        std::memcpy(pc.proj_inverse, v.proj_inverse.data(), 64);
        std::memcpy(pc.prev_view, v.prev_view.data(), 64);
        pc.frame_number = v.frame_number;
        pc.sample_number = v.sample_number;
        pbrt_trace_images im{};
        if (_g_buffer) {
            im.depth = _g_buffer->depth->as<float>();
            im.normal = _g_buffer->normal->as<float>();
            im.material = _g_buffer->material->as<uint8_t>();
            im.albedo = _g_buffer->albedo->as<uint8_t>();
        }
        im.illumination = _illumination_buffer->illumination_images[0]->as<float>();
        im.workspace = _workspace->data();
        im.workspace_bytes = _workspace->bytes();
        check(pbrt_trace(_tlas, &p, &pc, &im, counters, s));
    });
}

// ---------------------------------------------------------------------------
// Accumulator
// ---------------------------------------------------------------------------
Accumulator::Accumulator(GBufferPtr g_buffer, IlluminationBufferPtr illumination_buffer, bool separate_matrices, int,
                         int)
    : _width((int)g_buffer->depth->width()),
      _height((int)g_buffer->depth->height()),
      _g_buffer(g_buffer),
      _original_illumination(illumination_buffer) {
    accumulated_illumination = IlluminationBufferDemodulated(_width, _height);
    accumulation_buffer = AccumulationBuffer::create(_width, _height);
    if (!separate_matrices)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Accumulator: only the separate-matrices form (online tracing) is implemented");
    if (illumination_buffer->illumination_images[0]->format() != Format::RGBA32F)
        throw Exception(PBRT_ERR_UNSUPPORTED, "Accumulator: the source illumination must be RGBA32F");
    _params.width = _width;
    _params.height = _height;
}

void Accumulator::compile_images(Context& context) const {
    accumulation_buffer->compile(context);
    accumulated_illumination->compile(context);
}

void Accumulator::add_dispatch_to_command_graph(const CommandsPtr& command_graph) {
    command_graph->addChild([this](hipStream_t s) {
        pbrt_accumulate_images im{};
        im.illumination = _original_illumination->illumination_images[0]->as<float>();
        im.depth = _g_buffer->depth->as<float>();
        im.prev_depth = accumulation_buffer->prev_depth->as<float>();
        im.prev_spp = accumulation_buffer->prev_spp->as<uint8_t>();
        im.prev_accumulated = accumulation_buffer->prev_illu->as<uint16_t>();
        im.motion = accumulation_buffer->motion->as<uint16_t>();
        im.spp = accumulation_buffer->spp->as<uint8_t>();
        im.accumulated = accumulated_illumination->illumination_images[0]->as<uint16_t>();
        check(pbrt_accumulate(&_params, &im, s));
    });
}

// Accumulator.cpp:85-117 (separate matrices)
void Accumulator::set_camera_matrices(int frame_index, const CameraMatrices& cur, const CameraMatrices& prev) {
    if (!cur.proj || !cur.inv_proj)
        throw Exception(PBRT_ERR_INVALID_ARG,
                        "Accumulator::setDoubleMatrix(...) Accumulator was created with separateMatrices = true, but "
                        "DubleMatrix is missing seperate matrices");
    std::memcpy(_params.inv_proj, cur.inv_proj->data(), 64);
    std::memcpy(_params.inv_view, cur.inv_view.data(), 64);
    if (frame_index != 0) {
        std::memcpy(_params.prev_view, prev.view.data(), 64);
        std::array<double, 16> pv;
        for (int i = 0; i < 16; ++i) pv[i] = prev.view[i];
        std::array<double, 16> inv = inverse(pv);
        for (int r = 0; r < 3; ++r) _params.prev_origin[r] = (float)inv[12 + r];
        _params.prev_origin[3] = 1.0f;
    }
    _params.frame_number = (uint32_t)frame_index;
}

// ---------------------------------------------------------------------------
// BMFR
// ---------------------------------------------------------------------------
// fitting_kernel is implied by the block size in the C ABI (64 threads for 8x8
// blocks, 256 otherwise, as DenoiserUtils.cpp passes it).
BMFR::BMFR(uint32_t width, uint32_t height, uint32_t work_width, uint32_t work_height, GBufferPtr g_buffer,
           IlluminationBufferPtr illu_buffer, AccumulationBufferPtr acc_buffer, uint32_t)
    : _width(width),
      _height(height),
      _work_width(work_width),
      _g_buffer(g_buffer),
      _illumination(illu_buffer),
      _acc(acc_buffer) {
    if (illu_buffer->kind != IlluminationKind::Demodulated)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Illumination Buffer type is required to be IlluminationBufferDemodulated for BMFR");
    if (work_width != work_height || (work_width != 8 && work_width != 16 && work_width != 32))
        throw Exception(PBRT_ERR_UNSUPPORTED, "BMFR: square 8, 16 or 32 pixel blocks");
    _accumulated_illumination = std::make_shared<DeviceImage>(width, height, Format::RGBA16F, 2);
    _final_illumination = std::make_shared<DeviceImage>(width, height, Format::BGRA8);
}

void BMFR::compile(Context&) {
    _accumulated_illumination->compile();
    _final_illumination->compile();
    pbrt_bmfr_params p{};
    p.width = _width;
    p.height = _height;
    p.block_size = _work_width;
    p.fit = fit;
    size_t bytes = 0;
    check(pbrt_bmfr_workspace_size(&p, &bytes));
    _workspace = std::make_unique<DeviceBuffer>(bytes);
    _workspace->compile();
}

void BMFR::add_dispatch_to_command_graph(const CommandsPtr& command_graph, const PushConstantsPtr& push_constants) {
    if (!_workspace) throw Exception(PBRT_ERR_INVALID_ARG, "BMFR: compile() was not called");
    command_graph->push_constants(push_constants);
    command_graph->addChild([this, push_constants](hipStream_t s) {
        pbrt_bmfr_params p{};
        p.width = _width;
        p.height = _height;
        p.block_size = _work_width;
        p.frame_number = push_constants->value().frame_number;
        p.fit = fit;
        pbrt_bmfr_images im{};
        im.depth = _g_buffer->depth->as<float>();
        im.normal = _g_buffer->normal->as<float>();
        im.albedo = _g_buffer->albedo->as<uint8_t>();
        im.motion = _acc->motion->as<uint16_t>();
        im.spp = _acc->spp->as<uint8_t>();
        im.noisy = _illumination->illumination_images[0]->as<uint16_t>();
        im.denoised = _accumulated_illumination->as<uint16_t>();
        im.final_image = _final_illumination->as<uint8_t>();
        im.workspace = _workspace->data();
        im.workspace_bytes = _workspace->bytes();
        check(pbrt_bmfr(&p, &im, s));
    });
}

// ---------------------------------------------------------------------------
// BFR / BFRBlender (BFR.cpp, BFRBlender.cpp)
// ---------------------------------------------------------------------------
BFR::BFR(uint32_t width, uint32_t height, uint32_t work_width, uint32_t work_height, GBufferPtr g_buffer,
         IlluminationBufferPtr illu_buffer, AccumulationBufferPtr acc_buffer)
    : _width(width), _height(height), _work_width(work_width), _g_buffer(g_buffer), _illumination(illu_buffer),
      _acc(acc_buffer) {
    if (illu_buffer->kind != IlluminationKind::Demodulated)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Illumination Buffer type is required to be IlluminationBufferDemodulated for BFR");
    if (work_width != work_height || (work_width != 8 && work_width != 16 && work_width != 32))
        throw Exception(PBRT_ERR_UNSUPPORTED, "BFR: square 8, 16 or 32 pixel blocks");
    _accumulated_illumination = std::make_shared<DeviceImage>(width, height, Format::RGBA16F, 2);
    _final_illumination = std::make_shared<DeviceImage>(width, height, Format::BGRA8);
}

void BFR::compile(Context&) {
    _accumulated_illumination->compile();
    _final_illumination->compile();
}

void BFR::add_dispatch_to_command_graph(const CommandsPtr& command_graph, const PushConstantsPtr& push_constants) {
    command_graph->push_constants(push_constants);
    command_graph->addChild([this, push_constants](hipStream_t s) {
        pbrt_bfr_params p{};
        p.width = _width;
        p.height = _height;
        p.block_size = _work_width;
        p.frame_number = push_constants->value().frame_number;
        pbrt_bfr_images im{};
        im.depth = _g_buffer->depth->as<float>();
        im.normal = _g_buffer->normal->as<float>();
        im.albedo = _g_buffer->albedo->as<uint8_t>();
        im.motion = _acc->motion->as<uint16_t>();
        im.spp = _acc->spp->as<uint8_t>();
        im.noisy = _illumination->illumination_images[0]->as<uint16_t>();
        im.denoised = _accumulated_illumination->as<uint16_t>();
        im.final_image = _final_illumination->as<uint8_t>();
        check(pbrt_bfr(&p, &im, s));
    });
}

BFRBlender::BFRBlender(uint32_t width, uint32_t height, DeviceImagePtr average_image,
                       DeviceImagePtr average_squared_image, DeviceImagePtr denoised0, DeviceImagePtr denoised1,
                       DeviceImagePtr denoised2, uint32_t, uint32_t, uint32_t filter_radius)
    : _width(width), _height(height), _filter_radius(filter_radius), _average(average_image),
      _average_squared(average_squared_image), _den8(denoised0), _den16(denoised1), _den32(denoised2) {
    _final = std::make_shared<DeviceImage>(width, height, Format::BGRA8);
}

void BFRBlender::compile(Context&) { _final->compile(); }

void BFRBlender::add_dispatch_to_command_graph(const CommandsPtr& command_graph) {
    command_graph->addChild([this](hipStream_t s) {
        pbrt_bfr_blend_images im{};
        im.average = _average->as<uint16_t>();
        im.average_squared = _average_squared->as<uint16_t>();
        im.denoised8 = _den8->as<uint8_t>();
        im.denoised16 = _den16->as<uint8_t>();
        im.denoised32 = _den32->as<uint8_t>();
        im.final_image = _final->as<uint8_t>();
        check(pbrt_bfr_blend(_width, _height, _filter_radius, &im, s));
    });
}

// ---------------------------------------------------------------------------
// LearnedDenoiser
// ---------------------------------------------------------------------------
LearnedDenoiser::LearnedDenoiser(uint32_t width, uint32_t height, GBufferPtr g_buffer,
                                 IlluminationBufferPtr illu_buffer, AccumulationBufferPtr acc_buffer,
                                 std::string weights_path)
    : _width(width), _height(height), _g_buffer(g_buffer), _illumination(illu_buffer), _acc(acc_buffer),
      _weights_path(std::move(weights_path)) {
    if (illu_buffer->kind != IlluminationKind::Demodulated)
        throw Exception(PBRT_ERR_UNSUPPORTED,
                        "Illumination Buffer type is required to be IlluminationBufferDemodulated for the learned "
                        "denoiser");
    _final_illumination = std::make_shared<DeviceImage>(width, height, Format::BGRA8);
}

LearnedDenoiser::~LearnedDenoiser() {
    if (_net) (void)pbrt_cnn_destroy(_net);
}

std::string LearnedDenoiser::default_weights_path() {
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void*>(&pbrt_cnn_create), &info) || !info.dli_fname)
        throw Exception(PBRT_ERR_INTERNAL, "LearnedDenoiser: cannot locate libpbrt.so");
    std::string lib = info.dli_fname;
    const size_t slash = lib.find_last_of('/');
    return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/weights/kpn6.bin";
}

void LearnedDenoiser::compile(Context&) {
    _final_illumination->compile();
    // the parameter file's length names the network: the layer count whose
    // canonical parameter count it holds
    std::ifstream f(_weights_path, std::ios::binary | std::ios::ate);
    if (!f) throw Exception(PBRT_ERR_INVALID_ARG, "LearnedDenoiser: cannot open " + _weights_path);
    const std::streamsize bytes_in_file = f.tellg();
    uint32_t layers = 0, count = 0;
    for (uint32_t l = 2; l <= 6 && !layers; ++l) {
        uint32_t n = 0;
        check(pbrt_cnn_layers_parameter_count(l, &n));
        if (bytes_in_file == (std::streamsize)(n * sizeof(float))) layers = l, count = n;
    }
    if (!layers)
        throw Exception(PBRT_ERR_INVALID_ARG,
                        "LearnedDenoiser: " + _weights_path + " is not the float32 parameters of a 2..6-layer network");
    std::vector<float> params(count);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(params.data()), (std::streamsize)(count * sizeof(float)));
    if (f.gcount() != (std::streamsize)(count * sizeof(float)))
        throw Exception(PBRT_ERR_INVALID_ARG, "LearnedDenoiser: cannot read " + _weights_path);
    if (_net) (void)pbrt_cnn_destroy(_net);
    _net = nullptr;
    // two layers whose second convolution has only its centre tap (weights/kpnc.bin) run
    // as the compact network: the same function on its faster fused kernel
    bool pointwise = layers == 2;
    if (pointwise) {
        uint32_t first = 0;
        check(pbrt_cnn_layers_parameter_count(2, &first));
        const size_t second = 25 * 32 * 9 + 25;  // W[25][32][3][3], b[25]
        first -= (uint32_t)second;
        for (size_t i = 0; i < 25 * 32 * 9 && pointwise; ++i)
            if (i % 9 != 4 && params[first + i] != 0.0f) pointwise = false;
    }
    if (pointwise)
        check(pbrt_cnn_create_compact(1, &_net));
    else
        check(pbrt_cnn_create_layers(layers, 1, &_net));
    check(pbrt_cnn_set_parameters(_net, params.data()));
    size_t bytes = 0;
    check(pbrt_cnn_workspace_size(_width, _height, 1, 0, &bytes));
    _workspace = std::make_unique<DeviceBuffer>(bytes);
    _workspace->compile();
}

void LearnedDenoiser::add_dispatch_to_command_graph(const CommandsPtr& command_graph) {
    if (!_workspace) throw Exception(PBRT_ERR_INVALID_ARG, "LearnedDenoiser: compile() was not called");
    command_graph->addChild([this](hipStream_t s) {
        pbrt_cnn_images im{};
        im.width = _width;
        im.height = _height;
        im.count = 1;
        im.noisy = _illumination->illumination_images[0]->as<uint16_t>();
        im.depth = _g_buffer->depth->as<float>();
        im.normal = _g_buffer->normal->as<float>();
        im.albedo = _g_buffer->albedo->as<uint8_t>();
        im.spp = _acc->spp->as<uint8_t>();
        im.motion = _acc->motion->as<uint16_t>();
        im.final_image = _final_illumination->as<uint8_t>();
        im.workspace = _workspace->data();
        im.workspace_bytes = _workspace->bytes();
        check(pbrt_cnn_denoise(_net, &im, s));
    });
}

// ---------------------------------------------------------------------------
// Taa
// ---------------------------------------------------------------------------
Taa::Taa(uint32_t width, uint32_t height, uint32_t, uint32_t, GBufferPtr, AccumulationBufferPtr acc_buffer,
         DeviceImagePtr denoised)
    : _width(width), _height(height), _acc(acc_buffer), _denoised(denoised) {
    _accumulation_image = std::make_shared<DeviceImage>(width, height, Format::RGBA8);
    _final_image = std::make_shared<DeviceImage>(width, height, Format::BGRA8);
}

void Taa::compile(Context&) {
    _final_image->compile();
    _accumulation_image->compile();
}

void Taa::add_dispatch_to_command_graph(const CommandsPtr& command_graph) {
    // every block the modules push is a PushConstants<RayTracingPushConstants>
    PushConstantsPtr push_constants =
        std::static_pointer_cast<PushConstants<RayTracingPushConstants>>(command_graph->pushed_constants());
    if (!push_constants)
        throw Exception(PBRT_ERR_INVALID_ARG,
                        "Taa: no push constants were recorded before it (taa.comp reads frameNumber from the "
                        "block the denoiser pushes)");
    command_graph->addChild([this, push_constants](hipStream_t s) {
        pbrt_taa_images im{};
        im.motion = _acc->motion->as<uint16_t>();
        im.denoised = _denoised->as<uint8_t>();
        im.history = _accumulation_image->as<uint8_t>();
        im.final_image = _final_image->as<uint8_t>();
        check(pbrt_taa(_width, _height, push_constants->value().frame_number, &im, s));
    });
}

// ---------------------------------------------------------------------------
// FormatConverter
// ---------------------------------------------------------------------------
FormatConverter::FormatConverter(DeviceImagePtr src_image, Format dst_format, int, int) : _src(src_image) {
    if (src_image->format() != Format::RGBA32F || dst_format != Format::BGRA8)
        throw Exception(PBRT_ERR_UNSUPPORTED, "FormatConverter: RGBA32F -> B8G8R8A8 only");
    final_image = std::make_shared<DeviceImage>(src_image->width(), src_image->height(), Format::BGRA8);
}

void FormatConverter::compile_images(Context&) const { final_image->compile(); }

void FormatConverter::add_dispatch_to_command_graph(const CommandsPtr& command_graph) {
    command_graph->addChild([this](hipStream_t s) {
        check(pbrt_convert(_src->width(), _src->height(), _src->as<float>(), final_image->as<uint8_t>(), s));
    });
}

// ---------------------------------------------------------------------------
// DenoiserUtils.cpp:8-130
// ---------------------------------------------------------------------------
void add_denoiser_to_commands(DenoisingType denoising_type, DenoisingBlockSize denoising_size,
                              const CommandsPtr& commands, Context& compile, int width, int height,
                              const PushConstantsPtr& compute_constants, const GBufferPtr& g_buffer,
                              const IlluminationBufferPtr& illumination_buffer,
                              const AccumulationBufferPtr& accumulation_buffer, DeviceImagePtr& final_descriptor_image) {
    switch (denoising_type) {
        case DenoisingType::NONE:
            break;
        case DenoisingType::BMFR:
        case DenoisingType::BFR: {
            const bool bmfr = denoising_type == DenoisingType::BMFR;
            // one denoiser of block size b; returns its final image
            auto add = [&](uint32_t b) -> DeviceImagePtr {
                if (bmfr) {
                    auto d = BMFR::create(width, height, b, b, g_buffer, illumination_buffer, accumulation_buffer,
                                          b == 8 ? 64 : 256);
                    d->compile(compile);
                    d->update_image_layouts(compile);
                    d->add_dispatch_to_command_graph(commands, compute_constants);
                    commands->keep(d);
                    return d->get_final_descriptor_image();
                }
                auto d = BFR::create(width, height, b, b, g_buffer, illumination_buffer, accumulation_buffer);
                d->compile(compile);
                d->update_image_layouts(compile);
                d->add_dispatch_to_command_graph(commands, compute_constants);
                commands->keep(d);
                return d->get_final_descriptor_image();
            };
            switch (denoising_size) {
                case DenoisingBlockSize::X8: final_descriptor_image = add(8); break;
                case DenoisingBlockSize::X16: final_descriptor_image = add(16); break;
                case DenoisingBlockSize::X32: final_descriptor_image = add(32); break;
                case DenoisingBlockSize::X8X16X32: {
                    // DenoiserUtils.cpp:48-70 / 103-123.  The BMFR branch of the reference
                    // passes illumination_images[1] and [2] of a two-image buffer; both
                    // branches here blend on the accumulated illumination and its square.
                    auto d8 = add(8), d16 = add(16), d32 = add(32);
                    auto blender = BFRBlender::create(width, height, illumination_buffer->illumination_images[0],
                                                      illumination_buffer->illumination_images[1], d8, d16, d32);
                    blender->compile(compile);
                    blender->update_image_layouts(compile);
                    blender->add_dispatch_to_command_graph(commands);
                    commands->keep(blender);
                    final_descriptor_image = blender->get_final_descriptor_image();
                    break;
                }
                default:
                    throw Exception(PBRT_ERR_UNSUPPORTED, "add_denoiser_to_commands: block size not supported");
            }
            break;
        }
        case DenoisingType::LEARNED: {
            auto d = LearnedDenoiser::create(width, height, g_buffer, illumination_buffer, accumulation_buffer);
            d->compile(compile);
            d->update_image_layouts(compile);
            // the denoiser slot's compute constants stay in force for the TAA after it,
            // as the BMFR and BFR branches push them
            commands->push_constants(compute_constants);
            d->add_dispatch_to_command_graph(commands);
            commands->keep(d);
            final_descriptor_image = d->get_final_descriptor_image();
            break;
        }
        case DenoisingType::SVG:
            std::cout << "Not yet implemented" << std::endl;
            break;
    }
}

// ---------------------------------------------------------------------------
// Camera helpers (vsg transform.h:148-183), column-major
// ---------------------------------------------------------------------------
std::array<double, 16> perspective(double fovy, double aspect, double n, double f) {
    double ff = 1.0 / std::tan(fovy * 0.5);
    double r = 1.0 / (n - f);
    return {ff / aspect, 0, 0, 0, 0, -ff, 0, 0, 0, 0, f * r, -1, 0, 0, f * n * r, 0};
}

static void norm3(double* v) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

std::array<double, 16> look_at(const double* eye, const double* center, const double* up) {
    double fwd[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    norm3(fwd);
    double upn[3] = {up[0], up[1], up[2]};
    norm3(upn);
    double side[3] = {fwd[1] * upn[2] - fwd[2] * upn[1], fwd[2] * upn[0] - fwd[0] * upn[2],
                      fwd[0] * upn[1] - fwd[1] * upn[0]};
    norm3(side);
    double u[3] = {side[1] * fwd[2] - side[2] * fwd[1], side[2] * fwd[0] - side[0] * fwd[2],
                   side[0] * fwd[1] - side[1] * fwd[0]};
    norm3(u);
    std::array<double, 16> m{};
    // rows: side, u, -forward; then * translate(-eye)
    for (int c = 0; c < 3; ++c) {
        m[4 * c + 0] = side[c];
        m[4 * c + 1] = u[c];
        m[4 * c + 2] = -fwd[c];
    }
    m[15] = 1.0;
    for (int r = 0; r < 3; ++r) m[12 + r] = -(m[r] * eye[0] + m[4 + r] * eye[1] + m[8 + r] * eye[2]);
    return m;
}

std::array<double, 16> inverse(const std::array<double, 16>& m) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[4 * c + r];
            a[r][4 + c] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
        double piv = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= piv;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    std::array<double, 16> out;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[4 * c + r] = a[r][4 + c];
    return out;
}

std::array<float, 16> to_float(const std::array<double, 16>& m) {
    std::array<float, 16> f;
    for (int i = 0; i < 16; ++i) f[i] = (float)m[i];
    return f;
}

}  // namespace vkpbrt
