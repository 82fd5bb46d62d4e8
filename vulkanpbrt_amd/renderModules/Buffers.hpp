// Frame buffers with the reference's image sets and formats
// (source/buffers/GBuffer.cpp:53-125, IlluminationBuffer.cpp:73-308,
// AccumulationBuffer.cpp:245-339).
#pragma once

#include <memory>

#include "Core.hpp"

namespace vkpbrt {

class GBuffer {
public:
    GBuffer(uint32_t width, uint32_t height);
    static std::shared_ptr<GBuffer> create(uint32_t w, uint32_t h) { return std::make_shared<GBuffer>(w, h); }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    DeviceImagePtr depth;     // R32F
    DeviceImagePtr normal;    // RG32F spherical (acos z, atan(y, x))
    DeviceImagePtr material;  // RGBA8 category / 255
    DeviceImagePtr albedo;    // RGBA8
};
using GBufferPtr = std::shared_ptr<GBuffer>;

enum class IlluminationKind {
    FinalFloat,          // IlluminationBufferFinalFloat: RGBA32F progressive final image
    DemodulatedFloat,    // IlluminationBufferDemodulatedFloat: RGBA32F demodulated radiance
    Demodulated,         // IlluminationBufferDemodulated: RGBA16F illumination + illumination^2
};

class IlluminationBuffer {
public:
    IlluminationBuffer(IlluminationKind kind, uint32_t width, uint32_t height);
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    IlluminationKind kind;
    std::vector<DeviceImagePtr> illumination_images;
};
using IlluminationBufferPtr = std::shared_ptr<IlluminationBuffer>;

inline IlluminationBufferPtr IlluminationBufferFinalFloat(uint32_t w, uint32_t h) {
    return std::make_shared<IlluminationBuffer>(IlluminationKind::FinalFloat, w, h);
}
inline IlluminationBufferPtr IlluminationBufferDemodulatedFloat(uint32_t w, uint32_t h) {
    return std::make_shared<IlluminationBuffer>(IlluminationKind::DemodulatedFloat, w, h);
}
inline IlluminationBufferPtr IlluminationBufferDemodulated(uint32_t w, uint32_t h) {
    return std::make_shared<IlluminationBuffer>(IlluminationKind::Demodulated, w, h);
}

class AccumulationBuffer {
public:
    AccumulationBuffer(uint32_t width, uint32_t height);
    static std::shared_ptr<AccumulationBuffer> create(uint32_t w, uint32_t h) {
        return std::make_shared<AccumulationBuffer>(w, h);
    }
    void compile(Context& context);
    void update_image_layouts(Context&) {}
    // AccumulationBuffer::copy_to_back_images (AccumulationBuffer.cpp:72-244)
    void copy_to_back_images(const CommandsPtr& commands, const GBufferPtr& g_buffer,
                             const IlluminationBufferPtr& illumination_buffer);
    DeviceImagePtr spp, prev_spp;          // R8
    DeviceImagePtr prev_depth;             // R32F
    DeviceImagePtr prev_normal;            // RG32F
    DeviceImagePtr motion;                 // RG16F
    DeviceImagePtr prev_illu, prev_illu_squared;  // RGBA16F
    uint32_t width, height;
};
using AccumulationBufferPtr = std::shared_ptr<AccumulationBuffer>;

}  // namespace vkpbrt
