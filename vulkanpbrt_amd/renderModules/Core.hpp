// Thin replacements for the VSG types the render modules are written against
// (source/renderModules/*): vsg::Context -> Context (a HIP stream),
// vsg::Commands -> Commands (a recorded op list that can be replayed eagerly or
// captured into a hipGraph), vsg::DescriptorImage -> DeviceImage (a device
// allocation with a format), vsg::PushConstants -> PushConstants (a host value
// read when the ops run).  No Vulkan types appear anywhere.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbrt_c.h"

namespace vkpbrt {

// Project exception carrying the C ABI status and message (the reference throws
// vsg::Exception{message}).
struct Exception : std::runtime_error {
    int32_t status;
    Exception(int32_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void check(int32_t status);  // throws Exception with pbrt_last_error() on failure

enum class Format {
    R32F,
    RG32F,
    RGBA32F,
    RG16F,
    RGBA16F,
    R8,
    RGBA8,
    BGRA8,
};
size_t format_bytes(Format f);

// Device image: width x height x layers texels of one format, tightly packed.
class DeviceImage {
public:
    DeviceImage(uint32_t width, uint32_t height, Format format, uint32_t layers = 1);
    ~DeviceImage();
    DeviceImage(const DeviceImage&) = delete;
    DeviceImage& operator=(const DeviceImage&) = delete;
    void compile();  // allocates (idempotent), zero-filled
    void* data() const { return _ptr; }
    template <class T>
    T* as() const {
        return static_cast<T*>(_ptr);
    }
    uint32_t width() const { return _width; }
    uint32_t height() const { return _height; }
    uint32_t layers() const { return _layers; }
    Format format() const { return _format; }
    size_t bytes() const { return (size_t)_width * _height * _layers * format_bytes(_format); }

private:
    uint32_t _width, _height, _layers;
    Format _format;
    void* _ptr = nullptr;
};
using DeviceImagePtr = std::shared_ptr<DeviceImage>;

// Device scratch owned by a module.
class DeviceBuffer {
public:
    explicit DeviceBuffer(size_t bytes) : _bytes(bytes) {}
    ~DeviceBuffer();
    void compile();
    void* data() const { return _ptr; }
    size_t bytes() const { return _bytes; }

private:
    size_t _bytes;
    void* _ptr = nullptr;
};

struct Context {
    hipStream_t stream = nullptr;
};

// A push-constant block shared by the ops that read it (the value is read when
// the recorded ops run, so mutating it between frames works like vsg).
template <class T>
class PushConstants {
public:
    T& value() { return _value; }
    const T& value() const { return _value; }

private:
    T _value{};
};

// Recorded command list (replaces vsg::Commands).  Ops run in order on the
// stream; submit() replays them eagerly, capture()/launch() use a hipGraph
// rebuilt whenever the op list or the push-constant values change.
class Commands {
public:
    using Op = std::function<void(hipStream_t)>;
    void addChild(Op op) { _ops.push_back(std::move(op)); }
    // keeps a module alive as long as the recorded ops that reference it
    void keep(std::shared_ptr<void> owner) { _owned.push_back(std::move(owner)); }
    // vkCmdPushConstants: records `block` as the push-constant block in force for
    // the ops recorded after it, as a command buffer keeps the last pushed values
    // for later dispatches (taa.comp reads frameNumber from the block the denoiser
    // pushed, Taa.cpp:99-107).  The value is still read when the ops run.
    void push_constants(std::shared_ptr<void> block) { _pushed = std::move(block); }
    const std::shared_ptr<void>& pushed_constants() const { return _pushed; }
    size_t size() const { return _ops.size(); }
    void submit(hipStream_t stream) const;
    void launch_graph(hipStream_t stream);
    ~Commands();

private:
    std::vector<Op> _ops;
    std::vector<std::shared_ptr<void>> _owned;
    std::shared_ptr<void> _pushed;
    hipGraphExec_t _exec = nullptr;
};
using CommandsPtr = std::shared_ptr<Commands>;

}  // namespace vkpbrt
