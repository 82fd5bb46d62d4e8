#include "Core.hpp"

namespace vkpbrt {

void check(int32_t status) {
    if (status != PBRT_OK) throw Exception(status, pbrt_last_error());
}

static void hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Exception(PBRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t format_bytes(Format f) {
    switch (f) {
        case Format::R32F: return 4;
        case Format::RG32F: return 8;
        case Format::RGBA32F: return 16;
        case Format::RG16F: return 4;
        case Format::RGBA16F: return 8;
        case Format::R8: return 1;
        case Format::RGBA8: return 4;
        case Format::BGRA8: return 4;
    }
    return 0;
}

DeviceImage::DeviceImage(uint32_t width, uint32_t height, Format format, uint32_t layers)
    : _width(width), _height(height), _layers(layers), _format(format) {
    if (width == 0 || height == 0 || layers == 0) throw Exception(PBRT_ERR_INVALID_ARG, "DeviceImage: zero extent");
}

DeviceImage::~DeviceImage() {
    if (_ptr) (void)hipFree(_ptr);
}

void DeviceImage::compile() {
    if (_ptr) return;
    hip(hipMalloc(&_ptr, bytes()), "DeviceImage::compile hipMalloc");
    hip(hipMemset(_ptr, 0, bytes()), "DeviceImage::compile hipMemset");
}

DeviceBuffer::~DeviceBuffer() {
    if (_ptr) (void)hipFree(_ptr);
}

void DeviceBuffer::compile() {
    if (_ptr || _bytes == 0) return;
    hip(hipMalloc(&_ptr, _bytes), "DeviceBuffer::compile hipMalloc");
    hip(hipMemset(_ptr, 0, _bytes), "DeviceBuffer::compile hipMemset");
}

void Commands::submit(hipStream_t stream) const {
    for (const Op& op : _ops) op(stream);
}

// Captures the op list into a graph and launches it.  The graph is recaptured
// every call (host-side argument values such as push constants are baked into
// kernel nodes) and the executable is updated in place when the topology is
// unchanged, which avoids re-instantiation.
void Commands::launch_graph(hipStream_t stream) {
    hipStream_t cap = stream;
    hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (const Op& op : _ops) op(cap);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hipGraph_t graph = nullptr;
    hip(hipStreamEndCapture(cap, &graph), "hipStreamEndCapture");
    bool updated = false;
    if (_exec) {
        hipGraphExecUpdateResult res;
        hipGraphNode_t err_node = nullptr;
        updated = hipGraphExecUpdate(_exec, graph, &err_node, &res) == hipSuccess &&
                  res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGraphExecDestroy(_exec);
            _exec = nullptr;
        }
    }
    if (!_exec) hip(hipGraphInstantiate(&_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    hip(hipGraphDestroy(graph), "hipGraphDestroy");
    hip(hipGraphLaunch(_exec, stream), "hipGraphLaunch");
}

Commands::~Commands() {
    if (_exec) (void)hipGraphExecDestroy(_exec);
}

}  // namespace vkpbrt
