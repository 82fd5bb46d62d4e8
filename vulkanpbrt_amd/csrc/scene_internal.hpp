// Host-side packed scene shared by the scene builder and the kernels' launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/pbrt_c.h"
#include "scene_types.hpp"

namespace pbrt {

// Per-thread traversal stack entries (LDS, 4 B each).  The BVH builders bound
// the tree depth by this value (median splits once SAH splits could overrun
// it), and a traversal holds at most depth - 1 deferred children.  40 entries
// keep the stack at 10 KB per 64-lane wave: 16 resident waves per CU (four per
// SIMD, the persistent and wavefront kernels' and the deep-tree megakernel's
// register budget) fill the 160 KB of LDS exactly.
constexpr int kTraversalStackSize = 40;
// Trees of depth <= 32 (the host builders' on the atrium and the city) run the
// megakernel with a 32-entry stack: 8 KB per wave, so five waves per SIMD fit the
// LDS, and the kernel is built for five (at most 96 registers, spilling a few
// path-state values around the bounce loop, none inside the walk).
constexpr int kShortStackSize = 32;

// Largest leaf the GPU PLOC builder's SAH collapse may form (the node format holds 16).
constexpr int kPlocMaxLeaf = 8;

struct PackedScene {
    std::vector<DeviceTexture> textures;
    std::vector<uint32_t> texels;
    std::vector<DeviceMesh> meshes;
    std::vector<DeviceMaterial> materials;
    std::vector<float> positions, normals, uvs;
    std::vector<uint32_t> index_words;
    std::vector<bool> mesh_opaque;
    std::vector<DeviceInstance> instances;
    std::vector<DeviceLight> lights;
    float light_strength_sum = 0.0f;
    std::vector<float> world_tris;  // 9 floats per triangle, instance order
    std::vector<uint32_t> tri_inst, tri_prim;
};

void pack_scene(const pbrt_scene_desc& d, PackedScene& ps);
const DeviceScene& scene_device_view(const pbrt_scene* s);
uint32_t scene_bvh_depth(const pbrt_scene* s);

// Opt-in per-stage timing of pbrt_trace (pbrt_trace_timing_begin/_end): HIP
// events recorded on the launch stream at every stage boundary; each pair of
// consecutive events of a frame is one segment charged to a stage.
enum TimingKind { kTimeExtend = 0, kTimeShade, kTimeShadow, kTimeMegakernel, kTimeOther, kTimeKinds };

class TraceTiming {
public:
    ~TraceTiming();
    void start();                          // clears the log and activates recording
    bool active() const { return active_; }
    void frame_begin(hipStream_t st);      // first event of a pbrt_trace call
    void mark(int kind, hipStream_t st);   // closes the segment since the previous event
    void finish(pbrt_trace_timing* out);   // waits for the last event, sums, deactivates

private:
    uint32_t record(hipStream_t st);
    struct Segment {
        int kind;
        uint32_t a, b;
    };
    bool active_ = false;
    std::vector<hipEvent_t> events_;
    uint32_t used_ = 0, last_ = 0;
    uint64_t frames_ = 0;
    std::vector<Segment> segments_;
};

// The scene's timing log when recording is active, else nullptr.
TraceTiming* scene_timing(const pbrt_scene* s);

// Work counters of the persistent tracer: each call returns the next slot of a
// ring of kWorkSlots zeroed pairs (next pixel, finished waves); the launch's last
// wave sets its slot back to zero, so traces of one scene that run at the same
// time (on several streams) use different slots while fewer than kWorkSlots are
// in flight.
constexpr uint32_t kWorkSlots = 64;
constexpr uint32_t kWorkSlotWords = 16;
uint32_t* scene_work_counter(const pbrt_scene* s);

}  // namespace pbrt
