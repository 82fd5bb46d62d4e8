// Frame graph: one hipGraph per frame, captured from the frame entry points and
// launched with a single call (the recordAndSubmit of source/VulkanPBRT.cpp:588
// over the command graph assembled at source/VulkanPBRT.cpp:543).
//
// The camera, frame number and ping-pong image pointers are kernel arguments, so
// every frame is captured again; hipGraphExecUpdate patches the arguments of the
// existing executable graph when the node topology matches (the usual case: the
// same launches every frame) and the graph is instantiated only when it changes.
#include "capi_common.hpp"

struct pbrt_frame_graph {
    hipGraphExec_t exec = nullptr;
    hipStream_t capturing = nullptr;
    bool active = false;
};

namespace {

void abandon_capture(pbrt_frame_graph* g) {
    hipGraph_t partial = nullptr;
    (void)hipStreamEndCapture(g->capturing, &partial);
    if (partial) (void)hipGraphDestroy(partial);
    g->active = false;
    g->capturing = nullptr;
}

}  // namespace

extern "C" {

int32_t pbrt_frame_graph_create(pbrt_frame_graph** out_graph) {
    return pbrt::guard([&] {
        PBRT_REQUIRE(out_graph, "pbrt_frame_graph_create: out_graph is NULL");
        *out_graph = new pbrt_frame_graph();
    });
}

int32_t pbrt_frame_graph_destroy(pbrt_frame_graph* graph) {
    return pbrt::guard([&] {
        if (!graph) return;
        if (graph->active) abandon_capture(graph);
        if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
        delete graph;
    });
}

int32_t pbrt_frame_graph_abort(pbrt_frame_graph* graph) {
    return pbrt::guard([&] {
        PBRT_REQUIRE(graph, "pbrt_frame_graph_abort: graph is NULL");
        if (graph->active) abandon_capture(graph);
    });
}

int32_t pbrt_frame_graph_begin(pbrt_frame_graph* graph, void* stream) {
    return pbrt::guard([&] {
        PBRT_REQUIRE(graph, "pbrt_frame_graph_begin: graph is NULL");
        PBRT_REQUIRE(stream, "pbrt_frame_graph_begin: the null stream cannot be captured");
        PBRT_REQUIRE(!graph->active, "pbrt_frame_graph_begin: a capture is already open");
        hipStream_t st = (hipStream_t)stream;
        PBRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        graph->capturing = st;
        graph->active = true;
    });
}

int32_t pbrt_frame_graph_end(pbrt_frame_graph* graph, void* stream, uint32_t* node_count, uint32_t* instantiated) {
    return pbrt::guard([&] {
        PBRT_REQUIRE(graph, "pbrt_frame_graph_end: graph is NULL");
        PBRT_REQUIRE(graph->active, "pbrt_frame_graph_end: no capture is open");
        if ((hipStream_t)stream != graph->capturing) {
            abandon_capture(graph);
            pbrt::fail(PBRT_ERR_INVALID_ARG, "pbrt_frame_graph_end: stream differs from the one passed to begin");
        }
        hipStream_t st = graph->capturing;
        graph->active = false;
        graph->capturing = nullptr;
        hipGraph_t g = nullptr;
        PBRT_HIP(hipStreamEndCapture(st, &g));
        struct GraphHolder {
            hipGraph_t g;
            ~GraphHolder() {
                if (g) (void)hipGraphDestroy(g);
            }
        } holder{g};
        if (node_count) {
            size_t n = 0;
            PBRT_HIP(hipGraphGetNodes(g, nullptr, &n));
            *node_count = (uint32_t)n;
        }
        bool updated = false;
        if (graph->exec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(graph->exec, g, &err_node, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();  // a refused update is not an error of this call
                (void)hipGraphExecDestroy(graph->exec);
                graph->exec = nullptr;
            }
        }
        if (!graph->exec) PBRT_HIP(hipGraphInstantiate(&graph->exec, g, nullptr, nullptr, 0));
        if (instantiated) *instantiated = updated ? 0u : 1u;
        PBRT_HIP(hipGraphLaunch(graph->exec, st));
    });
}

}  // extern "C"
