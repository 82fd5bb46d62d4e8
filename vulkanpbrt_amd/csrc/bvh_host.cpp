// Binned-SAH BVH2 builder (host).  Produces the two-children-per-node layout
// described in scene_types.hpp, in depth-first order so a parent and its first
// child share cache lines.  The reference gets its BVH from the Vulkan driver
// (vkCmdBuildAccelerationStructuresKHR with PREFER_FAST_TRACE,
// external/vsg/src/vsg/raytracing/AccelerationStructure.cpp:44); this builder
// is the project's own.
#include "bvh_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace pbrt {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f && dy >= 0.0f && dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BuildNode {
    Box box;
    int32_t left = -1, right = -1;  // children in the build array
    uint32_t first = 0, count = 0;  // leaf range (count > 0 => leaf)
};

constexpr int kBins = 32;

}  // namespace

BvhHost build_bvh_sah(const float* tri_lo, const float* tri_hi, uint32_t n, int max_leaf, int max_depth) {
    BvhHost out;
    out.tri_order.resize(n);
    for (uint32_t i = 0; i < n; ++i) out.tri_order[i] = i;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > kMaxLeaf) max_leaf = kMaxLeaf;

    std::vector<float> centroid(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) centroid[3 * i + a] = 0.5f * (tri_lo[3 * i + a] + tri_hi[3 * i + a]);

    std::vector<BuildNode> nodes;
    nodes.reserve(n > 0 ? 2 * (size_t)n : 1);
    struct Task {
        int32_t node;
        uint32_t begin, end;
        int depth;
    };
    std::vector<Task> stack;
    nodes.push_back(BuildNode{});
    stack.push_back({0, 0, n, 1});
    uint32_t* order = out.tri_order.data();
    uint32_t max_seen_depth = 1;

    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        BuildNode& nd = nodes[t.node];
        Box box, cbox;
        box.reset();
        cbox.reset();
        for (uint32_t i = t.begin; i < t.end; ++i) {
            uint32_t tri = order[i];
            Box b;
            std::memcpy(b.lo, tri_lo + 3 * (size_t)tri, sizeof(b.lo));
            std::memcpy(b.hi, tri_hi + 3 * (size_t)tri, sizeof(b.hi));
            box.grow(b);
            cbox.grow(&centroid[3 * (size_t)tri]);
        }
        nd.box = box;
        uint32_t count = t.end - t.begin;
        max_seen_depth = std::max<uint32_t>(max_seen_depth, (uint32_t)t.depth);
        if (count <= (uint32_t)max_leaf && count > 0 && (count <= 2 || t.depth >= max_depth)) {
            nd.first = t.begin;
            nd.count = count;
            continue;
        }
        // choose axis of largest centroid extent
        int axis = 0;
        float ext[3];
        for (int a = 0; a < 3; ++a) ext[a] = cbox.hi[a] - cbox.lo[a];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t mid = t.begin + count / 2;
        bool make_leaf = false;
        // Levels that median splits still need to bring this range down to leaves;
        // once SAH splits could overrun the depth budget, split at the median.
        int need = 0;
        for (uint32_t c = count; c > (uint32_t)max_leaf; c = (c + 1) / 2) ++need;
        if (ext[axis] <= 0.0f || t.depth + need >= max_depth) {
            // degenerate centroid spread or depth budget reached: median split
            if (count <= (uint32_t)max_leaf) make_leaf = true;
            if (!make_leaf && ext[axis] > 0.0f) {
                std::nth_element(order + t.begin, order + mid, order + t.end, [&](uint32_t a, uint32_t b) {
                    return centroid[3 * (size_t)a + axis] < centroid[3 * (size_t)b + axis];
                });
            }
        } else {
            Box bin_box[kBins];
            uint32_t bin_cnt[kBins];
            for (int b = 0; b < kBins; ++b) {
                bin_box[b].reset();
                bin_cnt[b] = 0;
            }
            float scale = (float)kBins / ext[axis];
            float base = cbox.lo[axis];
            auto bin_of = [&](uint32_t tri) {
                int b = (int)((centroid[3 * (size_t)tri + axis] - base) * scale);
                return std::min(std::max(b, 0), kBins - 1);
            };
            for (uint32_t i = t.begin; i < t.end; ++i) {
                uint32_t tri = order[i];
                int b = bin_of(tri);
                Box tb;
                std::memcpy(tb.lo, tri_lo + 3 * (size_t)tri, sizeof(tb.lo));
                std::memcpy(tb.hi, tri_hi + 3 * (size_t)tri, sizeof(tb.hi));
                bin_box[b].grow(tb);
                bin_cnt[b]++;
            }
            float right_area[kBins];
            uint32_t right_cnt[kBins];
            Box acc;
            acc.reset();
            uint32_t c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bin_box[b]);
                c += bin_cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = c;
            }
            acc.reset();
            c = 0;
            float best = std::numeric_limits<float>::infinity();
            int best_split = -1;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bin_box[b]);
                c += bin_cnt[b];
                if (c == 0 || right_cnt[b + 1] == 0) continue;
                float cost = acc.area() * (float)c + right_area[b + 1] * (float)right_cnt[b + 1];
                if (cost < best) {
                    best = cost;
                    best_split = b;
                }
            }
            float parent_area = box.area();
            // SAH with unit traversal and intersection costs
            float split_cost = parent_area > 0.0f ? 1.0f + best / parent_area : std::numeric_limits<float>::infinity();
            float leaf_cost = (float)count;
            if (best_split < 0) {
                if (count <= (uint32_t)max_leaf) make_leaf = true;
            } else if (count <= (uint32_t)max_leaf && leaf_cost <= split_cost) {
                make_leaf = true;
            } else {
                uint32_t* p = std::partition(order + t.begin, order + t.end,
                                             [&](uint32_t tri) { return bin_of(tri) <= best_split; });
                mid = (uint32_t)(p - order);
                if (mid == t.begin || mid == t.end) mid = t.begin + count / 2;
            }
        }
        if (make_leaf) {
            nd.first = t.begin;
            nd.count = count;
            continue;
        }
        int32_t l = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        int32_t r = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        nodes[t.node].left = l;
        nodes[t.node].right = r;
        // push right first so the left subtree is processed (and laid out) first
        stack.push_back({r, mid, t.end, t.depth + 1});
        stack.push_back({l, t.begin, mid, t.depth + 1});
    }
    out.depth = max_seen_depth;

    // Emit the two-children layout in depth-first order.
    auto leaf_code = [](const BuildNode& b) -> int32_t {
        return ~(int32_t)((b.first << kLeafBits) | (b.count - 1));
    };
    const float inf = std::numeric_limits<float>::infinity();
    auto write_child = [&](float* dst, int slot, const Box& b) {
        // slot 0: floats 0..3 and 8..9 ; slot 1: floats 4..7 and 10..11
        float* xy = dst + 4 * slot;
        xy[0] = b.lo[0];
        xy[1] = b.hi[0];
        xy[2] = b.lo[1];
        xy[3] = b.hi[1];
        dst[8 + 2 * slot] = b.lo[2];
        dst[9 + 2 * slot] = b.hi[2];
    };
    std::vector<int32_t> out_index(nodes.size(), -1);
    // count inner nodes & assign DFS order
    std::vector<int32_t> order_stack;
    std::vector<int32_t> dfs;
    if (n > 0) order_stack.push_back(0);
    while (!order_stack.empty()) {
        int32_t b = order_stack.back();
        order_stack.pop_back();
        if (nodes[b].count > 0) continue;  // leaves are folded into their parent
        out_index[b] = (int32_t)dfs.size();
        dfs.push_back(b);
        order_stack.push_back(nodes[b].right);
        order_stack.push_back(nodes[b].left);
    }
    if (n == 0 || nodes[0].count > 0) {
        // whole scene is one leaf (or empty): a root with one real and one empty child
        out.nodes.assign(16, 0.0f);
        Box empty;
        empty.reset();
        if (n > 0) {
            write_child(out.nodes.data(), 0, nodes[0].box);
            int32_t code = leaf_code(nodes[0]);
            std::memcpy(&out.nodes[12], &code, 4);
        } else {
            write_child(out.nodes.data(), 0, empty);
            int32_t code = ~0;
            std::memcpy(&out.nodes[12], &code, 4);
        }
        write_child(out.nodes.data(), 1, empty);
        int32_t code = ~0;
        std::memcpy(&out.nodes[13], &code, 4);
        out.depth = 1;
        (void)inf;
        return out;
    }
    out.nodes.assign(16 * dfs.size(), 0.0f);
    for (size_t k = 0; k < dfs.size(); ++k) {
        const BuildNode& b = nodes[dfs[k]];
        float* dst = &out.nodes[16 * k];
        const BuildNode& c0 = nodes[b.left];
        const BuildNode& c1 = nodes[b.right];
        write_child(dst, 0, c0.box);
        write_child(dst, 1, c1.box);
        int32_t i0 = c0.count > 0 ? leaf_code(c0) : out_index[b.left];
        int32_t i1 = c1.count > 0 ? leaf_code(c1) : out_index[b.right];
        std::memcpy(&dst[12], &i0, 4);
        std::memcpy(&dst[13], &i1, 4);
    }
    return out;
}

}  // namespace pbrt
