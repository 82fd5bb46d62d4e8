// Binned-SAH BVH2 builder (host): 32 bins on each of the three axes.  Produces the two-children-per-node layout
// described in scene_types.hpp, in depth-first order so a parent and its first
// child share cache lines.  The reference gets its BVH from the Vulkan driver
// (vkCmdBuildAccelerationStructuresKHR with PREFER_FAST_TRACE,
// external/vsg/src/vsg/raytracing/AccelerationStructure.cpp:44); this builder
// is the project's own.
#include "bvh_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

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
// SAH cost of visiting an inner node (both child boxes tested), in units of one
// triangle test.
constexpr float kTraversalCost = 1.5f;

// Emit the two-children layout in depth-first order (leaves folded into their
// parent's child references).
void emit_layout(const std::vector<BuildNode>& nodes, bool nonempty, BvhHost& out) {

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
    if (nonempty) order_stack.push_back(0);
    while (!order_stack.empty()) {
        int32_t b = order_stack.back();
        order_stack.pop_back();
        if (nodes[b].count > 0) continue;  // leaves are folded into their parent
        out_index[b] = (int32_t)dfs.size();
        dfs.push_back(b);
        order_stack.push_back(nodes[b].right);
        order_stack.push_back(nodes[b].left);
    }
    if (!nonempty || nodes[0].count > 0) {
        // whole scene is one leaf (or empty): a root with one real and one empty child
        out.nodes.assign(16, 0.0f);
        Box empty;
        empty.reset();
        if (nonempty) {
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
        return;
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
}

}  // namespace

BvhHost build_bvh_sah(const float* tri_lo, const float* tri_hi, uint32_t n, int max_leaf, int max_depth,
                      const float* weight) {
    BvhHost out;
    out.tri_order.resize(n);
    for (uint32_t i = 0; i < n; ++i) out.tri_order[i] = i;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > kMaxLeaf) max_leaf = kMaxLeaf;

    std::vector<float> centroid(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) centroid[3 * i + a] = 0.5f * (tri_lo[3 * i + a] + tri_hi[3 * i + a]);

    struct Task {
        int32_t node;
        uint32_t begin, end;
        int depth;
    };
    uint32_t* order = out.tri_order.data();
    // One node: its bounds, then a leaf or a split of its range (in place in `order`)
    // into two child nodes appended to `nodes`, whose tasks go to `children`.
    auto step = [&](const Task& t, std::vector<BuildNode>& nodes, std::vector<Task>& children,
                    uint32_t& max_depth_seen) {
        BuildNode& nd = nodes[(size_t)t.node];
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
        max_depth_seen = std::max<uint32_t>(max_depth_seen, (uint32_t)t.depth);
        if (count <= (uint32_t)max_leaf && count > 0 && (count <= 2 || t.depth >= max_depth)) {
            nd.first = t.begin;
            nd.count = count;
            return;
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
            // binned SAH over all three axes; the best (axis, bin boundary) wins
            float best = std::numeric_limits<float>::infinity();
            int best_split = -1, best_axis = axis;
            Box bins[3][kBins];
            uint32_t counts[3][kBins];
            float weights[3][kBins];
            auto bin_axis = [&](int ax) {
                Box* bin_box = bins[ax];
                uint32_t* bin_cnt = counts[ax];
                float* bin_w = weights[ax];
                for (int b = 0; b < kBins; ++b) {
                    bin_box[b].reset();
                    bin_cnt[b] = 0;
                    bin_w[b] = 0.0f;
                }
                const float scale = (float)kBins / ext[ax], base = cbox.lo[ax];
                for (uint32_t i = t.begin; i < t.end; ++i) {
                    uint32_t tri = order[i];
                    int b = std::min(std::max((int)((centroid[3 * (size_t)tri + ax] - base) * scale), 0), kBins - 1);
                    Box tb;
                    std::memcpy(tb.lo, tri_lo + 3 * (size_t)tri, sizeof(tb.lo));
                    std::memcpy(tb.hi, tri_hi + 3 * (size_t)tri, sizeof(tb.hi));
                    bin_box[b].grow(tb);
                    bin_cnt[b]++;
                    bin_w[b] += weight ? weight[tri] : 1.0f;
                }
            };
            for (int ax = 0; ax < 3; ++ax)
                if (ext[ax] > 0.0f) bin_axis(ax);
            for (int ax = 0; ax < 3; ++ax) {
                if (!(ext[ax] > 0.0f)) continue;
                const Box* bin_box = bins[ax];
                const uint32_t* bin_cnt = counts[ax];
                const float* bin_w = weights[ax];
                float right_area[kBins], right_w[kBins];
                uint32_t right_cnt[kBins];
                Box acc;
                acc.reset();
                uint32_t c = 0;
                float wsum = 0.0f;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bin_box[b]);
                    c += bin_cnt[b];
                    wsum += bin_w[b];
                    right_area[b] = acc.area();
                    right_cnt[b] = c;
                    right_w[b] = wsum;
                }
                acc.reset();
                c = 0;
                wsum = 0.0f;
                for (int b = 0; b < kBins - 1; ++b) {
                    acc.grow(bin_box[b]);
                    c += bin_cnt[b];
                    wsum += bin_w[b];
                    if (c == 0 || right_cnt[b + 1] == 0) continue;
                    float cost = acc.area() * wsum + right_area[b + 1] * right_w[b + 1];
                    if (cost < best) {
                        best = cost;
                        best_split = b;
                        best_axis = ax;
                    }
                }
            }
            float parent_area = box.area();
            float split_cost = parent_area > 0.0f ? kTraversalCost + best / parent_area : std::numeric_limits<float>::infinity();
            float leaf_cost = (float)count;
            if (best_split < 0) {
                if (count <= (uint32_t)max_leaf) make_leaf = true;
            } else if (count <= (uint32_t)max_leaf && leaf_cost <= split_cost) {
                make_leaf = true;
            } else {
                const float scale = (float)kBins / ext[best_axis], base = cbox.lo[best_axis];
                uint32_t* p = std::partition(order + t.begin, order + t.end, [&](uint32_t tri) {
                    int b = std::min(std::max((int)((centroid[3 * (size_t)tri + best_axis] - base) * scale), 0),
                                     kBins - 1);
                    return b <= best_split;
                });
                mid = (uint32_t)(p - order);
                if (mid == t.begin || mid == t.end) mid = t.begin + count / 2;
            }
        }
        if (make_leaf) {
            nd.first = t.begin;
            nd.count = count;
            return;
        }
        int32_t l = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        int32_t r = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        nodes[t.node].left = l;
        nodes[t.node].right = r;
        // right first: a LIFO consumer builds (and lays out) the left subtree first
        children.push_back({r, mid, t.end, t.depth + 1});
        children.push_back({l, t.begin, mid, t.depth + 1});
    };
    // Serial from the root until the pending ranges are small, then those subtrees in
    // parallel, each into its own node array, merged at the end.  Splits depend on the
    // data only, so the tree is the same for any thread count.
    std::vector<BuildNode> nodes;
    nodes.reserve(n > 0 ? 2 * (size_t)n : 1);
    nodes.push_back(BuildNode{});
    uint32_t max_seen_depth = 1;
    const uint32_t small = std::max<uint32_t>(4096, n / 64);
    std::vector<Task> pending{{0, 0, n, 1}}, subtrees, children;
    while (!pending.empty()) {
        const Task t = pending.back();
        pending.pop_back();
        if (t.end - t.begin <= small) {
            subtrees.push_back(t);
            continue;
        }
        children.clear();
        step(t, nodes, children, max_seen_depth);
        pending.insert(pending.end(), children.begin(), children.end());
    }
    struct Part {
        std::vector<BuildNode> nodes;
        uint32_t depth = 1;
    };
    std::vector<Part> parts(subtrees.size());
    auto build_subtree = [&](size_t k) {
        Part& part = parts[k];
        part.nodes.push_back(BuildNode{});
        std::vector<Task> stack{{0, subtrees[k].begin, subtrees[k].end, subtrees[k].depth}}, kids;
        while (!stack.empty()) {
            const Task t = stack.back();
            stack.pop_back();
            kids.clear();
            step(t, part.nodes, kids, part.depth);
            stack.insert(stack.end(), kids.begin(), kids.end());
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nthreads = (unsigned)std::min<size_t>(hw, subtrees.size());
    if (nthreads <= 1) {
        for (size_t k = 0; k < subtrees.size(); ++k) build_subtree(k);
    } else {
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < nthreads; ++i)
            pool.emplace_back([&] {
                for (size_t k; (k = next.fetch_add(1)) < subtrees.size();) build_subtree(k);
            });
        for (std::thread& th : pool) th.join();
    }
    // merge: a part's node 0 is the subtree root already allocated in `nodes`
    for (size_t k = 0; k < parts.size(); ++k) {
        const int32_t root = subtrees[k].node;
        const int32_t base = (int32_t)nodes.size() - 1;
        auto map = [&](int32_t c) { return c <= 0 ? (c == 0 ? root : c) : c + base; };
        for (size_t i = 0; i < parts[k].nodes.size(); ++i) {
            BuildNode b = parts[k].nodes[i];
            if (b.count == 0) {
                b.left = map(b.left);
                b.right = map(b.right);
            }
            if (i == 0)
                nodes[(size_t)root] = b;
            else
                nodes.push_back(b);
        }
        max_seen_depth = std::max(max_seen_depth, parts[k].depth);
    }
    out.depth = max_seen_depth;
    emit_layout(nodes, n > 0, out);
    return out;
}

// ---------------------------------------------------------------------------
// Spatial-split BVH (Stich, Friedrich and Dietrich, "Spatial Splits in Bounding
// Volume Hierarchies", HPG 2009).  Each node weighs the binned object split above
// against a spatial split: 32 equal bins per axis, every reference chopped at the
// bin planes (its triangle clipped to each slab), then the best plane by SAH.
// A reference straddling the chosen plane goes to both children with its box
// clipped to each side.  Spatial splits are only tried where the object split's
// children overlap by more than kSpatialAlpha of the root's surface, and stop
// once the references exceed kSpatialBudget times the triangles.
// ---------------------------------------------------------------------------
namespace {

// alpha 1e-4 (1e-3 before): on the atrium spatial splits then reach 12 % more nodes and
// the frame's trace 1 % less time; 1e-5 with a budget of 2 gains nothing more
constexpr float kSpatialAlpha = 1e-4f;
constexpr float kSpatialBudget = 1.5f;

struct Ref {
    uint32_t tri;
    Box box;
};

Box overlap(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = std::max(a.lo[k], b.lo[k]);
        r.hi[k] = std::min(a.hi[k], b.hi[k]);
    }
    return r;
}

bool nonempty_box(const Box& b) { return b.lo[0] <= b.hi[0] && b.lo[1] <= b.hi[1] && b.lo[2] <= b.hi[2]; }

// boxes of the parts of reference r's triangle below and above the plane
// x[axis] = pos, each clipped to the reference's box
void split_reference(const float* tris, const Ref& r, int axis, float pos, Box& below, Box& above) {
    below.reset();
    above.reset();
    const float* v = tris + 9 * (size_t)r.tri;
    for (int i = 0; i < 3; ++i) {
        const float* a = v + 3 * i;
        const float* b = v + 3 * ((i + 1) % 3);
        if (a[axis] <= pos) below.grow(a);
        if (a[axis] >= pos) above.grow(a);
        if ((a[axis] < pos && b[axis] > pos) || (a[axis] > pos && b[axis] < pos)) {
            const float t = (pos - a[axis]) / (b[axis] - a[axis]);
            float p[3];
            for (int k = 0; k < 3; ++k) p[k] = a[k] + (b[k] - a[k]) * t;
            p[axis] = pos;
            below.grow(p);
            above.grow(p);
        }
    }
    below = overlap(below, r.box);
    above = overlap(above, r.box);
    below.hi[axis] = std::min(below.hi[axis], pos);
    above.lo[axis] = std::max(above.lo[axis], pos);
}

float centre(const Box& b, int axis) { return 0.5f * (b.lo[axis] + b.hi[axis]); }

struct ObjectSplit {
    float cost = std::numeric_limits<float>::infinity();
    int axis = -1, bin = -1;
    Box left, right;
};

ObjectSplit best_object_split(const std::vector<Ref>& refs, const Box& cbox) {
    ObjectSplit best;
    for (int ax = 0; ax < 3; ++ax) {
        const float ext = cbox.hi[ax] - cbox.lo[ax];
        if (!(ext > 0.0f)) continue;
        Box bin_box[kBins], right_box[kBins];
        uint32_t bin_cnt[kBins] = {}, right_cnt[kBins];
        for (Box& b : bin_box) b.reset();
        const float scale = (float)kBins / ext, base = cbox.lo[ax];
        for (const Ref& r : refs) {
            const int b = std::min(std::max((int)((centre(r.box, ax) - base) * scale), 0), kBins - 1);
            bin_box[b].grow(r.box);
            bin_cnt[b]++;
        }
        Box acc;
        acc.reset();
        uint32_t c = 0;
        for (int b = kBins - 1; b > 0; --b) {
            acc.grow(bin_box[b]);
            c += bin_cnt[b];
            right_box[b] = acc;
            right_cnt[b] = c;
        }
        acc.reset();
        c = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            acc.grow(bin_box[b]);
            c += bin_cnt[b];
            if (c == 0 || right_cnt[b + 1] == 0) continue;
            const float cost = acc.area() * (float)c + right_box[b + 1].area() * (float)right_cnt[b + 1];
            if (cost < best.cost) {
                best.cost = cost;
                best.axis = ax;
                best.bin = b;
                best.left = acc;
                best.right = right_box[b + 1];
            }
        }
    }
    return best;
}

struct SpatialSplit {
    float cost = std::numeric_limits<float>::infinity();
    int axis = -1;
    float pos = 0.0f;
};

SpatialSplit best_spatial_split(const float* tris, const std::vector<Ref>& refs, const Box& box) {
    SpatialSplit best;
    for (int ax = 0; ax < 3; ++ax) {
        const float ext = box.hi[ax] - box.lo[ax];
        if (!(ext > 0.0f)) continue;
        Box bin_box[kBins], right_box[kBins];
        uint32_t enter[kBins] = {}, leave[kBins] = {}, right_cnt[kBins];
        for (Box& b : bin_box) b.reset();
        const float width = ext / (float)kBins, base = box.lo[ax];
        auto bin_of = [&](float x) { return std::min(std::max((int)((x - base) / width), 0), kBins - 1); };
        for (const Ref& r : refs) {
            const int b0 = bin_of(r.box.lo[ax]), b1 = bin_of(r.box.hi[ax]);
            enter[b0]++;
            leave[b1]++;
            Ref rest = r;
            for (int b = b0; b < b1; ++b) {
                Box below, above;
                split_reference(tris, rest, ax, base + (float)(b + 1) * width, below, above);
                bin_box[b].grow(below);
                rest.box = above;
            }
            bin_box[b1].grow(rest.box);
        }
        Box acc;
        acc.reset();
        uint32_t c = 0;
        for (int b = kBins - 1; b > 0; --b) {
            acc.grow(bin_box[b]);
            c += leave[b];
            right_box[b] = acc;
            right_cnt[b] = c;
        }
        acc.reset();
        c = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            acc.grow(bin_box[b]);
            c += enter[b];
            if (c == 0 || right_cnt[b + 1] == 0) continue;
            const float cost = acc.area() * (float)c + right_box[b + 1].area() * (float)right_cnt[b + 1];
            if (cost < best.cost) {
                best.cost = cost;
                best.axis = ax;
                best.pos = base + (float)(b + 1) * width;
            }
        }
    }
    return best;
}

// halves at the median reference centre on `axis`
void median_split(std::vector<Ref>& refs, int axis, std::vector<Ref>& left, std::vector<Ref>& right) {
    const size_t mid = refs.size() / 2;
    std::nth_element(refs.begin(), refs.begin() + (std::ptrdiff_t)mid, refs.end(),
                     [&](const Ref& a, const Ref& b) { return centre(a.box, axis) < centre(b.box, axis); });
    left.assign(refs.begin(), refs.begin() + (std::ptrdiff_t)mid);
    right.assign(refs.begin() + (std::ptrdiff_t)mid, refs.end());
}

}  // namespace

BvhHost build_bvh_sbvh(const float* tris, uint32_t n, int max_leaf, int max_depth) {
    BvhHost out;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > kMaxLeaf) max_leaf = kMaxLeaf;
    std::vector<BuildNode> nodes;
    nodes.reserve(2 * (size_t)n + 1);
    std::vector<uint32_t> leaf_refs;
    leaf_refs.reserve((size_t)(kSpatialBudget * (float)n) + 1);
    struct Task {
        int32_t node;
        std::vector<Ref> refs;
        int depth;
    };
    std::vector<Task> stack;
    Box root;
    root.reset();
    {
        std::vector<Ref> refs(n);
        for (uint32_t i = 0; i < n; ++i) {
            refs[i].tri = i;
            refs[i].box.reset();
            for (int k = 0; k < 3; ++k) refs[i].box.grow(tris + 9 * (size_t)i + 3 * k);
            root.grow(refs[i].box);
        }
        nodes.push_back(BuildNode{});
        stack.push_back({0, std::move(refs), 1});
    }
    const float root_area = root.area();
    size_t total_refs = n;
    const size_t max_refs = (size_t)(kSpatialBudget * (float)n);
    uint32_t max_seen_depth = 1;
    while (!stack.empty()) {
        Task t = std::move(stack.back());
        stack.pop_back();
        std::vector<Ref>& refs = t.refs;
        const uint32_t count = (uint32_t)refs.size();
        Box box, cbox;
        box.reset();
        cbox.reset();
        for (const Ref& r : refs) {
            box.grow(r.box);
            const float c[3] = {centre(r.box, 0), centre(r.box, 1), centre(r.box, 2)};
            cbox.grow(c);
        }
        nodes[(size_t)t.node].box = box;
        max_seen_depth = std::max<uint32_t>(max_seen_depth, (uint32_t)t.depth);
        auto make_leaf = [&] {
            nodes[(size_t)t.node].first = (uint32_t)leaf_refs.size();
            nodes[(size_t)t.node].count = count;
            for (const Ref& r : refs) leaf_refs.push_back(r.tri);
        };
        if (count <= (uint32_t)max_leaf && count > 0 && (count <= 2 || t.depth >= max_depth)) {
            make_leaf();
            continue;
        }
        int widest = 0;
        for (int a = 1; a < 3; ++a)
            if (cbox.hi[a] - cbox.lo[a] > cbox.hi[widest] - cbox.lo[widest]) widest = a;
        std::vector<Ref> left, right;
        // median splits once SAH splits could overrun the depth budget
        int need = 0;
        for (uint32_t c = count; c > (uint32_t)max_leaf; c = (c + 1) / 2) ++need;
        if (t.depth + need >= max_depth || !(cbox.hi[widest] - cbox.lo[widest] > 0.0f)) {
            if (count <= (uint32_t)max_leaf) {
                make_leaf();
                continue;
            }
            median_split(refs, widest, left, right);
        } else {
            const ObjectSplit obj = best_object_split(refs, cbox);
            SpatialSplit sp;
            const Box ov = overlap(obj.left, obj.right);
            if (total_refs < max_refs && obj.axis >= 0 && nonempty_box(ov) && ov.area() > kSpatialAlpha * root_area)
                sp = best_spatial_split(tris, refs, box);
            const float best = std::min(obj.cost, sp.cost);
            const float parent_area = box.area();
            const float split_cost = parent_area > 0.0f ? kTraversalCost + best / parent_area : std::numeric_limits<float>::infinity();
            if (obj.axis < 0 && sp.axis < 0) {
                if (count <= (uint32_t)max_leaf) {
                    make_leaf();
                    continue;
                }
                median_split(refs, widest, left, right);
            } else if (count <= (uint32_t)max_leaf && (float)count <= split_cost) {
                make_leaf();
                continue;
            } else if (sp.cost < obj.cost) {
                for (const Ref& r : refs) {
                    if (r.box.hi[sp.axis] <= sp.pos) {
                        left.push_back(r);
                    } else if (r.box.lo[sp.axis] >= sp.pos) {
                        right.push_back(r);
                    } else {
                        Box below, above;
                        split_reference(tris, r, sp.axis, sp.pos, below, above);
                        const bool b = nonempty_box(below), a = nonempty_box(above);
                        if (b) left.push_back({r.tri, below});
                        if (a) right.push_back({r.tri, above});
                        if (a && b) ++total_refs;
                    }
                }
                if (left.empty() || right.empty()) {
                    left.clear();
                    right.clear();
                    median_split(refs, sp.axis, left, right);
                }
            } else {
                const float scale = (float)kBins / (cbox.hi[obj.axis] - cbox.lo[obj.axis]), base = cbox.lo[obj.axis];
                for (const Ref& r : refs) {
                    const int b = std::min(std::max((int)((centre(r.box, obj.axis) - base) * scale), 0), kBins - 1);
                    (b <= obj.bin ? left : right).push_back(r);
                }
                if (left.empty() || right.empty()) {
                    left.clear();
                    right.clear();
                    median_split(refs, obj.axis, left, right);
                }
            }
        }
        const int32_t l = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        const int32_t r = (int32_t)nodes.size();
        nodes.push_back(BuildNode{});
        nodes[(size_t)t.node].left = l;
        nodes[(size_t)t.node].right = r;
        std::vector<Ref>().swap(refs);
        // push right first so the left subtree is processed (and laid out) first
        stack.push_back({r, std::move(right), t.depth + 1});
        stack.push_back({l, std::move(left), t.depth + 1});
    }
    out.depth = max_seen_depth;
    out.tri_order = std::move(leaf_refs);
    emit_layout(nodes, n > 0, out);
    return out;
}

}  // namespace pbrt
