// CPU binned-SAH BVH constructor emitting the reference's flattened buffers.
//
// Restates Source/Core/BVH/BVHConstructor.cpp so that, for the same mesh input,
// the stackless node array (FlattenBVH, :783-845), the stack node array
// (FlattenStackBVH, :847-930) and the triangle array (GenerateTriangles,
// :630-650) are byte-identical to the reference's.
//
// Decisions that fix the bytes, and where they come from:
//   * per-triangle bounds start at +-1e7 and fold the three positions with
//     glm::min/max (x < y ? x : y, x > y ? x : y), the fold value as x
//     (:408-427, BVHConstructor.h:20-27); centroid = (min + max) / 2 (h:33-35)
//   * binned SAH (:276-363): 64 bins per axis over the node bounds, axes with
//     zero extent skipped, bin = min(63, int((c - min) * (64 / extent))),
//     prefix/suffix half-area x count, first strict minimum wins (axis order
//     x, y, z; bins ascending), border = min + (extent / 64) * (i + 1)
//   * Lomuto partition `centroid[axis] < border` (:532-549), degenerate split ->
//     midpoint (:553-556), child bounds rescanned with the element as x
//     (:573-597)
//   * leaves have <= 2 triangles (:49, :456) and receive their start index in
//     the order a LIFO stack pops them (left pushed first, so the right subtree
//     is emitted first, :443-472, :624-625)
//   * stackless builds swap the two children of every split with p = 1/2
//     (:599-609); only the stored child order changes, never the processing
//     order.  CRT_SWAP_SEEDED draws from std::mt19937(seed) through
//     std::uniform_int_distribution<unsigned long>(1, 4) exactly as the
//     reference draws from mt19937(random_device())
//
// Unlike the reference the tree is built with subtree parallelism: the split
// of a node only reads and permutes its own triangle range, so siblings are
// independent.  The order-dependent parts (leaf start indices, the random
// draws, MaxBVHDepth) are produced afterwards by replaying the reference's
// LIFO processing order over the finished tree.  Chunked parallel folds use
// the same comparison operator as the sequential fold, which keeps the
// "which of two equal values (+0 / -0) wins" choice identical.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "candela_rt.h"
#include "crt_internal.h"

namespace crt {
namespace {

constexpr float kArbitraryMax = 10000000.0f;   // BVHConstructor.h:20
constexpr float kArbitraryMin = -10000000.0f;  // BVHConstructor.h:21
constexpr float kInfCost = 1e29f;              // BVHConstructor.cpp:56

// glm::min / glm::max (func_common.inl:15-28)
inline float gmin(float x, float y) { return x < y ? x : y; }
inline float gmax(float x, float y) { return x > y ? x : y; }

struct Box {
    float mn[3];
    float mx[3];
};

inline Box empty_box() {
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = kArbitraryMax;
        b.mx[k] = kArbitraryMin;
    }
    return b;
}

// Bounds::GetArea (BVHConstructor.h:41-44): half surface area.
inline float box_area(const Box& b) {
    const float ex = b.mx[0] - b.mn[0];
    const float ey = b.mx[1] - b.mn[1];
    const float ez = b.mx[2] - b.mn[2];
    return ex * ey + ey * ez + ez * ex;
}

// One reference to a triangle, carried through the partition together with
// its cached bounds and centroid (the reference keeps these in side arrays
// indexed by the reference; the values are the same).
struct Item {
    float bmn[3];
    float bmx[3];
    float cen[3];
    int32_t tri;
};

struct TreeNode {
    Box box;
    uint32_t start;   // range in the item array (inner + leaf)
    uint32_t length;
    int32_t left = -1;   // child as split (processing order)
    int32_t right = -1;
    bool leaf = false;
};

// --- small parallel helpers --------------------------------------------------

template <typename Fn>
void parallel_chunks(size_t n, int threads, Fn fn) {
    if (threads <= 1 || n < 2) {
        fn(0, size_t(0), n);
        return;
    }
    const size_t t = std::min<size_t>(size_t(threads), n);
    std::vector<std::thread> pool;
    pool.reserve(t - 1);
    const size_t per = (n + t - 1) / t;
    for (size_t c = 1; c < t; ++c) {
        const size_t b = c * per, e = std::min(n, b + per);
        if (b >= e) break;
        pool.emplace_back([=, &fn] { fn(c, b, e); });
    }
    fn(0, size_t(0), std::min(n, per));
    for (auto& th : pool) th.join();
}

struct BinSet {
    int count[3][kBinCount];
    Box box[3][kBinCount];
    void reset() {
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < kBinCount; ++i) {
                count[a][i] = 0;
                box[a][i] = empty_box();
            }
    }
};

class Builder {
public:
    Builder(std::vector<Item>& items, int threads) : items_(items), threads_(std::max(1, threads)) {}

    std::vector<TreeNode> nodes;
    uint64_t split_fails = 0;

    void build(const Box& root_box, uint32_t tri_count) {
        // A binary tree with leaves of >= 1 triangle has < 2N nodes.
        nodes.resize(std::max<size_t>(1, 2 * size_t(tri_count)));
        TreeNode& root = nodes[0];
        root.box = root_box;
        root.start = 0;
        root.length = tri_count;
        next_.store(1);
        active_.store(1);
        build_subtree(0);
        nodes.resize(next_.load());
        split_fails = fails_.load();
    }

private:
    std::vector<Item>& items_;
    int threads_;
    std::atomic<uint32_t> next_{1};
    std::atomic<uint64_t> fails_{0};
    std::atomic<int> active_{1};

    static constexpr uint32_t kParallelFoldMin = 1u << 18;   // elements
    static constexpr uint32_t kSpawnMin = 1u << 12;          // elements

    // SearchSAHPlaneBinned (BVHConstructor.cpp:276-363).  Returns false when no
    // axis has extent (the reference leaves axis/border unset there).
    bool find_split(const TreeNode& node, int& o_axis, float& o_border) {
        const uint32_t b = node.start, e = node.start + node.length;
        float scale[3], mn_axis[3];
        bool axis_ok[3];
        for (int a = 0; a < 3; ++a) {
            mn_axis[a] = node.box.mn[a];
            const float mx = node.box.mx[a];
            axis_ok[a] = !(mn_axis[a] == mx);
            const float extent = mx - mn_axis[a];
            scale[a] = float(kBinCount) / extent;
        }
        if (!axis_ok[0] && !axis_ok[1] && !axis_ok[2]) return false;

        auto fill = [&](BinSet& bins, size_t lo, size_t hi) {
            bins.reset();
            for (size_t i = lo; i < hi; ++i) {
                const Item& it = items_[i];
                for (int a = 0; a < 3; ++a) {
                    if (!axis_ok[a]) continue;
                    const float rel = (it.cen[a] - mn_axis[a]) * scale[a];
                    int bi = bin_index(rel);
                    bins.count[a][bi] += 1;
                    Box& bb = bins.box[a][bi];
                    for (int k = 0; k < 3; ++k) {
                        bb.mn[k] = gmin(bb.mn[k], it.bmn[k]);
                        bb.mx[k] = gmax(bb.mx[k], it.bmx[k]);
                    }
                }
            }
        };

        BinSet bins;
        const uint32_t n = node.length;
        const int par = (n >= kParallelFoldMin) ? threads_ : 1;
        if (par > 1) {
            const size_t t = std::min<size_t>(size_t(par), n);
            std::vector<BinSet> part(t);
            parallel_chunks(n, int(t), [&](size_t c, size_t lo, size_t hi) { fill(part[c], b + lo, b + hi); });
            // Merge chunk folds left to right with the fold's own operator
            // (fold value as x), which reproduces the sequential fold exactly.
            bins = part[0];
            for (size_t c = 1; c < t; ++c)
                for (int a = 0; a < 3; ++a)
                    for (int i = 0; i < kBinCount; ++i) {
                        bins.count[a][i] += part[c].count[a][i];
                        Box& bb = bins.box[a][i];
                        const Box& pb = part[c].box[a][i];
                        for (int k = 0; k < 3; ++k) {
                            bb.mn[k] = gmin(bb.mn[k], pb.mn[k]);
                            bb.mx[k] = gmax(bb.mx[k], pb.mx[k]);
                        }
                    }
        } else {
            fill(bins, b, e);
        }

        float best = kInfCost;
        bool found = false;
        for (int a = 0; a < 3; ++a) {
            if (!axis_ok[a]) continue;
            float left_area[kBinCount - 1], right_area[kBinCount - 1];
            int left_count[kBinCount - 1], right_count[kBinCount - 1];
            Box lb = empty_box(), rb = empty_box();
            int lsum = 0, rsum = 0;
            for (int i = 0; i < kBinCount - 1; ++i) {
                lsum += bins.count[a][i];
                left_count[i] = lsum;
                for (int k = 0; k < 3; ++k) {
                    lb.mn[k] = gmin(lb.mn[k], bins.box[a][i].mn[k]);
                    lb.mx[k] = gmax(lb.mx[k], bins.box[a][i].mx[k]);
                }
                left_area[i] = box_area(lb);
                const int r = kBinCount - 1 - i;
                rsum += bins.count[a][r];
                right_count[r - 1] = rsum;
                for (int k = 0; k < 3; ++k) {
                    rb.mn[k] = gmin(rb.mn[k], bins.box[a][r].mn[k]);
                    rb.mx[k] = gmax(rb.mx[k], bins.box[a][r].mx[k]);
                }
                right_area[r - 1] = box_area(rb);
            }
            const float extent = node.box.mx[a] - node.box.mn[a];
            const float step = extent / float(kBinCount);
            for (int i = 0; i < kBinCount - 1; ++i) {
                const float lc = float(left_count[i]) * left_area[i];
                const float rc = float(right_count[i]) * right_area[i];
                const float cost = lc + rc;
                if (cost < best) {
                    best = cost;
                    o_axis = a;
                    o_border = mn_axis[a] + step * float(i + 1);
                    found = true;
                }
            }
        }
        return found;
    }

    // (int)x for the bin index, clamped to [0, 63].  The reference computes
    // min(63, int(x)); for finite x in [0, 64] the result is identical.  NaN or
    // out-of-range x (zero-width float extents) would index outside the bin
    // array in the reference; here it lands in bin 0 / 63.
    static inline int bin_index(float rel) {
        if (!(rel >= 0.0f)) return 0;
        if (rel >= float(kBinCount)) return kBinCount - 1;
        const int bi = int(rel);
        return bi < kBinCount - 1 ? bi : kBinCount - 1;
    }

    Box range_bounds(uint32_t lo, uint32_t hi) {
        // Child bounds rescan (BVHConstructor.cpp:573-597): element as x, so an
        // equal value keeps the earlier one.
        auto fold = [&](Box& acc, size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) {
                const Item& it = items_[i];
                for (int k = 0; k < 3; ++k) {
                    acc.mn[k] = gmin(it.bmn[k], acc.mn[k]);
                    acc.mx[k] = gmax(it.bmx[k], acc.mx[k]);
                }
            }
        };
        const uint32_t n = hi - lo;
        if (n >= kParallelFoldMin && threads_ > 1) {
            const size_t t = std::min<size_t>(size_t(threads_), n);
            std::vector<Box> part(t, empty_box());
            parallel_chunks(n, int(t), [&](size_t c, size_t a, size_t b) { fold(part[c], lo + a, lo + b); });
            Box acc = part[0];
            for (size_t c = 1; c < t; ++c)
                for (int k = 0; k < 3; ++k) {
                    acc.mn[k] = gmin(part[c].mn[k], acc.mn[k]);
                    acc.mx[k] = gmax(part[c].mx[k], acc.mx[k]);
                }
            return acc;
        }
        Box acc = empty_box();
        fold(acc, lo, hi);
        return acc;
    }

    void build_subtree(uint32_t root_index) {
        std::vector<uint32_t> stack;
        stack.push_back(root_index);
        std::vector<std::thread> spawned;
        while (!stack.empty()) {
            const uint32_t ni = stack.back();
            stack.pop_back();
            TreeNode node = nodes[ni];
            if (node.length <= kMaxLeafTris) {   // ShouldBeLeaf || Length <= 1
                nodes[ni].leaf = true;
                continue;
            }
            int axis = 0;
            float border = 0.0f;
            uint32_t split;
            const uint32_t b = node.start, e = node.start + node.length;
            if (find_split(node, axis, border)) {
                uint32_t mid = b;
                for (uint32_t i = b; i < e; ++i) {
                    if (items_[i].cen[axis] < border) {
                        std::swap(items_[i], items_[mid]);
                        ++mid;
                    }
                }
                split = mid;
            } else {
                split = b;   // no usable plane: the degenerate-split path below
            }
            if (split == b || split == e) {
                split = b + node.length / 2;
                fails_.fetch_add(1, std::memory_order_relaxed);
            }
            const uint32_t li = next_.fetch_add(2);
            const uint32_t ri = li + 1;
            TreeNode& L = nodes[li];
            TreeNode& R = nodes[ri];
            L.start = b;
            L.length = split - b;
            R.start = split;
            R.length = e - split;
            L.box = range_bounds(L.start, L.start + L.length);
            R.box = range_bounds(R.start, R.start + R.length);
            nodes[ni].left = int32_t(li);
            nodes[ni].right = int32_t(ri);
            // Independent subtrees: hand the larger one to a new worker while
            // workers are available.
            const uint32_t big = L.length >= R.length ? li : ri;
            const uint32_t small = big == li ? ri : li;
            if (nodes[big].length >= kSpawnMin && try_acquire_worker()) {
                spawned.emplace_back([this, big] {
                    build_subtree(big);
                    active_.fetch_sub(1);
                });
                stack.push_back(small);
            } else {
                stack.push_back(li);
                stack.push_back(ri);
            }
        }
        for (auto& th : spawned) th.join();
    }

    bool try_acquire_worker() {
        int cur = active_.load();
        while (cur < threads_) {
            if (active_.compare_exchange_weak(cur, cur + 1)) return true;
        }
        return false;
    }
};

// Writes vec4(bounds, 0) with a packed int in one lane (glm::intBitsToFloat).
inline float int_bits(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

inline int32_t pack_leaf(uint32_t start, uint32_t length) {
    // (StartIndex << 4) | (Length & 0xF) in uint arithmetic, stored as int.
    return int32_t((start << 4) | (length & 0xFu));
}

}  // namespace

// ---------------------------------------------------------------------------

struct BuiltTree {
    std::vector<TreeNode> nodes;
    std::vector<int32_t> sorted_refs;    // triangle (input order) per output slot
    std::vector<uint32_t> leaf_start;    // per node: global start for leaves
    uint64_t leaf_count = 0;
    uint64_t split_fails = 0;
    uint32_t max_depth = 0;
};

// Runs ConstructTree (BVHConstructor.cpp:385-627) over `indices`.
static void construct_tree(const std::vector<crt_vertex>& verts, const std::vector<uint32_t>& indices,
                           int32_t t_offset, int format, const crt_build_options& opt, int threads,
                           BuiltTree& out) {
    const uint32_t ntri = uint32_t(indices.size() / 3);
    std::vector<Item> items(ntri);
    // Bounds/centroid cache and root bounds (:404-430).
    parallel_chunks(ntri, ntri >= (1u << 16) ? threads : 1, [&](size_t, size_t lo, size_t hi) {
        for (size_t t = lo; t < hi; ++t) {
            Item& it = items[t];
            float mn[3] = {kArbitraryMax, kArbitraryMax, kArbitraryMax};
            float mx[3] = {kArbitraryMin, kArbitraryMin, kArbitraryMin};
            for (int c = 0; c < 3; ++c) {
                const float* p = verts[indices[3 * t + c]].position;
                for (int k = 0; k < 3; ++k) {
                    mn[k] = gmin(mn[k], p[k]);
                    mx[k] = gmax(mx[k], p[k]);
                }
            }
            for (int k = 0; k < 3; ++k) {
                it.bmn[k] = mn[k];
                it.bmx[k] = mx[k];
                it.cen[k] = (mn[k] + mx[k]) / 2.0f;
            }
            it.tri = int32_t(t);
        }
    });
    Box root = empty_box();
    for (uint32_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            root.mn[k] = gmin(root.mn[k], items[t].bmn[k]);
            root.mx[k] = gmax(root.mx[k], items[t].bmx[k]);
        }

    Builder builder(items, threads);
    builder.build(root, ntri);
    out.nodes = std::move(builder.nodes);
    out.split_fails = builder.split_fails;

    // Replay the reference's LIFO processing order: leaves get consecutive
    // start indices as they are popped; every split draws one swap decision.
    std::vector<TreeNode>& nodes = out.nodes;
    out.sorted_refs.clear();
    out.sorted_refs.reserve(ntri);
    out.leaf_start.assign(nodes.size(), 0);
    const bool swap = (format == CRT_FORMAT_STACKLESS) && opt.child_swap == CRT_SWAP_SEEDED;
    std::mt19937 rng(opt.seed);
    std::uniform_int_distribution<std::mt19937::result_type> dist(1, 4);
    std::vector<uint32_t> stack;
    stack.push_back(0);
    uint32_t max_depth = 0;
    uint64_t leaves = 0;
    while (!stack.empty()) {
        max_depth = std::max<uint32_t>(max_depth, uint32_t(stack.size()));
        const uint32_t ni = stack.back();
        stack.pop_back();
        TreeNode& n = nodes[ni];
        if (n.leaf) {
            ++leaves;
            const uint32_t fin = uint32_t(out.sorted_refs.size());
            for (uint32_t i = n.start; i < n.start + n.length; ++i) out.sorted_refs.push_back(items[i].tri);
            out.leaf_start[ni] = fin + uint32_t(t_offset);
            continue;
        }
        const int32_t l = n.left, r = n.right;
        if (swap) {
            const uint32_t h = uint32_t(dist(rng));
            if (h <= 2) std::swap(n.left, n.right);
        }
        stack.push_back(uint32_t(l));
        stack.push_back(uint32_t(r));
    }
    out.leaf_count = leaves;
    out.max_depth = max_depth;
}

// FlattenBVH (BVHConstructor.cpp:783-845): pre-order, miss links in max.w.
static void flatten_stackless(const BuiltTree& tree, std::vector<crt_stackless_node>& out) {
    const auto& nodes = tree.nodes;
    out.assign(nodes.size(), crt_stackless_node{});
    std::vector<int32_t> cache_x(nodes.size());
    std::vector<int32_t> cache_y(nodes.size(), 0);
    // Iterative pre-order; the flattened index of a node is its visit number.
    struct Frame {
        uint32_t node;
        int32_t flat;
        int stage;
    };
    std::vector<Frame> st;
    int32_t processed = 0;
    st.push_back({0u, -1, 0});
    while (!st.empty()) {
        Frame& f = st.back();
        const TreeNode& n = nodes[f.node];
        if (f.stage == 0) {
            f.flat = processed++;
            crt_stackless_node& fn = out[size_t(f.flat)];
            for (int k = 0; k < 3; ++k) {
                fn.min[k] = n.box.mn[k];
                fn.max[k] = n.box.mx[k];
            }
            fn.min[3] = 0.0f;
            fn.max[3] = 0.0f;
            if (n.leaf) {
                cache_y[size_t(f.flat)] = pack_leaf(tree.leaf_start[f.node], n.length);
                cache_x[size_t(f.flat)] = -1;
                st.pop_back();
                continue;
            }
            f.stage = 1;
            st.push_back({uint32_t(n.left), -1, 0});
            continue;
        }
        if (f.stage == 1) {
            // Right child's flat index is the next visit number.
            cache_x[size_t(f.flat)] = processed;
            f.stage = 2;
            st.push_back({uint32_t(n.right), -1, 0});
            continue;
        }
        st.pop_back();
    }
    const float neg1 = int_bits(-1);
    out[0].max[3] = neg1;
    for (size_t i = 0; i < out.size(); ++i) {
        if (cache_x[i] != -1) {
            out[i + 1].max[3] = int_bits(cache_x[i]);
            out[size_t(cache_x[i])].max[3] = out[i].max[3];
        }
    }
    for (size_t i = 0; i < out.size(); ++i) out[i].min[3] = cache_x[i] == -1 ? int_bits(cache_y[i]) : neg1;
}

// FlattenStackBVH (BVHConstructor.cpp:847-930): BFS over inner nodes.
static void flatten_stack(const BuiltTree& tree, std::vector<crt_stack_node>& out) {
    const auto& nodes = tree.nodes;
    out.assign(nodes.size(), crt_stack_node{});
    const float neg1 = int_bits(-1);
    std::vector<std::pair<uint32_t, int32_t>> queue;
    queue.reserve(nodes.size() / 2 + 1);
    queue.push_back({0u, 0});
    size_t head = 0;
    int32_t counter = 0;
    auto write_child = [&](crt_bounds& dst, uint32_t child, int32_t work) {
        const TreeNode& c = nodes[child];
        for (int k = 0; k < 3; ++k) {
            dst.min[k] = c.box.mn[k];
            dst.max[k] = c.box.mx[k];
        }
        dst.min[3] = 0.0f;
        dst.max[3] = 0.0f;
        if (!c.leaf) {
            dst.min[3] = neg1;
            queue.push_back({child, work});
        } else {
            dst.min[3] = int_bits(pack_leaf(tree.leaf_start[child], c.length));
        }
    };
    while (head < queue.size()) {
        const auto cur = queue[head++];
        crt_stack_node& e = out[size_t(counter++)];
        const TreeNode& n = nodes[cur.first];
        write_child(e.left, uint32_t(n.left), counter);
        write_child(out[size_t(counter - 1)].right, uint32_t(n.right), -counter);
        if (cur.second > 0)
            out[size_t(cur.second - 1)].left.max[3] = int_bits(counter - 1);
        else if (cur.second < 0)
            out[size_t(-cur.second - 1)].right.max[3] = int_bits(counter - 1);
    }
}

int build_bvh(int format, const crt_mesh* meshes, int mesh_count, int32_t t_offset,
              const crt_build_options* options, BvhResult& res) {
    if (format != CRT_FORMAT_STACKLESS && format != CRT_FORMAT_STACK)
        return set_error(CRT_E_INVALID, "crt_build_bvh: unknown format");
    if (mesh_count < 0 || (mesh_count > 0 && !meshes))
        return set_error(CRT_E_INVALID, "crt_build_bvh: bad mesh list");
    crt_build_options opt{CRT_SWAP_NONE, 0u, 0};
    if (options) opt = *options;
    if (opt.child_swap != CRT_SWAP_NONE && opt.child_swap != CRT_SWAP_SEEDED)
        return set_error(CRT_E_INVALID, "crt_build_bvh: unknown child_swap");
    int threads = opt.threads > 0 ? opt.threads : int(std::thread::hardware_concurrency());
    threads = std::max(1, std::min(threads, 256));

    // Mesh concatenation (BVHConstructor.cpp:981-1002).
    std::vector<uint32_t> indices;
    std::vector<int32_t> mesh_refs;
    res.vertices.clear();
    uint32_t index_offset = 0;
    size_t total_idx = 0, total_vtx = 0;
    for (int m = 0; m < mesh_count; ++m) {
        if (meshes[m].index_count % 3 != 0)
            return set_error(CRT_E_INVALID, "crt_build_bvh: index_count not a multiple of 3");
        if ((meshes[m].index_count && !meshes[m].indices) || (meshes[m].vertex_count && !meshes[m].vertices))
            return set_error(CRT_E_INVALID, "crt_build_bvh: null mesh data");
        total_idx += meshes[m].index_count;
        total_vtx += meshes[m].vertex_count;
    }
    if (total_idx / 3 >= (1ull << 27))
        return set_error(CRT_E_INVALID, "crt_build_bvh: too many triangles for the 28-bit leaf start");
    indices.reserve(total_idx);
    mesh_refs.reserve(total_idx / 3);
    res.vertices.reserve(total_vtx);
    for (int m = 0; m < mesh_count; ++m) {
        const crt_mesh& mesh = meshes[m];
        for (uint32_t x = 0; x < mesh.index_count; ++x) {
            const uint32_t idx = mesh.indices[x];
            if (idx >= mesh.vertex_count)
                return set_error(CRT_E_INVALID, "crt_build_bvh: vertex index out of range");
            indices.push_back(idx + index_offset);
            if (x % 3 == 0) mesh_refs.push_back(mesh.global_mesh_number);
        }
        res.vertices.insert(res.vertices.end(), mesh.vertices, mesh.vertices + mesh.vertex_count);
        index_offset += mesh.vertex_count;
    }
    const uint32_t ntri = uint32_t(indices.size() / 3);
    if (format == CRT_FORMAT_STACK && ntri <= kMaxLeafTris)
        return set_error(CRT_E_INVALID, "crt_build_bvh: the stack format needs an inner root (>= 3 triangles)");
    if (uint64_t(ntri) + uint64_t(uint32_t(t_offset)) >= (1ull << 28) || t_offset < 0)
        return set_error(CRT_E_INVALID, "crt_build_bvh: t_offset + triangles exceeds the 28-bit leaf start");

    BuiltTree tree;
    try {
        construct_tree(res.vertices, indices, t_offset, format, opt, threads, tree);
    } catch (const std::bad_alloc&) {
        return set_error(CRT_E_NOMEM, "crt_build_bvh: out of host memory");
    }

    // GenerateTriangles (BVHConstructor.cpp:630-650).
    res.triangles.resize(tree.sorted_refs.size());
    for (size_t i = 0; i < tree.sorted_refs.size(); ++i) {
        const size_t t = size_t(tree.sorted_refs[i]);
        crt_triangle& tr = res.triangles[i];
        tr.packed[0] = int32_t(indices[3 * t + 0]);
        tr.packed[1] = int32_t(indices[3 * t + 1]);
        tr.packed[2] = int32_t(indices[3 * t + 2]);
        tr.packed[3] = mesh_refs[t];
    }

    res.format = format;
    if (format == CRT_FORMAT_STACKLESS) {
        flatten_stackless(tree, res.stackless);
        res.stack.clear();
    } else {
        flatten_stack(tree, res.stack);
        res.stackless.clear();
    }
    res.stats.triangle_count = ntri;
    res.stats.node_count = tree.nodes.size() - 1;
    res.stats.leaf_count = tree.leaf_count;
    res.stats.split_fails = tree.split_fails;
    res.stats.max_depth = tree.max_depth;
    res.stats.array_length = tree.nodes.size();
    res.stats.vertex_count = res.vertices.size();
    return CRT_OK;
}

}  // namespace crt

// --------------------------------------------------------------------------
// C ABI
// --------------------------------------------------------------------------

extern "C" {

int crt_build_bvh(int format, const crt_mesh* meshes, int mesh_count, int32_t t_offset,
                  const crt_build_options* options, crt_bvh** out) {
    if (!out) return crt::set_error(CRT_E_INVALID, "crt_build_bvh: out is null");
    *out = nullptr;
    crt_bvh* b = new (std::nothrow) crt_bvh();
    if (!b) return crt::set_error(CRT_E_NOMEM, "crt_build_bvh: out of host memory");
    const int rc = crt::build_bvh(format, meshes, mesh_count, t_offset, options, b->r);
    if (rc != CRT_OK) {
        delete b;
        return rc;
    }
    b->t_offset = t_offset;
    if (options) b->options = *options;
    *out = b;
    return CRT_OK;
}

void crt_bvh_free(crt_bvh* bvh) { delete bvh; }

int crt_bvh_get_stats(const crt_bvh* bvh, crt_bvh_stats* out) {
    if (!bvh || !out) return crt::set_error(CRT_E_INVALID, "crt_bvh_get_stats: null argument");
    *out = bvh->r.stats;
    return CRT_OK;
}

const void* crt_bvh_nodes(const crt_bvh* bvh) {
    if (!bvh) return nullptr;
    return bvh->r.format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(bvh->r.stackless.data())
                                                 : static_cast<const void*>(bvh->r.stack.data());
}
const crt_triangle* crt_bvh_triangles(const crt_bvh* bvh) { return bvh ? bvh->r.triangles.data() : nullptr; }
const crt_vertex* crt_bvh_vertices(const crt_bvh* bvh) { return bvh ? bvh->r.vertices.data() : nullptr; }
int crt_bvh_format(const crt_bvh* bvh) { return bvh ? bvh->r.format : CRT_E_INVALID; }
int32_t crt_bvh_t_offset(const crt_bvh* bvh) { return bvh ? bvh->t_offset : 0; }

}  // extern "C"
