// Wide BVH (BVH4 / BVH8) derived from the reference's BVH2 buffers.
//
// The reference flattens its binary SAH tree into two layouts
// (BVHConstructor.cpp:783-930): pre-order nodes with miss links (stackless)
// and BFS entries holding both children (stack).  Either is decoded here back
// into a binary tree and collapsed into nodes of 4 or 8 children: starting
// from a BVH2 node's two children, the child of largest surface area is
// replaced by its own two children until the node holds `width` of them.
// The child boxes are the BVH2 boxes, bit for bit, and leaves keep their
// triangle ranges, so a closest-hit walk of the wide tree reports the same
// nearest triangle as any walk of the BVH2 (up to equal-distance ties).
//
// Triangles become 48-byte records in the same leaf order: the three vertex
// positions plus the triangle index and mesh id, so that a triangle test is
// three independent 16-byte loads instead of an index load followed by three
// dependent vertex loads.
//
// Memory of one wide node: width / 4 groups of 128 bytes (crt_wide_group).
// The inner children of a node occupy consecutive indices; nodes of one
// object are contiguous, so an entity's [root, root + count) range guards its
// walk exactly as node_offset / node_count guard the BVH2 walks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "candela_rt.h"
#include "crt_internal.h"

struct crt_wide {
    int32_t width = 4;
    std::vector<crt_wide_group> groups;
    std::vector<crt_tri_record> records;
    std::vector<crt_entity> entities;
    crt_wide_stats stats{};
};

namespace crt {
namespace {

struct Node2 {
    float lo[3], hi[3];
    int32_t left = -1, right = -1;  // decoded child ids; -1 for a leaf
    int32_t start = 0, count = 0;   // leaf triangles
    int32_t first = 0, last = 0;    // smallest leaf start / largest leaf end in the subtree
    int64_t tris = 0;               // triangles in the subtree
};

float half_area(const Node2& n) {
    const float dx = n.hi[0] - n.lo[0], dy = n.hi[1] - n.lo[1], dz = n.hi[2] - n.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

int32_t as_int(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}

struct Decoded {
    std::vector<Node2> nodes;  // nodes[0] is the root
};

// Fills subtree sums bottom-up (children always have larger ids than parents
// in both decoders below).
void finish_sums(Decoded& d) {
    for (int64_t i = int64_t(d.nodes.size()) - 1; i >= 0; --i) {
        Node2& n = d.nodes[size_t(i)];
        if (n.left < 0) {
            n.tris = n.count;
            n.first = n.start;
            n.last = n.start + n.count;
        } else {
            const Node2& a = d.nodes[size_t(n.left)];
            const Node2& b = d.nodes[size_t(n.right)];
            n.tris = a.tris + b.tris;
            n.first = std::min(a.first, b.first);
            n.last = std::max(a.last, b.last);
        }
    }
}

int leaf_ok(int32_t packed, int64_t ntris, const char* what) {
    const int32_t start = packed >> 4, len = packed & 0xF;
    if (start < 0 || int64_t(start) + len > ntris)
        return set_error(CRT_E_INVALID, std::string("crt_wide_build: ") + what + " leaf range outside the triangles");
    return CRT_OK;
}

// Stackless object [off, off + n): node i inner -> left i + 1, right = miss
// link of i + 1 (BVHConstructor.cpp:815-830); leaf -> (start << 4 | len).
int decode_stackless(const crt_stackless_node* nodes, int64_t off, int64_t n, int64_t ntris, Decoded& d) {
    std::vector<int32_t> seen(size_t(n), 0);
    std::vector<std::pair<int32_t, int32_t>> work;  // (local index, decoded id)
    d.nodes.clear();
    d.nodes.emplace_back();
    work.emplace_back(0, 0);
    while (!work.empty()) {
        const auto [li, id] = work.back();
        work.pop_back();
        if (li < 0 || li >= n) return set_error(CRT_E_INVALID, "crt_wide_build: stackless link out of range");
        if (seen[size_t(li)]++) return set_error(CRT_E_INVALID, "crt_wide_build: stackless links do not form a tree");
        const crt_stackless_node& s = nodes[off + li];
        Node2& out = d.nodes[size_t(id)];
        for (int k = 0; k < 3; ++k) {
            out.lo[k] = s.min[k];
            out.hi[k] = s.max[k];
        }
        const int32_t packed = as_int(s.min[3]);
        if (packed != -1) {
            if (int rc = leaf_ok(packed, ntris, "stackless")) return rc;
            out.start = packed >> 4;
            out.count = packed & 0xF;
            continue;
        }
        const int32_t l = li + 1;
        if (l >= n) return set_error(CRT_E_INVALID, "crt_wide_build: stackless inner node without children");
        const int32_t r = as_int(nodes[off + l].max[3]);
        // FlattenBVH: a right child escapes where its parent escapes (the
        // root's escape is -1), so the links are consistent with the tree.
        if (r >= 0 && r < n && as_int(nodes[off + r].max[3]) != as_int(s.max[3]))
            return set_error(CRT_E_INVALID, "crt_wide_build: stackless miss links inconsistent with the tree");
        const int32_t lid = int32_t(d.nodes.size()), rid = lid + 1;
        d.nodes[size_t(id)].left = lid;
        d.nodes[size_t(id)].right = rid;
        d.nodes.emplace_back();
        d.nodes.emplace_back();
        work.emplace_back(r, rid);
        work.emplace_back(l, lid);
    }
    finish_sums(d);
    return CRT_OK;
}

// Stack object [off, off + n): entry k holds both children of an inner node
// (BVHConstructor.cpp:847-930); X.min.w = leaf or -1, X.max.w = the inner
// child's own entry.  The root's box is the union of entry 0's children.
int decode_stack(const crt_stack_node* nodes, int64_t off, int64_t n, int64_t ntris, Decoded& d) {
    if (n < 1) return set_error(CRT_E_INVALID, "crt_wide_build: empty stack object");
    std::vector<int32_t> seen(size_t(n), 0);
    std::vector<std::pair<int32_t, int32_t>> work;  // (entry, decoded id of the node whose children it holds)
    d.nodes.clear();
    d.nodes.emplace_back();
    {
        const crt_stack_node& e = nodes[off];
        for (int k = 0; k < 3; ++k) {
            d.nodes[0].lo[k] = std::min(e.left.min[k], e.right.min[k]);
            d.nodes[0].hi[k] = std::max(e.left.max[k], e.right.max[k]);
        }
    }
    work.emplace_back(0, 0);
    while (!work.empty()) {
        const auto [entry, id] = work.back();
        work.pop_back();
        if (entry < 0 || entry >= n) return set_error(CRT_E_INVALID, "crt_wide_build: stack link out of range");
        if (seen[size_t(entry)]++) return set_error(CRT_E_INVALID, "crt_wide_build: stack links do not form a tree");
        const crt_stack_node& e = nodes[off + entry];
        const int32_t lid = int32_t(d.nodes.size()), rid = lid + 1;
        d.nodes[size_t(id)].left = lid;
        d.nodes[size_t(id)].right = rid;
        d.nodes.emplace_back();
        d.nodes.emplace_back();
        const crt_bounds* cb[2] = {&e.left, &e.right};
        for (int c = 1; c >= 0; --c) {
            Node2& out = d.nodes[size_t(lid + c)];
            for (int k = 0; k < 3; ++k) {
                out.lo[k] = cb[c]->min[k];
                out.hi[k] = cb[c]->max[k];
            }
            const int32_t packed = as_int(cb[c]->min[3]);
            if (packed != -1) {
                if (int rc = leaf_ok(packed, ntris, "stack")) return rc;
                out.start = packed >> 4;
                out.count = packed & 0xF;
            } else {
                work.emplace_back(as_int(cb[c]->max[3]), lid + c);
            }
        }
    }
    finish_sums(d);
    return CRT_OK;
}

struct Collapser {
    int32_t width;
    int32_t max_leaf;
    std::vector<crt_wide_group>& groups;
    int32_t mode = CRT_COLLAPSE_GREEDY;
    float node_cost = 1.0f;   // CRT_COLLAPSE_SAH: cost of one wide-node visit ...
    float tri_cost = 0.5f;    // ... and of one triangle test, both per unit of half area
    int64_t leaves = 0;
    int64_t used_slots = 0;
    int32_t max_depth = 0;
    // CRT_COLLAPSE_SAH tables, per decoded node: best[n * width + i - 1] is the
    // least cost of covering the subtree with at most i slots; as_leaf[n] says
    // whether one slot holds it best as a leaf (else as a wide node)
    std::vector<float> best;
    std::vector<uint8_t> as_leaf;

    int32_t groups_per_node() const { return width / 4; }

    bool is_leaf(const Node2& n) const {
        if (n.left < 0) return true;
        return n.tris <= max_leaf && int64_t(n.last) - n.first == n.tris;
    }

    int64_t alloc(int64_t count) {
        const int64_t first = int64_t(groups.size()) / groups_per_node();
        crt_wide_group empty;
        for (int k = 0; k < 4; ++k) {
            empty.lo_x[k] = empty.lo_y[k] = empty.lo_z[k] = 0.0f;
            empty.hi_x[k] = empty.hi_y[k] = empty.hi_z[k] = 0.0f;
            empty.child[k] = 0;
            empty.count[k] = -1;
        }
        groups.insert(groups.end(), size_t(count * groups_per_node()), empty);
        return first;
    }

    bool slot_is_leaf(const Decoded& d, int32_t k) const {
        if (mode == CRT_COLLAPSE_SAH) return as_leaf[size_t(k)] != 0;
        return is_leaf(d.nodes[size_t(k)]);
    }

    // Least cost of covering the subtree of n with at most j >= 2 slots when n
    // itself takes none (its two children share the j slots); k_out = the
    // left child's share.
    float split_cost(const Node2& n, int32_t j, int32_t* k_out) const {
        float c = std::numeric_limits<float>::infinity();
        int32_t bk = 1;
        for (int32_t k = 1; k < j; ++k) {
            const float v = best[size_t(n.left) * size_t(width) + size_t(k - 1)] +
                            best[size_t(n.right) * size_t(width) + size_t(j - k - 1)];
            if (v < c) {
                c = v;
                bk = k;
            }
        }
        if (k_out) *k_out = bk;
        return c;
    }

    // Fills best / as_leaf bottom-up (children have larger ids than parents).
    // A subtree in one slot costs area * tri_cost * triangles as a leaf (only
    // where the greedy collapse could make it one: a BVH2 leaf, or at most
    // max_leaf contiguous triangles) or area * node_cost plus its best cover
    // by `width` slots as a wide node; with more slots it may instead hand
    // them to its two children.
    void plan_sah(const Decoded& d) {
        const size_t n = d.nodes.size();
        best.assign(n * size_t(width), 0.0f);
        as_leaf.assign(n, 0);
        for (int64_t i = int64_t(n) - 1; i >= 0; --i) {
            const Node2& c = d.nodes[size_t(i)];
            float* b = &best[size_t(i) * size_t(width)];
            const float area = std::max(half_area(c), 0.0f);
            if (c.left < 0) {
                const float leaf = area * tri_cost * float(c.count);
                for (int32_t j = 0; j < width; ++j) b[j] = leaf;
                as_leaf[size_t(i)] = 1;
                continue;
            }
            const float leaf = is_leaf(c) ? area * tri_cost * float(c.tris) : std::numeric_limits<float>::infinity();
            const float node = area * node_cost + split_cost(c, width, nullptr);
            as_leaf[size_t(i)] = leaf <= node ? 1 : 0;
            b[0] = std::min(leaf, node);
            for (int32_t j = 2; j <= width; ++j) b[j - 1] = std::min(b[0], split_cost(c, j, nullptr));
        }
    }

    // The slots of a wide node at n under the SAH plan: the best cover of n's
    // subtree by `width` slots with n itself opened.
    void split_sah(const Decoded& d, const Node2& r, int32_t j, std::vector<int32_t>& out) const {
        int32_t k = 1;
        split_cost(r, j, &k);
        cover_sah(d, r.left, k, out);
        cover_sah(d, r.right, j - k, out);
    }

    void cover_sah(const Decoded& d, int32_t id, int32_t j, std::vector<int32_t>& out) const {
        const Node2& c = d.nodes[size_t(id)];
        const float* b = &best[size_t(id) * size_t(width)];
        if (j == 1 || c.left < 0 || b[j - 1] >= b[0]) {
            out.push_back(id);
            return;
        }
        split_sah(d, c, j, out);
    }

    // Collapses the decoded tree into wide nodes; returns the root's index.
    int collapse(const Decoded& d, int64_t& root) {
        struct Item {
            int64_t wide;
            int32_t node;
            int32_t depth;
        };
        std::vector<Item> work;
        if (mode == CRT_COLLAPSE_SAH) plan_sah(d);
        root = alloc(1);
        work.push_back({root, 0, 1});
        std::vector<int32_t> kids;
        while (!work.empty()) {
            const Item it = work.back();
            work.pop_back();
            max_depth = std::max(max_depth, it.depth);
            const Node2& r = d.nodes[size_t(it.node)];
            kids.clear();
            if (mode == CRT_COLLAPSE_SAH) {
                if (r.left < 0) {
                    kids.push_back(it.node);
                } else {
                    split_sah(d, r, width, kids);
                }
            } else if (is_leaf(r)) {
                kids.push_back(it.node);
            } else {
                kids.push_back(r.left);
                kids.push_back(r.right);
            }
            while (mode == CRT_COLLAPSE_GREEDY && int32_t(kids.size()) < width) {
                int best = -1;
                float best_area = -1.0f;
                for (int k = 0; k < int(kids.size()); ++k) {
                    const Node2& c = d.nodes[size_t(kids[size_t(k)])];
                    if (is_leaf(c)) continue;
                    const float a = half_area(c);
                    if (a > best_area) {
                        best_area = a;
                        best = k;
                    }
                }
                if (best < 0) break;
                const Node2& c = d.nodes[size_t(kids[size_t(best)])];
                kids[size_t(best)] = c.left;
                kids.insert(kids.begin() + best + 1, c.right);
            }
            int64_t inner = 0;
            for (int32_t k : kids) inner += slot_is_leaf(d, k) ? 0 : 1;
            const int64_t base = inner ? alloc(inner) : 0;
            int64_t next = base;
            for (int slot = 0; slot < int(kids.size()); ++slot) {
                const Node2& c = d.nodes[size_t(kids[size_t(slot)])];
                crt_wide_group& g = groups[size_t(it.wide * groups_per_node() + slot / 4)];
                const int j = slot % 4;
                g.lo_x[j] = c.lo[0];
                g.lo_y[j] = c.lo[1];
                g.lo_z[j] = c.lo[2];
                g.hi_x[j] = c.hi[0];
                g.hi_y[j] = c.hi[1];
                g.hi_z[j] = c.hi[2];
                if (slot_is_leaf(d, kids[size_t(slot)])) {
                    g.child[j] = c.left < 0 ? c.start : c.first;
                    g.count[j] = int32_t(c.left < 0 ? c.count : c.tris);
                    ++leaves;
                } else {
                    if (next > std::numeric_limits<int32_t>::max())
                        return set_error(CRT_E_INVALID, "crt_wide_build: more than 2^31 wide nodes");
                    g.child[j] = int32_t(next);
                    g.count[j] = 0;
                    work.push_back({next, kids[size_t(slot)], it.depth + 1});
                    ++next;
                }
                ++used_slots;
            }
        }
        return CRT_OK;
    }
};

}  // namespace
}  // namespace crt

extern "C" {

int crt_wide_build(int src_format, const void* nodes, int64_t node_count, const crt_triangle* triangles,
                   int64_t triangle_count, const crt_vertex* vertices, int64_t vertex_count,
                   const crt_entity* entities, int32_t entity_count, int32_t width, int32_t max_leaf,
                   crt_wide** out) {
    const crt_wide_options o{width, max_leaf, CRT_COLLAPSE_GREEDY, 0.0f, 0.0f};
    return crt_wide_build_ex(src_format, nodes, node_count, triangles, triangle_count, vertices, vertex_count,
                             entities, entity_count, &o, out);
}

int crt_wide_build_ex(int src_format, const void* nodes, int64_t node_count, const crt_triangle* triangles,
                      int64_t triangle_count, const crt_vertex* vertices, int64_t vertex_count,
                      const crt_entity* entities, int32_t entity_count, const crt_wide_options* options,
                      crt_wide** out) {
    using namespace crt;
    if (!out) return set_error(CRT_E_INVALID, "crt_wide_build: out is null");
    *out = nullptr;
    if (!options) return set_error(CRT_E_INVALID, "crt_wide_build: options is null");
    const int32_t width = options->width, max_leaf = options->max_leaf;
    if (options->collapse != CRT_COLLAPSE_GREEDY && options->collapse != CRT_COLLAPSE_SAH)
        return set_error(CRT_E_INVALID, "crt_wide_build: collapse must be CRT_COLLAPSE_GREEDY or CRT_COLLAPSE_SAH");
    const float node_cost = options->node_cost == 0.0f ? 1.0f : options->node_cost;
    const float tri_cost = options->tri_cost == 0.0f ? 0.5f : options->tri_cost;
    if (!(node_cost > 0.0f) || !(tri_cost > 0.0f) || !std::isfinite(node_cost) || !std::isfinite(tri_cost))
        return set_error(CRT_E_INVALID, "crt_wide_build: node_cost and tri_cost must be positive and finite");
    if (src_format != CRT_FORMAT_STACKLESS && src_format != CRT_FORMAT_STACK)
        return set_error(CRT_E_INVALID, "crt_wide_build: source format must be STACKLESS or STACK");
    if (width != 4 && width != 8) return set_error(CRT_E_INVALID, "crt_wide_build: width must be 4 or 8");
    if (max_leaf < 2 || max_leaf > 15) return set_error(CRT_E_INVALID, "crt_wide_build: max_leaf must be in 2..15");
    if (node_count < 0 || triangle_count < 0 || vertex_count < 0 || entity_count < 0 ||
        (node_count && !nodes) || (triangle_count && !triangles) || (vertex_count && !vertices) ||
        (entity_count && !entities))
        return set_error(CRT_E_INVALID, "crt_wide_build: bad buffers");
    if (triangle_count > std::numeric_limits<int32_t>::max())
        return set_error(CRT_E_INVALID, "crt_wide_build: more than 2^31 triangles");
    crt_wide* w = new (std::nothrow) crt_wide();
    if (!w) return set_error(CRT_E_NOMEM, "crt_wide_build: out of host memory");
    try {
        w->width = width;
        w->records.resize(size_t(triangle_count));
        for (int64_t i = 0; i < triangle_count; ++i) {
            const int32_t* p = triangles[i].packed;
            crt_tri_record& r = w->records[size_t(i)];
            for (int k = 0; k < 3; ++k) {
                if (p[k] < 0 || p[k] >= vertex_count) {
                    delete w;
                    return set_error(CRT_E_INVALID, "crt_wide_build: vertex index out of range");
                }
            }
            std::memcpy(r.v0, vertices[p[0]].position, 12);
            std::memcpy(r.v1, vertices[p[1]].position, 12);
            std::memcpy(r.v2, vertices[p[2]].position, 12);
            r.tri = int32_t(i);
            r.mesh = p[3];
            r.pad = 0;
        }
        Collapser col{width, max_leaf, w->groups};
        col.mode = options->collapse;
        col.node_cost = node_cost;
        col.tri_cost = tri_cost;
        std::map<std::pair<int32_t, int32_t>, std::pair<int32_t, int32_t>> built;  // (offset, count) -> (root, nodes)
        w->entities.assign(entities, entities + entity_count);
        Decoded d;
        for (int32_t e = 0; e < entity_count; ++e) {
            crt_entity& ent = w->entities[size_t(e)];
            const auto key = std::make_pair(ent.node_offset, ent.node_count);
            auto hit = built.find(key);
            if (hit == built.end()) {
                if (ent.node_offset < 0 || ent.node_count <= 0 || int64_t(ent.node_offset) + ent.node_count > node_count) {
                    delete w;
                    return set_error(CRT_E_INVALID, "crt_wide_build: entity node range outside the node buffer");
                }
                int rc = src_format == CRT_FORMAT_STACKLESS
                             ? decode_stackless(static_cast<const crt_stackless_node*>(nodes), ent.node_offset,
                                                ent.node_count, triangle_count, d)
                             : decode_stack(static_cast<const crt_stack_node*>(nodes), ent.node_offset,
                                            ent.node_count, triangle_count, d);
                int64_t root = 0;
                const int64_t before = int64_t(w->groups.size()) / (width / 4);
                if (!rc) rc = col.collapse(d, root);
                if (rc) {
                    delete w;
                    return rc;
                }
                const int64_t after = int64_t(w->groups.size()) / (width / 4);
                if (after > std::numeric_limits<int32_t>::max()) {
                    delete w;
                    return set_error(CRT_E_INVALID, "crt_wide_build: more than 2^31 wide nodes");
                }
                hit = built.emplace(key, std::make_pair(int32_t(root), int32_t(after - before))).first;
            }
            ent.node_offset = hit->second.first;
            ent.node_count = hit->second.second;
        }
        crt_wide_stats& st = w->stats;
        st.node_count = int64_t(w->groups.size()) / (width / 4);
        st.leaf_count = col.leaves;
        st.record_count = triangle_count;
        st.width = width;
        st.max_depth = col.max_depth;
        st.mean_fill = st.node_count ? double(col.used_slots) / double(st.node_count) : 0.0;
    } catch (const std::bad_alloc&) {
        delete w;
        return set_error(CRT_E_NOMEM, "crt_wide_build: out of host memory");
    }
    *out = w;
    return CRT_OK;
}

void crt_wide_free(crt_wide* wide) { delete wide; }

int crt_wide_get_stats(const crt_wide* wide, crt_wide_stats* out) {
    if (!wide || !out) return crt::set_error(CRT_E_INVALID, "crt_wide_get_stats: null argument");
    *out = wide->stats;
    return CRT_OK;
}

const crt_wide_group* crt_wide_groups(const crt_wide* wide) { return wide ? wide->groups.data() : nullptr; }
const crt_tri_record* crt_wide_records(const crt_wide* wide) { return wide ? wide->records.data() : nullptr; }
const crt_entity* crt_wide_entities(const crt_wide* wide) { return wide ? wide->entities.data() : nullptr; }

}  // extern "C"
