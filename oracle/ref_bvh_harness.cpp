// ref_bvh_harness.cpp -- drives the reference's own CPU BVH constructor.
//
// TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) against the
// unmodified Source/Core/BVH/BVHConstructor.cpp of the reference tree, it
// produces the reference's flattened buffers for a mesh given on disk so the
// tests can compare them byte for byte with libcandela_rt's constructor.
//
//   ref_bvh <stackless|stack> <in.bin> <out.bin>
//
// in.bin : u32 nverts, u32 nindices, Vertex[nverts] (32 B), u32 idx[nindices],
//          i32 mesh_id[nindices/3]   (already concatenated, BVHConstructor.cpp:981-1002)
// out.bin: u64 array_len, u64 ntri, u64 node_count, u64 leaf_count,
//          u64 split_fails, u64 max_depth, node bytes, Triangle[ntri]
//
// It calls ConstructHierarchy / ConstructHierarchy_StackBVH
// (BVHConstructor.cpp:702-767), the functions BuildBVH calls after mesh
// concatenation, because BuildBVH takes an Object whose constructor needs a
// GL context.  Consequences, both covered by the tests:
//   * the constructor's counters are file-static (BVHConstructor.cpp:58-66) and
//     only BuildBVH resets them, so a process performs one build;
//   * TriangleOffset_ and UsesStackless keep their zero initial values:
//     t_offset = 0 and no random child swap (the swap is only enabled by the
//     stackless BuildBVH, BVHConstructor.cpp:961).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "BVH/BVHConstructor.h"

namespace Candela {
namespace BVH {
void ConstructHierarchy(const std::vector<Vertex>& Vertices, const std::vector<GLuint>& OriginalIndices,
                        std::vector<FlattenedNode>& FlattenedNodes, std::vector<Triangle>& oTriangles,
                        const std::vector<int>& MeshIDs, Node* RootNode);
void ConstructHierarchy_StackBVH(const std::vector<Vertex>& Vertices, const std::vector<GLuint>& OriginalIndices,
                                 std::vector<FlattenedStackNode>& FlattenedNodes, std::vector<Triangle>& oTriangles,
                                 const std::vector<int>& MeshIDs, Node* RootNode);
}  // namespace BVH
}  // namespace Candela

static bool read_all(FILE* f, void* p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

static uint64_t count_leaves(const Candela::BVH::Node* n) {
    if (n->IsLeafNode) return 1;
    return count_leaves(n->LeftChildPtr) + count_leaves(n->RightChildPtr);
}

int main(int argc, char** argv) {
    using namespace Candela;
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_bvh <stackless|stack> <in.bin> <out.bin>\n");
        return 2;
    }
    const bool stackless = std::strcmp(argv[1], "stackless") == 0;
    FILE* in = std::fopen(argv[2], "rb");
    if (!in) return 3;
    uint32_t nv = 0, ni = 0;
    if (!read_all(in, &nv, 4) || !read_all(in, &ni, 4)) return 4;
    static_assert(sizeof(Vertex) == 32, "Vertex layout");
    std::vector<Vertex> verts(nv);
    std::vector<GLuint> idx(ni);
    std::vector<int> mesh(ni / 3);
    if (!read_all(in, verts.data(), size_t(nv) * 32) || !read_all(in, idx.data(), size_t(ni) * 4) ||
        !read_all(in, mesh.data(), size_t(ni / 3) * 4))
        return 5;
    std::fclose(in);

    BVH::Node* root = new BVH::Node;
    root->LeftChildPtr = nullptr;
    root->RightChildPtr = nullptr;
    root->StartIndex = 0;
    root->Length = ni / 3;

    std::vector<BVH::Triangle> tris;
    std::vector<char> node_bytes;
    uint64_t array_len = 0;
    if (stackless) {
        std::vector<BVH::FlattenedNode> nodes;
        BVH::ConstructHierarchy(verts, idx, nodes, tris, mesh, root);
        array_len = nodes.size();
        node_bytes.resize(nodes.size() * sizeof(BVH::FlattenedNode));
        std::memcpy(node_bytes.data(), nodes.data(), node_bytes.size());
    } else {
        std::vector<BVH::FlattenedStackNode> nodes;
        BVH::ConstructHierarchy_StackBVH(verts, idx, nodes, tris, mesh, root);
        array_len = nodes.size();
        node_bytes.resize(nodes.size() * sizeof(BVH::FlattenedStackNode));
        std::memcpy(node_bytes.data(), nodes.data(), node_bytes.size());
    }
    FILE* out = std::fopen(argv[3], "wb");
    if (!out) return 6;
    const uint64_t hdr[6] = {array_len, uint64_t(tris.size()), array_len - 1, count_leaves(root), 0, 0};
    std::fwrite(hdr, 8, 6, out);
    std::fwrite(node_bytes.data(), 1, node_bytes.size(), out);
    std::fwrite(tris.data(), sizeof(BVH::Triangle), tris.size(), out);
    std::fclose(out);
    return 0;
}
