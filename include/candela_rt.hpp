// candela_rt.hpp -- C++ host mirror of the reference's RayIntersector<T>
// (Source/Core/BVH/Intersector.h:60-124) over the C ABI of candela_rt.h.
//
// Same class and member names as the reference so a caller switches by
// changing the include and the namespace:
//
//   reference                                 here
//   Candela::RayIntersector<StacklessNode>    candela::RayIntersector<candela::StacklessTraversalNode>
//   AddObject(const Object&)                  AddObject(object_id, meshes)
//   PushEntity(const Entity&)                 PushEntity(object_id, model, emissive, translucency)
//   BufferEntities() / BufferData(bool)       same
//   IntersectPrimary(out, W, H, camera)       IntersectPrimary(rays, hits, W, H, inv_view, inv_proj)
//   (GLSL) IntersectRay / IntersectScene       Intersect(query, rays, hits, n)
//   m_BVHNodes / m_BVHVertices / ...          Nodes() / Vertices() / Triangles() / Entities()
//
// The reference throws string literals (Intersector.h:147,204); this header
// throws candela::Error (a std::runtime_error carrying crt_last_error()).
// Header-only; link with libcandela_rt.so.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "candela_rt.h"

namespace candela {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(int rc, const char* what) {
    if (rc != CRT_OK) throw Error(std::string(what) + ": " + crt_last_error());
}

// BVH::StacklessTraversalNode / BVH::StackTraversalNode (Intersector.h:39-40).
using StacklessTraversalNode = crt_stackless_node;
using StackTraversalNode = crt_stack_node;
using Vertex = crt_vertex;
using Triangle = crt_triangle;
using BVHEntity = crt_entity;

// One mesh of an object: vertices, indices and GlobalMeshNumber (Mesh.h:30-31,58).
struct MeshView {
    const Vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count;
    int32_t global_mesh_number;
};

template <typename T>
class RayIntersector {
    static_assert(std::is_same<T, StacklessTraversalNode>::value || std::is_same<T, StackTraversalNode>::value,
                  "RayIntersector<T>: T is StacklessTraversalNode or StackTraversalNode");

public:
    static constexpr int kFormat = std::is_same<T, StacklessTraversalNode>::value ? CRT_FORMAT_STACKLESS
                                                                                  : CRT_FORMAT_STACK;

    RayIntersector() { check(crt_create(kFormat, &ctx_), "RayIntersector"); }
    ~RayIntersector() { crt_destroy(ctx_); }
    RayIntersector(const RayIntersector&) = delete;
    RayIntersector& operator=(const RayIntersector&) = delete;

    // Initialize() compiled the GLSL kernels; the HIP kernels are in the library.
    void Initialize() {}

    void AddObject(int32_t object_id, const std::vector<MeshView>& meshes, const crt_build_options* opt = nullptr) {
        std::vector<crt_mesh> m(meshes.size());
        for (size_t i = 0; i < meshes.size(); ++i)
            m[i] = crt_mesh{meshes[i].vertices, meshes[i].vertex_count, meshes[i].indices, meshes[i].index_count,
                            meshes[i].global_mesh_number};
        check(crt_add_object(ctx_, object_id, m.data(), int(m.size()), opt), "AddObject");
    }

    // AddObject for buffers produced elsewhere (e.g. the reference's own BuildBVH).
    void AddObject(int32_t object_id, const std::vector<T>& nodes, const std::vector<Triangle>& tris,
                   const std::vector<Vertex>& verts) {
        check(crt_add_object_prebuilt(ctx_, object_id, nodes.data(), int64_t(nodes.size()), tris.data(),
                                      int64_t(tris.size()), verts.data(), int64_t(verts.size())),
              "AddObject");
    }

    // model: column-major 4x4 (glm::mat4 memory layout).
    void PushEntity(int32_t object_id, const std::array<float, 16>& model, float emissive = 0.0f,
                    float translucency = 0.0f) {
        check(crt_push_entity(ctx_, object_id, model.data(), emissive, translucency), "PushEntity");
    }

    void BufferEntities() { check(crt_buffer_entities(ctx_), "BufferEntities"); }
    void BufferData(bool clear_cpu_data) { check(crt_buffer_data(ctx_, clear_cpu_data ? 1 : 0), "BufferData"); }

    // IntersectRay / IntersectRayIgnoreTransparent (closest) and IntersectRay(ro, rd) (any) over
    // device buffers, stream-ordered.
    void Intersect(int query, const crt_ray* d_rays, crt_hit* d_hits, int64_t n, uint32_t flags = 0,
                   void* stream = nullptr) {
        check(crt_intersect(ctx_, query, flags, d_rays, d_hits, n, stream), "Intersect");
    }

    void IntersectPrimary(crt_ray* d_rays, crt_hit* d_hits, int width, int height,
                          const std::array<float, 16>& inv_view, const std::array<float, 16>& inv_proj,
                          uint32_t flags = 0, void* stream = nullptr) {
        check(crt_intersect_primary(ctx_, d_rays, d_hits, width, height, inv_view.data(), inv_proj.data(), flags,
                                    stream),
              "IntersectPrimary");
    }

    // Collide (Intersector.h:353-357) is a stub returning false in the reference.
    bool Collide(const float* /*point*/) const { return false; }
    void Recompile() {}

    // m_BVHNodes / m_BVHTriangles / m_BVHVertices / m_BVHEntities (copies).
    std::vector<T> Nodes() const { return host<T>(0); }
    std::vector<Triangle> Triangles() const { return host<Triangle>(1); }
    std::vector<Vertex> Vertices() const { return host<Vertex>(2); }
    std::vector<BVHEntity> Entities() const { return host<BVHEntity>(3); }

    crt_scene_view Scene() const {
        crt_scene_view s;
        check(crt_context_scene(ctx_, &s), "Scene");
        return s;
    }

    crt_context* handle() const { return ctx_; }

private:
    template <typename U>
    std::vector<U> host(int which) const {
        const void* p[4] = {};
        int64_t n[4] = {};
        check(crt_context_host_buffers(ctx_, &p[0], &n[0], reinterpret_cast<const crt_triangle**>(&p[1]), &n[1],
                                       reinterpret_cast<const crt_vertex**>(&p[2]), &n[2],
                                       reinterpret_cast<const crt_entity**>(&p[3]), &n[3]),
              "host buffers");
        const U* b = static_cast<const U*>(p[which]);
        return std::vector<U>(b, b + n[which]);
    }

    crt_context* ctx_ = nullptr;
};

// BuildBVH (BVHConstructor.cpp:951-1108): the flattened buffers of one object.
template <typename T>
struct BuiltBVH {
    std::vector<T> nodes;
    std::vector<Triangle> triangles;
    std::vector<Vertex> vertices;
    crt_bvh_stats stats;
};

template <typename T>
BuiltBVH<T> BuildBVH(const std::vector<MeshView>& meshes, int32_t t_offset = 0, const crt_build_options* opt = nullptr) {
    constexpr int fmt = RayIntersector<T>::kFormat;
    std::vector<crt_mesh> m(meshes.size());
    for (size_t i = 0; i < meshes.size(); ++i)
        m[i] = crt_mesh{meshes[i].vertices, meshes[i].vertex_count, meshes[i].indices, meshes[i].index_count,
                        meshes[i].global_mesh_number};
    crt_bvh* b = nullptr;
    check(crt_build_bvh(fmt, m.data(), int(m.size()), t_offset, opt, &b), "BuildBVH");
    BuiltBVH<T> out;
    crt_bvh_get_stats(b, &out.stats);
    const T* nodes = static_cast<const T*>(crt_bvh_nodes(b));
    out.nodes.assign(nodes, nodes + out.stats.array_length);
    out.triangles.assign(crt_bvh_triangles(b), crt_bvh_triangles(b) + out.stats.triangle_count);
    out.vertices.assign(crt_bvh_vertices(b), crt_bvh_vertices(b) + out.stats.vertex_count);
    crt_bvh_free(b);
    return out;
}

}  // namespace candela
