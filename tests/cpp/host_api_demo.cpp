// Exercises include/candela_rt.hpp the way the reference's host code uses
// RayIntersector<T> (Pipeline.cpp:1019-1028, 1250-1251): add objects, push
// entities, buffer, and (with --gpu) trace a few rays.  Prints "OK" on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "candela_rt.hpp"

using namespace candela;

static void grid_mesh(int n, float y, std::vector<Vertex>& v, std::vector<uint32_t>& idx) {
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            Vertex vv{};
            vv.position[0] = float(i);
            vv.position[1] = y + 0.1f * std::sin(float(i * 7 + j * 3));
            vv.position[2] = float(j);
            vv.position[3] = 1.0f;
            v.push_back(vv);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            uint32_t a = j * (n + 1) + i, b = a + 1, c = a + (n + 1), d = c + 1;
            idx.insert(idx.end(), {a, b, d, a, d, c});
        }
}

#define EXPECT(c)                                         \
    do {                                                  \
        if (!(c)) {                                       \
            std::printf("FAILED: %s (line %d)\n", #c, __LINE__); \
            return 1;                                     \
        }                                                 \
    } while (0)

template <typename T>
static int run(bool gpu) {
    std::vector<Vertex> v0, v1;
    std::vector<uint32_t> i0, i1;
    grid_mesh(20, 0.0f, v0, i0);
    grid_mesh(12, 2.0f, v1, i1);
    std::vector<MeshView> obj0{{v0.data(), uint32_t(v0.size()), i0.data(), uint32_t(i0.size()), 0}};
    std::vector<MeshView> obj1{{v1.data(), uint32_t(v1.size()), i1.data(), uint32_t(i1.size()), 1}};

    auto b0 = BuildBVH<T>(obj0);
    auto b1 = BuildBVH<T>(obj1, int32_t(b0.triangles.size()));
    EXPECT(b0.triangles.size() == i0.size() / 3);
    EXPECT(b0.nodes.size() == b0.stats.array_length);

    RayIntersector<T> r;
    r.Initialize();
    r.AddObject(7, obj0);
    r.AddObject(9, obj1);
    std::array<float, 16> eye{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    r.PushEntity(7, eye);
    r.PushEntity(9, eye, 0.5f, 0.25f);
    bool threw = false;
    try {
        r.PushEntity(123, eye);
    } catch (const Error&) {
        threw = true;
    }
    EXPECT(threw);
    auto nodes = r.Nodes();
    EXPECT(nodes.size() == b0.nodes.size() + b1.nodes.size());
    EXPECT(std::memcmp(nodes.data(), b0.nodes.data(), b0.nodes.size() * sizeof(T)) == 0);
    EXPECT(std::memcmp(nodes.data() + b0.nodes.size(), b1.nodes.data(), b1.nodes.size() * sizeof(T)) == 0);
    auto tris = r.Triangles();
    EXPECT(tris[b0.triangles.size()].packed[0] == b1.triangles[0].packed[0] + int32_t(v0.size()));
    if (!gpu) return 0;

    r.BufferEntities();
    r.BufferData(false);
    auto ents = r.Entities();
    EXPECT(ents.size() == 2 && ents[1].node_offset == int32_t(b0.nodes.size()));
    const int n = 400;
    std::vector<crt_ray> rays(n);
    for (int k = 0; k < n; ++k) {
        rays[k] = crt_ray{{0.5f + (k % 20), 10.0f, 0.5f + (k / 20)}, 1e6f, {0.0f, -1.0f, 0.0f}, uint32_t(k)};
    }
    void *dr = nullptr, *dh = nullptr;
    check(crt_device_alloc(&dr, n * sizeof(crt_ray)), "alloc");
    check(crt_device_alloc(&dh, n * sizeof(crt_hit)), "alloc");
    check(crt_memcpy_h2d(dr, rays.data(), n * sizeof(crt_ray), nullptr), "h2d");
    r.Intersect(CRT_QUERY_CLOSEST, static_cast<crt_ray*>(dr), static_cast<crt_hit*>(dh), n);
    std::vector<crt_hit> hits(n);
    check(crt_memcpy_d2h(hits.data(), dh, n * sizeof(crt_hit), nullptr), "d2h");
    int upper = 0;
    for (const auto& h : hits) {
        EXPECT(h.t > 0.0f);
        if (h.entity == 1) ++upper;
    }
    EXPECT(upper > 100);   // the raised grid covers [0,12]^2 of the [0,20]^2 footprint
    crt_device_free(dr);
    crt_device_free(dh);
    return 0;
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    if (run<StacklessTraversalNode>(gpu)) return 1;
    if (run<StackTraversalNode>(gpu)) return 1;
    std::printf("OK\n");
    return 0;
}
