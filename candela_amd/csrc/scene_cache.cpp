// BVH cache files: one object's flattened buffers on disk, so that an
// unchanged object is not rebuilt at every launch.
//
// The reference has no such thing: RayIntersector::AddObject
// (Intersector.h:169-198) runs BuildBVH (BVHConstructor.cpp:951-1108) for
// every object every time the application starts, which dominates the set-up
// of large scenes (15.6M triangles: about 10 s on 8 cores).
//
// A file is the 256-byte crt_cache_header followed by the node, triangle and
// vertex sections, byte for byte as BuildBVH leaves them.  Each section and
// the header carry an XXH64 hash, so a truncated or altered file is refused
// rather than traced.  Leaf starts depend on where the object lands in the
// context's triangle array (t_offset); files are written at t_offset 0 and
// rebased on load, which gives the same bytes as a build at the new offset.
#include <cerrno>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "candela_rt.h"
#include "crt_internal.h"

namespace crt {
namespace {

static_assert(sizeof(crt_cache_header) == 256, "cache header is 256 bytes");

// ---------------------------------------------------------------------------
// XXH64, after Y. Collet's published specification of xxHash64 (streaming form).
// ---------------------------------------------------------------------------
constexpr uint64_t kP1 = 0x9E3779B185EBCA87ull;
constexpr uint64_t kP2 = 0xC2B2AE3D27D4EB4Full;
constexpr uint64_t kP3 = 0x165667B19E3779F9ull;
constexpr uint64_t kP4 = 0x85EBCA77C2B2AE63ull;
constexpr uint64_t kP5 = 0x27D4EB2F165667C5ull;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t read64(const uint8_t* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline uint32_t read32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint64_t round64(uint64_t acc, uint64_t lane) {
    acc += lane * kP2;
    return rotl(acc, 31) * kP1;
}
inline uint64_t merge64(uint64_t h, uint64_t v) {
    h ^= round64(0, v);
    return h * kP1 + kP4;
}

class Hash64 {
  public:
    explicit Hash64(uint64_t seed = 0) : seed_(seed) {
        v_[0] = seed + kP1 + kP2;
        v_[1] = seed + kP2;
        v_[2] = seed;
        v_[3] = seed - kP1;
    }
    void update(const void* data, size_t n) {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        total_ += n;
        if (fill_ + n < 32) {
            std::memcpy(buf_ + fill_, p, n);
            fill_ += n;
            return;
        }
        if (fill_) {
            const size_t take = 32 - fill_;
            std::memcpy(buf_ + fill_, p, take);
            stripe(buf_);
            p += take;
            n -= take;
            fill_ = 0;
        }
        while (n >= 32) {
            stripe(p);
            p += 32;
            n -= 32;
        }
        std::memcpy(buf_, p, n);
        fill_ = n;
    }
    template <class T>
    void value(const T& v) {
        update(&v, sizeof(T));
    }
    uint64_t digest() const {
        uint64_t h;
        if (total_ >= 32) {
            h = rotl(v_[0], 1) + rotl(v_[1], 7) + rotl(v_[2], 12) + rotl(v_[3], 18);
            for (int i = 0; i < 4; ++i) h = merge64(h, v_[i]);
        } else {
            h = seed_ + kP5;
        }
        h += total_;
        const uint8_t* p = buf_;
        size_t n = fill_;
        while (n >= 8) {
            h ^= round64(0, read64(p));
            h = rotl(h, 27) * kP1 + kP4;
            p += 8;
            n -= 8;
        }
        if (n >= 4) {
            h ^= uint64_t(read32(p)) * kP1;
            h = rotl(h, 23) * kP2 + kP3;
            p += 4;
            n -= 4;
        }
        while (n > 0) {
            h ^= uint64_t(*p) * kP5;
            h = rotl(h, 11) * kP1;
            ++p;
            --n;
        }
        h ^= h >> 33;
        h *= kP2;
        h ^= h >> 29;
        h *= kP3;
        h ^= h >> 32;
        return h;
    }

  private:
    void stripe(const uint8_t* p) {
        for (int i = 0; i < 4; ++i) v_[i] = round64(v_[i], read64(p + 8 * i));
    }
    uint64_t seed_;
    uint64_t v_[4];
    uint8_t buf_[32];
    size_t fill_ = 0;
    uint64_t total_ = 0;
};

uint64_t hash_bytes(const void* data, size_t n) {
    Hash64 h;
    h.update(data, n);
    return h.digest();
}

// The C++ library that drew CRT_SWAP_SEEDED's child swaps.
void toolchain_id(char out[32]) {
    std::memset(out, 0, 32);
#if defined(__GLIBCXX__)
    std::snprintf(out, 32, "libstdc++ %ld", long(__GLIBCXX__));
#elif defined(_LIBCPP_VERSION)
    std::snprintf(out, 32, "libc++ %ld", long(_LIBCPP_VERSION));
#else
    std::snprintf(out, 32, "unknown");
#endif
}

constexpr uint64_t align64(uint64_t x) { return (x + 63) & ~uint64_t(63); }

inline int32_t as_int(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}
inline float as_float(int32_t i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

// Adds `delta` to the start of one packed leaf word ((start << 4) | count).
inline bool shift_leaf(float& w, int64_t delta) {
    const int32_t packed = as_int(w);
    const int64_t start = int64_t(uint32_t(packed) >> 4) + delta;
    if (start < 0 || start >= (int64_t(1) << 28)) return false;
    w = as_float(int32_t((uint32_t(start) << 4) | (uint32_t(packed) & 0xFu)));
    return true;
}

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

// Leaf words: stackless min.w of every leaf (-1 marks an inner node); stack
// min.w of each leaf child in the used entries [0, inner count).  The zero
// tail of the stack array is not part of the tree and stays zero.
int rebase_bvh(crt_bvh& b, int32_t t_offset) {
    if (t_offset < 0) return set_error(CRT_E_INVALID, "crt_bvh_rebase: negative t_offset");
    const int64_t delta = int64_t(t_offset) - int64_t(b.t_offset);
    if (uint64_t(b.r.stats.triangle_count) + uint64_t(t_offset) >= (1ull << 28))
        return set_error(CRT_E_INVALID, "crt_bvh_rebase: t_offset + triangles exceeds the 28-bit leaf start");
    if (delta == 0) return CRT_OK;
    // Pass 0 checks every leaf, pass 1 moves them, so a failure changes nothing.
    for (int pass = 0; pass < 2; ++pass) {
        bool ok = true;
        auto visit = [&](float& w) {
            if (as_int(w) == -1) return;
            float copy = w;
            ok &= shift_leaf(pass ? w : copy, delta);
        };
        if (b.r.format == CRT_FORMAT_STACKLESS) {
            for (auto& n : b.r.stackless) visit(n.min[3]);
        } else {
            const size_t used = b.r.stack.empty() ? 0 : (b.r.stack.size() - 1) / 2;
            for (size_t k = 0; k < used; ++k) {
                visit(b.r.stack[k].left.min[3]);
                visit(b.r.stack[k].right.min[3]);
            }
        }
        if (!ok) return set_error(CRT_E_CORRUPT, "crt_bvh_rebase: a leaf start would leave [0, 2^28)");
    }
    b.t_offset = t_offset;
    return CRT_OK;
}

int bvh_key(int format, const crt_mesh* meshes, int mesh_count, const crt_build_options* options, uint64_t* key) {
    if (!key) return set_error(CRT_E_INVALID, "crt_bvh_key: key is null");
    if (format != CRT_FORMAT_STACKLESS && format != CRT_FORMAT_STACK)
        return set_error(CRT_E_INVALID, "crt_bvh_key: unknown format");
    if (mesh_count < 0 || (mesh_count > 0 && !meshes)) return set_error(CRT_E_INVALID, "crt_bvh_key: bad mesh list");
    const crt_build_options opt = options ? *options : crt_build_options{CRT_SWAP_NONE, 0u, 0};
    Hash64 h;
    static const char kTag[] = "crt-bvh-key/1";
    h.update(kTag, sizeof(kTag) - 1);
    h.value(int32_t(format));
    h.value(int32_t(opt.child_swap));
    h.value(uint32_t(opt.child_swap == CRT_SWAP_SEEDED ? opt.seed : 0u));
    h.value(int32_t(mesh_count));
    for (int m = 0; m < mesh_count; ++m) {
        const crt_mesh& mesh = meshes[m];
        if ((mesh.vertex_count && !mesh.vertices) || (mesh.index_count && !mesh.indices))
            return set_error(CRT_E_INVALID, "crt_bvh_key: null mesh data");
        h.value(uint32_t(mesh.vertex_count));
        h.value(uint32_t(mesh.index_count));
        h.value(int32_t(mesh.global_mesh_number));
        h.update(mesh.vertices, size_t(mesh.vertex_count) * sizeof(crt_vertex));
        h.update(mesh.indices, size_t(mesh.index_count) * sizeof(uint32_t));
    }
    *key = h.digest();
    return CRT_OK;
}

int save_bvh(const crt_bvh& b, uint64_t key, const std::string& path) {
    const BvhResult& r = b.r;
    crt_cache_header hd;
    std::memset(&hd, 0, sizeof(hd));
    std::memcpy(hd.magic, CRT_CACHE_MAGIC, sizeof(CRT_CACHE_MAGIC));
    hd.header_bytes = sizeof(crt_cache_header);
    hd.version = CRT_CACHE_VERSION;
    hd.format = r.format;
    hd.t_offset = b.t_offset;
    hd.key = key;
    hd.triangle_count = r.triangles.size();
    hd.array_length = r.stats.array_length;
    hd.vertex_count = r.vertices.size();
    hd.node_count = r.stats.node_count;
    hd.leaf_count = r.stats.leaf_count;
    hd.split_fails = r.stats.split_fails;
    hd.max_depth = r.stats.max_depth;
    hd.child_swap = b.options.child_swap;
    hd.seed = b.options.child_swap == CRT_SWAP_SEEDED ? b.options.seed : 0u;
    hd.bins = uint32_t(kBinCount);
    hd.max_leaf = kMaxLeafTris;
    toolchain_id(hd.toolchain);
    const void* node_data = r.format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(r.stackless.data())
                                                             : static_cast<const void*>(r.stack.data());
    const uint64_t node_bytes = r.format == CRT_FORMAT_STACKLESS ? r.stackless.size() * sizeof(crt_stackless_node)
                                                                 : r.stack.size() * sizeof(crt_stack_node);
    const void* data[3] = {node_data, r.triangles.data(), r.vertices.data()};
    const uint64_t bytes[3] = {node_bytes, r.triangles.size() * sizeof(crt_triangle),
                               r.vertices.size() * sizeof(crt_vertex)};
    uint64_t off = sizeof(crt_cache_header);
    for (int s = 0; s < 3; ++s) {
        off = align64(off);
        hd.section_offset[s] = off;
        hd.section_bytes[s] = bytes[s];
        hd.section_hash[s] = hash_bytes(data[s], size_t(bytes[s]));
        off += bytes[s];
    }
    hd.header_hash = hash_bytes(&hd, offsetof(crt_cache_header, header_hash));

    const std::string tmp = path + ".tmp";
    {
        File out;
        out.f = std::fopen(tmp.c_str(), "wb");
        if (!out.f) return set_error(CRT_E_IO, "crt_bvh_save: cannot open " + tmp + ": " + std::strerror(errno));
        static const uint8_t zeros[64] = {};
        bool ok = std::fwrite(&hd, sizeof(hd), 1, out.f) == 1;
        uint64_t pos = sizeof(hd);
        for (int s = 0; s < 3 && ok; ++s) {
            ok = std::fwrite(zeros, 1, size_t(hd.section_offset[s] - pos), out.f) == size_t(hd.section_offset[s] - pos);
            if (ok && bytes[s]) ok = std::fwrite(data[s], 1, size_t(bytes[s]), out.f) == size_t(bytes[s]);
            pos = hd.section_offset[s] + bytes[s];
        }
        ok = ok && std::fflush(out.f) == 0;
        const int closed = std::fclose(out.f);
        out.f = nullptr;
        if (!ok || closed != 0) {
            std::remove(tmp.c_str());
            return set_error(CRT_E_IO, "crt_bvh_save: write to " + tmp + " failed");
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        const std::string why = std::strerror(errno);
        std::remove(tmp.c_str());
        return set_error(CRT_E_IO, "crt_bvh_save: cannot rename " + tmp + " to " + path + ": " + why);
    }
    return CRT_OK;
}

int load_bvh(const std::string& path, uint64_t expect_key, crt_bvh& b, crt_cache_header* header_out) {
    File in;
    in.f = std::fopen(path.c_str(), "rb");
    if (!in.f) return set_error(CRT_E_IO, "crt_bvh_load: cannot open " + path + ": " + std::strerror(errno));
    if (std::fseek(in.f, 0, SEEK_END) != 0) return set_error(CRT_E_IO, "crt_bvh_load: cannot seek " + path);
    const long size = std::ftell(in.f);
    std::rewind(in.f);
    crt_cache_header hd;
    if (size < long(sizeof(hd)) || std::fread(&hd, sizeof(hd), 1, in.f) != 1)
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + " is shorter than its header");
    if (std::memcmp(hd.magic, CRT_CACHE_MAGIC, sizeof(CRT_CACHE_MAGIC)) != 0)
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + " is not a BVH cache file");
    if (hd.header_bytes != sizeof(hd) || hd.version != CRT_CACHE_VERSION)
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + " has an unknown header version");
    if (hd.header_hash != hash_bytes(&hd, offsetof(crt_cache_header, header_hash)))
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": header hash mismatch");
    if (hd.format != CRT_FORMAT_STACKLESS && hd.format != CRT_FORMAT_STACK)
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": unknown format");
    const uint64_t rec[3] = {hd.format == CRT_FORMAT_STACKLESS ? sizeof(crt_stackless_node) : sizeof(crt_stack_node),
                             sizeof(crt_triangle), sizeof(crt_vertex)};
    const uint64_t count[3] = {hd.array_length, hd.triangle_count, hd.vertex_count};
    uint64_t end = sizeof(hd);
    for (int s = 0; s < 3; ++s) {
        if (count[s] > (uint64_t(1) << 40) || hd.section_bytes[s] != count[s] * rec[s] ||
            hd.section_offset[s] < end || hd.section_offset[s] % 64 != 0)
            return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": inconsistent section table");
        end = hd.section_offset[s] + hd.section_bytes[s];
    }
    if (uint64_t(size) != end)
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": file size " + std::to_string(size) +
                                            " does not match its section table (" + std::to_string(end) + ")");
    if (hd.triangle_count >= (1ull << 28) || hd.t_offset < 0 ||
        hd.triangle_count + uint64_t(uint32_t(hd.t_offset)) >= (1ull << 28))
        return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": counts exceed the 28-bit leaf start");
    if (expect_key != 0 && hd.key != expect_key)
        return set_error(CRT_E_STALE, "crt_bvh_load: " + path + " was built from other inputs");

    BvhResult& r = b.r;
    r.format = hd.format;
    try {
        if (hd.format == CRT_FORMAT_STACKLESS)
            r.stackless.resize(size_t(hd.array_length));
        else
            r.stack.resize(size_t(hd.array_length));
        r.triangles.resize(size_t(hd.triangle_count));
        r.vertices.resize(size_t(hd.vertex_count));
    } catch (const std::bad_alloc&) {
        return set_error(CRT_E_NOMEM, "crt_bvh_load: out of host memory");
    }
    void* dst[3] = {hd.format == CRT_FORMAT_STACKLESS ? static_cast<void*>(r.stackless.data())
                                                      : static_cast<void*>(r.stack.data()),
                    r.triangles.data(), r.vertices.data()};
    for (int s = 0; s < 3; ++s) {
        if (!hd.section_bytes[s]) {
            if (hd.section_hash[s] != hash_bytes(nullptr, 0))
                return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": section hash mismatch");
            continue;
        }
        if (std::fseek(in.f, long(hd.section_offset[s]), SEEK_SET) != 0 ||
            std::fread(dst[s], 1, size_t(hd.section_bytes[s]), in.f) != size_t(hd.section_bytes[s]))
            return set_error(CRT_E_IO, "crt_bvh_load: cannot read " + path);
        if (hash_bytes(dst[s], size_t(hd.section_bytes[s])) != hd.section_hash[s])
            return set_error(CRT_E_CORRUPT, "crt_bvh_load: " + path + ": section " + std::to_string(s) +
                                                " hash mismatch");
    }
    r.stats.triangle_count = hd.triangle_count;
    r.stats.node_count = hd.node_count;
    r.stats.leaf_count = hd.leaf_count;
    r.stats.split_fails = hd.split_fails;
    r.stats.max_depth = hd.max_depth;
    r.stats.array_length = hd.array_length;
    r.stats.vertex_count = hd.vertex_count;
    b.t_offset = hd.t_offset;
    b.options = crt_build_options{hd.child_swap, hd.seed, 0};
    if (header_out) *header_out = hd;
    return CRT_OK;
}

bool same_toolchain(const crt_cache_header& hd) {
    char here[32];
    toolchain_id(here);
    return std::memcmp(here, hd.toolchain, 32) == 0;
}

}  // namespace crt

extern "C" {

uint64_t crt_hash64(const void* data, size_t bytes, uint64_t seed) {
    crt::Hash64 h(seed);
    if (data && bytes) h.update(data, bytes);
    return h.digest();
}

int crt_bvh_key(int format, const crt_mesh* meshes, int mesh_count, const crt_build_options* options,
                uint64_t* key) {
    return crt::bvh_key(format, meshes, mesh_count, options, key);
}

int crt_bvh_save(const crt_bvh* bvh, uint64_t key, const char* path) {
    if (!bvh || !path) return crt::set_error(CRT_E_INVALID, "crt_bvh_save: null argument");
    return crt::save_bvh(*bvh, key, path);
}

int crt_bvh_load(const char* path, uint64_t expect_key, crt_bvh** out, crt_cache_header* header) {
    if (!path || !out) return crt::set_error(CRT_E_INVALID, "crt_bvh_load: null argument");
    *out = nullptr;
    crt_bvh* b = new (std::nothrow) crt_bvh();
    if (!b) return crt::set_error(CRT_E_NOMEM, "crt_bvh_load: out of host memory");
    const int rc = crt::load_bvh(path, expect_key, *b, header);
    if (rc != CRT_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return CRT_OK;
}

int crt_bvh_rebase(crt_bvh* bvh, int32_t t_offset) {
    if (!bvh) return crt::set_error(CRT_E_INVALID, "crt_bvh_rebase: bvh is null");
    return crt::rebase_bvh(*bvh, t_offset);
}

}  // extern "C"
