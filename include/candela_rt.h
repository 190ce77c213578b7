/*
 * candela_rt.h -- C ABI of the MI355X ray-traversal engine for Candela.
 *
 * Everything that crosses this boundary is a plain C type: fixed-layout structs,
 * pointers and sizes.  Device pointers are plain `void*` / typed pointers that were
 * obtained from crt_device_alloc() (or any HIP allocation of the same process).
 * Streams are passed as `void*` holding a hipStream_t (NULL = the null stream).
 *
 * Every function returns CRT_OK (0) or a negative CRT_E_* code; crt_last_error()
 * gives the message of the last failure on the calling thread.  No C++ exception
 * crosses the ABI.
 *
 * Reference interfaces replaced (paths relative to the Candela source tree):
 *   buffer layouts ......... Source/Core/Utils/Vertex.h:7-12,
 *                            Source/Core/BVH/BVHConstructor.h:61-84,
 *                            Source/Core/BVH/Intersector.h:43-49
 *   CPU constructor ........ Source/Core/BVH/BVHConstructor.cpp:951-1108 (BuildBVH x2)
 *   host Ray Intersection API Source/Core/BVH/Intersector.h:60-124 (RayIntersector<T>)
 *   device Ray Intersection API
 *                            Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl:175-579
 *                            Source/Core/Shaders/Intersectors/Include/TraverseBVHStack.glsl:168-681
 */
#ifndef CANDELA_RT_H
#define CANDELA_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRT_ABI_VERSION 2

/* ------------------------------------------------------------------------- */
/* Status codes                                                              */
/* ------------------------------------------------------------------------- */
enum {
    CRT_OK = 0,
    CRT_E_INVALID = -1,   /* bad argument (null pointer, out-of-range enum, size) */
    CRT_E_HIP = -2,       /* a HIP runtime call failed (message names it)         */
    CRT_E_NOMEM = -3,     /* host or device allocation failed                     */
    CRT_E_STATE = -4,     /* call out of order (e.g. trace before buffer_data)    */
    CRT_E_NODEVICE = -5,  /* no HIP device visible                                */
    CRT_E_IO = -6,        /* a file could not be opened, read or written          */
    CRT_E_CORRUPT = -7,   /* a cache file failed its magic, size or hash checks   */
    CRT_E_STALE = -8      /* a cache file was built from other inputs             */
};

/* ------------------------------------------------------------------------- */
/* Buffer layouts -- byte-identical to the reference's std430 SSBO contents    */
/* ------------------------------------------------------------------------- */

/* Vertex.h:7-12 -- vec4 position; uvec3 normal_tangent_data; uint texcoords (32 B).
 * Only position.xyz is read by traversal. */
typedef struct crt_vertex {
    float position[4];
    uint32_t normal_tangent_data[3];
    uint32_t texcoords;
} crt_vertex;

/* BVHConstructor.h:79-84 -- {vertex index 0,1,2 ; mesh id} (16 B).  The
 * triangle's position in the flattened array is its "triangle id". */
typedef struct crt_triangle {
    int32_t packed[4];
} crt_triangle;

/* BVHConstructor.h:66-70 -- stackless node (32 B), pre-order.
 *   min.w : int bits; -1 for an inner node, (start<<4 | count) for a leaf
 *   max.w : int bits; object-local index of the escape ("miss") node, -1 = end */
typedef struct crt_stackless_node {
    float min[4];
    float max[4];
} crt_stackless_node;

/* BVHConstructor.h:61-64 */
typedef struct crt_bounds {
    float min[4];
    float max[4];
} crt_bounds;

/* BVHConstructor.h:72-77 -- stack node (64 B), BFS order.  Entry k holds both
 * children of inner node k:
 *   X.min.w : int bits; -1 if child X is inner, (start<<4 | count) if a leaf
 *   X.max.w : int bits; object-local entry index of child X (inner children) */
typedef struct crt_stack_node {
    crt_bounds left;
    crt_bounds right;
} crt_stack_node;

/* Intersector.h:43-49 -- per-instance record (192 B).  Matrices are column-major
 * (glm / GLSL).  data[0] = float bits of emissive amount, data[1] = float bits
 * of alpha = 1 - translucency. */
typedef struct crt_entity {
    float model[16];
    float inverse[16];
    int32_t node_offset;
    int32_t node_count;
    int32_t data[14];
} crt_entity;

/* Ray record (32 B).  tmax is the TMax the reference's scene loop starts from
 * (1e6 in the reference, TraverseBVHStackless.glsl:284).  `id` is not read by
 * traversal; callers may keep a pixel index there. */
typedef struct crt_ray {
    float org[3];
    float tmax;
    float dir[3];
    uint32_t id;
} crt_ray;

/* Hit record (32 B): the reference's TUVW + Mesh + TriangleIdx + Entity outs
 * (TraverseBVHStackless.glsl:280-319).
 *   closest-hit : t,u,v,w = TUVW (u<->vertex A, v<->B, w<->C); all -1 on a miss
 *   any-hit     : t = first accepted distance or -1; u=v=w=-1
 *   status      : bits 0-23 node visits summed over entities (saturating),
 *                 bits 24-31 CRT_STATUS_* flags */
typedef struct crt_hit {
    float t, u, v, w;
    int32_t tri;
    int32_t mesh;
    int32_t entity;
    uint32_t status;
} crt_hit;

/* STACKLESS and STACK are the reference's two layouts (FlattenedNode,
 * FlattenedStackNode).  WIDE4 and WIDE8 are layouts derived from either of
 * them by crt_wide_build (no reference counterpart; the reference only has
 * BVH2): 4- or 8-wide nodes and triangle records in the same leaf order. */
enum { CRT_FORMAT_STACKLESS = 0, CRT_FORMAT_STACK = 1, CRT_FORMAT_WIDE4 = 2, CRT_FORMAT_WIDE8 = 3 };
enum { CRT_QUERY_CLOSEST = 0, CRT_QUERY_ANY = 1 };

/* Trace flags */
enum {
    /* IntersectSceneIgnoreTransparent: skip entities whose alpha < 0.99
     * (TraverseBVHStackless.glsl:321-349) */
    CRT_FLAG_IGNORE_TRANSPARENT = 1u << 0,
    /* Reproduce the scene-level test `TriangleIdx > 0`
     * (TraverseBVHStackless.glsl:304): a closest hit on flattened triangle 0 is
     * reported with t=u=v=w=-1 (tri/mesh/entity still set). */
    CRT_FLAG_TRI0_IS_MISS = 1u << 1,
    /* Child ordering of any-hit walks of the wide layouts (no reference
     * counterpart): the hit children of a node are taken in slot order
     * instead of nearest first, which saves the sort.  Closest-hit walks, the
     * BVH2 formats and the speculative kernels (always nearest first) ignore
     * it. */
    CRT_FLAG_SLOT_ORDER = 1u << 2
};

/* Per-ray status flags (hit.status bits 24-31) */
enum {
    CRT_STATUS_ITER_CAP = 1u << 24,       /* an entity walk stopped at 1024 iterations  */
    CRT_STATUS_STACK_OVERFLOW = 1u << 25, /* stack walk aborted at 63 pushed entries    */
    CRT_STATUS_BAD_INDEX = 1u << 26       /* node index left the entity's range         */
};
#define CRT_STATUS_VISITS_MASK 0x00FFFFFFu

/* ------------------------------------------------------------------------- */
/* Library                                                                   */
/* ------------------------------------------------------------------------- */
const char* crt_last_error(void);
int crt_abi_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int crt_device_count(void);

/* ------------------------------------------------------------------------- */
/* CPU SAH constructor -- BVHConstructor.cpp:951-1108                        */
/* ------------------------------------------------------------------------- */

/* One mesh of an Object (Object.h:62, Mesh.h:30-31,58). */
typedef struct crt_mesh {
    const crt_vertex* vertices;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t index_count; /* multiple of 3 */
    int32_t global_mesh_number;
} crt_mesh;

/* Child order of the stackless build.  The reference swaps the two children of
 * every split with p = 1/2 from mt19937(random_device())
 * (BVHConstructor.cpp:389-391,599-609).  SEEDED draws the same sequence from
 * mt19937(seed); NONE keeps the split order. */
enum { CRT_SWAP_NONE = 0, CRT_SWAP_SEEDED = 1 };

typedef struct crt_build_options {
    int32_t child_swap;  /* CRT_SWAP_* (stackless format only) */
    uint32_t seed;       /* mt19937 seed for CRT_SWAP_SEEDED   */
    int32_t threads;     /* worker threads; 0 = hardware concurrency */
} crt_build_options;

typedef struct crt_bvh_stats {
    uint64_t triangle_count;
    uint64_t node_count;     /* LastNodeIndex: tree nodes excluding the root */
    uint64_t leaf_count;
    uint64_t split_fails;
    uint32_t max_depth;      /* max size of the build stack (MaxBVHDepth) */
    uint64_t array_length;   /* flattened node array length = node_count + 1 */
    uint64_t vertex_count;
} crt_bvh_stats;

typedef struct crt_bvh crt_bvh;

/* BuildBVH(object, nodes, vertices, tris, t_offset): concatenates the meshes,
 * builds the binned-SAH tree and flattens it into `format`.  Fails (rather than
 * crashing) where the reference would dereference a leaf root in the stack
 * flattener (BVHConstructor.cpp:873): a stack-format object needs >= 3
 * triangles. */
int crt_build_bvh(int format, const crt_mesh* meshes, int mesh_count, int32_t t_offset,
                  const crt_build_options* options, crt_bvh** out);
void crt_bvh_free(crt_bvh* bvh);
int crt_bvh_get_stats(const crt_bvh* bvh, crt_bvh_stats* out);
/* Node array: crt_stackless_node[array_length] or crt_stack_node[array_length]. */
const void* crt_bvh_nodes(const crt_bvh* bvh);
const crt_triangle* crt_bvh_triangles(const crt_bvh* bvh);
const crt_vertex* crt_bvh_vertices(const crt_bvh* bvh);
int crt_bvh_format(const crt_bvh* bvh);
/* The t_offset the buffers were built (or rebased) with. */
int32_t crt_bvh_t_offset(const crt_bvh* bvh);

/* ------------------------------------------------------------------------- */
/* BVH cache files.  No reference counterpart: AddObject rebuilds every      */
/* object at every launch (Intersector.h:169-198 -> BuildBVH,                */
/* BVHConstructor.cpp:951-1108).  A cache file holds one object's flattened  */
/* buffers, byte for byte as BuildBVH leaves them, behind this header.       */
/* ------------------------------------------------------------------------- */

/* File layout (little-endian): the 256-byte header, then the node, triangle
 * and vertex sections, each at a 64-byte aligned offset.  Every hash is XXH64
 * (seed 0); header_hash covers bytes [0, 248) of the header. */
#define CRT_CACHE_MAGIC "CRTBVH1"
#define CRT_CACHE_VERSION 1
typedef struct crt_cache_header {
    char magic[8];               /* CRT_CACHE_MAGIC, NUL-terminated        */
    uint32_t header_bytes;       /* 256                                    */
    uint32_t version;            /* CRT_CACHE_VERSION                      */
    int32_t format;              /* CRT_FORMAT_STACKLESS or _STACK         */
    int32_t t_offset;            /* leaf starts are relative to this       */
    uint64_t key;                /* crt_bvh_key of the build's inputs      */
    uint64_t triangle_count;     /* N                                      */
    uint64_t array_length;       /* node records (2L - 1)                  */
    uint64_t vertex_count;       /* V                                      */
    uint64_t node_count;         /* crt_bvh_stats.node_count               */
    uint64_t leaf_count;         /* L                                      */
    uint64_t split_fails;
    uint32_t max_depth;
    int32_t child_swap;          /* crt_build_options.child_swap           */
    uint32_t seed;               /* crt_build_options.seed                 */
    uint32_t bins;               /* binned-SAH bin count (64)              */
    uint32_t max_leaf;           /* triangles per leaf (2)                 */
    uint32_t reserved0;
    uint64_t section_offset[3];  /* nodes, triangles, vertices             */
    uint64_t section_bytes[3];
    uint64_t section_hash[3];
    char toolchain[32];          /* C++ library of the build ("libstdc++ 20220324"):
                                    it fixes mt19937's uniform_int draws for
                                    CRT_SWAP_SEEDED                         */
    uint8_t reserved1[40];
    uint64_t header_hash;
} crt_cache_header;

/* XXH64 of `bytes` bytes (Y. Collet's published xxHash64). */
uint64_t crt_hash64(const void* data, size_t bytes, uint64_t seed);
/* The cache key of a build: XXH64 over "crt-bvh-key/1", then format,
 * child_swap, seed (0 unless CRT_SWAP_SEEDED), mesh_count as int32, then per
 * mesh its vertex_count, index_count and global_mesh_number as 32-bit ints,
 * its vertex records and its indices.  t_offset and threads do not enter it:
 * they do not change the tree. */
int crt_bvh_key(int format, const crt_mesh* meshes, int mesh_count, const crt_build_options* options,
                uint64_t* key);
/* Writes `bvh` to `path` (a temporary file renamed over it, so readers never
 * see a partial file).  `key` is stored in the header. */
int crt_bvh_save(const crt_bvh* bvh, uint64_t key, const char* path);
/* Reads a cache file.  CRT_E_IO if it cannot be read, CRT_E_CORRUPT if a
 * check fails, CRT_E_STALE if expect_key != 0 and differs from the stored
 * key.  `header` (may be NULL) receives the header. */
int crt_bvh_load(const char* path, uint64_t expect_key, crt_bvh** out, crt_cache_header* header);
/* Moves every leaf start from the bvh's t_offset to `t_offset`; the result is
 * byte-identical to a build with that t_offset. */
int crt_bvh_rebase(crt_bvh* bvh, int32_t t_offset);

/* ------------------------------------------------------------------------- */
/* Device BVH builder.  No reference counterpart (BuildBVH runs on the CPU,   */
/* BVHConstructor.cpp:951-1108): a linear BVH over 30-bit Morton codes of the */
/* triangle centroids (Karras 2012), for scenes rebuilt every frame.  The     */
/* tree is NOT BuildBVH's, so the buffers are not byte-identical to it; every */
/* walk over them still reports the nearest hit.                              */
/* ------------------------------------------------------------------------- */

/* Bytes of device workspace crt_build_bvh_gpu needs for ntris triangles. */
size_t crt_build_bvh_gpu_workspace(int64_t ntris);
/* Builds the stack layout of one object on the device, asynchronously on
 * `stream`: nodes [ntris - 1] (entry i holds both children of inner node i;
 * a child covering two triangles is written as a 2-triangle leaf) and
 * triangles [ntris] ({i0, i1, i2, mesh} in leaf order; leaf starts count from
 * t_offset).  vertices / indices (3 per triangle, indices into vertices) are
 * device arrays; mesh_ids (device, one per triangle) or, when NULL, mesh_id
 * for every triangle.  Every index must be < nverts (device data is not
 * checked).  workspace: 256-byte aligned device memory of
 * crt_build_bvh_gpu_workspace(ntris) bytes.  Needs ntris >= 2. */
int crt_build_bvh_gpu(const crt_vertex* vertices, int64_t nverts, const uint32_t* indices, int64_t ntris,
                      const int32_t* mesh_ids, int32_t mesh_id, int32_t t_offset, crt_stack_node* nodes,
                      crt_triangle* triangles, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------- */
/* Device memory (thin wrappers so callers need no HIP headers)              */
/* ------------------------------------------------------------------------- */
int crt_device_alloc(void** ptr, size_t bytes);
int crt_device_free(void* ptr);
int crt_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int crt_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int crt_memset_d(void* dst, int value, size_t bytes, void* stream);
/* Device-to-device copy by a streaming kernel (16-byte loads and stores, one
 * pass): the copy whose counters calibrate the HBM-traffic measurements.
 * `bytes` must be a multiple of 16 and both pointers 16-byte aligned. */
int crt_device_copy(void* dst, const void* src, size_t bytes, void* stream);
int crt_stream_create(void** stream);
int crt_stream_destroy(void* stream);
int crt_stream_sync(void* stream);
int crt_device_sync(void);
/* Stream capture into a replayable graph, for callers without HIP headers
 * (no reference counterpart).  crt_graph_begin starts capturing `stream`
 * (thread-local mode): the asynchronous calls made on it afterwards are
 * recorded, not run.  Calls that synchronize or read back (crt_wide_build_gpu,
 * crt_context_* uploads) fail while capturing.  crt_graph_end ends the
 * capture and instantiates it into *graph; crt_graph_launch replays it on a
 * stream.  A replay reads the buffers' contents at that time. */
int crt_graph_begin(void* stream);
int crt_graph_end(void* stream, void** graph);
int crt_graph_launch(void* graph, void* stream);
int crt_graph_destroy(void* graph);
/* HIP events for timing work on a stream. */
int crt_event_create(void** event);
int crt_event_destroy(void* event);
int crt_event_record(void* event, void* stream);
/* Work enqueued on `stream` after this call waits for `event` (recorded on
 * another stream): independent ray batches can then run on two streams and
 * fill each other's launch tails (no reference counterpart). */
int crt_stream_wait_event(void* stream, void* event);
/* Waits for `end` and returns the milliseconds between the two events. */
int crt_event_elapsed_ms(void* start, void* end, float* ms);

/* ------------------------------------------------------------------------- */
/* Raw device scene + trace (SSBO bindings 16..19, Intersector.h:275-281)    */
/* ------------------------------------------------------------------------- */

/* Device-resident scene: the four buffers the traversal include binds plus
 * u_EntityCount / u_TotalNodes.  For the wide formats `nodes` holds
 * crt_wide_group[total_nodes * width / 4], `tri_records` the triangle records
 * and `entities` the entity copy whose node_offset / node_count name the
 * object's wide root and wide node count; `triangles` and `vertices` are not
 * read. */
typedef struct crt_scene_view {
    int32_t format;              /* CRT_FORMAT_*                              */
    int32_t entity_count;        /* u_EntityCount                             */
    int64_t total_nodes;         /* u_TotalNodes (entries of `nodes`)         */
    const void* nodes;           /* crt_stackless_node[] or crt_stack_node[]  */
    const crt_triangle* triangles;
    const crt_vertex* vertices;
    const crt_entity* entities;
    const struct crt_tri_record* tri_records;  /* wide formats only, else NULL */
} crt_scene_view;

/* Kernel choice and launch shape. */
enum {
    CRT_KERNEL_AUTO = 0,        /* the wave kernel for stackless and wide formats,
                                   thread-per-ray for the stack format              */
    CRT_KERNEL_THREAD_PER_RAY = 1,  /* one lane per ray, grid = ceil(n / block)     */
    CRT_KERNEL_PERSISTENT = 2,  /* device-filling waves claiming rays from a counter:
                                   stackless, 64 at a time (leaf batching + node
                                   window); wide formats, one per idle lane whenever
                                   enough lanes are idle                            */
    CRT_KERNEL_WAVE = 3,        /* one ray per lane, leaf batching, XCD-aware block
                                   order (stackless: + node window; wide formats:
                                   leaves of a node parked and tested by the wave) */
    CRT_KERNEL_SPECULATIVE = 4, /* every format: one ray per lane, lanes never
                                   wait for leaf tests: leaves go to a per-lane
                                   queue, and after a round of node steps the wave
                                   tests the queued leaves 64 at a time, with
                                   culling by the tmax of the last such phase.
                                   Grid launch, 64 consecutive rays per wave.
                                   Records are deterministic, counters included;
                                   nearest t and hit / miss equal the other walks'
                                   of the format, the triangle up to exact
                                   distance ties (stackless format: where boxes
                                   touch, a hit the reference walk's culling skips
                                   can be found, so t is then nearer); visit and
                                   test counts follow the preset's schedule
                                   (variant 0..5 over the wide formats, 0..3 over
                                   the stack and stackless formats) */
    CRT_KERNEL_SPECULATIVE_REFILL = 5 /* the same walk in device-filling waves that
                                   refill idle lanes from a work queue (uses
                                   config->queue like CRT_KERNEL_PERSISTENT); a
                                   ray's record does not depend on which rays
                                   share its wave, so records are deterministic
                                   too (variants as kernel 4, with their own
                                   schedules) */
};

typedef struct crt_trace_config {
    int32_t block_size;    /* thread-per-ray: threads per workgroup, 64..1024, multiple
                              of 64; 0 = 256 (persistent kernels use 256) */
    int32_t lds_stack;     /* stack format: entries kept in LDS, one of 8, 16, 32, 48,
                              64 (0 = 16); deeper entries live in per-lane scratch */
    int32_t kernel;        /* CRT_KERNEL_* */
    int32_t variant;       /* persistent / wave preset: 0..5 over the BVH2 (inner steps, leaf
                              batch, window); 0..9 over the wide layouts (stack depth, inner
                              steps, leaf batch, refill threshold; 6..9 trade register spills
                              for occupancy) */
    int32_t grid;          /* persistent: workgroups; 0 = enough to fill every CU */
    int32_t reserved;
    uint32_t* tri_tests;   /* optional device array [nrays]: the counting build of the
                              kernel writes the triangle tests of each ray (node
                              visits are in hit.status); NULL = no counting */
    void* queue;           /* persistent kernels and crt_trace_paths: optional device
                              memory of CRT_QUEUE_BYTES, 256-byte aligned, for the
                              launch's work-queue counters (the launch zeroes it on
                              its stream; launches that share one must not overlap).
                              NULL = a block leased from the library's per-device
                              pool, reused only after the launch that last held it
                              has completed.  Required under stream capture: a
                              graph owns its counters, and NULL there is
                              CRT_E_INVALID. */
} crt_trace_config;

/* Bytes of one work-queue block (crt_trace_config.queue). */
#define CRT_QUEUE_BYTES 2048

/* IntersectScene / IntersectSceneIgnoreTransparent (closest) or the occlusion
 * IntersectScene (any) for `nrays` rays; all pointers are device pointers.
 * Asynchronous on `stream`. */
int crt_trace(const crt_scene_view* scene, int query, uint32_t flags,
              const crt_ray* rays, crt_hit* hits, int64_t nrays,
              const crt_trace_config* config, void* stream);

/* crt_trace that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid allocation after start-up).
 * The persistent kernels' work-queue counters come from one static
 * CRT_QUEUE_BYTES block per GPU, held from module load to process end,
 * instead of a pool block or config->queue (which must be NULL).  The kernels
 * zero the counters on the call's stream, so nothing needs clearing between
 * calls; calls whose persistent launches overlap on one GPU share that block
 * and race.  At most CRT_NOALLOC_MAX_RAYS rays per call (CRT_E_INVALID
 * beyond).  Otherwise identical to crt_trace: same kernels, same records. */
#define CRT_NOALLOC_MAX_RAYS 2147483647LL
int crt_trace_noalloc(const crt_scene_view* scene, int query, uint32_t flags,
                      const crt_ray* rays, crt_hit* hits, int64_t nrays,
                      const crt_trace_config* config, void* stream);

/* ------------------------------------------------------------------------- */
/* Wide BVH (BVH4 / BVH8) derived from the reference's BVH2                   */
/* ------------------------------------------------------------------------- */

/* Four children of a wide node (128 B), structure of arrays so each field of
 * the four children is one 16-byte load.  A BVH4 node is one group, a BVH8
 * node two consecutive groups.  count[k]: 0 = inner child (child[k] = wide
 * node index, absolute), n > 0 = leaf of n triangles starting at record
 * child[k], -1 = empty slot.  The child boxes are the BVH2 node boxes,
 * unchanged.  Leaves hold at most 15 triangles (crt_wide_build's max_leaf).
 * The walks refuse inner children outside the entity's node range
 * (CRT_STATUS_BAD_INDEX) and treat a negative inner child as an empty slot;
 * the sequential kernels read leaf records unchecked, the speculative ones
 * refuse a leaf whose start lies outside 0 .. 2^26 - n (CRT_STATUS_BAD_INDEX,
 * its records unread). */
typedef struct crt_wide_group {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    int32_t child[4];
    int32_t count[4];
} crt_wide_group;

/* One triangle in leaf order (48 B): the three vertex positions, the
 * triangle's own index (= its index in the reference's Triangle array) and
 * its mesh id, so a test needs no index or vertex fetch. */
typedef struct crt_tri_record {
    float v0[3];
    int32_t tri;
    float v1[3];
    int32_t mesh;
    float v2[3];
    int32_t pad;
} crt_tri_record;

typedef struct crt_wide crt_wide;

typedef struct crt_wide_stats {
    int64_t node_count;      /* wide nodes (each width / 4 groups)              */
    int64_t leaf_count;      /* leaf slots                                      */
    int64_t record_count;    /* triangle records (= triangles)                  */
    int32_t width;           /* 4 or 8                                          */
    int32_t max_depth;       /* deepest wide node (root = 1)                    */
    double  mean_fill;       /* mean used slots per wide node                   */
} crt_wide_stats;

/* Collapses the BVH2 of every distinct object the entities reference
 * (src_format STACKLESS or STACK buffers, object-local node links, global
 * triangle starts as AddObject concatenates them) into `width`-wide nodes:
 * each wide node repeatedly opens the child BVH2 node of largest surface
 * area until it holds `width` children.  A subtree of at most `max_leaf`
 * triangles (2..15) whose triangles are contiguous becomes one leaf (2 keeps
 * the BVH2's own leaves).  Validates every link, leaf range and vertex index
 * of the input. */
int crt_wide_build(int src_format, const void* nodes, int64_t node_count, const crt_triangle* triangles,
                   int64_t triangle_count, const crt_vertex* vertices, int64_t vertex_count,
                   const crt_entity* entities, int32_t entity_count, int32_t width, int32_t max_leaf,
                   crt_wide** out);

/* Collapse policies of crt_wide_build_ex. */
#define CRT_COLLAPSE_GREEDY 0   /* crt_wide_build's: open the child of largest area */
#define CRT_COLLAPSE_SAH 1      /* least surface-area cost over all collapses */

typedef struct crt_wide_options {
    int32_t width;      /* 4 or 8 */
    int32_t max_leaf;   /* 2..15 */
    int32_t collapse;   /* CRT_COLLAPSE_GREEDY or CRT_COLLAPSE_SAH */
    float node_cost;    /* SAH: cost of visiting one wide node (0 = 1.0) */
    float tri_cost;     /* SAH: cost of testing one triangle (0 = 0.5) */
} crt_wide_options;

/* crt_wide_build with a choice of collapse (no reference counterpart).
 * CRT_COLLAPSE_SAH: dynamic programming over the BVH2, bottom-up, finds the
 * collapse of least expected cost, where a wide node costs
 * half_area * node_cost and a leaf half_area * tri_cost * triangles.  A
 * subtree may still become a leaf only where the greedy collapse allows it
 * (a BVH2 leaf, or at most max_leaf contiguous triangles), but the cost
 * decides whether it does; each wide node's slots are the cheapest cover of
 * its subtree by `width` subtrees.  Child boxes, leaf ranges and records are
 * those of crt_wide_build, so the walks report the same nearest hits. */
int crt_wide_build_ex(int src_format, const void* nodes, int64_t node_count, const crt_triangle* triangles,
                      int64_t triangle_count, const crt_vertex* vertices, int64_t vertex_count,
                      const crt_entity* entities, int32_t entity_count, const crt_wide_options* options,
                      crt_wide** out);
void crt_wide_free(crt_wide* wide);
int crt_wide_get_stats(const crt_wide* wide, crt_wide_stats* out);
const crt_wide_group* crt_wide_groups(const crt_wide* wide);       /* node_count * width / 4 */
const crt_tri_record* crt_wide_records(const crt_wide* wide);      /* record_count */
const crt_entity* crt_wide_entities(const crt_wide* wide);         /* entity_count */

/* Bytes of device workspace crt_wide_build_gpu needs (ntris >= 2, width 4 or 8). */
size_t crt_wide_build_gpu_workspace(int64_t ntris, int32_t width);
/* crt_wide_build of one object on the device (no reference counterpart; for
 * trees built every frame by crt_build_bvh_gpu): the object's stack-layout
 * entries nodes [node_count] (links local to them, entry 0 the root),
 * triangles [ntris] and vertices [nverts], all device arrays, give the wide
 * layout with 2-triangle leaves.  Every wide node has crt_wide_build's slots,
 * boxes and slot order (max_leaf 2); nodes are numbered breadth-first rather
 * than depth-first, so walks of either layout give identical hit records and
 * counters.  groups: device room for ntris * width / 4 groups; records
 * [ntris]; *wide_nodes (host) receives the node count (the entity's
 * node_count, root 0).  The tree must be built with t_offset 0: every leaf range
 * must lie in [0, ntris), else CRT_E_INVALID.  Runs crt_wide_build_gpu_async with 32 levels (again
 * with CRT_WIDE_GPU_MAX_LEVELS for a deeper tree) and reads the result back,
 * so it synchronizes `stream`; links and vertex indices are checked on the
 * device (CRT_E_INVALID).  workspace: 256-byte aligned,
 * crt_wide_build_gpu_workspace bytes. */
int crt_wide_build_gpu(const crt_stack_node* nodes, int64_t node_count, const crt_triangle* triangles, int64_t ntris,
                       const crt_vertex* vertices, int64_t nverts, int32_t width, crt_wide_group* groups,
                       crt_tri_record* records, int64_t* wide_nodes, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Status bits of crt_wide_build_gpu_async (0: the layout is complete). */
#define CRT_WIDE_GPU_BAD_LINK 1u     /* a stack link leaves the node array */
#define CRT_WIDE_GPU_NOT_TREE 2u     /* more wide nodes than triangles */
#define CRT_WIDE_GPU_BAD_VERTEX 4u   /* a vertex index leaves the vertex array */
#define CRT_WIDE_GPU_TOO_DEEP 8u     /* the wide tree has more than max_levels levels */
#define CRT_WIDE_GPU_BAD_LEAF 16u    /* a leaf range is empty or leaves [0, ntris): the tree
                                        was built with a triangle offset, or is damaged */
#define CRT_WIDE_GPU_MAX_LEVELS 4096

/* Bytes of device workspace crt_wide_build_gpu_async needs (0 for invalid
 * arguments). */
size_t crt_wide_build_gpu_async_workspace(int64_t ntris, int32_t width, int32_t max_levels);
/* crt_wide_build_gpu without any host readback (no reference counterpart; for
 * a per-frame rebuild captured into a graph with crt_build_bvh_gpu and the
 * traces).  It launches max_levels tree levels (1 to CRT_WIDE_GPU_MAX_LEVELS);
 * a level past the tree's depth returns at once.  The output is
 * crt_wide_build_gpu's, byte for byte.  wide_nodes (int64 [2]) and status
 * (uint32) are DEVICE pointers written at the end: the node count and the
 * number of levels the tree needs (max_levels at least that gives status 0;
 * a caller that rebuilds every frame can size the next frame's launch from
 * it), and the CRT_WIDE_GPU_* bits.  Nothing is read back, so errors in the data show only
 * in *status; argument errors return CRT_E_INVALID.  When any status bit is set the
 * root is written with every slot empty, so a trace launched after the collapse (in
 * the same graph, before the status is read) ends at the root with no hit and reads
 * nothing else of the incomplete layout.  A trace can run before
 * the node count is known: the scene's node count may be ntris (the capacity
 * of `groups`), which bounds every valid link.  workspace: 256-byte aligned,
 * crt_wide_build_gpu_async_workspace bytes. */
int crt_wide_build_gpu_async(const crt_stack_node* nodes, int64_t node_count, const crt_triangle* triangles,
                             int64_t ntris, const crt_vertex* vertices, int64_t nverts, int32_t width,
                             int32_t max_levels, crt_wide_group* groups, crt_tri_record* records,
                             int64_t* wide_nodes, uint32_t* status, void* workspace, size_t workspace_bytes,
                             void* stream);

/* ------------------------------------------------------------------------- */
/* RayIntersector<T> -- Intersector.h:60-351                                  */
/* ------------------------------------------------------------------------- */
typedef struct crt_context crt_context;

/* RayIntersector<T>() + Initialize(); format picks T. */
int crt_create(int format, crt_context** out);
void crt_destroy(crt_context* ctx);
int crt_context_format(const crt_context* ctx);

/* AddObject (Intersector.h:169-198): builds the object's BVH on the CPU with
 * t_offset = triangles already added, appends nodes (object-local links),
 * triangles (vertex indices + running vertex offset) and vertices. */
int crt_add_object(crt_context* ctx, int32_t object_id, const crt_mesh* meshes, int mesh_count,
                   const crt_build_options* options);
/* AddObject for buffers that were already built (e.g. by the reference
 * constructor): same offset rules, no build. */
int crt_add_object_prebuilt(crt_context* ctx, int32_t object_id, const void* nodes,
                            int64_t node_count, const crt_triangle* triangles,
                            int64_t triangle_count, const crt_vertex* vertices,
                            int64_t vertex_count);
/* crt_add_object through a cache directory: loads `<cache_dir>/<key>.crtbvh`
 * (key as 16 hex digits) when it holds this object's build, else builds with
 * t_offset 0 and stores the file.  Either way the appended buffers are
 * byte-identical to crt_add_object's.  *cache_hit (may be NULL): 1 loaded,
 * 0 built and stored, -1 built but the file could not be stored (the object
 * is still added; crt_last_error() says why).  A corrupt or stale file, or a
 * CRT_SWAP_SEEDED file written under another C++ library, is rebuilt and
 * overwritten. */
int crt_add_object_cached(crt_context* ctx, int32_t object_id, const crt_mesh* meshes, int mesh_count,
                          const crt_build_options* options, const char* cache_dir, int32_t* cache_hit);
/* PushEntity (Intersector.h:200-216): model matrix (column-major), emissive,
 * translucency; the record is crt_make_entity's with the object's node
 * offset and count.  A singular model is CRT_E_INVALID (the reference pushes
 * its inf / NaN inverse). */
int crt_push_entity(crt_context* ctx, int32_t object_id, const float model[16], float emissive,
                    float translucency);
/* The BVHEntity record PushEntity builds (Intersector.h:207-213), for callers
 * that assemble their own entity buffers for crt_trace: model, inverse =
 * glm::inverse(model) computed with glm 0.9.8.5's float steps (byte-identical
 * to the reference's), node offset and count, data[0] = float bits of
 * emissive, data[1] = float bits of 1 - translucency, the rest zero.
 * CRT_E_INVALID for a singular model. */
int crt_make_entity(const float model[16], float emissive, float translucency, int32_t node_offset,
                    int32_t node_count, crt_entity* out);
/* BufferEntities (Intersector.h:226-239) */
int crt_buffer_entities(crt_context* ctx);
/* BufferData (Intersector.h:321-351) */
int crt_buffer_data(crt_context* ctx, int clear_cpu_data);

/* Host copies (the public m_BVHNodes / m_BVHVertices / m_BVHTriangles /
 * m_BVHEntities members).  Sizes in elements; pointers valid until the next
 * mutating call. */
int crt_context_host_buffers(const crt_context* ctx, const void** nodes, int64_t* node_count,
                             const crt_triangle** tris, int64_t* tri_count,
                             const crt_vertex** verts, int64_t* vert_count,
                             const crt_entity** entities, int64_t* entity_count);
/* Device scene of the context (valid after buffer_data + buffer_entities). */
int crt_context_scene(const crt_context* ctx, crt_scene_view* out);
/* Builds the wide layout (crt_wide_build) from the context's host copies of
 * the buffered data and entities and uploads it (needs buffer_data without
 * clearing the CPU data, and buffer_entities).  No reference counterpart. */
int crt_context_build_wide(crt_context* ctx, int32_t width, int32_t max_leaf);
/* The same with a choice of collapse (crt_wide_build_ex's options). */
int crt_context_build_wide_ex(crt_context* ctx, const crt_wide_options* options);
/* Device scene of the wide layout (CRT_FORMAT_WIDE4 / WIDE8). */
int crt_context_wide_scene(const crt_context* ctx, crt_scene_view* out);

/* IntersectRay / IntersectRayIgnoreTransparent / IntersectRay(ro,rd) over a
 * device ray buffer. */
int crt_intersect(crt_context* ctx, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits,
                  int64_t nrays, void* stream);

/* IntersectPrimary (Intersector.h:241-266 + TraverseBVHStackless.glsl:371-388):
 * one pinhole ray per pixel, pixel-corner uv = pixel/dims, written to
 * `rays` (width*height device records, row-major) and traced closest-hit into
 * `hits`.  inv_view / inv_proj are column-major.  Unlike the reference
 * dispatch (floor(W/16) x floor(H/16) groups), every pixel is covered. */
int crt_intersect_primary(crt_context* ctx, crt_ray* rays, crt_hit* hits, int width, int height,
                          const float inv_view[16], const float inv_proj[16], uint32_t flags,
                          void* stream);

/* ------------------------------------------------------------------------- */
/* Hit attributes -- GetData (TraverseBVHStackless.glsl:375-408)              */
/* ------------------------------------------------------------------------- */

/* Per-mesh material record, indexed by the hit's mesh id (BVH::TextureReferences,
 * TraverseBVHStackless.glsl:44-49, SSBO binding 20; 32 B). albedo / normal:
 * index into the texture table, -1 = none. */
typedef struct crt_texture_ref {
    float model_color[4];
    int32_t albedo;
    int32_t normal;
    int32_t pad[2];
} crt_texture_ref;

/* One texture of the table (the reference's bindless sampler2D Textures[]):
 * width * height RGBA8 texels, sRGB-encoded color, row-major from v = 0. */
typedef struct crt_texture {
    const uint32_t* texels;   /* device pointer */
    int32_t width;
    int32_t height;
} crt_texture;

/* GetData's outputs for one hit (32 B). */
typedef struct crt_hit_attr {
    float normal[3];      /* normalize(sum of half-unpacked vertex normals x barycentrics); -1 on a miss */
    float emissivity;     /* entity Data[0]                                                            */
    float albedo[3];      /* albedo texture at the interpolated UV, else the mesh's model color        */
    float alpha;          /* entity Data[1] = 1 - translucency (0 on a miss)                           */
} crt_hit_attr;

/* GetData for n closest-hit records (crt_trace / crt_intersect output of
 * this scene): interpolated normal, albedo and the entity's emissivity and
 * alpha.  `refs` [nrefs] and `textures` [ntextures] are device arrays (the
 * table itself and every texel array).  A miss (t < 0 or mesh < 0) gives
 * normal -1, albedo 0, emissivity 0, alpha 0.  Texture sampling: level 0,
 * bilinear, repeat wrap, sRGB decoded per texel before filtering.  A mesh id
 * outside the table gives albedo 0; a texture index outside the table, or a
 * texture without texels, falls back to the model color. */
int crt_hit_attributes(const crt_scene_view* scene, const crt_texture_ref* refs, int32_t nrefs,
                       const crt_texture* textures, int32_t ntextures, const crt_hit* hits, int64_t n,
                       crt_hit_attr* out, void* stream);

/* ------------------------------------------------------------------------- */
/* Multi-bounce diffuse paths with wavefront regeneration (one persistent      */
/* kernel; no reference counterpart -- the reference's nearest analogue is the */
/* fixed two-bounce loop of DiffuseTrace.glsl:508-533)                         */
/* ------------------------------------------------------------------------- */

typedef struct crt_path_params {
    float inv_view[16];     /* column-major camera inverses (IntersectPrimary's)   */
    float inv_proj[16];
    int32_t width, height;  /* image                                               */
    int32_t spp;            /* samples per pixel: path p is sample p % spp of pixel p / spp */
    uint32_t seed;
    int32_t max_bounces;    /* diffuse bounces after the camera segment            */
    float rr;               /* continue probability from the second bounce on (0, 1] */
    float tmax;             /* every segment's TMax                                */
    float sky[3];           /* radiance of rays that leave the scene               */
    const float* albedo;    /* device, nmesh * 3 (per mesh id)                     */
    int32_t nmesh;
    int32_t pad;
} crt_path_params;

/* One path's result (32 B). */
typedef struct crt_path_sample {
    float radiance[3];
    int32_t segments;       /* rays traced for this path (camera + bounces)  */
    uint32_t visits;        /* node visits over all its segments             */
    uint32_t tests;         /* triangle tests over all its segments          */
    uint32_t status;        /* CRT_STATUS_* of any of its walks              */
    uint32_t pad;
} crt_path_sample;

/* Traces paths [0, npaths) over a wide scene (CRT_FORMAT_WIDE4 / WIDE8): per
 * path a jittered camera ray, closest hits, diffuse bounces with Russian
 * roulette, sky radiance on escape; out[p] per path.  Device-filling waves
 * claim new paths whenever enough lanes are idle, so segments of different
 * paths and bounces share the waves.  config: variant 0..5 (the sequential
 * wide walk: stack depth, inner steps, leaf batch, refill threshold) or 6..9
 * (the speculative walk of CRT_KERNEL_SPECULATIVE: stack depth, inner steps,
 * queue slots, refill threshold; radiance, segments and status equal the
 * other presets', visit and test counts follow the wave schedule), grid
 * (0 = fill the device). */
int crt_trace_paths(const crt_scene_view* scene, const crt_path_params* params, crt_path_sample* out,
                    int64_t npaths, const crt_trace_config* config, void* stream);

/* ------------------------------------------------------------------------- */
/* Wavefront ray generation for the reference's consumer passes: the rays     */
/* DiffuseTrace.glsl:441-446, SpecularTrace.glsl:509-513 and                  */
/* UpdateRadianceProbes.glsl:408-423 build per pixel / probe before calling   */
/* IntersectRay, produced as dense ray buffers for crt_trace.                 */
/* ------------------------------------------------------------------------- */

typedef struct crt_raygen_params {
    uint32_t seed;            /* per-ray hash chain h0 = pcg(id ^ seed * 0x9E3779B9), h(k+1) = pcg(hk);
                                 uniform k = (hk >> 8) * 2^-24 (replaces the reference's fract(sin) hash2) */
    float offset;             /* diffuse / AO origin offset along the normal (DiffuseTrace.glsl:445: 0.05) */
    float ao_tmax;            /* AO any-hit TMax                                                           */
    float gi_tmax;            /* GI closest-hit TMax                                                       */
    float spec_offset[2];     /* specular origin offset mix(lo, hi, clamp(roughness * 1.4, 0, 1))
                                 (SpecularTrace.glsl:512: 0.05, 0.1)                                       */
    float spec_tmax;          /* specular closest-hit TMax                                                 */
    float roughness_scale;    /* roughness passed to the sampler = table value * scale (0.9 at :513)       */
    const float* roughness;   /* device, per mesh id (the G-buffer's PBR.x); NULL = every mesh 1.0         */
    int32_t nmesh;            /* entries in `roughness`; a mesh id outside gives 1.0                       */
    int32_t pad;
} crt_raygen_params;

/* Bytes of device workspace crt_gen_diffuse_rays / crt_gen_specular_rays need
 * for n input records. */
size_t crt_raygen_workspace_bytes(int64_t n);

/* For every closest-hit record i with t > 0 (in input order, compacted):
 * P = org + dir * t, N = the hit triangle's face normal in world space
 * (inverse-transpose of the entity's model matrix) turned against the ray,
 * origin = P + N * offset; AO ray (u0, u1) and GI ray (u2, u3) with
 * CosWeightedHemisphere (Include/Sampling.glsl:1-12) directions, hash key and
 * output id = primary[i].id.  *d_count (device int64) receives the number of
 * rays written to each of ao_out / gi_out (either may be NULL).  The scene
 * view is the one the hits were traced on (BVH2, or wide with tri_records). */
int crt_gen_diffuse_rays(const crt_scene_view* scene, const crt_ray* primary, const crt_hit* hits, int64_t n,
                         const crt_raygen_params* params, crt_ray* ao_out, crt_ray* gi_out, int64_t* d_count,
                         void* workspace, size_t workspace_bytes, void* stream);

/* Reflection rays (StochasticReflectionDirection, SpecularTrace.glsl:102-135):
 * incident I = normalize(dir), roughness r = table[mesh] * roughness_scale;
 * r < 0.01 reflects about N, else up to 12 GGX half-vector draws
 * (SampleGGXVNDF, Include/Sampling.glsl:63-83, with uniforms scaled by
 * (0.8, 0.7)) until one has N.H > 0.001, then reflect(I, H).  Same
 * compaction, origin offset and outputs as crt_gen_diffuse_rays. */
int crt_gen_specular_rays(const crt_scene_view* scene, const crt_ray* primary, const crt_hit* hits, int64_t n,
                          const crt_raygen_params* params, crt_ray* out, int64_t* d_count, void* workspace,
                          size_t workspace_bytes, void* stream);

/* One ray per probe of a res[0] x res[1] x res[2] grid
 * (UpdateRadianceProbes.glsl:408-423): probe (x, y, z), index
 * x + res[0] * (y + res[1] * z) = the ray's id; origin = box_origin +
 * (probe / res * 2 - 1) * size; direction uniform on the sphere (LambertBRDF
 * at :351-362 normalized: azimuth 2 pi u0, cos(theta) = 2 u1 - 1). */
typedef struct crt_probe_grid {
    int32_t res[3];
    uint32_t seed;
    float box_origin[3];
    float tmax;
    float size[3];
    int32_t pad;
} crt_probe_grid;
int crt_gen_probe_rays(const crt_probe_grid* grid, crt_ray* out, void* stream);

/* ------------------------------------------------------------------------- */
/* Batched traversal: many scenes and ray batches in one persistent launch    */
/* ------------------------------------------------------------------------- */

/* One ray batch: a device scene (any format; every batch of a plan uses the
 * same one), a query, device rays and hits.  The batch is traced exactly as
 * crt_trace would trace it: the wave walk for the stackless and wide
 * formats, the one-ray-per-lane stack walk for the stack format.  Node width
 * is tuned by planning the same batches over each layout. */
typedef struct crt_batch {
    crt_scene_view scene;
    int32_t query;          /* CRT_QUERY_* */
    uint32_t flags;         /* CRT_FLAG_*  */
    const crt_ray* rays;
    crt_hit* hits;
    int64_t count;
    uint32_t* tri_tests;    /* optional counting output [count] */
} crt_batch;

/* Tuning of the batched kernel (crt_autotune_batched searches these). */
typedef struct crt_batch_config {
    int32_t variant;        /* walk preset 0..7 (stackless: inner steps, leaf batch, window;
                               wide: LDS stack depth, inner steps, leaf batch; stack: LDS stack
                               depth; 6 and 7 are compiled for at least 5 or 6 waves per SIMD);
                               8..11: the speculative walk with crt_trace's
                               CRT_KERNEL_SPECULATIVE presets 0, 1, 3 and 4 (wide formats) or
                               0..3 (stack and stackless formats); a tile is one wave of that
                               launch, so its records equal crt_trace's */
    int32_t tiles_per_claim;/* 64-ray tiles a wave claims per atomic: 1, 2 or 4 */
    int32_t waves_per_group;/* 1, 2 or 4 waves per workgroup */
    int32_t grid;           /* workgroups; 0 = enough to fill every CU */
    int32_t order;          /* 0 = batch order, 1 = heaviest tiles first (cost estimate) */
    int32_t shards;         /* tile-queue counters: 0 or 1 = one, 16 = sixteen interleaved shards
                               (shard k holds tiles k, k + 16, ...) */
    int32_t child_order;    /* 0 = nearest child first; 1 = the any-hit batches walk children in
                               slot order (CRT_FLAG_SLOT_ORDER added to their flags; the wide
                               formats honour it, the stackless format has no choice to make) */
} crt_batch_config;

typedef struct crt_batch_plan crt_batch_plan;

/* Uploads the batch descriptors and the tile table once.  The plan keeps
 * the device pointers of the batches; their contents may change between
 * launches (new rays, same sizes). */
int crt_batch_plan_create(const crt_batch* batches, int nbatches, const crt_batch_config* config,
                          crt_batch_plan** out);
void crt_batch_plan_destroy(crt_batch_plan* plan);
/* One persistent launch over every batch (a counter-reset kernel + the
 * batched kernel; no memset, no allocation -- capturable). */
int crt_batch_plan_launch(crt_batch_plan* plan, void* stream);
/* Captures crt_batch_plan_launch into a hipGraph (once) and replays it. */
int crt_batch_plan_launch_graph(crt_batch_plan* plan, void* stream);
/* Total rays of the plan. */
int64_t crt_batch_plan_rays(const crt_batch_plan* plan);

/* Times every candidate configuration on the batches (`repeats` launches
 * each, after one warm-up) and returns the fastest in *best and its time per
 * launch in *best_ms.  candidates == NULL searches the built-in grid:
 * variants 0..7 x tiles_per_claim {1,2,4} x waves_per_group {1,4} x order {0,1}
 * x shards {1,16}, x child_order {0,1} when the batches use a wide format and
 * hold any-hit queries. */
int crt_autotune_batched(const crt_batch* batches, int nbatches, const crt_batch_config* candidates,
                         int ncandidates, int repeats, void* stream, crt_batch_config* best, float* best_ms);

/* The node-width axis of the autotuner: `batches` holds nlayouts consecutive
 * copies of the same nbatches batches, copy l over one layout of the same
 * scenes (say the BVH2, its BVH4 and its BVH8 collapse: same queries, flags,
 * rays, hits and counts, only the scene views differ; CRT_E_INVALID
 * otherwise).  Each layout is tuned as crt_autotune_batched tunes it
 * (candidates == NULL: its built-in grid); *best_layout receives the index of
 * the fastest layout, *best and *best_ms its configuration and time. */
int crt_autotune_batched_layouts(const crt_batch* batches, int nbatches, int nlayouts,
                                 const crt_batch_config* candidates, int ncandidates, int repeats, void* stream,
                                 int32_t* best_layout, crt_batch_config* best, float* best_ms);

/* ------------------------------------------------------------------------- */
/* Differentiable shading: Cook-Torrance direct light + one diffuse bounce    */
/* ------------------------------------------------------------------------- */

/* Per-pixel record of the forward pass (80 B): the primary hit and its bounce
 * hit, the geometry the backward pass reuses.  Visibility is not reused: the
 * backward pass re-traces every shadow ray; `vis` holds the forward pass's
 * visibility bits for inspection only (bit l: light l visible from p, bit
 * 16+l: from q). */
typedef struct crt_shade_record {
    float p[3];     int32_t tri0;   /* primary hit point; hit triangle, -1 on a miss  */
    float n[3];     int32_t tri1;   /* face normal at p toward the eye; bounce triangle */
    float eye[3];   uint32_t vis;   /* primary ray origin; forward visibility bits      */
    float q[3];     float pad0;     /* bounce hit point                                  */
    float nq[3];    float pad1;     /* face normal at q toward p                         */
} crt_shade_record;

/* Shading parameters (device pointers).  Cook-Torrance as in
 * Source/Core/Shaders/Include/CookTorranceBRDF.glsl:37-79 with metalness 0,
 * per-triangle albedo and roughness, white point lights of radiance I/d^2. */
typedef struct crt_shade_params {
    const float* albedo;      /* [ntri * 3] */
    const float* roughness;   /* [ntri]     */
    const float* light_pos;   /* [nlights * 3] */
    const float* intensity;   /* [nlights]  */
    int32_t ntri;
    int32_t nlights;          /* 1..16 */
} crt_shade_params;

/* Forward: for every pixel with a primary hit, traces one cosine-weighted
 * bounce ray (closest hit, per-pixel PCG seeded by `seed`) and the shadow rays
 * toward every light from p and from the bounce point (any hit); writes the
 * RGB color [npix * 3] and the per-pixel records.  Pixels whose primary ray
 * missed are black. */
int crt_shade_forward(const crt_scene_view* scene, const crt_ray* primary_rays, const crt_hit* primary_hits,
                      int64_t npix, const crt_shade_params* params, uint32_t seed, float* color,
                      crt_shade_record* records, void* stream);

/* Backward (adjoint) of loss = sum over pixels and channels of (color -
 * target)^2 with geometry and visibility held fixed: re-traces the shadow
 * rays of every record, then ADDS d(loss)/d(albedo) [ntri*3],
 * d(loss)/d(roughness) [ntri] and d(loss)/d(intensity) [nlights] into the
 * gradient buffers (zero them first). */
int crt_shade_backward(const crt_scene_view* scene, const crt_shade_record* records, int64_t npix,
                       const crt_shade_params* params, const float* color, const float* target,
                       float* grad_albedo, float* grad_roughness, float* grad_intensity, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CANDELA_RT_H */
