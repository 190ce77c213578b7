// Host build of the device BVH4 / BVH8 collapse (candela_amd/csrc/wide_gpu.hip).
//
// Runs the per-node functions of wide_collapse.h -- the ones the expand,
// assign and root kernels call -- level by level on one thread, in the order
// the device numbers nodes (breadth-first, items of a level in order), so the
// layout equals the device's.  tests/test_wide_collapse_host.py builds it with
// clang++ three ways -- under -fsanitize=undefined with every local set to a
// pattern (-ftrivial-auto-var-init=pattern), with every local zeroed, and
// plainly optimized -- and compares the layouts with crt_wide_build's: a value
// read before it is written, on some path only, would make them differ.
//
// stdin:  int64 node_count, int64 ntris, int32 width, then node_count
//         crt_stack_node entries (64 B each)
// stdout: int64 wide_nodes, uint32 status, then wide_nodes * width / 4
//         crt_wide_group (128 B each)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

// What the HIP headers provide to the device build, for one host thread.
#define __device__
#define __forceinline__ inline
static inline unsigned atomicOr(unsigned* p, unsigned v) {
    const unsigned old = *p;
    *p = old | v;
    return old;
}
static inline int32_t __float_as_int(float f) {
    int32_t i;
    std::memcpy(&i, &f, sizeof(i));
    return i;
}
using std::max;
using std::min;

#include "wide_collapse.h"

using namespace crt::wg;

template <int W>
static int run(const std::vector<crt_stack_node>& nodes, int64_t ntris) {
    const int64_t node_count = int64_t(nodes.size());
    const int64_t cap = ntris;   // the device's output capacity: one wide node per triangle
    std::vector<crt_wide_group> groups(size_t(cap) * (W / 4));
    unsigned err = 0;
    int64_t total = 1;
    Kid k;
    if (root_one_slot(nodes[0], k)) {
        const bool ok = leaf_ok(k.start, k.count, ntris);
        if (!ok) err = kErrLeaf;
        write_one_slot_root<W>(groups.data(), k, ok);
    } else {
        std::vector<int32_t> entries{0}, wides{0};
        while (!entries.empty()) {
            const size_t nf = entries.size();
            std::vector<int32_t> kids(nf * W);
            std::vector<uint32_t> counts(nf);
            for (size_t i = 0; i < nf; ++i)
                counts[i] = expand_one<W>(nodes.data(), node_count, ntris, entries[i], wides[i], groups.data(),
                                          kids.data() + i * W, &err);
            uint64_t all = 0;
            for (uint32_t c : counts) all += c;
            if (int64_t(total + int64_t(all)) > cap) {   // assign_async's cap check
                err |= kErrCap;
                break;
            }
            std::vector<int32_t> next_entries(all), next_wides(all);
            int64_t off = 0;
            for (size_t i = 0; i < nf; ++i) {
                (void)assign_one<W>(wides[i], kids.data() + i * W, off, total, cap, groups.data(),
                                    next_entries.data(), next_wides.data(), &err);
                off += counts[i];
            }
            total += int64_t(all);
            entries.swap(next_entries);
            wides.swap(next_wides);
        }
    }
    if (err) {   // finish_async: any status bit leaves the root empty
        for (int q = 0; q < W / 4; ++q) {
            int32_t* g = reinterpret_cast<int32_t*>(&groups[size_t(q)]);
            for (int word = 0; word < 32; ++word) g[word] = word >= 28 ? -1 : 0;
        }
    }
    const uint32_t status = err;
    std::fwrite(&total, sizeof(total), 1, stdout);
    std::fwrite(&status, sizeof(status), 1, stdout);
    std::fwrite(groups.data(), sizeof(crt_wide_group), size_t(total) * (W / 4), stdout);
    return 0;
}

int main() {
    int64_t node_count = 0, ntris = 0;
    int32_t width = 0;
    if (std::fread(&node_count, 8, 1, stdin) != 1 || std::fread(&ntris, 8, 1, stdin) != 1 ||
        std::fread(&width, 4, 1, stdin) != 1 || node_count < 1 || ntris < 2 || (width != 4 && width != 8))
        return 2;
    std::vector<crt_stack_node> nodes(static_cast<size_t>(node_count));
    if (std::fread(nodes.data(), sizeof(crt_stack_node), nodes.size(), stdin) != nodes.size()) return 2;
    return width == 4 ? run<4>(nodes, ntris) : run<8>(nodes, ntris);
}
