// broad.hpp — LBVH broad phase: buffers and launcher.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/diomedes_hip.h"
#include "shapes.hpp"

namespace dio {

// Body poses and shapes (SoA) as the collision kernels read them.
struct BodyView {
    const float4* pos;
    const float4* rot;
    const uint32_t* tag;
    const float4* corners;  // 4 per body (GpuShape corners)
    const float4* hull;     // hull vertex pool (may be null)
    uint32_t hull_count = 0;  // vertices in the pool; 0 when not known (the narrow phase then reads it in place)
};

// Device buffers of one broad-phase pipeline (capacity n_cap bodies).
struct BroadBuffers {
    float4* aabb_min;  // n_cap
    float4* aabb_max;
    float4* bounds;    // 2
    uint32_t *keys0, *keys1, *vals1;          // n_cap each (radix sort ping-pong)
    uint32_t* perm;                           // n_cap: body at each Morton-sorted position
    uint32_t* sort_scratch;                   // sort_scratch_words(n_cap)
    float4 *node_min, *node_max;              // tree_nodes(n_cap) nodes of the implicit 8-ary tree
    uint32_t* counts;                         // n_cap: partners of higher index per body
    uint32_t* offsets;                        // n_cap + 1
    uint32_t* scan_partials;                  // scan_partials_size(n_cap + 1)
    uint32_t* pair_cache;                     // kPairCache * n_cap: partners found by the counting pass
    uint2* pairs;                             // p_max
    uint32_t* p_dev;                          // pair count min(total, p_max)
    uint32_t* overflow;                       // largest total that exceeded p_max
    uint32_t* err;                            // traversal stack overflow flag
    uint32_t* clear_words[2];                 // words the emit launch zeroes for the next stage (may be null)
    static size_t sort_scratch_words(uint32_t n_cap);
    static size_t tree_nodes(uint32_t n_cap);
};

constexpr uint32_t kPairCache = 32;  // partners per body kept between the two traversal passes
constexpr uint32_t kMaxBroadBodies = 1u << 20;  // tree depth <= 7: the top refit workgroup covers 4 levels

// A batched world's scene table: bodies of scene s are offsets off[s] ..
// off[s+1] - 1 and never pair with another scene's.  bits = bits of the scene
// id in the sort key (0: a single scene, no table).
struct SceneTable {
    const uint32_t* of = nullptr;   // scene of each body
    const uint32_t* off = nullptr;  // num_scenes + 1 body offsets
    uint32_t bits = 0;
};

void launch_unpack_gpu_bodies(const dio_gpu_body* rec, uint32_t n, float4* pos, float4* rot, uint32_t* tag,
                              float4* corners, hipStream_t s);

// AABBs, BVH and the sorted pair list of n bodies; the pair count stays on
// the device (b.p_dev).  resort: recompute Morton codes and the sorted order
// (otherwise the tree is refit over the previous order, which keeps the pair
// set exact and only changes the traversal cost).  lanes: lanes per query in
// the traversal (1, 2, 4, 8, 16 or 32).  scenes: with a table, the sort key's
// high bits are the scene id (so each scene is a contiguous range of sorted
// positions, equal to its body range) and each query walks only its scene's
// range.  refit_done (refit steps only): the first refit launch already ran
// on this stream as part of another launch (tree_refit.hpp).
void run_broad(const BodyView& bv, uint32_t n, const BroadBuffers& b, uint32_t p_max, hipStream_t s, bool resort,
               int lanes, const SceneTable& scenes = SceneTable{}, bool refit_done = false);

}  // namespace dio
