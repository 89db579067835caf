// tree_refit.hpp — the first refit launch of the broad phase's implicit
// 8-ary tree (the bodies' AABBs, the leaves and the levels of each chunk of
// 512 sorted positions) as a workgroup-level part, so that a launch of
// independent work can carry it: on steps that refit the tree over the
// previous Morton order, the contact build's first launch does (the refit
// reads only poses and the sorted order and writes only boxes).
#pragma once

#include "broad.hpp"

namespace dio {

constexpr uint32_t kChunkLeaves = 512;  // 8^3 leaves per refit workgroup of 256 threads

__host__ __device__ __forceinline__ uint32_t level_start(uint32_t L) { return ((1u << (3 * L)) - 1u) / 7u; }

__host__ __device__ __forceinline__ uint32_t tree_depth(uint32_t n) {
    uint32_t d = 0;
    while ((1u << (3 * d)) < n) ++d;
    return d;
}

__device__ __forceinline__ float4 box_union_lo(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ __forceinline__ float4 box_union_hi(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// One body's shape record as the collision code reads it.
__device__ __forceinline__ ShapeRef load_shape(const BodyView& bv, uint32_t i) {
    ShapeRef s;
    s.pos = xyz(bv.pos[i]);
    s.rot = as_q(bv.rot[i]);
    s.tag = bv.tag[i];
    s.c[0] = bv.corners[4 * i];
    s.c[1] = bv.corners[4 * i + 1];
    s.c[2] = bv.corners[4 * i + 2];
    s.c[3] = bv.corners[4 * i + 3];
    return s;
}

// Arguments of the part; `blocks` is its workgroup count (0: no refit here).
struct TreeLower {
    BodyView bv;
    const uint32_t* perm;
    uint32_t n, D, chunk, blocks;
    float4 *amn, *amx, *nmn, *nmx;
};

inline TreeLower tree_lower_args(const BodyView& bv, uint32_t n, const BroadBuffers& b) {
    const uint32_t D = tree_depth(n);
    const uint32_t full = 1u << (3 * D);
    const uint32_t chunk = full < kChunkLeaves ? full : kChunkLeaves;
    return TreeLower{bv, b.perm, n, D, chunk, n > 0 ? (n + chunk - 1) / chunk : 0u,
                     b.aabb_min, b.aabb_max, b.node_min, b.node_max};
}

// Leaves (the bodies' AABBs, computed here and also stored per body for the
// queries) and the levels of chunk `bid`, on the 256 threads of a workgroup.
__device__ __forceinline__ void tree_lower_part(uint32_t bid, const TreeLower& tl) {
    const BodyView& bv = tl.bv;
    const uint32_t* perm = tl.perm;
    const uint32_t n = tl.n, D = tl.D, chunk = tl.chunk;
    float4 *amn = tl.amn, *amx = tl.amx, *nmn = tl.nmn, *nmx = tl.nmx;
    __shared__ float4 lo[kChunkLeaves], hi[kChunkLeaves];
    const uint32_t base = bid * chunk;
    const uint32_t leaf0 = level_start(D);
    for (uint32_t t = threadIdx.x; t < chunk; t += 256) {
        const uint32_t p = base + t;
        float4 a = make_float4(3.4e38f, 3.4e38f, 3.4e38f, 0.0f), b = make_float4(-3.4e38f, -3.4e38f, -3.4e38f, 0.0f);
        uint32_t body = 0xFFFFFFFFu;  // past the last body
        if (p < n) {
            body = perm[p];
            f3 u, w;
            shape_aabb(load_shape(bv, body), HullPool{bv.hull}, u, w);
            a = f4(u, 0.0f);
            b = f4(w, 0.0f);
            amn[body] = a;
            amx[body] = b;
        }
        lo[t] = a;
        hi[t] = b;
        // a leaf carries its body's index in the unused w, so the traversal
        // gets it with the box instead of by a second, dependent load
        nmn[leaf0 + p] = make_float4(a.x, a.y, a.z, __uint_as_float(body));
        nmx[leaf0 + p] = b;
    }
    __syncthreads();
    uint32_t width = chunk / 8;  // this chunk's nodes on the level being built
    uint32_t L = D;
    while (width >= 1) {
        --L;
        float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f), w = u;
        const uint32_t t = threadIdx.x;  // width <= 64 < 256
        if (t < width) {
            u = lo[8 * t];
            w = hi[8 * t];
#pragma unroll
            for (int j = 1; j < 8; ++j) {
                u = box_union_lo(u, lo[8 * t + j]);
                w = box_union_hi(w, hi[8 * t + j]);
            }
            nmn[level_start(L) + bid * width + t] = u;
            nmx[level_start(L) + bid * width + t] = w;
        }
        __syncthreads();
        if (t < width) {
            lo[t] = u;
            hi[t] = w;
        }
        __syncthreads();
        width /= 8;
    }
}

// Levels above the chunk roots (level D - 3), one workgroup; children whose
// first leaf is past n were never built and are skipped.
}  // namespace dio
