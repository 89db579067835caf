// broad.hip — LBVH broad phase for gfx950.
//
// Replaces the reference's dense O(n^2) sweep (broadphase/src/lib.rs:154-216)
// with: support-based AABBs (the same six support queries, same float
// order), 30-bit Morton codes of the AABB centres, a stable LSD radix sort,
// a Karras (2012) binary radix tree, an atomic bottom-up refit, and a
// two-pass traversal (count, scan, write) in which each body reports the
// overlapping bodies of higher index.  Leaves are tested with the reference's
// closed-interval test on the reference's exact AABBs, so the pair set equals
// the dense sweep's; each body's pairs are sorted, so the list is in
// lexicographic (a, b) order without a global sort.
#include <hip/hip_runtime.h>

#include "broad.hpp"
#include "prim.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kStack = 64;

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

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

__global__ void __launch_bounds__(kBlock) k_aabb(BodyView bv, uint32_t n, float4* mn, float4* mx) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const ShapeRef s = load_shape(bv, i);
    f3 a, b;
    shape_aabb(s, HullPool{bv.hull}, a, b);
    mn[i] = f4(a, 0.0f);
    mx[i] = f4(b, 0.0f);
}

// scene bounds of the AABB centres (one block)
__global__ void __launch_bounds__(1024) k_bounds(const float4* mn, const float4* mx, uint32_t n, float4* bounds) {
    __shared__ float red[6][1024];
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        const float4 a = mn[i], b = mx[i];
        const float c[3] = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], c[k]);
            hi[k] = fmaxf(hi[k], c[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = lo[k];
        red[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bounds[0] = make_float4(red[0][0], red[1][0], red[2][0], 0.0f);
        bounds[1] = make_float4(red[3][0], red[4][0], red[5][0], 0.0f);
    }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// one scale for all three axes (cubic Morton cells): a flat or elongated scene
// keeps spatially compact runs of codes, hence compact tree nodes
__device__ __forceinline__ uint32_t quantize(float c, float lo, float e) {
    float t = e > 0.0f ? (c - lo) / e * 1024.0f : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1023.0f);
    return (uint32_t)t;
}

__global__ void __launch_bounds__(kBlock) k_morton(const float4* mn, const float4* mx, uint32_t n,
                                                   const float4* bounds, uint32_t* keys, uint32_t* vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = bounds[0], hi = bounds[1];
    const float4 a = mn[i], b = mx[i];
    const float e = fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    const uint32_t qx = quantize((a.x + b.x) * 0.5f, lo.x, e);
    const uint32_t qy = quantize((a.y + b.y) * 0.5f, lo.y, e);
    const uint32_t qz = quantize((a.z + b.z) * 0.5f, lo.z, e);
    keys[i] = (expand_bits(qx) << 2) | (expand_bits(qy) << 1) | expand_bits(qz);
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// Implicit LBVH: a complete binary tree in heap layout over the bodies in
// Morton order (perm[p] = body at sorted position p).  With N2 = the next
// power of two >= n, node 0 is the root, node i has children 2i+1, 2i+2, and
// sorted position p is the leaf N2 - 1 + p; positions p >= n are empty leaves
// (inverted boxes that overlap nothing).  Boxes are refit bottom-up by
// LDS reductions: chunks of C leaves per workgroup, then one workgroup for the
// top levels -- no atomics, no inter-workgroup fences.  Leaf boxes are copies
// of the bodies' AABBs, so the leaf test is the reference's exact closed-
// interval test and the pair set equals the dense sweep's for any tree shape.

__device__ __forceinline__ int ilog2(uint32_t x) { return 31 - __clz(x); }

// leaves and the lower levels of one chunk of C <= 1024 sorted positions
constexpr uint32_t kMaxChunk = 1024;
__global__ void __launch_bounds__(256) k_bvh_lower(const uint32_t* perm, uint32_t n, uint32_t n2, uint32_t chunk,
                                                   const float4* mn, const float4* mx, float4* nmn, float4* nmx) {
    __shared__ float4 lo[kMaxChunk], hi[kMaxChunk];
    const uint32_t base = blockIdx.x * chunk;
    const uint32_t leaf0 = n2 - 1;
    for (uint32_t t = threadIdx.x; t < chunk; t += 256) {
        const uint32_t p = base + t;
        float4 a = make_float4(3.4e38f, 3.4e38f, 3.4e38f, 0.0f), b = make_float4(-3.4e38f, -3.4e38f, -3.4e38f, 0.0f);
        if (p < n) {
            const uint32_t body = perm[p];
            a = mn[body];
            b = mx[body];
        }
        lo[t] = a;
        hi[t] = b;
        nmn[leaf0 + p] = a;
        nmx[leaf0 + p] = b;
    }
    __syncthreads();
    // level above the leaves: `width` nodes of this chunk starting at heap index
    // level_first + blockIdx.x * width
    uint32_t width = chunk >> 1;
    uint32_t level_first = (leaf0 + 1) / 2 - 1;
    while (width >= 1) {
        float4 r0[kMaxChunk / 512], r1[kMaxChunk / 512];
        uint32_t k = 0;
        for (uint32_t t = threadIdx.x; t < width; t += 256, ++k) {
            const float4 a0 = lo[2 * t], a1 = hi[2 * t], b0 = lo[2 * t + 1], b1 = hi[2 * t + 1];
            r0[k] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
            r1[k] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
            nmn[level_first + blockIdx.x * width + t] = r0[k];
            nmx[level_first + blockIdx.x * width + t] = r1[k];
        }
        __syncthreads();
        k = 0;
        for (uint32_t t = threadIdx.x; t < width; t += 256, ++k) {
            lo[t] = r0[k];
            hi[t] = r1[k];
        }
        __syncthreads();
        width >>= 1;
        level_first = (level_first + 1) / 2 - 1;
    }
}

// the top levels: reduce the n2 / chunk chunk roots to the root (one workgroup)
__global__ void __launch_bounds__(1024) k_bvh_top(uint32_t n2, uint32_t chunk, float4* nmn, float4* nmx) {
    __shared__ float4 lo[1024], hi[1024];
    uint32_t width = n2 / chunk;  // number of chunk roots
    if (width <= 1) return;
    uint32_t first = width - 1;  // heap index of the first chunk root
    for (uint32_t t = threadIdx.x; t < width; t += 1024) {
        lo[t] = nmn[first + t];
        hi[t] = nmx[first + t];
    }
    __syncthreads();
    while (width > 1) {
        width >>= 1;
        first = (first + 1) / 2 - 1;
        float4 u0, u1;
        const bool act = threadIdx.x < width;
        if (act) {
            const float4 a0 = lo[2 * threadIdx.x], a1 = hi[2 * threadIdx.x];
            const float4 b0 = lo[2 * threadIdx.x + 1], b1 = hi[2 * threadIdx.x + 1];
            u0 = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
            u1 = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
            nmn[first + threadIdx.x] = u0;
            nmx[first + threadIdx.x] = u1;
        }
        __syncthreads();
        if (act) {
            lo[threadIdx.x] = u0;
            hi[threadIdx.x] = u1;
        }
        __syncthreads();
    }
}

constexpr uint32_t kTopCached = 511;  // nodes of the top 9 levels kept in LDS
constexpr int kTravBlock = 128;

// Traverse the subtree of one start node for body a; emit(b) for each
// overlapping body b > a.  False if the stack would overflow (never for the
// <= 22-level trees this library builds; the check keeps accesses in bounds).
template <class Emit>
__device__ __forceinline__ bool traverse_sub(uint32_t a, uint32_t start, uint32_t n, uint32_t n2, const uint32_t* perm,
                                             const float4* mn, const float4* mx, const float4* nmn,
                                             const float4* nmx, const float4* clo, const float4* chi,
                                             uint32_t ncached, Emit&& emit) {
    const float4 qa = mn[a], qb = mx[a];
    const uint32_t leaf0 = n2 - 1;
    uint32_t stack[kStack];
    int sp = 0;
    auto test = [&](uint32_t node) -> bool {
        const float4 lo = node < ncached ? clo[node] : nmn[node];
        const float4 hi = node < ncached ? chi[node] : nmx[node];
        return aabb_overlap(qa, qb, lo, hi);
    };
    auto visit = [&](uint32_t node) -> bool {
        if (node >= leaf0) {
            const uint32_t p = node - leaf0;
            if (p < n) {
                const uint32_t b = perm[p];
                if (b > a && aabb_overlap(qa, qb, mn[b], mx[b])) emit(b);
            }
            return true;
        }
        if (!test(node)) return true;
        if (sp >= kStack) return false;
        stack[sp++] = node;
        return true;
    };
    if (!visit(start)) return false;
    while (sp > 0) {
        const uint32_t node = stack[--sp];
        if (!visit(2 * node + 1)) return false;
        if (!visit(2 * node + 2)) return false;
    }
    return true;
}

// S lanes per body: lane s of the group walks the subtree of node S-1+s (the
// S nodes of level log2 S).  Queries run in Morton order (group q takes body
// perm[q]) so neighbouring lanes walk similar paths.
//
// Pass 1 (k_pairs_count) counts into counts[a*S + s] and keeps the first
// kCache / S bodies it finds in cache[a*kCache + s*(kCache/S) ...].  Pass 2
// (k_pairs_emit), after the scan, places each body's pairs by rank (b is
// distinct within a body's list, so the rank among the list is its sorted
// position): from the cache in LDS when every lane of the body fit, otherwise
// by walking again and sorting the body's segment in place.
template <int S>
__device__ __forceinline__ bool lane_start(uint32_t n, uint32_t n2, uint32_t sub, uint32_t& start) {
    // the subtree this lane walks; a tree smaller than S nodes is walked by lane 0
    start = 0;
    if (n <= 1) return false;
    if (S > 1) {
        if (n2 >= (uint32_t)S) start = (uint32_t)S - 1 + sub;
        else return sub == 0;
    }
    return true;
}

__device__ __forceinline__ void load_top(const float4* nmn, const float4* nmx, uint32_t n2, float4* clo, float4* chi,
                                         uint32_t& ncached) {
    const uint32_t nnodes = 2 * n2 - 1;
    ncached = nnodes < kTopCached ? nnodes : kTopCached;
    for (uint32_t t = threadIdx.x; t < ncached; t += kTravBlock) {
        clo[t] = nmn[t];
        chi[t] = nmx[t];
    }
    __syncthreads();
}

template <int S>
__global__ void __launch_bounds__(kTravBlock) k_pairs_count(uint32_t n, uint32_t n2, const uint32_t* perm,
                                                            const float4* mn, const float4* mx, const float4* nmn,
                                                            const float4* nmx, uint32_t* counts, uint32_t* cache,
                                                            uint32_t* err) {
    constexpr uint32_t C = kPairCache / S;
    __shared__ float4 clo[kTopCached], chi[kTopCached];
    uint32_t ncached;
    load_top(nmn, nmx, n2, clo, chi, ncached);
    const uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    const uint32_t q = gid / S, sub = gid % S;
    if (q >= n) return;
    const uint32_t a = perm[q];
    uint32_t start;
    uint32_t c = 0;
    uint32_t* mine = cache + (size_t)a * kPairCache + sub * C;
    if (lane_start<S>(n, n2, sub, start) &&
        !traverse_sub(a, start, n, n2, perm, mn, mx, nmn, nmx, clo, chi, ncached, [&](uint32_t b) {
            if (c < C) mine[c] = b;
            ++c;
        }))
        atomicOr(err, 1u);
    counts[a * S + sub] = c;
}

template <int S>
__global__ void __launch_bounds__(kTravBlock) k_pairs_emit(uint32_t n, uint32_t n2, const uint32_t* perm,
                                                           const float4* mn, const float4* mx, const float4* nmn,
                                                           const float4* nmx, const uint32_t* offsets,
                                                           const uint32_t* cache, uint32_t p_max, uint2* pairs) {
    constexpr uint32_t C = kPairCache / S;
    constexpr int kBodies = kTravBlock / S;
    __shared__ float4 clo[kTopCached], chi[kTopCached];
    __shared__ uint32_t list[kBodies][kPairCache];
    __shared__ uint32_t spill[kBodies];
    const uint32_t slot = threadIdx.x / S;
    if (threadIdx.x < (uint32_t)kBodies) spill[threadIdx.x] = 0u;
    uint32_t ncached;
    load_top(nmn, nmx, n2, clo, chi, ncached);  // includes a barrier
    const uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    const uint32_t q = gid / S, sub = gid % S;
    const bool valid = q < n;
    const uint32_t a = valid ? perm[q] : 0u;
    uint32_t seg = 0, base = 0, cnt = 0;
    if (valid) {
        seg = offsets[a * S];
        base = offsets[a * S + sub];
        cnt = offsets[a * S + sub + 1] - base;
        if (cnt > C) spill[slot] = 1u;
    }
    __syncthreads();
    const bool fast = valid && spill[slot] == 0u;
    const uint32_t* mine = cache + (size_t)a * kPairCache + sub * C;
    if (fast) {
        for (uint32_t j = 0; j < cnt; ++j) list[slot][base - seg + j] = mine[j];
    } else if (valid) {
        uint32_t start;
        if (cnt > C) {
            uint32_t k = 0;
            if (lane_start<S>(n, n2, sub, start))
                traverse_sub(a, start, n, n2, perm, mn, mx, nmn, nmx, clo, chi, ncached, [&](uint32_t b) {
                    if (base + k < p_max) pairs[base + k] = make_uint2(a, b);
                    ++k;
                });
        } else {
            for (uint32_t j = 0; j < cnt; ++j)
                if (base + j < p_max) pairs[base + j] = make_uint2(a, mine[j]);
        }
    }
    __syncthreads();
    if (fast) {
        const uint32_t total = offsets[a * S + S] - seg;
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t b = list[slot][base - seg + j];
            uint32_t rank = 0;
            for (uint32_t x = 0; x < total; ++x) rank += list[slot][x] < b ? 1u : 0u;
            if (seg + rank < p_max) pairs[seg + rank] = make_uint2(a, b);
        }
    } else if (valid && sub == 0) {
        const uint32_t b0 = seg;
        uint32_t e = offsets[a * S + S];
        if (e > p_max) e = p_max;
        for (uint32_t x = b0 + 1; x < e; ++x) {
            const uint2 v = pairs[x];
            uint32_t y = x;
            while (y > b0 && pairs[y - 1].y > v.y) {
                pairs[y] = pairs[y - 1];
                --y;
            }
            pairs[y] = v;
        }
    }
}

// pair count = min(total, p_max); the overflow word records a total past p_max
__global__ void k_pairs_finish(const uint32_t* total, uint32_t p_max, uint32_t* p_dev, uint32_t* overflow) {
    const uint32_t t = *total;
    *p_dev = t < p_max ? t : p_max;
    if (t > p_max) atomicMax(overflow, t);
}

}  // namespace

__global__ void __launch_bounds__(kBlock) k_unpack_gpu_bodies(const dio_gpu_body* rec, uint32_t n, float4* pos,
                                                              float4* rot, uint32_t* tag, float4* corners) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const dio_gpu_body r = rec[i];
    pos[i] = make_float4(r.position[0], r.position[1], r.position[2], 0.0f);
    rot[i] = make_float4(r.orientation[0], r.orientation[1], r.orientation[2], r.orientation[3]);
    tag[i] = r.shape.tag;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        corners[4 * i + k] = make_float4(r.shape.corners[k][0], r.shape.corners[k][1], r.shape.corners[k][2], r.shape.corners[k][3]);
}

void launch_unpack_gpu_bodies(const dio_gpu_body* rec, uint32_t n, float4* pos, float4* rot, uint32_t* tag,
                              float4* corners, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_unpack_gpu_bodies, dim3(blocks_for(n)), dim3(kBlock), 0, s, rec, n, pos, rot, tag, corners);
}

size_t BroadBuffers::sort_scratch_words(uint32_t n_cap) { return radix_scratch_size(n_cap); }

uint32_t BroadBuffers::pow2_ceil(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

namespace {
template <int S>
void launch_pairs(int mode, uint32_t n, uint32_t n2, const BroadBuffers& b, uint32_t p_max, hipStream_t s) {
    const uint32_t blocks = (n * S + kTravBlock - 1) / kTravBlock;
    if (mode == 0)
        hipLaunchKernelGGL((k_pairs_count<S>), dim3(blocks), dim3(kTravBlock), 0, s, n, n2, (const uint32_t*)b.perm,
                           (const float4*)b.aabb_min, (const float4*)b.aabb_max, (const float4*)b.node_min,
                           (const float4*)b.node_max, b.counts, b.pair_cache, b.err);
    else
        hipLaunchKernelGGL((k_pairs_emit<S>), dim3(blocks), dim3(kTravBlock), 0, s, n, n2, (const uint32_t*)b.perm,
                           (const float4*)b.aabb_min, (const float4*)b.aabb_max, (const float4*)b.node_min,
                           (const float4*)b.node_max, (const uint32_t*)b.offsets, (const uint32_t*)b.pair_cache, p_max,
                           b.pairs);
}
void pairs_pass(int lanes, int mode, uint32_t n, uint32_t n2, const BroadBuffers& b, uint32_t p_max, hipStream_t s) {
    switch (lanes) {
        case 1: launch_pairs<1>(mode, n, n2, b, p_max, s); break;
        case 2: launch_pairs<2>(mode, n, n2, b, p_max, s); break;
        case 8: launch_pairs<8>(mode, n, n2, b, p_max, s); break;
        case 16: launch_pairs<16>(mode, n, n2, b, p_max, s); break;
        default: launch_pairs<4>(mode, n, n2, b, p_max, s); break;
    }
}
}  // namespace

void run_broad(const BodyView& bv, uint32_t n, const BroadBuffers& b, uint32_t p_max, hipStream_t s, bool resort,
               int lanes) {
    if (n == 0) {
        launch_fill_u32(b.offsets, 0u, 1, s);
        hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, b.offsets, p_max, b.p_dev, b.overflow);
        return;
    }
    if (lanes != 1 && lanes != 2 && lanes != 8 && lanes != 16) lanes = 4;
    hipLaunchKernelGGL(k_aabb, dim3(blocks_for(n)), dim3(kBlock), 0, s, bv, n, b.aabb_min, b.aabb_max);
    if (resort) {
        hipLaunchKernelGGL(k_bounds, dim3(1), dim3(1024), 0, s, b.aabb_min, b.aabb_max, n, b.bounds);
        hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.aabb_min, b.aabb_max, n, b.bounds,
                           b.keys0, b.perm);
        // 30-bit keys: four 8-bit passes, so the sorted values land back in perm
        radix_sort_pairs(b.keys0, b.perm, b.keys1, b.vals1, n, 30, b.sort_scratch, s);
    }
    const uint32_t n2 = BroadBuffers::pow2_ceil(n);
    uint32_t chunk = n2 < 512u ? n2 : 512u;
    while (n2 / chunk > 1024u) chunk <<= 1;
    hipLaunchKernelGGL(k_bvh_lower, dim3(n2 / chunk), dim3(256), 0, s, (const uint32_t*)b.perm, n, n2, chunk,
                       (const float4*)b.aabb_min, (const float4*)b.aabb_max, b.node_min, b.node_max);
    hipLaunchKernelGGL(k_bvh_top, dim3(1), dim3(1024), 0, s, n2, chunk, b.node_min, b.node_max);
    const uint32_t nc = n * (uint32_t)lanes;
    pairs_pass(lanes, 0, n, n2, b, p_max, s);
    launch_scan_exclusive(b.counts, b.offsets, nc, nullptr, b.scan_partials, s);
    pairs_pass(lanes, 1, n, n2, b, p_max, s);
    hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, b.offsets + nc, p_max, b.p_dev, b.overflow);
}

}  // namespace dio
