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

__device__ __forceinline__ uint32_t quantize(float c, float lo, float hi) {
    const float e = hi - lo;
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
    const uint32_t qx = quantize((a.x + b.x) * 0.5f, lo.x, hi.x);
    const uint32_t qy = quantize((a.y + b.y) * 0.5f, lo.y, hi.y);
    const uint32_t qz = quantize((a.z + b.z) * 0.5f, lo.z, hi.z);
    keys[i] = (expand_bits(qx) << 2) | (expand_bits(qy) << 1) | expand_bits(qz);
    vals[i] = i;
}

// Karras: common-prefix length of sorted keys i and j, ties broken by index
__device__ __forceinline__ int delta(const uint32_t* codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t ki = codes[i], kj = codes[j];
    if (ki == kj) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(ki ^ kj);
}

__global__ void __launch_bounds__(kBlock) k_lbvh_build(const uint32_t* codes, int n, uint32_t* left,
                                                       uint32_t* right, uint32_t* parent_int,
                                                       uint32_t* parent_leaf) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t lc, rc;
    if (lo == gamma) {
        lc = (uint32_t)gamma | kLeafBit;
        parent_leaf[gamma] = (uint32_t)i;
    } else {
        lc = (uint32_t)gamma;
        parent_int[gamma] = (uint32_t)i;
    }
    if (hi == gamma + 1) {
        rc = (uint32_t)(gamma + 1) | kLeafBit;
        parent_leaf[gamma + 1] = (uint32_t)i;
    } else {
        rc = (uint32_t)(gamma + 1);
        parent_int[gamma + 1] = (uint32_t)i;
    }
    left[i] = lc;
    right[i] = rc;
    if (i == 0) parent_int[0] = kNone;
}

__device__ __forceinline__ float ld_coherent(const float* p) {
    return __builtin_bit_cast(float, __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void child_box(uint32_t c, const uint32_t* vals, const float4* mn, const float4* mx,
                                          const float4* nmn, const float4* nmx, float4& a, float4& b) {
    if (c & kLeafBit) {
        const uint32_t body = vals[c & ~kLeafBit];
        a = mn[body];
        b = mx[body];
    } else {
        const float* pa = reinterpret_cast<const float*>(nmn + c);
        const float* pb = reinterpret_cast<const float*>(nmx + c);
        a = make_float4(ld_coherent(pa), ld_coherent(pa + 1), ld_coherent(pa + 2), 0.0f);
        b = make_float4(ld_coherent(pb), ld_coherent(pb + 1), ld_coherent(pb + 2), 0.0f);
    }
}

// bottom-up refit: the second child to arrive at a node computes its box
__global__ void __launch_bounds__(kBlock) k_lbvh_refit(int n, const uint32_t* vals, const float4* mn,
                                                       const float4* mx, const uint32_t* left, const uint32_t* right,
                                                       const uint32_t* parent_int, const uint32_t* parent_leaf,
                                                       uint32_t* flags, float4* nmn, float4* nmx, uint32_t* err) {
    const int leaf = blockIdx.x * kBlock + threadIdx.x;
    if (leaf >= n) return;
    uint32_t node = parent_leaf[leaf];
    // a tree of n leaves is at most n - 1 levels deep; any longer walk or
    // out-of-range index means the build went wrong: flag it and stop
    for (int level = 0; node != kNone; ++level) {
        if (node >= (uint32_t)(n - 1) || level >= n) {
            atomicOr(err, 2u);
            return;
        }
        __threadfence();
        const uint32_t old = atomicAdd(&flags[node], 1u);
        if (old == 0) return;
        __threadfence();
        float4 a0, a1, b0, b1;
        child_box(left[node], vals, mn, mx, nmn, nmx, a0, a1);
        child_box(right[node], vals, mn, mx, nmn, nmx, b0, b1);
        nmn[node] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
        nmx[node] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
        node = parent_int[node];
    }
}

// Traverse the tree for body `a`; call emit(b) for each overlapping body b > a.
// Returns false if the stack would overflow (the tree is never that deep for
// 30-bit codes with index tie-breaks; the check keeps every access in bounds).
template <class Emit>
__device__ __forceinline__ bool traverse(uint32_t a, int n, const uint32_t* vals, const float4* mn, const float4* mx,
                                         const uint32_t* left, const uint32_t* right, const float4* nmn,
                                         const float4* nmx, Emit&& emit) {
    if (n < 2) return true;
    const float4 qa = mn[a], qb = mx[a];
    uint32_t stack[kStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const uint32_t node = stack[--sp];
        const uint32_t ch[2] = {left[node], right[node]};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint32_t x = ch[c];
            if (x & kLeafBit) {
                if ((x & ~kLeafBit) >= (uint32_t)n) return false;
                const uint32_t b = vals[x & ~kLeafBit];
                if (b >= (uint32_t)n) return false;
                if (b > a && aabb_overlap(qa, qb, mn[b], mx[b])) emit(b);
            } else {
                if (x >= (uint32_t)(n - 1)) return false;
                if (!aabb_overlap(qa, qb, nmn[x], nmx[x])) continue;
                if (sp >= kStack) return false;
                stack[sp++] = x;
            }
        }
    }
    return true;
}

__global__ void __launch_bounds__(kBlock) k_pairs_count(int n, const uint32_t* vals, const float4* mn,
                                                        const float4* mx, const uint32_t* left, const uint32_t* right,
                                                        const float4* nmn, const float4* nmx, uint32_t* counts,
                                                        uint32_t* err) {
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    uint32_t c = 0;
    if (!traverse((uint32_t)a, n, vals, mn, mx, left, right, nmn, nmx, [&](uint32_t) { ++c; })) atomicOr(err, 1u);
    counts[a] = c;
}

__global__ void __launch_bounds__(kBlock) k_pairs_write(int n, const uint32_t* vals, const float4* mn,
                                                        const float4* mx, const uint32_t* left, const uint32_t* right,
                                                        const float4* nmn, const float4* nmx, const uint32_t* offsets,
                                                        uint32_t p_max, uint2* pairs) {
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    const uint32_t base = offsets[a];
    uint32_t k = 0;
    traverse((uint32_t)a, n, vals, mn, mx, left, right, nmn, nmx, [&](uint32_t b) {
        if (base + k < p_max) pairs[base + k] = make_uint2((uint32_t)a, b);
        ++k;
    });
    // sort this body's pairs by b (insertion sort; lists are short)
    uint32_t end = base + k;
    if (end > p_max) end = p_max;
    for (uint32_t x = base + 1; x < end; ++x) {
        const uint2 v = pairs[x];
        uint32_t y = x;
        while (y > base && pairs[y - 1].y > v.y) {
            pairs[y] = pairs[y - 1];
            --y;
        }
        pairs[y] = v;
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

void run_broad(const BodyView& bv, uint32_t n, const BroadBuffers& b, uint32_t p_max, hipStream_t s) {
    if (n == 0) {
        launch_fill_u32(b.offsets, 0u, 1, s);
        hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, b.offsets, p_max, b.p_dev, b.overflow);
        return;
    }
    hipLaunchKernelGGL(k_aabb, dim3(blocks_for(n)), dim3(kBlock), 0, s, bv, n, b.aabb_min, b.aabb_max);
    hipLaunchKernelGGL(k_bounds, dim3(1), dim3(1024), 0, s, b.aabb_min, b.aabb_max, n, b.bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.aabb_min, b.aabb_max, n, b.bounds,
                       b.keys0, b.vals0);
    const int which = radix_sort_pairs(b.keys0, b.vals0, b.keys1, b.vals1, n, 30, b.sort_scratch, s);
    const uint32_t* codes = which ? b.keys1 : b.keys0;
    const uint32_t* vals = which ? b.vals1 : b.vals0;
    const int ni = (int)n;
    if (n >= 2) {
        hipLaunchKernelGGL(k_lbvh_build, dim3(blocks_for(n - 1)), dim3(kBlock), 0, s, codes, ni, b.left, b.right,
                           b.parent_int, b.parent_leaf);
        launch_fill_u32(b.flags, 0u, n - 1, s);
        hipLaunchKernelGGL(k_lbvh_refit, dim3(blocks_for(n)), dim3(kBlock), 0, s, ni, vals, b.aabb_min, b.aabb_max,
                           b.left, b.right, b.parent_int, b.parent_leaf, b.flags, b.node_min, b.node_max, b.err);
    }
    hipLaunchKernelGGL(k_pairs_count, dim3(blocks_for(n)), dim3(kBlock), 0, s, ni, vals, b.aabb_min, b.aabb_max,
                       b.left, b.right, b.node_min, b.node_max, b.counts, b.err);
    launch_scan_exclusive(b.counts, b.offsets, n, nullptr, b.scan_partials, s);
    hipLaunchKernelGGL(k_pairs_write, dim3(blocks_for(n)), dim3(kBlock), 0, s, ni, vals, b.aabb_min, b.aabb_max,
                       b.left, b.right, b.node_min, b.node_max, b.offsets, p_max, b.pairs);
    hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, b.offsets + n, p_max, b.p_dev, b.overflow);
}

}  // namespace dio
