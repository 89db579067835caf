// broad.hip — broad phase for gfx950.
//
// Replaces the reference's dense O(n^2) sweep (broadphase/src/lib.rs:154-216)
// with: support-based AABBs (the same six support queries, same float
// order), 30-bit Morton codes of the AABB centres on cubic cells, a stable LSD
// radix sort (redone every few steps; in between the tree is refit over the
// previous order), an implicit 8-ary BVH over the sorted bodies, and a
// two-pass traversal (count, scan, write) in which each body reports the
// overlapping bodies of higher index.  Leaves are tested with the reference's
// closed-interval test on the reference's exact AABBs, so the pair set equals
// the dense sweep's; each body's pairs are sorted, so the list is in
// lexicographic (a, b) order without a global sort.
#include <hip/hip_runtime.h>

#include "broad.hpp"
#include "tree_refit.hpp"
#include "prim.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

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
                                                   const float4* bounds, uint32_t* keys, uint32_t* vals,
                                                   SceneTable scenes) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = bounds[0], hi = bounds[1];
    const float4 a = mn[i], b = mx[i];
    const float e = fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    const uint32_t qx = quantize((a.x + b.x) * 0.5f, lo.x, e);
    const uint32_t qy = quantize((a.y + b.y) * 0.5f, lo.y, e);
    const uint32_t qz = quantize((a.z + b.z) * 0.5f, lo.z, e);
    const uint32_t code = (expand_bits(qx) << 2) | (expand_bits(qy) << 1) | expand_bits(qz);
    // with a scene table: scene id in the high bits, the code's top bits below
    keys[i] = scenes.bits ? (scenes.of[i] << (30 - scenes.bits)) | (code >> scenes.bits) : code;
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// Implicit 8-ary BVH over the bodies in Morton order (perm[p] = body at sorted
// position p).  With depth D (8^D >= n), node h has children 8h+1 .. 8h+8,
// level L starts at heap index (8^L - 1) / 7, and sorted position p is leaf
// (8^D - 1) / 7 + p.  A node whose first leaf position is >= n is empty: the
// refit skips it and the traversal never reads it.  Leaf boxes are the bodies'
// exact AABBs, so the leaf test is the reference's closed-interval test and
// the pair set equals the dense sweep's for any tree shape; a stale Morton
// order only loosens the boxes.  Refit: one workgroup per chunk of 512 leaves
// (three levels reduced in LDS), then one workgroup for the levels above -- no
// atomics.  Traversal: a group of LANES lanes per query tests the 8 children
// of a node together (8 / LANES each; with 16 or 32 lanes, each 8 lanes take
// a node of their own, so a step pops 2 or 4 nodes) and keeps the group's
// stack in LDS, so each step costs one round of independent loads; a shallow
// tree (5 levels at 8192 bodies) keeps the chain of dependent steps short.

constexpr uint32_t kTopNodes = 585;     // levels 0..3, kept in LDS by the traversal
constexpr int kTravBlock = 128;
constexpr int kStackDepth = 64;         // >= 7 D + 1 for D <= 7
// popping P nodes a step needs at most 7 P D + P entries: 64 P covers D <= 7
template <int LANES>
constexpr int stack_depth() {
    return kStackDepth * (LANES > 8 ? LANES / 8 : 1);
}

// Leaves and the levels of each chunk of kChunkLeaves sorted positions.
__global__ void __launch_bounds__(256) k_tree_lower(TreeLower tl) { tree_lower_part(blockIdx.x, tl); }

__global__ void __launch_bounds__(1024) k_tree_top(uint32_t n, uint32_t D, float4* nmn, float4* nmx) {
    for (int L = (int)D - 4; L >= 0; --L) {
        const uint32_t count = 1u << (3 * L);
        const uint32_t span = 1u << (3 * (D - L - 1));  // leaves under one child
        for (uint32_t k = threadIdx.x; k < count; k += 1024) {
            const uint32_t c0 = level_start(L + 1) + 8 * k;
            if ((uint64_t)(8 * k) * span >= n) continue;
            float4 u = nmn[c0], w = nmx[c0];
            for (uint32_t j = 1; j < 8; ++j) {
                if ((uint64_t)(8 * k + j) * span >= n) break;
                u = box_union_lo(u, nmn[c0 + j]);
                w = box_union_hi(w, nmx[c0 + j]);
            }
            nmn[level_start(L) + k] = u;
            nmx[level_start(L) + k] = w;
        }
        __syncthreads();
    }
}

// lanes of one wavefront: order LDS / memory accesses of the group's lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the LANES bits of this lane's group in a wavefront ballot
template <int LANES>
__device__ __forceinline__ uint32_t group_bits(unsigned long long m) {
    const uint32_t g = (threadIdx.x % 64u) / (uint32_t)LANES;
    return (uint32_t)(m >> (g * LANES)) & (LANES >= 32 ? 0xFFFFFFFFu : ((1u << (LANES % 32)) - 1u));
}

// Walk the tree for query body a (box qa..qb) with the LANES lanes of this
// group, over the sorted positions lo_pos .. hi_pos - 1 (all bodies, or the
// query's scene).  hit(k, b, rank, total) is called on every lane for each child slot k
// after each step: b is the partner found by this lane (or kNone), rank its
// position among the group's hits of that slot, total their count.  Returns
// false if the stack would overflow (never for D <= 7).  The partners come in
// no particular order; both passes place them by rank.
template <int LANES, class Hit>
__device__ __forceinline__ bool walk(uint32_t a, float4 qa, float4 qb, uint32_t lo_pos, uint32_t hi_pos, uint32_t D,
                                     const float4* nmn, const float4* nmx, const float4* clo, const float4* chi,
                                     uint32_t ncached, uint32_t* stk, Hit&& hit) {
    constexpr int kPer = LANES >= 8 ? 1 : 8 / LANES;        // child slots per lane
    constexpr uint32_t kPop = LANES > 8 ? LANES / 8 : 1u;  // nodes popped per step
    constexpr uint32_t kDepth = (uint32_t)stack_depth<LANES>();
    const uint32_t gl = threadIdx.x % (uint32_t)LANES;
    const uint32_t oct = LANES > 8 ? gl / 8u : 0u;  // which popped node this lane tests
    const uint32_t slot = LANES > 8 ? gl % 8u : gl;
    const uint32_t below = (1u << gl) - 1u;
    bool ok = true;
    if (D == 0) return ok;
    if (gl == 0) stk[0] = 0u;  // root, level 0
    uint32_t sp = 1;
    wave_sync();
    while (sp > 0) {
        const uint32_t popped = sp < kPop ? sp : kPop;
        const bool live = oct < popped;
        const uint32_t e = stk[live ? sp - 1 - oct : sp - 1];
        sp -= popped;
        wave_sync();
        const uint32_t h = e & 0x00FFFFFFu, cl = (e >> 24) + 1u;  // children's level
        const uint32_t span = 1u << (3 * (D - cl));
        const uint32_t cs = level_start(cl);
        const bool leaves = cl == D;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t c = 8 * h + 1 + slot + LANES * k;
            const uint32_t first = (c - cs) * span;
            bool ov = false;
            uint32_t leaf_body = kNone;
            if (live && first < hi_pos && first + span > lo_pos) {
                const float4 lo = c < ncached ? clo[c] : nmn[c];
                const float4 hi = c < ncached ? chi[c] : nmx[c];
                ov = aabb_overlap(qa, qb, lo, hi);
                leaf_body = __float_as_uint(lo.w);  // a leaf's body (perm[first]); 0 for inner nodes
            }
            uint32_t b = kNone;
            if (leaves && ov) {
                b = leaf_body;
                if (b <= a) b = kNone;
            }
            const uint32_t hm = group_bits<LANES>(__ballot(b != kNone));
            hit(k, b, __popc(hm & below), __popc(hm));
            const bool push = !leaves && ov;
            const uint32_t pm = group_bits<LANES>(__ballot(push));
            const uint32_t pos = sp + __popc(pm & below);
            if (push) {
                if (pos < kDepth) stk[pos] = c | (cl << 24);
                else ok = false;
            }
            sp += __popc(pm);
            if (sp > kDepth) sp = kDepth;
        }
        wave_sync();
    }
    return ok;
}

// The top levels 0..3 into LDS.  The refit's chunks write every level from
// D - 3 down; above that, k_tree_top writes the levels when D > kLocalTopMax,
// and for D = 4 or 5 each traversal workgroup forms them here from level
// D - 3 (9 nodes at most), which saves that launch.  Unions are exact, so
// both give the same boxes; nodes past the last body stay unset and are
// never read.
constexpr uint32_t kLocalTopMax = 5;
__device__ __forceinline__ void load_top(const float4* nmn, const float4* nmx, uint32_t n, uint32_t D, float4* clo,
                                         float4* chi, uint32_t& ncached) {
    const uint32_t nodes = level_start(D + 1);
    ncached = nodes < kTopNodes ? nodes : kTopNodes;
    const bool local_top = D >= 4 && D <= kLocalTopMax;
    const uint32_t from = local_top ? level_start(D - 3) : 0u;
    for (uint32_t t = from + threadIdx.x; t < ncached; t += kTravBlock) {
        clo[t] = nmn[t];
        chi[t] = nmx[t];
    }
    __syncthreads();
    if (!local_top) return;
    for (int L = (int)D - 4; L >= 0; --L) {
        const uint32_t count = 1u << (3 * L);            // 8 or 1 < kTravBlock
        const uint32_t span = 1u << (3 * (D - L - 1));  // leaves under one child
        const uint32_t k = threadIdx.x;
        if (k < count && (uint64_t)(8 * k) * span < n) {
            const uint32_t c0 = level_start(L + 1) + 8 * k;
            float4 u = clo[c0], w = chi[c0];
            for (uint32_t j = 1; j < 8; ++j) {
                if ((uint64_t)(8 * k + j) * span >= n) break;
                u = box_union_lo(u, clo[c0 + j]);
                w = box_union_hi(w, chi[c0 + j]);
            }
            clo[level_start(L) + k] = u;
            chi[level_start(L) + k] = w;
        }
        __syncthreads();
    }
}

// Pass 1: count body a's partners b > a into counts[a] and keep the first
// kPairCache of them in cache[a * kPairCache ...].  Groups take bodies in
// Morton order, so the groups of a wavefront walk similar paths.
__device__ __forceinline__ void query_range(const SceneTable& st, uint32_t a, uint32_t n, uint32_t& lo,
                                            uint32_t& hi) {
    lo = 0;
    hi = n;
    if (st.bits) {
        const uint32_t s = st.of[a];
        lo = st.off[s];
        hi = st.off[s + 1];
    }
}

template <int LANES>
__global__ void __launch_bounds__(kTravBlock) k_tree_count(uint32_t n, uint32_t D, const uint32_t* perm,
                                                           const float4* amn, const float4* amx, const float4* nmn,
                                                           const float4* nmx, uint32_t* counts, uint32_t* cache,
                                                           uint32_t* err, SceneTable scenes) {
    __shared__ float4 clo[kTopNodes], chi[kTopNodes];
    __shared__ uint32_t stack[kTravBlock / LANES][stack_depth<LANES>()];
    uint32_t ncached;
    load_top(nmn, nmx, n, D, clo, chi, ncached);
    const uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    const uint32_t q = gid / LANES;
    if (q >= n) return;  // whole groups leave together (LANES divides 64)
    const uint32_t a = perm[q];
    uint32_t cnt = 0;
    uint32_t* mine = cache + (size_t)a * kPairCache;
    uint32_t lo, hi;
    query_range(scenes, a, n, lo, hi);
    const bool ok = walk<LANES>(a, amn[a], amx[a], lo, hi, D, nmn, nmx, clo, chi, ncached,
                                stack[threadIdx.x / LANES], [&](int, uint32_t b, uint32_t rank, uint32_t total) {
                                    if (b != kNone && cnt + rank < kPairCache) mine[cnt + rank] = b;
                                    cnt += total;
                                });
    if (!ok) atomicOr(err, 1u);
    if (threadIdx.x % LANES == 0) counts[a] = cnt;
}

// Pass 2: write body a's pairs at offsets[a], sorted by b.  When all of them
// are cached, each one is placed at its rank in the list (the b's of one body
// are distinct); otherwise the group walks again and sorts the segment.
template <int LANES>
__global__ void __launch_bounds__(kTravBlock) k_tree_emit(uint32_t n, uint32_t D, const uint32_t* perm,
                                                          const float4* amn, const float4* amx, const float4* nmn,
                                                          const float4* nmx, const uint32_t* offsets,
                                                          const uint32_t* cache, uint32_t p_max, uint2* pairs,
                                                          SceneTable scenes, uint32_t* p_dev, uint32_t* overflow,
                                                          uint32_t* clear0, uint32_t* clear1) {
    __shared__ float4 clo[kTopNodes], chi[kTopNodes];
    __shared__ uint32_t stack[kTravBlock / LANES][stack_depth<LANES>()];
    __shared__ uint32_t list[kTravBlock / LANES][kPairCache];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the pair count min(total, p_max) for the narrow phase, and its counters
        const uint32_t t = offsets[n];
        *p_dev = t < p_max ? t : p_max;
        if (t > p_max) atomicMax(overflow, t);
        if (clear0) *clear0 = 0u;
        if (clear1) *clear1 = 0u;
    }
    const uint32_t gid = blockIdx.x * kTravBlock + threadIdx.x;
    const uint32_t q = gid / LANES, gl = threadIdx.x % LANES;
    const bool active = q < n;
    const uint32_t a = active ? perm[q] : 0u;
    const uint32_t seg = active ? offsets[a] : 0u;
    const uint32_t total = active ? offsets[a + 1] - seg : 0u;
    // only a body with more than kPairCache partners walks the tree again:
    // a workgroup without one skips the top levels' load
    uint32_t ncached = 0;
    if (__syncthreads_or(active && total > kPairCache)) load_top(nmn, nmx, n, D, clo, chi, ncached);
    if (!active) return;
    const uint32_t grp = threadIdx.x / LANES;
    if (total <= kPairCache) {
        const uint32_t* mine = cache + (size_t)a * kPairCache;
        for (uint32_t j = gl; j < total; j += LANES) list[grp][j] = mine[j];
        wave_sync();
        for (uint32_t j = gl; j < total; j += LANES) {
            const uint32_t b = list[grp][j];
            uint32_t rank = 0;
            for (uint32_t x = 0; x < total; ++x) rank += list[grp][x] < b ? 1u : 0u;
            if (seg + rank < p_max) pairs[seg + rank] = make_uint2(a, b);
        }
        return;
    }
    uint32_t k = 0;
    uint32_t lo, hi;
    query_range(scenes, a, n, lo, hi);
    walk<LANES>(a, amn[a], amx[a], lo, hi, D, nmn, nmx, clo, chi, ncached, stack[grp],
                [&](int, uint32_t b, uint32_t rank, uint32_t cnt) {
                    if (b != kNone && seg + k + rank < p_max) pairs[seg + k + rank] = make_uint2(a, b);
                    k += cnt;
                });
    __threadfence_block();
    wave_sync();
    if (gl == 0) {
        uint32_t e = seg + total;
        if (e > p_max) e = p_max;
        for (uint32_t x = seg + 1; x < e; ++x) {
            const uint2 v = pairs[x];
            uint32_t y = x;
            while (y > seg && pairs[y - 1].y > v.y) {
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

size_t BroadBuffers::tree_nodes(uint32_t n_cap) { return level_start(tree_depth(n_cap > 0 ? n_cap : 1) + 1); }

namespace {
template <int LANES>
void launch_tree_pairs(int mode, uint32_t n, uint32_t D, const BroadBuffers& b, uint32_t p_max, hipStream_t s,
                       const SceneTable& st) {
    const uint32_t blocks = (n * LANES + kTravBlock - 1) / kTravBlock;
    if (mode == 0)
        hipLaunchKernelGGL((k_tree_count<LANES>), dim3(blocks), dim3(kTravBlock), 0, s, n, D, (const uint32_t*)b.perm,
                           (const float4*)b.aabb_min, (const float4*)b.aabb_max, (const float4*)b.node_min,
                           (const float4*)b.node_max, b.counts, b.pair_cache, b.err, st);
    else
        hipLaunchKernelGGL((k_tree_emit<LANES>), dim3(blocks), dim3(kTravBlock), 0, s, n, D, (const uint32_t*)b.perm,
                           (const float4*)b.aabb_min, (const float4*)b.aabb_max, (const float4*)b.node_min,
                           (const float4*)b.node_max, (const uint32_t*)b.offsets, (const uint32_t*)b.pair_cache,
                           p_max, b.pairs, st, b.p_dev, b.overflow, b.clear_words[0], b.clear_words[1]);
}
void tree_pairs(int lanes, int mode, uint32_t n, uint32_t D, const BroadBuffers& b, uint32_t p_max, hipStream_t s,
                const SceneTable& st) {
    switch (lanes) {
        case 1: launch_tree_pairs<1>(mode, n, D, b, p_max, s, st); break;
        case 2: launch_tree_pairs<2>(mode, n, D, b, p_max, s, st); break;
        case 4: launch_tree_pairs<4>(mode, n, D, b, p_max, s, st); break;
        case 16: launch_tree_pairs<16>(mode, n, D, b, p_max, s, st); break;
        case 32: launch_tree_pairs<32>(mode, n, D, b, p_max, s, st); break;
        default: launch_tree_pairs<8>(mode, n, D, b, p_max, s, st); break;
    }
}
}  // namespace

void run_broad(const BodyView& bv, uint32_t n, const BroadBuffers& b, uint32_t p_max, hipStream_t s, bool resort,
               int lanes, const SceneTable& scenes, bool refit_done) {
    if (n == 0) {
        launch_fill_u32(b.offsets, 0u, 1, s);
        hipLaunchKernelGGL(k_pairs_finish, dim3(1), dim3(1), 0, s, b.offsets, p_max, b.p_dev, b.overflow);
        if (b.clear_words[0]) launch_fill_u32(b.clear_words[0], 0u, 1, s);
        if (b.clear_words[1]) launch_fill_u32(b.clear_words[1], 0u, 1, s);
        return;
    }
    if (resort) {
        hipLaunchKernelGGL(k_aabb, dim3(blocks_for(n)), dim3(kBlock), 0, s, bv, n, b.aabb_min, b.aabb_max);
        hipLaunchKernelGGL(k_bounds, dim3(1), dim3(1024), 0, s, b.aabb_min, b.aabb_max, n, b.bounds);
        hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.aabb_min, b.aabb_max, n, b.bounds,
                           b.keys0, b.perm, scenes);
        // 30-bit keys: four 8-bit passes, so the sorted values land back in perm
        radix_sort_pairs(b.keys0, b.perm, b.keys1, b.vals1, n, 30, b.sort_scratch, s);
    }
    const TreeLower tl = tree_lower_args(bv, n, b);
    const uint32_t D = tl.D;
    if (resort || !refit_done) hipLaunchKernelGGL(k_tree_lower, dim3(tl.blocks), dim3(256), 0, s, tl);
    if (D > kLocalTopMax) hipLaunchKernelGGL(k_tree_top, dim3(1), dim3(1024), 0, s, n, D, b.node_min, b.node_max);
    tree_pairs(lanes, 0, n, D, b, p_max, s, scenes);
    launch_scan_exclusive(b.counts, b.offsets, n, nullptr, b.scan_partials, s);
    tree_pairs(lanes, 1, n, D, b, p_max, s, scenes);  // also sets p_dev and clears clear_words
}

}  // namespace dio
