// narrow.hip — narrow phase for gfx950.
//
// Counterpart of narrowphase_main (narrowphase/src/lib.rs:89-158).  Pairs are
// dispatched by class:
//  - box / box (CUBE or BOX): the OBB SAT manifold of cuboid.rs, up to 4
//    contacts, one lane per pair, in registers;
//  - sphere / sphere: the analytic contact of lib.rs:125-144;
//  - everything else: GJK (gjk.rs) runs on one lane per pair; intersecting
//    pairs are appended (with their final simplex) to a list that the EPA
//    kernel consumes: a first pass with 16 lanes per pair (four pairs per
//    wavefront) and a polytope of 24 vertices and 32 faces, then the pairs
//    that outgrow it again with one 64-lane wavefront per pair at the caps of
//    epa.rs (256 vertices, 512 faces; ~26 KB).  The polytopes live in LDS,
//    and the face loops of epa.rs — closest face, visibility, horizon search,
//    compaction, stitching — are split across a pair's lanes with ballot
//    prefix sums that keep the reference's sequential orders (first minimal
//    face, horizon edges in face-then-edge order, kept faces in index order).
#include <hip/hip_runtime.h>

#include "narrow.hpp"
#include "prim.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
// EPA caps of epa.rs (256 vertices, 512 faces), and the smaller polytope of
// the first pass: a pair that would outgrow it is run again at full size, so
// every pair's result is the full-size algorithm's
constexpr int kMaxV = 256;
constexpr int kMaxF = 512;
constexpr int kSmallV = 24;
constexpr int kSmallF = 32;
constexpr int kEpaLanes = 16;  // lanes per pair in the first pass (four pairs per wave)
// With at most kEpaFewPairs GJK-intersecting pairs the first pass runs them a
// wavefront each at 64 vertices and 128 faces instead: then the pairs are in
// flight together either way, and a pair's own latency (shorter with 64 lanes)
// and the retry launch (which the larger polytope leaves with next to nothing)
// set the time.
constexpr int kFewV = 64;
constexpr int kFewF = 128;
constexpr uint32_t kEpaFewPairs = 32768;
// A hull vertex pool of at most this many vertices (16 KB) is staged in LDS
// by every workgroup of the GJK and EPA launches: the support scans then read
// LDS instead of the L2 (the same values, so the same results).
constexpr uint32_t kLdsPoolMax = 1024;
// Workgroups of the EPA first pass (one wavefront each, taking its pairs a
// grid apart): a grid far wider than the resident wavefronts evens out the
// pairs' uneven iteration counts.  The retry pass takes up to kRetryBlocks
// pairs at once (one wavefront and 26 KB of LDS each).
constexpr uint32_t kEpaBlocks = 65536;
constexpr uint32_t kRetryBlocks = 2048;
extern __shared__ float4 pool_lds[];

// The pool the support functions read: staged in LDS (LDS_POOL) or in place.
template <bool LDS_POOL> __device__ __forceinline__ HullPool stage_pool(const BodyView& bv) {
    if (!LDS_POOL) return HullPool{bv.hull};
    for (uint32_t i = threadIdx.x; i < bv.hull_count; i += blockDim.x) pool_lds[i] = bv.hull[i];
    __syncthreads();
    return HullPool{pool_lds};
}

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

// ------------------------------------------------------------------ GJK
// gjk.rs:67-73
__device__ __forceinline__ void support_witness(const ShapeRef& a, const ShapeRef& b, f3 d, const HullPool& hp,
                                                f3& point, f3& witness) {
    const f3 ap = add(qrot(a.rot, support_local(a, qrot_inv(a.rot, d), hp)), a.pos);
    const f3 bp = add(qrot(b.rot, support_local(b, qrot_inv(b.rot, neg(d)), hp)), b.pos);
    point = sub(ap, bp);
    witness = ap;
}

// The shifts store every slot with its new or old value (unrolled selects): a
// runtime index, or conditional stores the compiler merges into one, would
// place the simplex in scratch memory.
struct Simplex {
    f3 p[4], w[4];
    uint32_t len;
    __device__ void push_front(f3 pt, f3 wt) {
#pragma unroll
        for (uint32_t i = 3; i > 0; --i) {
            p[i] = sel_f3(i <= len, p[i - 1], p[i]);
            w[i] = sel_f3(i <= len, w[i - 1], w[i]);
        }
        p[0] = pt;
        w[0] = wt;
        len = len + 1 < 4 ? len + 1 : 4;
    }
    __device__ void remove(uint32_t index) {
#pragma unroll
        for (uint32_t i = 0; i < 3; ++i) {
            const bool shift = i >= index && i + 1 < len;
            p[i] = sel_f3(shift, p[i + 1], p[i]);
            w[i] = sel_f3(shift, w[i + 1], w[i]);
        }
        len -= 1;
    }
    __device__ void swap12() {
        f3 t = p[1]; p[1] = p[2]; p[2] = t;
        t = w[1]; w[1] = w[2]; w[2] = t;
    }
};

__device__ __forceinline__ bool same_dir(f3 a, f3 b) { return dot3(a, b) > 0.0f; }
__device__ __forceinline__ f3 triple(f3 a, f3 b, f3 c) { return cross3(cross3(a, b), c); }
// gjk.rs:214-223
__device__ __forceinline__ f3 any_perpendicular(f3 v) {
    f3 axis;
    if (fabsf(v.x) <= fabsf(v.y) && fabsf(v.x) <= fabsf(v.z)) axis = f3{1.0f, 0.0f, 0.0f};
    else if (fabsf(v.y) <= fabsf(v.z)) axis = f3{0.0f, 1.0f, 0.0f};
    else axis = f3{0.0f, 0.0f, 1.0f};
    return normalize_or_zero3(cross3(axis, v));
}

// gjk.rs:112-201
__device__ bool contains_origin(Simplex& s, f3& dir) {
    if (s.len == 4) {
        const f3 a = s.p[0], b = s.p[1], c = s.p[2], d = s.p[3];
        const f3 ao = neg(a), ab = sub(b, a), ac = sub(c, a), ad = sub(d, a);
        const f3 abc = cross3(ab, ac), acd = cross3(ac, ad), adb = cross3(ad, ab);
        if (same_dir(abc, ao)) s.remove(3);
        else if (same_dir(acd, ao)) s.remove(1);
        else if (same_dir(adb, ao)) s.remove(2);
        else return true;
    }
    if (s.len == 1) {
        dir = neg(s.p[0]);
        return false;
    }
    if (s.len == 2) {
        const f3 a = s.p[0], b = s.p[1];
        const f3 ab = sub(b, a), ao = neg(a);
        if (same_dir(ab, ao)) {
            f3 d = triple(ab, ao, ab);
            if (dot3(d, d) < 1e-12f) d = any_perpendicular(ab);
            dir = d;
        } else {
            s.remove(1);
            dir = ao;
        }
        return false;
    }
    if (s.len == 3) {
        const f3 a = s.p[0], b = s.p[1], c = s.p[2];
        const f3 ab = sub(b, a), ac = sub(c, a), ao = neg(a);
        const f3 abc = cross3(ab, ac);
        if (same_dir(cross3(abc, ac), ao)) {
            if (same_dir(ac, ao)) { s.remove(1); dir = triple(ac, ao, ac); }
            else if (same_dir(ab, ao)) { s.remove(2); dir = triple(ab, ao, ab); }
            else { s.len = 1; dir = ao; }
        } else if (same_dir(cross3(ab, abc), ao)) {
            if (same_dir(ab, ao)) { s.remove(2); dir = triple(ab, ao, ab); }
            else { s.len = 1; dir = ao; }
        } else {
            if (dot3(abc, ao) > 0.0f) dir = abc;
            else { s.swap12(); dir = neg(abc); }
        }
        return false;
    }
    return false;
}

// gjk.rs:77-109
__device__ bool gjk(const ShapeRef& a, const ShapeRef& b, const HullPool& hp, Simplex& s) {
    f3 dir = sub(a.pos, b.pos);
    if (eq3(dir, f3{0.0f, 0.0f, 0.0f})) dir = f3{1.0f, 0.0f, 0.0f};
    s.len = 0;
    f3 pt, wt;
    support_witness(a, b, dir, hp, pt, wt);
    s.push_front(pt, wt);
    dir = neg(s.p[0]);
    for (int it = 0; it < 64; ++it) {
        support_witness(a, b, dir, hp, pt, wt);
        if (dot3(pt, dir) < 0.0f) return false;
        s.push_front(pt, wt);
        if (contains_origin(s, dir)) return true;
        if (dot3(dir, dir) < 1e-12f) dir = any_perpendicular(s.p[0]);
    }
    return false;
}

__device__ __forceinline__ void write_slot(const NarrowOut& o, uint32_t slot, f3 n, float depth, f3 pa, f3 pb) {
    o.n[slot] = f4(n, depth);
    o.pa[slot] = f4(pa, 0.0f);
    o.pb[slot] = f4(pb, 0.0f);
}

// One lane per pair: SAT, sphere, or GJK (intersecting pairs go to EPA).
// The workgroup takes the pair list 256 pairs at a time and first sorts its
// chunk by class in LDS (box / box and sphere / sphere pairs first, then the
// GJK pairs, each in list order), so that a wave runs one branch instead of
// both; every pair's result goes to its own slots, as before.  Four waves
// per SIMD (at most 128 VGPRs, a few spilled): faster than three without
// spills or five with more (684, 764 and 728 us on the 4096-scene table).
template <bool LDS_POOL>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_narrow(BodyView bv, NarrowOut o, float4* simplex_store) {
    constexpr int kWaves = kBlock / kWave;
    __shared__ uint32_t s_pair[kBlock];
    __shared__ uint32_t s_wave[2][kWaves];
    const uint32_t np = *o.p_dev;
    const HullPool hp = stage_pool<LDS_POOL>(bv);
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint64_t lt = lane == 0 ? 0ull : ((~0ull) >> (kWave - lane));
    for (uint32_t base = blockIdx.x * kBlock; base < np; base += gridDim.x * kBlock) {
        // the chunk's pairs by class: 0 SAT or sphere, 1 GJK, 2 past the list's end
        const uint32_t q = base + threadIdx.x;
        uint32_t cls = 2;
        if (q < np) {
            const uint2 pr = o.pairs[q];
            const uint32_t ta = bv.tag[pr.x], tb = bv.tag[pr.y];
            const bool bx = (ta == kCube || ta == kBox) && (tb == kCube || tb == kBox);
            cls = (bx || (ta == kSphere && tb == kSphere)) ? 0u : 1u;
        }
        const uint64_t b0 = __ballot(cls == 0), b1 = __ballot(cls == 1);
        if (lane == 0) {
            s_wave[0][wave] = (uint32_t)__popcll(b0);
            s_wave[1][wave] = (uint32_t)__popcll(b1);
        }
        __syncthreads();
        uint32_t n0 = 0, before0 = 0, before1 = 0;
        for (int w = 0; w < kWaves; ++w) {
            n0 += s_wave[0][w];
            if (w < (int)wave) { before0 += s_wave[0][w]; before1 += s_wave[1][w]; }
        }
        if (cls == 0) s_pair[before0 + (uint32_t)__popcll(b0 & lt)] = q;
        else if (cls == 1) s_pair[n0 + before1 + (uint32_t)__popcll(b1 & lt)] = q;
        __syncthreads();
        const uint32_t take = np - base < (uint32_t)kBlock ? np - base : (uint32_t)kBlock;
        if (threadIdx.x < take) {
            const uint32_t p = s_pair[threadIdx.x];
            const uint2 pr = o.pairs[p];
            const ShapeRef a = load_shape(bv, pr.x);
            const ShapeRef b = load_shape(bv, pr.y);
            uint32_t count = 0;
            const bool boxa = a.tag == kCube || a.tag == kBox, boxb = b.tag == kCube || b.tag == kBox;
            if (boxa && boxb) {
                Manifold m{};
                if (box_manifold(a, b, m)) {
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if (k < m.count) write_slot(o, p * 4 + k, m.normal, m.depth, m.pa[k], m.pb[k]);
                    count = m.count;
                }
            } else if (a.tag == kSphere && b.tag == kSphere) {
                const f3 delta = sub(b.pos, a.pos);
                const float distance = len3(delta);
                const float radius = a.c[0].x + b.c[0].x;
                if (!(distance >= radius || distance <= 1.1920929e-7f)) {
                    const f3 n = divs(delta, distance);
                    write_slot(o, p * 4, n, radius - distance, add(a.pos, scale(n, a.c[0].x)),
                               sub(b.pos, scale(n, b.c[0].x)));
                    count = 1;
                }
            } else {
                Simplex s;
                if (gjk(a, b, hp, s)) {
                    const uint32_t at = atomicAdd(o.gjk_count, 1u);
                    o.gjk_list[at] = p;
                    for (int k = 0; k < 4; ++k) {
                        simplex_store[8 * (size_t)p + k] = f4(s.p[k], 0.0f);
                        simplex_store[8 * (size_t)p + 4 + k] = f4(s.w[k], 0.0f);
                    }
                }
            }
            o.cnt[p] = count;
        }
        __syncthreads();  // s_pair and s_wave are rewritten by the next chunk
    }
}

// ------------------------------------------------------------------ EPA
template <int MV, int MF> struct EpaLds {
    float vx[MV], vy[MV], vz[MV];
    float wx[MV], wy[MV], wz[MV];
    uint32_t fa[MF], fb[MF], fc[MF];
    float fnx[MF], fny[MF], fnz[MF], fd[MF];
    uint8_t vis[MF];
    uint16_t vlist[MF];  // the visible faces, in index order
    uint32_t ha[MF], hb[MF];
    uint32_t count;  // broadcast word
};

// A group of GL lanes (16 or 64) of a 64-lane wave works on one pair.
template <int GL> struct LaneGroup {
    int lane;         // lane within the group
    int base;         // the group's first lane in the wave
    uint64_t mask;    // the group's lanes
    uint64_t below;   // the group's lanes below this one
};
template <int GL> __device__ __forceinline__ LaneGroup<GL> lane_group() {
    static_assert(GL == 16 || GL == 64, "16 or 64 lanes per group");
    LaneGroup<GL> g;
    const int wl = (int)(threadIdx.x % (unsigned)kWave);
    g.lane = wl % GL;
    g.base = wl - g.lane;
    g.mask = GL == 64 ? ~0ull : (0xFFFFull << g.base);
    g.below = g.lane == 0 ? 0ull : ((~0ull >> (64 - g.lane)) << g.base);
    return g;
}
template <int GL> __device__ __forceinline__ uint64_t group_ballot(const LaneGroup<GL>& g, bool x) {
    return __ballot(x) & g.mask;
}
template <int GL> __device__ __forceinline__ uint32_t group_first(const LaneGroup<GL>& g, uint64_t bal) {
    return (uint32_t)(__ffsll((long long)bal) - 1 - g.base);
}

// Reductions over a group, every lane of the group active.  Within each row
// of 16 lanes four DPP steps (quad_perm xor 1, quad_perm xor 2, half-row
// mirror, row mirror) leave every lane holding its row's result; for a
// 64-lane group the four rows' results are then read from lanes 0, 16, 32 and
// 48.  The combining op returns one of its operands (no NaNs reach it), so
// the result is one of the inputs.
template <int GL, class Op> __device__ __forceinline__ uint32_t group_reduce_bits(uint32_t x, Op op) {
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));
    if (GL == 16) return x;
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
    return op(op(r0, r1), op(r2, r3));
}
template <int GL> __device__ __forceinline__ float group_max(float x) {
    return __uint_as_float(group_reduce_bits<GL>(__float_as_uint(x), [](uint32_t a, uint32_t b) {
        return __float_as_uint(fmaxf(__uint_as_float(a), __uint_as_float(b)));
    }));
}
template <int GL> __device__ __forceinline__ float group_min(float x) {
    return __uint_as_float(group_reduce_bits<GL>(__float_as_uint(x), [](uint32_t a, uint32_t b) {
        return __float_as_uint(fminf(__uint_as_float(a), __uint_as_float(b)));
    }));
}
template <int GL> __device__ __forceinline__ uint32_t group_min_u32(uint32_t x) {
    return group_reduce_bits<GL>(x, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}
constexpr uint32_t kNoIndex = 0xFFFFFFFFu;

// support_local of a hull with the group's lanes splitting the vertices: the
// same dot products, and the vertex the sequential scan keeps (the first
// maximal one, skipping NaNs; a NaN at vertex 0 is never replaced).  Each
// lane keeps its first maximal vertex; the group's maximum value is reduced
// alone, and the first vertex holding it is the lowest lane's (up to GL
// vertices, where vertex i is lane i's) or the least index.  The shape is the
// same on every lane of a group, so each group takes one side of the branch.
template <int GL>
__device__ __forceinline__ f3 support_local_group(const ShapeRef& s, f3 d, const HullPool& hp,
                                                  const LaneGroup<GL>& g) {
    if (s.tag != kHull) return support_local(s, d, hp);
    const uint32_t first = (uint32_t)s.c[0].x, count = (uint32_t)s.c[0].y;
    float bd = -__builtin_inff();
    uint32_t bi = kNoIndex;
    for (uint32_t i = (uint32_t)g.lane; i < count; i += GL) {
        const float dd = dot3(xyz(hp.v[first + i]), d);
        if (dd == dd && (bi == kNoIndex || dd > bd)) { bd = dd; bi = i; }
    }
    const float m = group_max<GL>(bd);
    const bool at_max = bi != kNoIndex && bd == m;
    uint32_t win;
    if (count <= (uint32_t)GL) {
        const uint64_t bal = group_ballot(g, at_max);
        win = bal ? group_first(g, bal) : kNoIndex;
    } else {
        win = group_min_u32<GL>(at_max ? bi : kNoIndex);
    }
    const float d0 = dot3(xyz(hp.v[first]), d);
    if (win == kNoIndex || !(d0 == d0)) win = 0;
    return xyz(hp.v[first + win]);
}

// support_witness (gjk.rs:67-73) with group-parallel hull supports
template <int GL>
__device__ __forceinline__ void support_witness_group(const ShapeRef& a, const ShapeRef& b, f3 d,
                                                      const HullPool& hp, const LaneGroup<GL>& g, f3& point,
                                                      f3& witness) {
    const f3 ap = add(qrot(a.rot, support_local_group<GL>(a, qrot_inv(a.rot, d), hp, g)), a.pos);
    const f3 bp = add(qrot(b.rot, support_local_group<GL>(b, qrot_inv(b.rot, neg(d)), hp, g)), b.pos);
    point = sub(ap, bp);
    witness = ap;
}

template <class L_t> __device__ __forceinline__ f3 vtx(const L_t& L, uint32_t i) { return f3{L.vx[i], L.vy[i], L.vz[i]}; }
template <class L_t> __device__ __forceinline__ f3 wit(const L_t& L, uint32_t i) { return f3{L.wx[i], L.wy[i], L.wz[i]}; }

struct FaceR {
    uint32_t a, b, c;
    f3 n;
    float d;
    bool ok;
};

// Polytope::add_face (epa.rs:58-85) without the store: the face and whether
// it is non-degenerate
template <class L_t> __device__ __forceinline__ FaceR make_face(const L_t& L, uint32_t a, uint32_t b, uint32_t c) {
    FaceR f;
    const f3 va = vtx(L, a);
    f3 n = cross3(sub(vtx(L, b), va), sub(vtx(L, c), va));
    f.ok = dot3(n, n) >= 1e-12f;
    f.a = a; f.b = b; f.c = c;
    if (!f.ok) return f;
    n = normalize3(n);
    f.n = n;
    f.d = dot3(n, va);
    if (f.d < 0.0f) {
        f.a = c;
        f.c = a;
        f.n = neg(n);
        f.d = -f.d;
    }
    return f;
}
template <class L_t> __device__ __forceinline__ void store_face(L_t& L, uint32_t slot, const FaceR& f) {
    L.fa[slot] = f.a; L.fb[slot] = f.b; L.fc[slot] = f.c;
    L.fnx[slot] = f.n.x; L.fny[slot] = f.n.y; L.fnz[slot] = f.n.z; L.fd[slot] = f.d;
}

__device__ __forceinline__ bool edge_in(uint32_t ea, uint32_t eb, uint32_t fa, uint32_t fb, uint32_t fc) {
    return (fa == ea && fb == eb) || (fb == ea && fc == eb) || (fc == ea && fa == eb) || (fa == eb && fb == ea) ||
           (fb == eb && fc == ea) || (fc == eb && fa == ea);
}

// epa.rs:193-240
template <class L_t>
__device__ void face_contact(const L_t& L, uint32_t fi, f3& normal, float& depth, f3& pa, f3& pb) {
    const uint32_t ia = L.fa[fi], ib = L.fb[fi], ic = L.fc[fi];
    const f3 fn{L.fnx[fi], L.fny[fi], L.fnz[fi]};
    const float fdist = L.fd[fi];
    const f3 va = vtx(L, ia), vb = vtx(L, ib), vc = vtx(L, ic);
    const f3 closest = scale(fn, fdist);
    const f3 v0 = sub(vb, va), v1 = sub(vc, va), v2 = sub(closest, va);
    const float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
    const float denom = d00 * d11 - d01 * d01;
    float w1, w2, w3;
    if (fabsf(denom) > 1e-12f) {
        w2 = (d11 * d20 - d01 * d21) / denom;
        w3 = (d00 * d21 - d01 * d20) / denom;
        w1 = 1.0f - w2 - w3;
    } else {
        w1 = 1.0f; w2 = 0.0f; w3 = 0.0f;
    }
    w1 = rclamp(w1, 0.0f, 1.0f);
    w2 = rclamp(w2, 0.0f, 1.0f);
    w3 = rclamp(w3, 0.0f, 1.0f);
    const float sum = w1 + w2 + w3;
    if (sum > 1e-12f) { w1 = w1 / sum; w2 = w2 / sum; w3 = w3 / sum; }
    else { w1 = 1.0f; w2 = 0.0f; w3 = 0.0f; }
    const f3 pw = add(add(scale(wit(L, ia), w1), scale(wit(L, ib), w2)), scale(wit(L, ic), w3));
    const f3 mk = add(add(scale(va, w1), scale(vb, w2)), scale(vc, w3));
    normal = fn;
    depth = fdist;
    pa = pw;
    pb = sub(pw, mk);
}

// EPA (epa.rs:89-189) with a group of GL lanes per GJK-intersecting pair and
// 64 / GL pairs per one-wave workgroup, each with a polytope of MV vertices
// and MF faces in LDS.  FIRST: the pairs of the GJK list; a pair that would
// pass MV or MF (where the caps of epa.rs, 256 and 512, are not yet reached)
// goes to the retry list instead of giving a result.  Otherwise (MV, MF = the
// caps of epa.rs): the pairs of the retry list.  The groups of a wave run the
// loop together; a group whose pair is done (or that has none) stays in it
// with every update masked off until the wave's last pair is done.
template <int GL, int MV, int MF, bool FIRST, class Next>
__device__ __forceinline__ void epa_jobs(const BodyView& bv, const NarrowOut& o, const float4* simplex_store,
                                         const HullPool& hp, EpaLds<MV, MF>* polys, uint32_t jobs, uint32_t j0,
                                         Next next) {
    constexpr int kGroups = kWave / GL;
    constexpr int kFacesPerLane = MF / GL;
    static_assert(MF % GL == 0, "faces split evenly over the lanes");
    const LaneGroup<GL> g = lane_group<GL>();
    const int lane = g.lane;
    EpaLds<MV, MF>& L = polys[threadIdx.x / GL];
    for (; j0 < jobs; j0 = next()) {
        const uint32_t job = j0 + threadIdx.x / GL;
        const bool has = job < jobs;  // the same on every lane of a group
        // a group with no pair computes on the round's first pair, and stores nothing
        const uint32_t p = FIRST ? o.gjk_list[has ? job : j0] : o.epa_retry_list[has ? job : j0];
        const uint2 pr = o.pairs[p];
        const ShapeRef a = load_shape(bv, pr.x);
        const ShapeRef b = load_shape(bv, pr.y);
        if (has && lane < 4) {
            const float4 sp = simplex_store[8 * (size_t)p + lane];
            const float4 sw = simplex_store[8 * (size_t)p + 4 + lane];
            L.vx[lane] = sp.x; L.vy[lane] = sp.y; L.vz[lane] = sp.z;
            L.wx[lane] = sw.x; L.wy[lane] = sw.y; L.wz[lane] = sw.z;
        }
        __syncthreads();
        if (has && lane == 0) {
            // Polytope::new zeroes the faces; then the tetrahedron's four faces
            L.fa[0] = 0; L.fb[0] = 0; L.fc[0] = 0;
            L.fnx[0] = 0.0f; L.fny[0] = 0.0f; L.fnz[0] = 0.0f; L.fd[0] = 0.0f;
            uint32_t fc = 0;
            const uint32_t tri[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
            for (int t = 0; t < 4; ++t) {
                const FaceR f = make_face(L, tri[t][0], tri[t][1], tri[t][2]);
                if (f.ok) store_face(L, fc++, f);
            }
            L.count = fc;
        }
        __syncthreads();
        uint32_t face_count = has ? L.count : 0u;
        uint32_t vertex_count = 4;
        int result_face = -1;
        bool retry = false;
        bool live = has;
        for (int it = 0; it < 64; ++it) {
            if (__ballot(live) == 0ull) break;  // every lane is active here: uniform
            // closest face: first minimal distance in index order (strict <).
            // The minimum value is reduced alone; the first face holding it is
            // found by ballots, face group by face group (face k * GL + lane)
            float md = __builtin_inff();
            float fdk[kFacesPerLane];
            for (int k = 0; k < kFacesPerLane; ++k) {
                const uint32_t f = (uint32_t)(k * GL + lane);
                fdk[k] = f < face_count ? L.fd[f] : __builtin_nanf("");
                if (fdk[k] < md) md = fdk[k];
            }
            const float mval = group_min<GL>(md);
            uint32_t mi = kNoIndex;
            for (int k = 0; k < kFacesPerLane && mi == kNoIndex; ++k) {
                const uint64_t bal = group_ballot(g, fdk[k] == mval);
                if (bal) mi = (uint32_t)(k * GL) + group_first(g, bal);
            }
            // the sequential scan starts from face 0: a NaN there is never
            // replaced (nor is face 0 when no distance is a number)
            if (mi == kNoIndex || face_count == 0 || !(L.fd[0] == L.fd[0])) mi = 0;
            const f3 mn{L.fnx[mi], L.fny[mi], L.fnz[mi]};
            const float mdist = L.fd[mi];
            f3 sp, sw;
            support_witness_group<GL>(a, b, mn, hp, g, sp, sw);
            const float sd = dot3(mn, sp);
            const uint32_t ni = vertex_count;
            if (live) {
                if (sd - mdist < 1e-4f) {
                    result_face = (int)mi;
                    live = false;
                } else if (ni >= (uint32_t)MV) {
                    if (MV < kMaxV) retry = true;
                    else result_face = (int)mi;
                    live = false;
                }
            }
            if (__ballot(live) == 0ull) break;  // the wave's last pair is done (every lane is active here)
            if (live && lane == 0) {
                L.vx[ni] = sp.x; L.vy[ni] = sp.y; L.vz[ni] = sp.z;
                L.wx[ni] = sw.x; L.wy[ni] = sw.y; L.wz[ni] = sw.z;
            }
            if (live) vertex_count += 1;
            // visibility of each face from the new vertex, and the list of the
            // visible faces in index order
            uint32_t nvis = 0;
            for (int k = 0; k < kFacesPerLane; ++k) {
                const uint32_t f = (uint32_t)(k * GL + lane);
                bool v = false;
                if (live && f < face_count) {
                    const f3 fn{L.fnx[f], L.fny[f], L.fnz[f]};
                    v = dot3(fn, sub(sp, vtx(L, L.fa[f]))) > 0.0f;
                    L.vis[f] = v ? 1u : 0u;
                }
                const uint64_t bal = group_ballot(g, v);
                if (v) L.vlist[nvis + (uint32_t)__popcll(bal & g.below)] = (uint16_t)f;
                nvis += (uint32_t)__popcll(bal);
            }
            __syncthreads();
            // horizon: edges of visible faces that no other visible face has,
            // in (face, edge) order, at most MF of them.  Candidate c is edge
            // c % 3 of visible face c / 3; only visible faces can share it.
            uint32_t hbase = 0;
            for (uint32_t cb = 0; cb < 3 * nvis; cb += GL) {
                const uint32_t c = cb + (uint32_t)lane;
                bool hz = false;
                uint32_t ea = 0, eb = 0;
                if (c < 3 * nvis) {
                    const uint32_t f = L.vlist[c / 3], q = c % 3;
                    const uint32_t fa = L.fa[f], fb = L.fb[f], fcc = L.fc[f];
                    ea = q == 0 ? fa : (q == 1 ? fb : fcc);
                    eb = q == 0 ? fb : (q == 1 ? fcc : fa);
                    bool shared = false;
                    for (uint32_t t = 0; t < nvis && !shared; ++t) {
                        const uint32_t h = L.vlist[t];
                        if (h != f && edge_in(ea, eb, L.fa[h], L.fb[h], L.fc[h])) shared = true;
                    }
                    hz = !shared;
                }
                const uint64_t bal = group_ballot(g, hz);
                const uint32_t pos = hbase + (uint32_t)__popcll(bal & g.below);
                if (hz && pos < (uint32_t)MF) { L.ha[pos] = ea; L.hb[pos] = eb; }
                hbase += (uint32_t)__popcll(bal);
            }
            if (live && MF < kMaxF && hbase > (uint32_t)MF) {
                retry = true;
                live = false;
            }
            if (__ballot(live) == 0ull) break;
            const uint32_t hcount = hbase < (uint32_t)MF ? hbase : (uint32_t)MF;
            // compact the kept (non-visible) faces in index order
            FaceR keep[kFacesPerLane];
            uint32_t kflag = 0;
            for (int k = 0; k < kFacesPerLane; ++k) {
                const uint32_t f = (uint32_t)(k * GL + lane);
                if (live && f < face_count && !L.vis[f]) {
                    kflag |= 1u << k;
                    keep[k].a = L.fa[f]; keep[k].b = L.fb[f]; keep[k].c = L.fc[f];
                    keep[k].n = f3{L.fnx[f], L.fny[f], L.fnz[f]};
                    keep[k].d = L.fd[f];
                }
            }
            __syncthreads();
            uint32_t wbase = 0;
            for (int k = 0; k < kFacesPerLane; ++k) {
                const bool kf = (kflag >> k) & 1u;
                const uint64_t bal = group_ballot(g, kf);
                if (kf) store_face(L, wbase + (uint32_t)__popcll(bal & g.below), keep[k]);
                wbase += (uint32_t)__popcll(bal);
            }
            if (live) face_count = wbase;
            __syncthreads();
            // stitch the horizon to the new vertex (epa.rs:182-185)
            const uint32_t hn = live ? hcount : 0u;
            for (uint32_t hb0 = 0; hb0 < hn; hb0 += GL) {
                const uint32_t h = hb0 + lane;
                FaceR f;
                f.ok = false;
                if (h < hn) f = make_face(L, ni, L.ha[h], L.hb[h]);
                const uint64_t bal = group_ballot(g, f.ok);
                const uint32_t slot = face_count + (uint32_t)__popcll(bal & g.below);
                if (f.ok && slot < (uint32_t)MF) store_face(L, slot, f);
                const uint32_t added = face_count + (uint32_t)__popcll(bal);
                if (MF < kMaxF && added > (uint32_t)MF) retry = true;
                face_count = added < (uint32_t)MF ? added : (uint32_t)MF;
            }
            if (retry) live = false;
            __syncthreads();
        }
        if (retry && lane == 0) o.epa_retry_list[atomicAdd(o.epa_retry_count, 1u)] = p;
        if (result_face >= 0 && lane == 0) {
            f3 n, pa, pb;
            float depth;
            face_contact(L, (uint32_t)result_face, n, depth, pa, pb);
            write_slot(o, p * 4, n, depth, pa, pb);
            o.cnt[p] = 1;
        }
        __syncthreads();
    }
}

// The first pass over the GJK list.  With at most kEpaFewPairs pairs, a
// wavefront per pair and a polytope of kFewV vertices and kFewF faces;
// otherwise four pairs per wavefront, 16 lanes and a polytope of kSmallV
// vertices and kSmallF faces each.  Workgroup b takes job b (or jobs 4 b ..
// 4 b + 3), then the ones a grid further on, and so on.  (A queue claimed by
// atomics instead was slower: the claims on one word from every XCD
// serialise.)
template <bool LDS_POOL>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4)))
k_epa_first(BodyView bv, NarrowOut o, const float4* simplex_store) {
    union Polys {
        EpaLds<kFewV, kFewF> few;
        EpaLds<kSmallV, kSmallF> many[kWave / kEpaLanes];
    };
    __shared__ Polys polys;
    const uint32_t jobs = *o.gjk_count;
    const bool few = jobs <= kEpaFewPairs;
    const uint32_t per = few ? 1u : (uint32_t)(kWave / kEpaLanes);  // jobs per claim
    if (blockIdx.x * per >= jobs) return;  // uniform: a workgroup with no pair stages nothing
    const HullPool hp = stage_pool<LDS_POOL>(bv);
    uint32_t j = blockIdx.x * per;
    auto next = [&]() { return j += gridDim.x * per; };
    if (few)
        epa_jobs<kWave, kFewV, kFewF, true>(bv, o, simplex_store, hp, &polys.few, jobs, j, next);
    else
        epa_jobs<kEpaLanes, kSmallV, kSmallF, true>(bv, o, simplex_store, hp, polys.many, jobs, j, next);
}

// The retry pass: a wavefront per pair at the caps of epa.rs, workgroup b
// taking pairs b, b + grid, ...
template <bool LDS_POOL>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(1)))
k_epa_retry(BodyView bv, NarrowOut o, const float4* simplex_store) {
    __shared__ EpaLds<kMaxV, kMaxF> poly;
    const uint32_t jobs = *o.epa_retry_count;
    if (blockIdx.x >= jobs) return;
    const HullPool hp = stage_pool<LDS_POOL>(bv);
    uint32_t j = blockIdx.x;
    epa_jobs<kWave, kMaxV, kMaxF, false>(bv, o, simplex_store, hp, &poly, jobs, j, [&]() { return j += gridDim.x; });
}

__global__ void __launch_bounds__(kBlock) k_pack_contact_out(NarrowOut o, uint32_t np, dio_gpu_contact_out* rec) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= np * 4) return;
    const uint32_t p = s / 4, k = s % 4;
    dio_gpu_contact_out r;
    r.pad = 0;
    r.pad2[0] = r.pad2[1] = r.pad2[2] = 0.0f;
    if (k < o.cnt[p]) {
        const uint2 pr = o.pairs[p];
        const float4 n = o.n[s], pa = o.pa[s], pb = o.pb[s];
        r.valid = 1; r.a = pr.x; r.b = pr.y;
        r.normal[0] = n.x; r.normal[1] = n.y; r.normal[2] = n.z; r.normal[3] = 0.0f;
        r.depth = n.w;
        r.point_a[0] = pa.x; r.point_a[1] = pa.y; r.point_a[2] = pa.z; r.point_a[3] = 0.0f;
        r.point_b[0] = pb.x; r.point_b[1] = pb.y; r.point_b[2] = pb.z; r.point_b[3] = 0.0f;
    } else {
        r.valid = 0; r.a = 0; r.b = 0;
        for (int q = 0; q < 4; ++q) { r.normal[q] = 0.0f; r.point_a[q] = 0.0f; r.point_b[q] = 0.0f; }
        r.depth = 0.0f;
    }
    rec[s] = r;
}

}  // namespace

void run_narrow(const BodyView& bv, const NarrowOut& o, float4* simplex_store, hipStream_t s, bool counters_cleared,
                bool gjk_possible) {
    if (!counters_cleared) launch_fill2_u32(o.gjk_count, o.epa_retry_count, 0u, 1, s);
    uint32_t blocks = blocks_for(o.p_max > 0 ? o.p_max : 1);
    if (blocks > 4096) blocks = 4096;
    const bool lds = bv.hull != nullptr && bv.hull_count > 0 && bv.hull_count <= kLdsPoolMax;
    const size_t pool_bytes = lds ? (size_t)bv.hull_count * sizeof(float4) : 0;
    if (lds) hipLaunchKernelGGL(k_narrow<true>, dim3(blocks), dim3(kBlock), pool_bytes, s, bv, o, simplex_store);
    else hipLaunchKernelGGL(k_narrow<false>, dim3(blocks), dim3(kBlock), 0, s, bv, o, simplex_store);
    if (gjk_possible) {
        const uint32_t np = o.p_max > 0 ? o.p_max : 1;
        const uint32_t eblocks = np < kEpaBlocks ? np : kEpaBlocks;
        const uint32_t rblocks = np < kRetryBlocks ? np : kRetryBlocks;
        const float4* ss = simplex_store;
        if (lds) {
            hipLaunchKernelGGL(k_epa_first<true>, dim3(eblocks), dim3(kWave), pool_bytes, s, bv, o, ss);
            hipLaunchKernelGGL(k_epa_retry<true>, dim3(rblocks), dim3(kWave), pool_bytes, s, bv, o, ss);
        } else {
            hipLaunchKernelGGL(k_epa_first<false>, dim3(eblocks), dim3(kWave), 0, s, bv, o, ss);
            hipLaunchKernelGGL(k_epa_retry<false>, dim3(rblocks), dim3(kWave), 0, s, bv, o, ss);
        }
    }
    launch_scan_exclusive(o.cnt, o.off, o.p_max, o.p_dev, o.scan_partials, s);
}

void launch_pack_contact_out(const NarrowOut& o, uint32_t np, dio_gpu_contact_out* rec, hipStream_t s) {
    if (np == 0) return;
    hipLaunchKernelGGL(k_pack_contact_out, dim3(blocks_for(np * 4)), dim3(kBlock), 0, s, o, np, rec);
}

}  // namespace dio
