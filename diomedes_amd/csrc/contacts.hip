// contacts.hip — the contact build of one step, on the device.
//
// The reference does this on the CPU every frame: raw_contact for each
// decoded narrow contact (physics/gpu.rs:239-264), container_contacts for
// each dynamic body (:352-404), then finalize_contacts (:266-311): a HashMap
// warm start in which the LAST old contact of a key (a, b) is kept and the
// FIRST new contact of that key takes it, the alpha*gamma / gamma scaling,
// and the CSR in contact order.
//
// Here both contact lists are ordered runs of keys — narrow contacts in
// (sorted pair, slot) order and container contacts by body — so "first new
// contact of a key" is slot 0 of a pair (or the first wall contact of a
// body) and "last old contact of a key" is found by binary search of the old
// pair list (or from the old per-body container offsets).  The CSR is a
// histogram, a scan, an atomic fill and a per-body sort of each segment, which
// reproduces the reference's ascending contact order.
#include <hip/hip_runtime.h>

#include "contacts.hpp"
#include "prim.hpp"
#include "tree_refit.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

__device__ __forceinline__ void put_contact(const ContactArrays& c, uint32_t i, uint32_t a, uint32_t b, f3 n, f3 t1,
                                            f3 t2, f3 ra, f3 rb, float c0x, float friction, float4 lam, float4 pen) {
    c.ab[i] = make_uint2(a, b);
    c.n[i] = f4(n, friction);
    c.t1[i] = f4(t1, 0.0f);
    c.t2[i] = f4(t2, 0.0f);
    c.ra[i] = f4(ra, c0x);  // c0 rides in ra.w for the solve
    c.rb[i] = f4(rb, 0.0f);
    c.c0[i] = make_float4(c0x, 0.0f, 0.0f, 0.0f);
    c.lam[i] = lam;
    c.pen[i] = pen;
}

// finalize_contacts scaling (physics/gpu.rs:278-282)
__device__ __forceinline__ void warm_scale(const BuildOptions& o, float4& lam, float4& pen) {
    const float ag = o.alpha * o.gamma;
    lam = make_float4(lam.x * ag, lam.y * ag, lam.z * ag, 0.0f);
    pen = make_float4(rmin(rmax(pen.x * o.gamma, o.penalty_min), o.penalty_max),
                      rmin(rmax(pen.y * o.gamma, o.penalty_min), o.penalty_max),
                      rmin(rmax(pen.z * o.gamma, o.penalty_min), o.penalty_max), 0.0f);
}

__device__ __forceinline__ bool pair_less(uint2 q, uint32_t a, uint32_t b) { return q.x < a || (q.x == a && q.y < b); }

// index of the last old contact with key (a, b) among the narrow prefix, or -1.
// The old and new pair lists are both sorted and change little from step to
// step, so the lower bound of (a, b) among the old pairs is searched by
// galloping out from `hint` (the new pair's own index) and then bisecting the
// bracket found: a few dependent loads instead of log2(np).
__device__ __forceinline__ int64_t find_old_pair(const NarrowSource& old_src, uint32_t a, uint32_t b, uint32_t hint) {
    const uint32_t np = *old_src.p_dev;
    // bracket [lo, hi) with pairs[lo - 1] < (a, b) <= pairs[hi] (ends open)
    uint32_t lo = 0, hi = np;
    if (np > 0) {
        uint32_t h = hint < np ? hint : np - 1;
        if (pair_less(old_src.pairs[h], a, b)) {
            lo = h + 1;
            for (uint32_t step = 1; lo < np; step <<= 1) {
                const uint32_t probe = lo + step - 1 < np ? lo + step - 1 : np - 1;
                if (!pair_less(old_src.pairs[probe], a, b)) {
                    hi = probe;
                    break;
                }
                lo = probe + 1;
            }
        } else {
            hi = h;
            for (uint32_t step = 1; hi > 0; step <<= 1) {
                const uint32_t probe = hi >= step ? hi - step : 0;
                if (pair_less(old_src.pairs[probe], a, b)) {
                    lo = probe + 1;
                    break;
                }
                hi = probe;
            }
        }
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pair_less(old_src.pairs[mid], a, b)) lo = mid + 1;
        else hi = mid;
    }
    if (lo < np) {
        const uint2 q = old_src.pairs[lo];
        if (q.x == a && q.y == b) {
            const uint32_t s0 = old_src.off[lo], s1 = old_src.off[lo + 1];
            if (s1 > s0) return (int64_t)s1 - 1;
        }
    }
    return -1;
}

// The narrow-phase contacts, on workgroups bid = 0 .. nblk - 1 of the launch.
__device__ __forceinline__ void contacts_narrow_part(uint32_t bid, uint32_t nblk, const WorldView& w,
                                                     const BuildOptions& o, const NarrowSource& src,
                                                     const ContactSet& dst, const NarrowSource& old_src,
                                                     const ContactSet& old, uint32_t* overflow, uint32_t* csr_cnt) {
    const uint32_t np = *src.p_dev;
    if (bid == 0 && threadIdx.x == 0) {
        const uint32_t t = src.off[np];
        dst.m_narrow[0] = t < dst.m_cap ? t : dst.m_cap;
        dst.m_narrow[1] = t;
        if (t > dst.m_cap) atomicMax(overflow, t);
    }
    const uint32_t old_m = *old.m_dev;
    for (uint32_t p = bid * kBlock + threadIdx.x; p < np; p += nblk * kBlock) {
        const uint32_t base = src.off[p], cnt = src.off[p + 1] - base;
        if (cnt == 0) continue;
        const uint2 pr = src.pairs[p];
        const f3 pa_pos = xyz(w.pos[pr.x]), pb_pos = xyz(w.pos[pr.y]);
        const q4 qa = as_q(w.rot[pr.x]), qb = as_q(w.rot[pr.y]);
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t ci = base + k;
            if (ci >= dst.m_cap) break;
            const float4 n4 = src.n[p * 4 + k];
            const f3 normal = xyz(n4);
            f3 t1, t2;
            tangents(normal, t1, t2);
            const f3 ra = qrot_inv(qa, sub(xyz(src.pa[p * 4 + k]), pa_pos));
            const f3 rb = qrot_inv(qb, sub(xyz(src.pb[p * 4 + k]), pb_pos));
            float4 lam = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float4 pen = make_float4(o.penalty_min, o.penalty_min, o.penalty_min, 0.0f);
            if (k == 0) {
                const int64_t q = find_old_pair(old_src, pr.x, pr.y, p);
                if (q >= 0 && (uint32_t)q < old_m) {
                    lam = old.c.lam[q];
                    pen = old.c.pen[q];
                }
            }
            warm_scale(o, lam, pen);
            put_contact(dst.c, ci, pr.x, pr.y, neg(normal), t1, t2, ra, rb, o.margin - n4.w, o.friction, lam, pen);
        }
        // each written contact is a CSR entry of both bodies
        const uint32_t written = base < dst.m_cap ? min(cnt, dst.m_cap - base) : 0u;
        if (written) {
            atomicAdd(&csr_cnt[pr.x], written);
            atomicAdd(&csr_cnt[pr.y], written);
        }
    }
}

// container_contacts (physics/gpu.rs:352-404), one wall at a time: the
// wall's candidate points (a cube or box's face towards it, cube_support_face
// physics/gpu.rs:314-350; else one support point), each point's depth
// max(current, predicted) and the wall's contact frame.  A point is a contact
// when its depth is positive; a body's contacts run in wall, then point order.
struct WallPoints {
    f3 pts[4];
    float depth[4];
    uint32_t count;
    f3 normal, t1, t2;
};
__device__ __forceinline__ ShapeRef body_shape(const WorldView& w, uint32_t i) {
    ShapeRef s;
    s.pos = xyz(w.pos[i]);
    s.rot = as_q(w.rot[i]);
    s.tag = w.tag[i];
    s.c[0] = w.corners[4 * i];
    s.c[1] = w.corners[4 * i + 1];
    s.c[2] = w.corners[4 * i + 2];
    s.c[3] = w.corners[4 * i + 3];
    return s;
}
__device__ __forceinline__ void wall_points(const WorldView& w, const ShapeRef& s, f3 vel, const ContainerParams& ct,
                                            const BuildOptions& o, int wall, WallPoints& out) {
    const q4 crot{ct.qx, ct.qy, ct.qz, ct.qw};
    const f3 center{ct.cx, ct.cy, ct.cz};
    const float sign = (wall & 1) ? -1.0f : 1.0f;
    const int axis = wall >> 1;
    const f3 unit{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    const f3 outward = qrot(crot, scale(unit, sign));
    const float plane = dot3(center, outward) + ct.half;
    if (s.tag == kCube || s.tag == kBox) {
        const f3 ax0 = qrot(s.rot, f3{1.0f, 0.0f, 0.0f});
        const f3 ax1 = qrot(s.rot, f3{0.0f, 1.0f, 0.0f});
        const f3 ax2 = qrot(s.rot, f3{0.0f, 0.0f, 1.0f});
        uint32_t ai = 0;
        float best = fabsf(dot3(ax0, outward));
        const float al1 = fabsf(dot3(ax1, outward));
        if (al1 > best) { ai = 1; best = al1; }
        const float al2 = fabsf(dot3(ax2, outward));
        if (al2 > best) { ai = 2; best = al2; }
        const f3 fa = sel3(ax0, ax1, ax2, ai);
        const f3 fnrm = scale(fa, dot3(fa, outward) >= 0.0f ? 1.0f : -1.0f);
        const uint32_t iu = ai == 0 ? 1u : 0u;
        const uint32_t iv = ai == 2 ? 1u : 2u;
        const float hn = box_half(s, ai), hu = box_half(s, iu), hv = box_half(s, iv);
        const f3 c = add(s.pos, scale(fnrm, hn));
        const f3 u = sel3(ax0, ax1, ax2, iu), v = sel3(ax0, ax1, ax2, iv);
        out.pts[0] = add(add(c, scale(u, hu)), scale(v, hv));
        out.pts[1] = add(sub(c, scale(u, hu)), scale(v, hv));
        out.pts[2] = sub(sub(c, scale(u, hu)), scale(v, hv));
        out.pts[3] = sub(add(c, scale(u, hu)), scale(v, hv));
        out.count = 4;
    } else {
        const f3 local = qrot_inv(s.rot, outward);
        out.pts[0] = add(qrot(s.rot, support_host(s, local, HullPool{w.hull})), s.pos);
        // the unused slots too: a struct left partly undefined on one path
        // stays in scratch memory
        out.pts[1] = out.pts[0];
        out.pts[2] = out.pts[0];
        out.pts[3] = out.pts[0];
        out.count = 1;
    }
    out.normal = neg(outward);
    tangents(out.normal, out.t1, out.t2);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        out.depth[k] = 0.0f;
        if (k < out.count) {
            const float cur = dot3(out.pts[k], outward) - plane;
            const float pred = dot3(add(out.pts[k], scale(vel, o.dt)), outward) - plane;
            out.depth[k] = rmax(cur, pred);
        }
    }
}
__device__ __forceinline__ uint32_t wall_contact_count(const WallPoints& wp) {
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) c += (k < wp.count && wp.depth[k] > 0.0f) ? 1u : 0u;
    return c;
}

// Container contacts run kWallLanes lanes per body, lane l on wall l (lanes 6
// and 7 idle): six times the parallelism of a lane per body, with each wall's
// arithmetic unchanged.
constexpr int kWallLanes = 8;

// Each body's container contact count, on workgroup bid of its part.
__device__ __forceinline__ void container_count_part(uint32_t bid, const WorldView& w, const ContainerParams& ct,
                                                     const BuildOptions& o, uint32_t* cnt) {
    const uint32_t t = bid * kBlock + threadIdx.x;
    const uint32_t i = t / kWallLanes;
    const int wall = (int)(t % kWallLanes);
    if (i >= w.n) return;  // whole groups of kWallLanes lanes leave together
    uint32_t c = 0;
    const ContainerParams cp = w.scene_ct ? w.scene_ct[w.scene_of[i]] : ct;
    if (wall < 6 && cp.enabled && w.inv_mm[i].w != 0.0f) {
        WallPoints wp;
        wall_points(w, body_shape(w, i), xyz(w.vel[i]), cp, o, wall, wp);
        c = wall_contact_count(wp);
    }
    c += __shfl_xor(c, 1, kWallLanes);
    c += __shfl_xor(c, 2, kWallLanes);
    c += __shfl_xor(c, 4, kWallLanes);
    if (wall == 0) cnt[i] = c;
}

// The narrow-phase contacts, the container counts and (on refit steps) the
// tree refit's chunks are independent: one launch runs them all, the first pb
// workgroups on the pairs, the next nb on the bodies (kWallLanes lanes each),
// the rest on the refit chunks; whole workgroups per part.
__global__ void __launch_bounds__(kBlock) k_contacts_build(WorldView w, BuildOptions o, NarrowSource src,
                                                           ContactSet dst, NarrowSource old_src, ContactSet old,
                                                           uint32_t* overflow, uint32_t* csr_cnt, ContainerParams ct,
                                                           uint32_t pb, uint32_t nb, TreeLower refit) {
    static_assert(kBlock == 256, "the refit part runs on 256 threads");
    if (blockIdx.x < pb) contacts_narrow_part(blockIdx.x, pb, w, o, src, dst, old_src, old, overflow, csr_cnt);
    else if (blockIdx.x < pb + nb) container_count_part(blockIdx.x - pb, w, ct, o, dst.cont_cnt);
    else tree_lower_part(blockIdx.x - pb - nb, refit);
}

__global__ void __launch_bounds__(kBlock) k_container_write(WorldView w, ContainerParams ct, BuildOptions o,
                                                            ContactSet dst, ContactSet old, uint32_t* overflow,
                                                            uint32_t* csr_cnt) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t i = t / kWallLanes;
    const int wall = (int)(t % kWallLanes);
    if (i >= w.n) return;  // whole groups of kWallLanes lanes leave together
    const uint32_t mn = *dst.m_narrow;
    if (t == 0) {
        const uint32_t tot = mn + dst.cont_off[w.n];
        *dst.m_dev = tot < dst.m_cap ? tot : dst.m_cap;
        const uint32_t need = dst.m_narrow[1] + dst.cont_off[w.n];  // with the narrow contacts cut off by m_cap
        if (need > dst.m_cap) atomicMax(overflow, need);
    }
    if (dst.cont_off[i + 1] == dst.cont_off[i]) return;  // the whole group: body i has no wall contact
    WallPoints wp;
    wp.count = 0;
    const ShapeRef s = body_shape(w, i);
    const ContainerParams cp = w.scene_ct ? w.scene_ct[w.scene_of[i]] : ct;
    if (wall < 6) wall_points(w, s, xyz(w.vel[i]), cp, o, wall, wp);
    const uint32_t mine = wall < 6 ? wall_contact_count(wp) : 0u;
    // exclusive prefix of the counts over the group's walls
    uint32_t incl = mine;
#pragma unroll
    for (int d = 1; d < kWallLanes; d <<= 1) {
        const uint32_t up = __shfl_up(incl, d, kWallLanes);
        if (wall >= d) incl += up;
    }
    if (mine == 0) return;
    uint32_t before = incl - mine;
    // warm value of key (i, CONTAINER): the last old container contact of body
    // i, taken by the body's first new one
    float4 wl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 wpn = make_float4(o.penalty_min, o.penalty_min, o.penalty_min, 0.0f);
    if (before == 0) {
        const uint32_t o0 = old.cont_off[i], o1 = old.cont_off[i + 1];
        if (o1 > o0) {
            const uint32_t q = *old.m_narrow + o1 - 1;
            if (q < *old.m_dev) {
                wl = old.c.lam[q];
                wpn = old.c.pen[q];
            }
        }
    }
    uint32_t ci = mn + dst.cont_off[i] + before;
    uint32_t written = 0;  // this lane's CSR entries of body i
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= wp.count || !(wp.depth[k] > 0.0f)) continue;
        if (ci < dst.m_cap) {
            float4 lam = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float4 pen = make_float4(o.penalty_min, o.penalty_min, o.penalty_min, 0.0f);
            if (before == 0) { lam = wl; pen = wpn; }
            warm_scale(o, lam, pen);
            const f3 ra = qrot_inv(s.rot, sub(wp.pts[k], s.pos));
            put_contact(dst.c, ci, i, kContainer, wp.normal, wp.t1, wp.t2, ra, f3{0.0f, 0.0f, 0.0f},
                        o.margin - wp.depth[k], o.friction, lam, pen);
            ++written;
        }
        ++before;
        ++ci;
    }
    if (written) atomicAdd(&csr_cnt[i], written);
}

__global__ void __launch_bounds__(kBlock) k_csr_fill(ContactSet cs, const uint32_t* off, uint32_t* cur, uint32_t* tmp,
                                                     uint32_t idx_cap) {
    const uint32_t m = *cs.m_dev;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < m; c += gridDim.x * kBlock) {
        const uint2 ab = cs.c.ab[c];
        const uint32_t sa = off[ab.x] + atomicAdd(&cur[ab.x], 1u);
        if (sa < idx_cap) tmp[sa] = c;
        if (ab.y != kContainer) {
            const uint32_t sb = off[ab.y] + atomicAdd(&cur[ab.y], 1u);
            if (sb < idx_cap) tmp[sb] = c;
        }
    }
}

__device__ __forceinline__ uint32_t peer_code(uint2 ab, uint32_t body) {
    return ab.x == body ? ab.y : (ab.x | kPeerIsB);
}

// Each body's entries arrive in atomic order; an entry's sorted position is
// its rank among the body's entries (contact indices are distinct within a
// body), computed by kCsrLanes lanes per body.  Also writes the peer code of
// every entry (the other body, flagged when this body is the contact's b).
constexpr int kCsrLanes = 8;
__global__ void __launch_bounds__(kBlock) k_csr_rank(uint32_t n, const uint32_t* off, const uint32_t* tmp,
                                                     const uint2* ab, uint32_t* idx, uint32_t* peer, uint32_t idx_cap,
                                                     uint32_t* overflow, uint32_t* cnt, uint32_t* cur) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    if (g == 0 && off[n] > idx_cap) atomicMax(overflow, off[n]);
    const uint32_t i = g / kCsrLanes, lane = g % kCsrLanes;
    if (i >= n) return;  // whole groups of kCsrLanes lanes leave together
    if (lane == 0) {
        // the counters' last reader was the fill: zero them for the next step's writers
        cnt[i] = 0u;
        cur[i] = 0u;
    }
    const uint32_t b = off[i];
    uint32_t e = off[i + 1];
    if (e > idx_cap) e = idx_cap;
    // The group walks the body's entries kCsrLanes at a time: each lane loads
    // one and the others read it by shuffles, so a chunk costs one load.
    for (uint32_t x0 = b; x0 < e; x0 += kCsrLanes) {
        const uint32_t x = x0 + lane;
        const uint32_t c = x < e ? tmp[x] : 0xFFFFFFFFu;
        uint32_t r = 0;
        for (uint32_t y0 = b; y0 < e; y0 += kCsrLanes) {
            const uint32_t y = y0 + lane;
            const uint32_t mine = y < e ? tmp[y] : 0xFFFFFFFFu;  // past the end: ranks nothing
#pragma unroll
            for (int t = 0; t < kCsrLanes; ++t) r += __shfl(mine, t, kCsrLanes) < c ? 1u : 0u;
        }
        if (x < e) {
            idx[b + r] = c;
            peer[b + r] = peer_code(ab[c], i);
        }
    }
}

// peer codes of a CSR given in sorted form (the granular solve)
__global__ void __launch_bounds__(kBlock) k_csr_peers(uint32_t n, const uint32_t* off, const uint32_t* idx,
                                                      const uint2* ab, uint32_t* peer) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t i = g / kCsrLanes, lane = g % kCsrLanes;
    if (i >= n) return;
    const uint32_t e = off[i + 1];
    for (uint32_t x = off[i] + lane; x < e; x += kCsrLanes) peer[x] = peer_code(ab[idx[x]], i);
}

__global__ void __launch_bounds__(kBlock) k_unpack_contacts(const dio_avbd_contact* rec, uint32_t m, ContactArrays c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const dio_avbd_contact& r = rec[i];
    c.ab[i] = make_uint2(r.a, r.b);
    c.n[i] = make_float4(r.normal[0], r.normal[1], r.normal[2], r.friction);
    c.t1[i] = make_float4(r.tangent1[0], r.tangent1[1], r.tangent1[2], 0.0f);
    c.t2[i] = make_float4(r.tangent2[0], r.tangent2[1], r.tangent2[2], 0.0f);
    c.ra[i] = make_float4(r.r_a[0], r.r_a[1], r.r_a[2], 0.0f);
    c.rb[i] = make_float4(r.r_b[0], r.r_b[1], r.r_b[2], 0.0f);
    c.c0[i] = make_float4(r.c0[0], r.c0[1], r.c0[2], 0.0f);
    c.pen[i] = make_float4(r.penalty[0], r.penalty[1], r.penalty[2], 0.0f);
    c.lam[i] = make_float4(r.lambda[0], r.lambda[1], r.lambda[2], 0.0f);
}

}  // namespace

void run_contact_pairs(const WorldView& w, const ContainerParams& ct, const BuildOptions& o, const NarrowSource& src,
                       const ContactSet& dst, const NarrowSource& old_src, const ContactSet& old, uint32_t* overflow,
                       hipStream_t s, uint32_t* csr_cnt, const TreeLower* refit) {
    uint32_t pb = blocks_for(dst.m_cap > 0 ? dst.m_cap / 4 + 1 : 1);
    if (pb > 2048) pb = 2048;
    const uint32_t nb = blocks_for((w.n > 0 ? w.n : 1) * (uint32_t)kWallLanes);
    TreeLower tl{};
    if (refit) tl = *refit;
    hipLaunchKernelGGL(k_contacts_build, dim3(pb + nb + tl.blocks), dim3(kBlock), 0, s, w, o, src, dst, old_src, old,
                       overflow, csr_cnt, ct, pb, nb, tl);
}

void run_contact_walls(const WorldView& w, const ContainerParams& ct, const BuildOptions& o, const ContactSet& dst,
                       const ContactSet& old, uint32_t* scan_partials, uint32_t* overflow, hipStream_t s,
                       uint32_t* csr_cnt) {
    const uint32_t nb = blocks_for((w.n > 0 ? w.n : 1) * (uint32_t)kWallLanes);
    launch_scan_exclusive(dst.cont_cnt, dst.cont_off, w.n, nullptr, scan_partials, s);
    hipLaunchKernelGGL(k_container_write, dim3(nb), dim3(kBlock), 0, s, w, ct, o, dst, old, overflow, csr_cnt);
}

void run_csr(const ContactSet& cs, uint32_t n, const CsrBuffers& b, uint32_t* overflow, hipStream_t s) {
    uint32_t mb = blocks_for(cs.m_cap > 0 ? cs.m_cap : 1);
    if (mb > 2048) mb = 2048;
    launch_scan_exclusive(b.cnt, b.off, n, nullptr, b.scan_partials, s);
    hipLaunchKernelGGL(k_csr_fill, dim3(mb), dim3(kBlock), 0, s, cs, (const uint32_t*)b.off, b.cur, b.tmp, b.idx_cap);
    const uint32_t rb = blocks_for((n > 0 ? n : 1) * (uint32_t)kCsrLanes);
    hipLaunchKernelGGL(k_csr_rank, dim3(rb), dim3(kBlock), 0, s, n, (const uint32_t*)b.off, (const uint32_t*)b.tmp,
                       (const uint2*)cs.c.ab, b.idx, b.peer, b.idx_cap, overflow, b.cnt, b.cur);
}

void launch_csr_peers(uint32_t n, const uint32_t* off, const uint32_t* idx, const uint2* ab, uint32_t* peer,
                      hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_csr_peers, dim3(blocks_for(n * (uint32_t)kCsrLanes)), dim3(kBlock), 0, s, n, off, idx, ab,
                       peer);
}

void launch_unpack_contacts(const dio_avbd_contact* rec, uint32_t m, const ContactArrays& c, hipStream_t s) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_unpack_contacts, dim3(blocks_for(m)), dim3(kBlock), 0, s, rec, m, c);
}

}  // namespace dio
