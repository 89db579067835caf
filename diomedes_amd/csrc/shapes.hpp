// shapes.hpp — shape support functions and the box-box SAT manifold.
//
// Counterparts of engine/gpu_physics/narrowphase/src/{support,cuboid}.rs and
// engine/gpu_physics/broadphase/src/lib.rs:79-151, evaluated in the same
// float operation order.  BOX (per-axis half extents) and HULL (vertex pool)
// extend the reference's three shapes; a CUBE keeps the reference's scalar
// arithmetic exactly (SAT and walls read only its first half extent).
#pragma once

#include "dio_math.hpp"

namespace dio {

// A body as the collision code sees it: pose + shape.
struct ShapeRef {
    f3 pos;
    q4 rot;
    uint32_t tag;
    float4 c[4];  // GpuShape corners
};

struct HullPool {
    const float4* v;  // xyz per vertex
};

// narrowphase/src/support.rs:21-61: tetrahedra and hulls keep the first
// maximal vertex
DIO_HD f3 support_local(const ShapeRef& s, f3 d, const HullPool& hp) {
    switch (s.tag) {
        case kSphere:
            return scale(normalize_or_zero3(d), s.c[0].x);
        case kCube:
        case kBox:
            return f3{d.x >= 0.0f ? s.c[0].x : -s.c[0].x, d.y >= 0.0f ? s.c[0].y : -s.c[0].y,
                      d.z >= 0.0f ? s.c[0].z : -s.c[0].z};
        case kTetrahedron: {
            f3 best = xyz(s.c[0]);
            float bd = dot3(best, d);
            for (int i = 1; i < 4; ++i) {
                const f3 cc = xyz(s.c[i]);
                const float dd = dot3(cc, d);
                if (dd > bd) { bd = dd; best = cc; }
            }
            return best;
        }
        case kHull: {
            // the first maximal vertex (strict >, in vertex order; a NaN at
            // vertex 0 is never replaced).  Four vertices per round: their
            // loads and dot products are independent, and the comparisons
            // keep the scan's order.
            const uint32_t first = (uint32_t)s.c[0].x, count = (uint32_t)s.c[0].y;
            const float4* v = hp.v + first;
            uint32_t bi = 0;
            float bd = dot3(xyz(v[0]), d);
            uint32_t i = 1;
            for (; i + 4 <= count; i += 4) {
                const float d0 = dot3(xyz(v[i]), d), d1 = dot3(xyz(v[i + 1]), d);
                const float d2 = dot3(xyz(v[i + 2]), d), d3 = dot3(xyz(v[i + 3]), d);
                if (d0 > bd) { bd = d0; bi = i; }
                if (d1 > bd) { bd = d1; bi = i + 1; }
                if (d2 > bd) { bd = d2; bi = i + 2; }
                if (d3 > bd) { bd = d3; bi = i + 3; }
            }
            for (; i < count; ++i) {
                const float dd = dot3(xyz(v[i]), d);
                if (dd > bd) { bd = dd; bi = i; }
            }
            return xyz(v[bi]);
        }
        default:
            return f3{0.0f, 0.0f, 0.0f};
    }
}

// physics/shape.rs:20-48 (host Shape::support used by the container walls):
// the tetrahedron keeps the LAST maximal corner under f32::total_cmp.
DIO_HD int32_t total_key(float v) {
    int32_t i = __builtin_bit_cast(int32_t, v);
    return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
DIO_HD f3 support_host(const ShapeRef& s, f3 d, const HullPool& hp) {
    if (s.tag == kCube) {
        const float h = s.c[0].x;
        return f3{d.x >= 0.0f ? h : -h, d.y >= 0.0f ? h : -h, d.z >= 0.0f ? h : -h};
    }
    if (s.tag == kTetrahedron) {
        // the winning corner itself, not its index (a runtime index into the
        // corners would place the shape in scratch memory)
        f3 best = xyz(s.c[0]);
        int32_t bk = total_key(dot3(best, d));
        for (int i = 1; i < 4; ++i) {
            const f3 cc = xyz(s.c[i]);
            const int32_t k = total_key(dot3(cc, d));
            if (k >= bk) { bk = k; best = cc; }
        }
        return best;
    }
    return support_local(s, d, hp);
}

// broadphase/src/lib.rs:116-119
DIO_HD f3 world_support(const ShapeRef& s, f3 d, const HullPool& hp) {
    const f3 local = qrot_inv(s.rot, d);
    return add(qrot(s.rot, support_local(s, local, hp)), s.pos);
}

// broadphase/src/lib.rs:123-151
// A hull's six axis supports in one pass over its vertices: per direction
// the first maximal vertex, as six separate scans would find it.
DIO_HD void hull_axis_supports(const ShapeRef& s, const HullPool& hp, f3 out[6]) {
    const f3 axes[6] = {f3{1.0f, 0.0f, 0.0f}, neg(f3{1.0f, 0.0f, 0.0f}), f3{0.0f, 1.0f, 0.0f},
                        neg(f3{0.0f, 1.0f, 0.0f}), f3{0.0f, 0.0f, 1.0f}, neg(f3{0.0f, 0.0f, 1.0f})};
    const uint32_t first = (uint32_t)s.c[0].x, count = (uint32_t)s.c[0].y;
    const float4* v = hp.v + first;
    f3 dir[6];
    float bd[6];
    uint32_t bi[6];
    const f3 v0 = xyz(v[0]);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        dir[k] = qrot_inv(s.rot, axes[k]);
        bd[k] = dot3(v0, dir[k]);
        bi[k] = 0;
    }
    for (uint32_t i = 1; i < count; ++i) {
        const f3 vi = xyz(v[i]);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float dd = dot3(vi, dir[k]);
            if (dd > bd[k]) { bd[k] = dd; bi[k] = i; }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = add(qrot(s.rot, xyz(v[bi[k]])), s.pos);
}

DIO_HD void shape_aabb(const ShapeRef& s, const HullPool& hp, f3& mn, f3& mx) {
    mn = s.pos;
    mx = s.pos;
    if (s.tag == kHull) {
        f3 sp[6];
        hull_axis_supports(s, hp, sp);
        if (sp[0].x > mx.x) mx.x = sp[0].x;
        if (sp[1].x < mn.x) mn.x = sp[1].x;
        if (sp[2].y > mx.y) mx.y = sp[2].y;
        if (sp[3].y < mn.y) mn.y = sp[3].y;
        if (sp[4].z > mx.z) mx.z = sp[4].z;
        if (sp[5].z < mn.z) mn.z = sp[5].z;
        return;
    }
    const f3 sx = world_support(s, f3{1.0f, 0.0f, 0.0f}, hp);
    const f3 snx = world_support(s, neg(f3{1.0f, 0.0f, 0.0f}), hp);
    const f3 sy = world_support(s, f3{0.0f, 1.0f, 0.0f}, hp);
    const f3 sny = world_support(s, neg(f3{0.0f, 1.0f, 0.0f}), hp);
    const f3 sz = world_support(s, f3{0.0f, 0.0f, 1.0f}, hp);
    const f3 snz = world_support(s, neg(f3{0.0f, 0.0f, 1.0f}), hp);
    if (sx.x > mx.x) mx.x = sx.x;
    if (snx.x < mn.x) mn.x = snx.x;
    if (sy.y > mx.y) mx.y = sy.y;
    if (sny.y < mn.y) mn.y = sny.y;
    if (sz.z > mx.z) mx.z = sz.z;
    if (snz.z < mn.z) mn.z = snz.z;
}

// broadphase/src/lib.rs:195-201 (closed intervals)
DIO_HD bool aabb_overlap(float4 mna, float4 mxa, float4 mnb, float4 mxb) {
    return mna.x <= mxb.x && mnb.x <= mxa.x && mna.y <= mxb.y && mnb.y <= mxa.y && mna.z <= mxb.z &&
           mnb.z <= mxa.z;
}

// ------------------------------------------------------------ box SAT
DIO_HD void box_axes(q4 r, f3 ax[3]) {
    ax[0] = qrot(r, f3{1.0f, 0.0f, 0.0f});
    ax[1] = qrot(r, f3{0.0f, 1.0f, 0.0f});
    ax[2] = qrot(r, f3{0.0f, 0.0f, 1.0f});
}
// One of three values for a runtime k, by selects on the values (a runtime
// index into a local array would place the array in scratch memory).
DIO_HD f3 sel3(f3 a0, f3 a1, f3 a2, uint32_t k) {
    f3 r = a2;
    r.x = k == 1 ? a1.x : r.x; r.y = k == 1 ? a1.y : r.y; r.z = k == 1 ? a1.z : r.z;
    r.x = k == 0 ? a0.x : r.x; r.y = k == 0 ? a0.y : r.y; r.z = k == 0 ? a0.z : r.z;
    return r;
}
DIO_HD f3 pick3(const f3 a[3], uint32_t k) { return sel3(a[0], a[1], a[2], k); }
// The conditionals select values, not lvalues: a ?: over struct members is an
// lvalue in C++, compiled to a load through a selected address, which keeps
// the struct in scratch memory.
DIO_HD float box_half(const ShapeRef& s, uint32_t axis) {
    const float4 c0 = s.c[0];
    const float hx = c0.x, hy = c0.y, hz = c0.z;
    float r = hz;
    r = axis == 1 ? hy : r;
    r = axis == 0 ? hx : r;
    r = s.tag == kCube ? hx : r;
    return r;
}
// cuboid.rs:37-40
DIO_HD float box_radius(const ShapeRef& s, const f3 ax[3], f3 axis) {
    if (s.tag == kCube)
        return s.c[0].x * (fabsf(dot3(axis, ax[0])) + fabsf(dot3(axis, ax[1])) + fabsf(dot3(axis, ax[2])));
    return s.c[0].x * fabsf(dot3(axis, ax[0])) + s.c[0].y * fabsf(dot3(axis, ax[1])) +
           s.c[0].z * fabsf(dot3(axis, ax[2]));
}

struct SatBest {
    float overlap;
    f3 normal;
    uint32_t kind, index;
};

// cuboid.rs:19-52
DIO_HD bool sat_axis(f3 raw, uint32_t kind, uint32_t index, const ShapeRef& a, const f3 aa[3], const ShapeRef& b,
                     const f3 ba[3], f3 delta, SatBest& best) {
    if (dot3(raw, raw) <= 1.0e-8f) return true;
    const f3 axis = normalize3(raw);
    const float ra = box_radius(a, aa, axis);
    const float rb = box_radius(b, ba, axis);
    const float overlap = ra + rb - fabsf(dot3(delta, axis));
    if (overlap < 0.0f) return false;
    if (overlap < best.overlap) {
        best.overlap = overlap;
        best.normal = dot3(delta, axis) >= 0.0f ? axis : neg(axis);
        best.kind = kind;
        best.index = index;
    }
    return true;
}

// cuboid.rs:54-167
DIO_HD bool box_sat(const ShapeRef& a, const ShapeRef& b, SatBest& best) {
    f3 aa[3], ba[3];
    box_axes(a.rot, aa);
    box_axes(b.rot, ba);
    const f3 delta = sub(b.pos, a.pos);
    best.overlap = 3.40282347e+38f;
    best.normal = f3{1.0f, 0.0f, 0.0f};
    best.kind = 0;
    best.index = 0;
    for (uint32_t i = 0; i < 3; ++i)
        if (!sat_axis(aa[i], 0, i, a, aa, b, ba, delta, best)) return false;
    for (uint32_t i = 0; i < 3; ++i)
        if (!sat_axis(ba[i], 1, i, a, aa, b, ba, delta, best)) return false;
    for (uint32_t i = 0; i < 3; ++i)
        for (uint32_t j = 0; j < 3; ++j)
            if (!sat_axis(cross3(aa[i], ba[j]), 2, i * 3 + j, a, aa, b, ba, delta, best)) return false;
    return true;
}

// cuboid.rs:169-196 (vertices in the reference's order)
DIO_HD void box_face(const ShapeRef& s, uint32_t ai, f3 outward, f3 verts[4], f3& center, f3& u, f3& v, float& eu,
                     float& ev) {
    f3 ax[3];
    box_axes(s.rot, ax);
    const f3 axis = pick3(ax, ai);
    const float hn = box_half(s, ai);
    const f3 fnrm = scale(axis, dot3(axis, outward) >= 0.0f ? 1.0f : -1.0f);
    center = add(s.pos, scale(fnrm, hn));
    const uint32_t iu = ai == 0 ? 1u : 0u;
    const uint32_t iv = ai == 2 ? 1u : 2u;
    u = pick3(ax, iu);
    v = pick3(ax, iv);
    eu = box_half(s, iu);
    ev = box_half(s, iv);
    verts[0] = add(add(center, scale(u, eu)), scale(v, ev));
    verts[1] = add(sub(center, scale(u, eu)), scale(v, ev));
    verts[2] = sub(sub(center, scale(u, eu)), scale(v, ev));
    verts[3] = sub(add(center, scale(u, eu)), scale(v, ev));
}

// cuboid.rs:198-214
DIO_HD uint32_t box_incident_axis(const ShapeRef& s, f3 n) {
    f3 ax[3];
    box_axes(s.rot, ax);
    uint32_t best = 0;
    float bd = fabsf(dot3(ax[0], n));
    for (uint32_t i = 1; i < 3; ++i) {
        const float d = fabsf(dot3(ax[i], n));
        if (d > bd) { bd = d; best = i; }
    }
    return best;
}

// Fixed 8-slot polygons kept in registers: every loop below is unrolled so
// each slot index is static, and a point goes to slot k by selects (a runtime
// index into a local array would place the array in scratch memory).
struct Poly8 {
    f3 v[8];
};
DIO_HD float fsel(bool c, float a, float b) { return c ? a : b; }
DIO_HD f3 sel_f3(bool c, f3 a, f3 b) { return f3{fsel(c, a.x, b.x), fsel(c, a.y, b.y), fsel(c, a.z, b.z)}; }
// every slot is stored, with the new or its old value: conditional stores to
// the slots would be merged into one store at a runtime address
DIO_HD void put8(Poly8& p, uint32_t k, f3 x) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) p.v[j] = sel_f3(j == k, x, p.v[j]);
}

// cuboid.rs:216-250: Sutherland-Hodgman against one plane, at most 8 points out
DIO_HD uint32_t clip_plane(const Poly8& in, uint32_t in_count, f3 n, float offset, Poly8& out) {
    if (in_count == 0) return 0;
    uint32_t count = 0;
    f3 prev = in.v[0];
#pragma unroll
    for (uint32_t j = 1; j < 8; ++j)
        if (j == in_count - 1) prev = in.v[j];
    float pd = dot3(n, prev) - offset;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        if (i < in_count) {
            const f3 cur = in.v[i];
            const float cd = dot3(n, cur) - offset;
            const bool pin = pd <= 1.0e-5f;
            const bool cin = cd <= 1.0e-5f;
            if (pin != cin) {
                const float den = pd - cd;
                if (fabsf(den) > 1.0e-8f && count < 8) {
                    const float t = rclamp(pd / den, 0.0f, 1.0f);
                    put8(out, count, add(prev, scale(sub(cur, prev), t)));
                    ++count;
                }
            }
            if (cin && count < 8) {
                put8(out, count, cur);
                ++count;
            }
            prev = cur;
            pd = cd;
        }
    }
    return count;
}

struct Manifold {
    f3 normal;
    float depth;
    f3 pa[4], pb[4];
    uint32_t count;
};

// cuboid.rs:252-272
DIO_HD void centered_contact(const ShapeRef& a, const ShapeRef& b, f3 n, float depth, Manifold& m) {
    f3 aa[3], ba[3];
    box_axes(a.rot, aa);
    box_axes(b.rot, ba);
    const float ra = box_radius(a, aa, n);
    const float rb = box_radius(b, ba, n);
    m.pa[0] = add(a.pos, scale(n, ra));
    m.pb[0] = sub(b.pos, scale(n, rb));
    m.normal = n;
    m.depth = depth;
    m.count = 1;
}

// c ? a : b member by member, on values (a ?: over two struct references is
// an lvalue, compiled to a load through a selected address)
DIO_HD ShapeRef sel_shape(bool c, const ShapeRef& a, const ShapeRef& b) {
    ShapeRef r;
    r.pos = f3{fsel(c, a.pos.x, b.pos.x), fsel(c, a.pos.y, b.pos.y), fsel(c, a.pos.z, b.pos.z)};
    r.rot = q4{fsel(c, a.rot.x, b.rot.x), fsel(c, a.rot.y, b.rot.y), fsel(c, a.rot.z, b.rot.z),
               fsel(c, a.rot.w, b.rot.w)};
    r.tag = c ? a.tag + 0u : b.tag + 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r.c[k] = make_float4(fsel(c, a.c[k].x, b.c[k].x), fsel(c, a.c[k].y, b.c[k].y), fsel(c, a.c[k].z, b.c[k].z),
                             fsel(c, a.c[k].w, b.c[k].w));
    return r;
}

// cuboid.rs:276-377
DIO_HD bool box_manifold(const ShapeRef& a, const ShapeRef& b, Manifold& m) {
    SatBest best;
    if (!box_sat(a, b, best)) return false;
    const f3 normal = best.normal;
    const float depth = best.overlap;
    if (best.kind == 2) {
        centered_contact(a, b, normal, depth, m);
        return true;
    }
    const bool ref_is_a = best.kind == 0;
    const f3 rn = ref_is_a ? normal : neg(normal);
    const ShapeRef rb = sel_shape(ref_is_a, a, b);
    const ShapeRef ib = sel_shape(ref_is_a, b, a);
    const uint32_t inc = box_incident_axis(ib, rn);
    f3 rface[4], center, u, v;
    float eu, ev;
    box_face(rb, best.index, rn, rface, center, u, v, eu, ev);
    f3 iface[4], c2, u2, v2;
    float eu2, ev2;
    box_face(ib, inc, neg(rn), iface, c2, u2, v2, eu2, ev2);
    Poly8 pa, pb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        pa.v[i] = i < 4 ? iface[i & 3] : f3{0.0f, 0.0f, 0.0f};
        pb.v[i] = f3{0.0f, 0.0f, 0.0f};
    }
    uint32_t count = 4;
    count = clip_plane(pa, count, u, dot3(u, center) + eu, pb);
    count = clip_plane(pb, count, neg(u), dot3(neg(u), center) + eu, pa);
    count = clip_plane(pa, count, v, dot3(v, center) + ev, pb);
    count = clip_plane(pb, count, neg(v), dot3(neg(v), center) + ev, pa);
    if (count == 0) {
        centered_contact(a, b, normal, depth, m);
        return true;
    }
    const float rplane = dot3(rn, center);
    uint32_t cc = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) {
        // point i is kept when it exists, fewer than 4 are kept, and it lies
        // no more than 1e-5 above the reference face
        const f3 ip = pa.v[i];
        const float pd = rplane - dot3(rn, ip);
        const bool take = i < count && cc < 4 && !(pd < -1.0e-5f);
        const f3 rp = add(ip, scale(rn, rmax(pd, 0.0f)));
        const f3 xa = sel_f3(ref_is_a, rp, ip);
        const f3 xb = sel_f3(ref_is_a, ip, rp);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            m.pa[j] = sel_f3(take && j == cc, xa, m.pa[j]);
            m.pb[j] = sel_f3(take && j == cc, xb, m.pb[j]);
        }
        cc += take ? 1u : 0u;
    }
    if (cc == 0) {
        centered_contact(a, b, normal, depth, m);
        return true;
    }
    m.normal = normal;
    m.depth = depth;
    m.count = cc;
    return true;
}

}  // namespace dio
