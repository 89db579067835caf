// dio_math.hpp — vector / quaternion arithmetic for the gfx950 kernels.
//
// The formulas follow glam 0.33.3's scalar path, which the reference kernels
// use (engine/gpu_physics/*/Cargo.lock pins glam 0.33.3): dot is
// (x*x' + y*y') + z*z', normalize is v * (1 / |v|), Quat * Vec3 is
// v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w).  Every routine is written so that
// each float operation rounds once in the stated order; the library is built
// with -ffp-contract=off so no multiply-add is fused.  Functions are
// __host__ __device__ so the CPU emulation harness can run the same code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DIO_HD __host__ __device__ __forceinline__

namespace dio {

constexpr uint32_t kContainer = 0xFFFFFFFFu;
// CSR peer code: the other body of a contact entry, with this flag set when
// the entry's body is the contact's b; a container contact is kContainer
// (its body is always a).
constexpr uint32_t kPeerIsB = 0x80000000u;
constexpr uint32_t kSphere = 0, kCube = 1, kTetrahedron = 2, kBox = 3, kHull = 4;
constexpr uint32_t kSpring = 0, kRod = 1, kBallSocket = 2, kRope = 3;
// hard joints: rod, socket and rope geometry held by a multiplier and a
// ramped penalty per row (no counterpart in the reference's soft kinds)
constexpr uint32_t kHardRod = 4, kHardSocket = 5, kHardRope = 6;
// bending joints: b's orientation held at a's times a rest rotation (the rest
// quaternion in anchor_a), softly (kBend, stiffness in anchor_b.w) or with
// multipliers (kHardBend)
constexpr uint32_t kBend = 7, kHardBend = 8;
DIO_HD bool is_hard(uint32_t kind) { return (kind >= kHardRod && kind <= kHardRope) || kind == kHardBend; }
DIO_HD bool is_bend(uint32_t kind) { return kind == kBend || kind == kHardBend; }
// three rows per joint (socket and bend kinds), else one
DIO_HD bool three_rows(uint32_t kind) { return kind == kHardSocket || kind == kHardBend; }

struct f3 {
    float x, y, z;
};
struct q4 {
    float x, y, z, w;
};

DIO_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
DIO_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DIO_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DIO_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
DIO_HD f3 scale(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
DIO_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
DIO_HD f3 mulv(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DIO_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DIO_HD f3 cross3(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
DIO_HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
DIO_HD f3 normalize3(f3 a) { return scale(a, 1.0f / len3(a)); }
DIO_HD bool finite_f(float x) { return x - x == 0.0f; }
DIO_HD f3 normalize_or_zero3(f3 a) {
    float rcp = 1.0f / len3(a);
    if (finite_f(rcp) && rcp > 0.0f) return scale(a, rcp);
    return f3{0.0f, 0.0f, 0.0f};
}
DIO_HD bool eq3(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Rust f32::min / max / clamp semantics (a NaN operand yields the other)
DIO_HD float rmin(float a, float b) { return (b < a || a != a) ? b : a; }
DIO_HD float rmax(float a, float b) { return (b > a || a != a) ? b : a; }
DIO_HD float rclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

DIO_HD q4 qmul(q4 a, q4 b) {
    return q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
DIO_HD q4 qconj(q4 q) { return q4{-q.x, -q.y, -q.z, q.w}; }
DIO_HD q4 qnormalize(q4 q) {
    float len = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    float r = 1.0f / len;
    return q4{q.x * r, q.y * r, q.z * r, q.w * r};
}
DIO_HD f3 qrot(q4 q, f3 v) {
    float w = q.w;
    f3 b{q.x, q.y, q.z};
    float b2 = dot3(b, b);
    return add(add(scale(v, w * w - b2), scale(b, dot3(v, b) * 2.0f)), scale(cross3(b, v), w * 2.0f));
}
// Quat::inverse() * v for a unit quaternion (the conjugate)
DIO_HD f3 qrot_inv(q4 q, f3 v) { return qrot(qconj(q), v); }

// avbd/src/lib.rs:91-94: (q + (p * q) * 0.5).normalize(), p = (d, 0)
DIO_HD q4 quat_step(q4 q, f3 d) {
    q4 p{d.x, d.y, d.z, 0.0f};
    q4 pq = qmul(p, q);
    q4 s{q.x + pq.x * 0.5f, q.y + pq.y * 0.5f, q.z + pq.z * 0.5f, q.w + pq.w * 0.5f};
    return qnormalize(s);
}
// avbd/src/lib.rs:96-99: 2 * (to * from^-1).xyz
DIO_HD f3 ang_diff(q4 to, q4 from) {
    q4 rel = qmul(to, qconj(from));
    return f3{2.0f * rel.x, 2.0f * rel.y, 2.0f * rel.z};
}

// A bending joint's error: the rotation from b's orientation to a's times the
// rest rotation, 2 vec(qa rest qb^-1) on the shorter arc (+d when a turns by a
// small world rotation d, -d when b does).
DIO_HD f3 bend_error(q4 qa, q4 rest, q4 qb) {
    const q4 rel = qmul(qmul(qa, rest), qconj(qb));
    const float s = rel.w < 0.0f ? -1.0f : 1.0f;
    return f3{(2.0f * rel.x) * s, (2.0f * rel.y) * s, (2.0f * rel.z) * s};
}

// float4 <-> f3 / q4 views of the SoA arrays
DIO_HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
DIO_HD q4 as_q(float4 v) { return q4{v.x, v.y, v.z, v.w}; }
DIO_HD float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
DIO_HD float4 f4q(q4 q) { return make_float4(q.x, q.y, q.z, q.w); }

// physics/gpu.rs:406-414: tangent frame of a normal
DIO_HD void tangents(f3 n, f3& t1, f3& t2) {
    f3 c = fabsf(n.x) > 0.9f ? cross3(n, f3{0.0f, 1.0f, 0.0f}) : cross3(n, f3{1.0f, 0.0f, 0.0f});
    t1 = normalize3(c);
    t2 = cross3(n, t1);
}

}  // namespace dio
