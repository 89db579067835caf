// avbd.hip — the AVBD solver kernels for gfx950.
//
// Counterparts of the reference's rust-gpu entry points
// (engine/gpu_physics/avbd/src/lib.rs): avbd_init :202-254, avbd_primal
// :258-551, avbd_dual :565-663, avbd_recover :667-699, scheduled as in
// engine/src/render/gpu_physics.rs:541-550 (init; I x (primal, dual) with
// parity k % 2; recover with parity (I - 1) % 2).
//
// Jacobi mode: one lane per body for the primal (its 6x6 block is assembled in
// 27 registers and solved by an in-register LDL^T), one lane per contact for
// the dual.  Each Hessian / gradient entry is accumulated in the reference's
// order (CSR contacts in ascending contact index, rows n, t1, t2; then joints
// in ascending constraint index), so results are bit-identical to a
// sequential FP32 evaluation with no fused multiply-adds.  Joints are found
// through a per-body CSR instead of the reference's scan of every constraint
// per body (avbd/src/lib.rs:422-539), which visits the same joints in the
// same order.
#include <algorithm>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "dio_device.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;

// The 6x6 block system: lin / ang lower triangles (00,10,20,11,21,22), the
// full ang-by-lin cross block, and the two gradient halves.
struct Block6 {
    float l[6], a[6], c[9], rl[3], ra[3];

    __device__ __forceinline__ void add_lin(f3 j, float k) {
        l[0] = l[0] + (j.x * j.x) * k;
        l[1] = l[1] + (j.y * j.x) * k;
        l[2] = l[2] + (j.z * j.x) * k;
        l[3] = l[3] + (j.y * j.y) * k;
        l[4] = l[4] + (j.z * j.y) * k;
        l[5] = l[5] + (j.z * j.z) * k;
    }
    __device__ __forceinline__ void add_ang(f3 j, float k) {
        a[0] = a[0] + (j.x * j.x) * k;
        a[1] = a[1] + (j.y * j.x) * k;
        a[2] = a[2] + (j.z * j.x) * k;
        a[3] = a[3] + (j.y * j.y) * k;
        a[4] = a[4] + (j.z * j.y) * k;
        a[5] = a[5] + (j.z * j.z) * k;
    }
    // c[3 r + col] += ja[r] * jl[col] * k
    __device__ __forceinline__ void add_cross(f3 ja, f3 jl, float k) {
        c[0] = c[0] + (ja.x * jl.x) * k;
        c[1] = c[1] + (ja.x * jl.y) * k;
        c[2] = c[2] + (ja.x * jl.z) * k;
        c[3] = c[3] + (ja.y * jl.x) * k;
        c[4] = c[4] + (ja.y * jl.y) * k;
        c[5] = c[5] + (ja.y * jl.z) * k;
        c[6] = c[6] + (ja.z * jl.x) * k;
        c[7] = c[7] + (ja.z * jl.y) * k;
        c[8] = c[8] + (ja.z * jl.z) * k;
    }
    __device__ __forceinline__ void add_rhs(f3 jl, f3 ja, float f) {
        rl[0] = rl[0] + jl.x * f;
        rl[1] = rl[1] + jl.y * f;
        rl[2] = rl[2] + jl.z * f;
        ra[0] = ra[0] + ja.x * f;
        ra[1] = ra[1] + ja.y * f;
        ra[2] = ra[2] + ja.z * f;
    }
    // add_constraint_row (avbd/src/lib.rs:182-198)
    __device__ __forceinline__ void row(f3 jl, f3 ja, float force, float k) {
        add_lin(jl, k);
        add_ang(ja, k);
        add_cross(ja, jl, k);
        add_rhs(jl, ja, force);
    }
};

// solve6 (avbd/src/lib.rs:107-180): LDL^T of the block, reading the lower
// triangle; solves H x = -g with bl = -rhs_lin, ba = -rhs_ang.
__device__ __forceinline__ void solve6(const Block6& B, float x[6]) {
    const float a11 = B.l[0], a21 = B.l[1], a31 = B.l[2], a22 = B.l[3], a32 = B.l[4], a33 = B.l[5];
    const float a41 = B.c[0], a51 = B.c[3], a61 = B.c[6];
    const float a42 = B.c[1], a52 = B.c[4], a62 = B.c[7];
    const float a43 = B.c[2], a53 = B.c[5], a63 = B.c[8];
    const float a44 = B.a[0], a54 = B.a[1], a64 = B.a[2], a55 = B.a[3], a65 = B.a[4], a66 = B.a[5];
    const float l21 = a21 / a11, l31 = a31 / a11, l41 = a41 / a11, l51 = a51 / a11, l61 = a61 / a11;
    const float d1 = a11;
    const float d2 = a22 - l21 * l21 * d1;
    const float l32 = (a32 - l21 * l31 * d1) / d2;
    const float l42 = (a42 - l21 * l41 * d1) / d2;
    const float l52 = (a52 - l21 * l51 * d1) / d2;
    const float l62 = (a62 - l21 * l61 * d1) / d2;
    const float d3 = a33 - (l31 * l31 * d1 + l32 * l32 * d2);
    const float l43 = (a43 - l31 * l41 * d1 - l32 * l42 * d2) / d3;
    const float l53 = (a53 - l31 * l51 * d1 - l32 * l52 * d2) / d3;
    const float l63 = (a63 - l31 * l61 * d1 - l32 * l62 * d2) / d3;
    const float d4 = a44 - (l41 * l41 * d1 + l42 * l42 * d2 + l43 * l43 * d3);
    const float l54 = (a54 - l41 * l51 * d1 - l42 * l52 * d2 - l43 * l53 * d3) / d4;
    const float l64 = (a64 - l41 * l61 * d1 - l42 * l62 * d2 - l43 * l63 * d3) / d4;
    const float d5 = a55 - (l51 * l51 * d1 + l52 * l52 * d2 + l53 * l53 * d3 + l54 * l54 * d4);
    const float l65 = (a65 - l51 * l61 * d1 - l52 * l62 * d2 - l53 * l63 * d3 - l54 * l64 * d4) / d5;
    const float d6 = a66 - (l61 * l61 * d1 + l62 * l62 * d2 + l63 * l63 * d3 + l64 * l64 * d4 + l65 * l65 * d5);
    const float b1 = -B.rl[0], b2 = -B.rl[1], b3 = -B.rl[2], b4 = -B.ra[0], b5 = -B.ra[1], b6 = -B.ra[2];
    const float y1 = b1;
    const float y2 = b2 - l21 * y1;
    const float y3 = b3 - l31 * y1 - l32 * y2;
    const float y4 = b4 - l41 * y1 - l42 * y2 - l43 * y3;
    const float y5 = b5 - l51 * y1 - l52 * y2 - l53 * y3 - l54 * y4;
    const float y6 = b6 - l61 * y1 - l62 * y2 - l63 * y3 - l64 * y4 - l65 * y5;
    const float z1 = y1 / d1, z2 = y2 / d2, z3 = y3 / d3, z4 = y4 / d4, z5 = y5 / d5, z6 = y6 / d6;
    const float x6 = z6;
    const float x5 = z5 - l65 * x6;
    const float x4 = z4 - l54 * x5 - l64 * x6;
    const float x3 = z3 - l43 * x4 - l53 * x5 - l63 * x6;
    const float x2 = z2 - l32 * x3 - l42 * x4 - l52 * x5 - l62 * x6;
    const float x1 = z1 - l21 * x2 - l31 * x3 - l41 * x4 - l51 * x5 - l61 * x6;
    x[0] = x1; x[1] = x2; x[2] = x3; x[3] = x4; x[4] = x5; x[5] = x6;
}

// Clamped contact force (avbd/src/lib.rs:388-398 and :641-651).
struct Force3 {
    float f0, f1, f2, bounds, tang;
};
__device__ __forceinline__ Force3 contact_force(float4 pen, float4 lam, float friction, float c0, float c1,
                                                float c2) {
    Force3 r;
    r.f0 = rmin(pen.x * c0 + lam.x, 0.0f);
    const float f1r = pen.y * c1 + lam.y;
    const float f2r = pen.z * c2 + lam.z;
    r.bounds = fabsf(r.f0) * friction;
    r.tang = sqrtf(f1r * f1r + f2r * f2r);
    r.f1 = f1r;
    r.f2 = f2r;
    if (r.tang > r.bounds && r.tang > 0.0f) {
        const float s = r.bounds / r.tang;
        r.f1 = f1r * s;
        r.f2 = f2r * s;
    }
    return r;
}

// A contact's tangent frame and c0 (physics/gpu.rs:239-264,352-404).  For
// contacts the device built (v.derived_frame) they are recomputed instead of
// read: the frame is tangents() of the narrow normal, which the build stored
// negated, or for a wall contact of the stored normal itself; c0 is
// (margin - depth, 0, 0) with its first component in ra.w.  Both come out
// bit-identical to the stored values.
struct ContactFrame {
    f3 t1, t2;
    float4 c0;
};
__device__ __forceinline__ ContactFrame contact_frame(const SolveView& v, uint32_t ci, f3 n0, float4 ra4,
                                                      bool wall) {
    ContactFrame fr;
    if (v.derived_frame) {
        tangents(wall ? n0 : neg(n0), fr.t1, fr.t2);
        fr.c0 = make_float4(ra4.w, 0.0f, 0.0f, 0.0f);
    } else {
        fr.t1 = xyz(v.c_t1[ci]);
        fr.t2 = xyz(v.c_t2[ci]);
        fr.c0 = v.c_c0[ci];
    }
    return fr;
}

// avbd_dual (avbd/src/lib.rs:565-663) for one contact: the constraint value
// from both bodies' displacements and world lever arms (the b terms are zero
// for a container contact), the clamped force as the new lambda, and the
// penalty ramp.  r0..r2 are the contact's normal and tangents as stored (a's
// convention).
struct DualState {
    float4 lam, pen;
};
__device__ __forceinline__ DualState dual_update(const AvbdParams& p, float kk, f3 r0, f3 r1, f3 r2, float friction,
                                                 f3 raw, f3 rbw, f3 dal, f3 daa, f3 dbl, f3 dba, float4 c04,
                                                 float4 pen, float4 lam) {
    float c0 = c04.x * kk, c1 = c04.y * kk, c2 = c04.z * kk;
    c0 = c0 + (dot3(r0, dal) + dot3(cross3(raw, r0), daa) - dot3(r0, dbl) - dot3(cross3(rbw, r0), dba));
    c1 = c1 + (dot3(r1, dal) + dot3(cross3(raw, r1), daa) - dot3(r1, dbl) - dot3(cross3(rbw, r1), dba));
    c2 = c2 + (dot3(r2, dal) + dot3(cross3(raw, r2), daa) - dot3(r2, dbl) - dot3(cross3(rbw, r2), dba));
    const Force3 F = contact_force(pen, lam, friction, c0, c1, c2);
    DualState st;
    st.lam = make_float4(F.f0, F.f1, F.f2, 0.0f);
    if (F.f0 < 0.0f) pen.x = rmin(pen.x + p.beta_lin * fabsf(c0), p.penalty_max);
    if (F.tang <= F.bounds) {
        pen.y = rmin(pen.y + p.beta_lin * fabsf(c1), p.penalty_max);
        pen.z = rmin(pen.z + p.beta_lin * fabsf(c2), p.penalty_max);
    }
    st.pen = pen;
    return st;
}

// Hard joints (kHardRod, kHardSocket, kHardRope; no counterpart in the
// reference's soft kinds).  The stabilised row values are C(x) - alpha C(x0)
// from the anchor separation delta = p_a - p_b at the current poses and delta0
// at the step's initial poses (a rope, an inequality C = |delta| - L <= 0,
// subtracts only a violated start value); the row force is penalty * C^ +
// lambda (kept >= 0 for a rope, which only pulls); a dual update sets lambda
// to the row force and ramps an active row's penalty by beta |C^|, as contacts
// do.  A rod or rope row is defined only when |delta| > 1e-6.
struct HardRows {
    f3 chat;  // C^ per row (x alone for a rod or rope)
    f3 dir;   // delta / |delta| (rod, rope)
    bool ok;
};
__device__ __forceinline__ HardRows hard_rows(uint32_t kind, f3 delta, f3 delta0, float rest, float alpha) {
    HardRows r;
    r.dir = f3{0.0f, 0.0f, 0.0f};
    if (three_rows(kind)) {
        r.chat = f3{delta.x - alpha * delta0.x, delta.y - alpha * delta0.y, delta.z - alpha * delta0.z};
        r.ok = true;
    } else {
        const float len = len3(delta);
        const float len0 = len3(delta0);
        r.ok = len > 1.0e-6f;
        float c0 = len0 - rest;
        if (kind == kHardRope) c0 = rmax(c0, 0.0f);
        r.chat = f3{(len - rest) - alpha * c0, 0.0f, 0.0f};
        if (r.ok) r.dir = divs(delta, len);
    }
    return r;
}
__device__ __forceinline__ float hard_force(uint32_t kind, float pen, float chat, float lam) {
    const float f = pen * chat + lam;
    return kind == kHardRope ? rmax(f, 0.0f) : f;
}
__device__ __forceinline__ void hard_row_update(const AvbdParams& p, uint32_t kind, float chat, float& lam,
                                                float& pen) {
    const float f = hard_force(kind, pen, chat, lam);
    lam = f;
    if (kind != kHardRope || f > 0.0f) pen = rmin(pen + p.beta_lin * fabsf(chat), p.penalty_max);
}
__device__ __forceinline__ void hard_update(const AvbdParams& p, uint32_t kind, const HardRows& r, float4& lam,
                                            float4& pen) {
    if (!r.ok) return;
    hard_row_update(p, kind, r.chat.x, lam.x, pen.x);
    if (three_rows(kind)) {
        hard_row_update(p, kind, r.chat.y, lam.y, pen.y);
        hard_row_update(p, kind, r.chat.z, lam.z, pen.z);
    }
}
// the warm start of a step (the contacts' rule, physics/gpu.rs:277-283):
// lambda * alpha gamma, penalty * gamma clamped to [penalty_min, penalty_max]
__device__ __forceinline__ void hard_warm(const AvbdParams& p, float4 lam, float4 pen, float4& lam_out,
                                          float4& pen_out) {
    const float ag = p.alpha * p.gamma;
    lam_out = make_float4(lam.x * ag, lam.y * ag, lam.z * ag, 0.0f);
    pen_out = make_float4(rmin(rmax(pen.x * p.gamma, p.penalty_min), p.penalty_max),
                          rmin(rmax(pen.y * p.gamma, p.penalty_min), p.penalty_max),
                          rmin(rmax(pen.z * p.gamma, p.penalty_min), p.penalty_max), 0.0f);
}
// the value vector of a hard joint from a pose buffer pair: the anchor
// separation p_a - p_b, or a bending joint's error (aa4 = the rest rotation)
__device__ __forceinline__ f3 joint_delta(const float4* P, const float4* Q, uint4 abk, float4 aa4, f3 ab) {
    if (is_bend(abk.z)) return bend_error(as_q(Q[abk.x]), as_q(aa4), as_q(Q[abk.y]));
    return sub(add(xyz(P[abk.x]), qrot(as_q(Q[abk.x]), xyz(aa4))), add(xyz(P[abk.y]), qrot(as_q(Q[abk.y]), ab)));
}

// The multipliers after j dual updates live in buffer j % 2 (0: the contact
// set's own lam / pen, 1: the scratch pair).  Primal iteration k >= 1 first
// performs dual update k - 1 for each of its contacts (both bodies of a
// contact evaluate it identically; body a stores it, or body b when a is
// static), so a solve is init, I primal launches and one final dual +
// recover launch.  A contact between two static bodies is read by no primal;
// the final launch applies its I updates at once (its poses do not move).
__device__ __forceinline__ float4* lam_buf(const SolveView& v, uint32_t j) { return (j & 1u) ? v.c_lam_alt : v.c_lam; }
// the body of work item k: k itself, or entry k of a colour's body list (p.n
// past its end: no body)
__device__ __forceinline__ uint32_t body_at(const SolveView& v, const AvbdParams& p, uint32_t k) {
    if (!v.body_list) return k;
    if (v.list_range) {
        const uint32_t b0 = v.list_range[0], b1 = v.list_range[1];
        return k < b1 - b0 ? v.body_list[b0 + k] : p.n;
    }
    return k < v.list_count ? v.body_list[k] : p.n;
}
// The chunk of work items workgroup b takes: the XCD mapping below, or, for a
// colour whose length only the device knows (a grid sized for every body),
// chunk b itself, so the workgroups with work are spread over all the XCDs.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb);
__device__ __forceinline__ uint32_t work_chunk(const SolveView& v) {
    return v.list_range ? blockIdx.x : xcd_chunk(blockIdx.x, gridDim.x);
}
// Whether a workgroup's first item lies past the end of a device-sized colour
// (the whole workgroup then has nothing to do).
__device__ __forceinline__ bool past_colour(const SolveView& v, uint32_t first_item) {
    return v.list_range && first_item >= v.list_range[1] - v.list_range[0];
}
__device__ __forceinline__ float4* pen_buf(const SolveView& v, uint32_t j) { return (j & 1u) ? v.c_pen_alt : v.c_pen; }
// The chunk of work items workgroup b takes.  Workgroups are handed to the 8
// XCDs in turn (b % 8), each XCD with an L2 of its own; this mapping gives
// XCD x one contiguous range of chunks, so bodies that are neighbours in the
// work order (the Morton order of the broad phase) share an L2, and with it
// the contact records both bodies of a contact read and each other's poses.
// It is a bijection of [0, nb).
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
    constexpr uint32_t kXcds = 8;
    const uint32_t per = nb / kXcds, extra = nb % kXcds;
    const uint32_t x = b % kXcds, local = b / kXcds;
    return x < extra ? x * (per + 1) + local : extra * (per + 1) + (x - extra) * per + local;
}

// avbd_init of one body (avbd/src/lib.rs:202-254) from its pose and
// velocities, into target t
__device__ __forceinline__ void init_body(const InitTarget& t, const AvbdParams& p, uint32_t i, float4 pos4,
                                          float4 rot4, f3 vel, f3 angv, f3 prev, float inv_mass) {
    const float dt = p.dt;
    const f3 g{p.gx, p.gy, p.gz};
    const f3 pos = xyz(pos4);
    const q4 rot = as_q(rot4);
    t.init_lin[i] = f4(pos, 0.0f);
    t.init_ang[i] = rot4;
    f3 ip = add(pos, scale(vel, dt));
    if (inv_mass > 0.0f) ip = add(ip, scale(g, dt * dt));
    t.iner_lin[i] = f4(ip, 0.0f);
    const q4 irot = quat_step(rot, scale(angv, dt));
    t.iner_ang[i] = f4q(irot);
    // adaptive gravity warm start (avbd/src/lib.rs:235-251)
    f3 wp = pos;
    q4 wr = rot;
    if (inv_mass > 0.0f) {
        const f3 accel = divs(sub(vel, prev), dt);
        const float gm = len3(g);
        const float weight = gm > 0.0f ? rclamp(dot3(accel, divs(g, gm)) / gm, 0.0f, 1.0f) : 0.0f;
        wp = add(add(pos, scale(vel, dt)), scale(g, weight * dt * dt));
        wr = irot;
    }
    // both parity buffers start from S_k (gpu_physics.rs:477-482); buffer 0
    // then takes the warm start
    t.pos[1][i] = f4(pos, 0.0f);
    t.rot[1][i] = rot4;
    t.pos[0][i] = f4(wp, 0.0f);
    t.rot[0][i] = f4q(wr);
}

// the warm start of hard joint cj into buffer 0 (soft joints: nothing)
__device__ __forceinline__ void warm_joint(const SolveView& v, const AvbdParams& p, uint32_t cj) {
    if (!is_hard(v.j_abk[cj].z)) return;
    float4 lam, pen;
    hard_warm(p, v.j_st_lam[cj], v.j_st_pen[cj], lam, pen);
    v.j_lam[0][cj] = lam;
    v.j_pen[0][cj] = pen;
}

__global__ void __launch_bounds__(kBlock) k_avbd_init(SolveView v, AvbdParams p) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (v.j_hard)
        for (uint32_t cj = i; cj < v.j_count; cj += gridDim.x * kBlock) warm_joint(v, p, cj);
    if (i >= p.n) return;
    const InitTarget t{{v.pos[0], v.pos[1]}, {v.rot[0], v.rot[1]}, v.init_lin, v.init_ang, v.iner_lin, v.iner_ang};
    init_body(t, p, i, v.cur_pos[i], v.cur_rot[i], xyz(v.vel[i]), xyz(v.ang[i]), xyz(v.prev_vel[i]), v.inv_mm[i].w);
}

// The primal of body i (one lane per body): its 6x6 block over its contacts
// and joints, solved, and the pose written to the other parity buffer.
template <bool FUSED>
__device__ __forceinline__ void primal_one(const SolveView& v, const AvbdParams& p, uint32_t iter, uint32_t i) {
    const uint32_t parity = iter & 1u;
    const bool fuse = FUSED && iter > 0;
    const float4* __restrict__ lam_in = lam_buf(v, fuse ? iter - 1 : 0);
    const float4* __restrict__ pen_in = pen_buf(v, fuse ? iter - 1 : 0);
    float4* lam_out = lam_buf(v, iter);
    float4* pen_out = pen_buf(v, iter);
    const uint32_t jin = fuse ? (iter - 1) & 1u : 0u;
    const float4 imm = v.inv_mm[i];
    if (imm.w == 0.0f) return;
    const float4* __restrict__ rp = v.pos[parity];
    const float4* __restrict__ rr = v.rot[parity];
    const f3 snap_pos = xyz(rp[i]);
    const q4 snap_rot = as_q(rr[i]);
    const float dt = p.dt;
    const float mass = 1.0f / imm.w;
    const f3 moment{1.0f / imm.x, 1.0f / imm.y, 1.0f / imm.z};
    const float dt2 = dt * dt;
    const float m_dt2 = mass / dt2;
    const f3 i_dt2 = divs(moment, dt2);

    Block6 B;
#pragma unroll
    for (int k = 0; k < 6; ++k) { B.l[k] = 0.0f; B.a[k] = 0.0f; }
#pragma unroll
    for (int k = 0; k < 9; ++k) B.c[k] = 0.0f;
    B.l[0] = m_dt2; B.l[3] = m_dt2; B.l[5] = m_dt2;
    B.a[0] = i_dt2.x; B.a[3] = i_dt2.y; B.a[5] = i_dt2.z;
    const f3 rl = scale(sub(snap_pos, xyz(v.iner_lin[i])), 1.0f);
    B.rl[0] = m_dt2 * rl.x; B.rl[1] = m_dt2 * rl.y; B.rl[2] = m_dt2 * rl.z;
    const f3 ad = ang_diff(snap_rot, as_q(v.iner_ang[i]));
    B.ra[0] = i_dt2.x * ad.x; B.ra[1] = i_dt2.y * ad.y; B.ra[2] = i_dt2.z * ad.z;

    const f3 dsl = sub(snap_pos, xyz(v.init_lin[i]));
    const f3 dsa = ang_diff(snap_rot, as_q(v.init_ang[i]));
    const float kk = 1.0f - p.alpha;
    const f3 zero{0.0f, 0.0f, 0.0f};

    const uint32_t jb = v.offsets[i], je = v.offsets[i + 1];
    for (uint32_t j = jb; j < je; ++j) {
        const uint32_t ci = v.indices[j];
        const uint32_t pc = v.peers[j];
        const bool is_a = pc == kContainer || (pc & kPeerIsB) == 0u;
        const uint32_t other = pc == kContainer ? kContainer : (pc & ~kPeerIsB);
        const float4 n4 = v.c_n[ci];
        const f3 n0 = xyz(n4);
        const float4 ra4 = v.c_ra[ci];
        const f3 ra = xyz(ra4);
        const f3 rb = other != kContainer ? xyz(v.c_rb[ci]) : zero;  // a wall contact's r_b is zero
        const ContactFrame cf = contact_frame(v, ci, n0, ra4, other == kContainer);
        const f3 t1 = cf.t1, t2 = cf.t2;
        const f3 r0 = is_a ? n0 : neg(n0);
        const f3 r1 = is_a ? t1 : neg(t1);
        const f3 r2 = is_a ? t2 : neg(t2);
        const f3 r_world = qrot(snap_rot, is_a ? ra : rb);
        f3 dol = zero, doa = zero, r_other = zero;
        if (other != kContainer) {
            const q4 orr = as_q(rr[other]);
            dol = sub(xyz(rp[other]), xyz(v.init_lin[other]));
            doa = ang_diff(orr, as_q(v.init_ang[other]));
            r_other = qrot(orr, is_a ? rb : ra);
        }
        const float4 c04 = cf.c0;
        float4 pen = pen_in[ci];
        float4 lam = lam_in[ci];
        if (fuse) {
            // dual update k - 1 of the contact, in its own (a, b) order
            const DualState d = dual_update(p, kk, n0, t1, t2, n4.w, is_a ? r_world : r_other,
                                            is_a ? r_other : r_world, is_a ? dsl : dol, is_a ? dsa : doa,
                                            is_a ? dol : dsl, is_a ? doa : dsa, c04, pen, lam);
            pen = d.pen;
            lam = d.lam;
            if (is_a || v.inv_mm[other].w == 0.0f) {
                lam_out[ci] = lam;
                pen_out[ci] = pen;
            }
        }
        const f3 j0 = cross3(r_world, r0), j1 = cross3(r_world, r1), j2 = cross3(r_world, r2);
        float c0 = c04.x * kk, c1 = c04.y * kk, c2 = c04.z * kk;
        c0 = c0 + (dot3(r0, dsl) + dot3(j0, dsa) - dot3(r0, dol) - dot3(cross3(r_other, r0), doa));
        c1 = c1 + (dot3(r1, dsl) + dot3(j1, dsa) - dot3(r1, dol) - dot3(cross3(r_other, r1), doa));
        c2 = c2 + (dot3(r2, dsl) + dot3(j2, dsa) - dot3(r2, dol) - dot3(cross3(r_other, r2), doa));
        const Force3 F = contact_force(pen, lam, n4.w, c0, c1, c2);
        B.add_lin(r0, pen.x); B.add_lin(r1, pen.y); B.add_lin(r2, pen.z);
        B.add_ang(j0, pen.x); B.add_ang(j1, pen.y); B.add_ang(j2, pen.z);
        B.add_cross(j0, r0, pen.x); B.add_cross(j1, r1, pen.y); B.add_cross(j2, r2, pen.z);
        B.rl[0] = B.rl[0] + r0.x * F.f0; B.rl[1] = B.rl[1] + r0.y * F.f0; B.rl[2] = B.rl[2] + r0.z * F.f0;
        B.rl[0] = B.rl[0] + r1.x * F.f1; B.rl[1] = B.rl[1] + r1.y * F.f1; B.rl[2] = B.rl[2] + r1.z * F.f1;
        B.rl[0] = B.rl[0] + r2.x * F.f2; B.rl[1] = B.rl[1] + r2.y * F.f2; B.rl[2] = B.rl[2] + r2.z * F.f2;
        B.ra[0] = B.ra[0] + j0.x * F.f0; B.ra[1] = B.ra[1] + j0.y * F.f0; B.ra[2] = B.ra[2] + j0.z * F.f0;
        B.ra[0] = B.ra[0] + j1.x * F.f1; B.ra[1] = B.ra[1] + j1.y * F.f1; B.ra[2] = B.ra[2] + j1.z * F.f1;
        B.ra[0] = B.ra[0] + j2.x * F.f2; B.ra[1] = B.ra[1] + j2.y * F.f2; B.ra[2] = B.ra[2] + j2.z * F.f2;
    }

    // joints (avbd/src/lib.rs:420-539)
    if (v.j_off) {
        const uint32_t kb = v.j_off[i], ke = v.j_off[i + 1];
        for (uint32_t q = kb; q < ke; ++q) {
            const uint32_t cidx = v.j_idx[q];
            const uint4 abk = v.j_abk[cidx];
            const float4 aa4 = v.j_anc_a[cidx];
            const float4 ab4 = v.j_anc_b[cidx];
            const bool is_a = abk.x == i;
            const f3 aa = xyz(aa4), abn = xyz(ab4);
            const f3 pa = is_a ? snap_pos : xyz(rp[abk.x]);
            const f3 pb = !is_a ? snap_pos : xyz(rp[abk.y]);
            const q4 qa = is_a ? snap_rot : as_q(rr[abk.x]);
            const q4 qb = !is_a ? snap_rot : as_q(rr[abk.y]);
            const f3 p_a = add(pa, qrot(qa, aa));
            const f3 p_b = add(pb, qrot(qb, abn));
            const f3 self_anchor = is_a ? aa : abn;
            const float sign = is_a ? 1.0f : -1.0f;
            const f3 delta = sub(p_a, p_b);
            const f3 r_self = qrot(snap_rot, self_anchor);
            const float stiff = ab4.w;
            if (is_bend(abk.z)) {
                // three angular rows: j_lin = 0, j_ang = +-e_r
                const f3 err = bend_error(qa, as_q(aa4), qb);
                const f3 zl{0.0f, 0.0f, 0.0f};
                float4 lam = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pen = make_float4(stiff, stiff, stiff, 0.0f);
                f3 fr{err.x * stiff, err.y * stiff, err.z * stiff};
                if (abk.z == kHardBend) {
                    const f3 d0 = joint_delta(v.init_lin, v.init_ang, abk, aa4, abn);
                    const HardRows hr = hard_rows(abk.z, err, d0, aa4.w, p.alpha);
                    lam = v.j_lam[jin][cidx];
                    pen = v.j_pen[jin][cidx];
                    if (fuse) {
                        hard_update(p, abk.z, hr, lam, pen);
                        if (is_a || v.inv_mm[abk.x].w == 0.0f) {
                            v.j_lam[iter & 1u][cidx] = lam;
                            v.j_pen[iter & 1u][cidx] = pen;
                        }
                    }
                    fr = f3{hard_force(abk.z, pen.x, hr.chat.x, lam.x), hard_force(abk.z, pen.y, hr.chat.y, lam.y),
                            hard_force(abk.z, pen.z, hr.chat.z, lam.z)};
                }
                B.row(zl, scale(f3{1.0f, 0.0f, 0.0f}, sign), fr.x, pen.x);
                B.row(zl, scale(f3{0.0f, 1.0f, 0.0f}, sign), fr.y, pen.y);
                B.row(zl, scale(f3{0.0f, 0.0f, 1.0f}, sign), fr.z, pen.z);
            } else if (is_hard(abk.z)) {
                const f3 d0 = joint_delta(v.init_lin, v.init_ang, abk, aa4, abn);
                const HardRows hr = hard_rows(abk.z, delta, d0, aa4.w, p.alpha);
                float4 lam = v.j_lam[jin][cidx];
                float4 pen = v.j_pen[jin][cidx];
                if (fuse) {
                    // dual update k - 1 of the joint (the same on both bodies)
                    hard_update(p, abk.z, hr, lam, pen);
                    if (is_a || v.inv_mm[abk.x].w == 0.0f) {
                        v.j_lam[iter & 1u][cidx] = lam;
                        v.j_pen[iter & 1u][cidx] = pen;
                    }
                }
                if (abk.z == kHardSocket) {
                    const f3 jx = scale(f3{1.0f, 0.0f, 0.0f}, sign);
                    const f3 jy = scale(f3{0.0f, 1.0f, 0.0f}, sign);
                    const f3 jz = scale(f3{0.0f, 0.0f, 1.0f}, sign);
                    B.row(jx, cross3(r_self, jx), hard_force(abk.z, pen.x, hr.chat.x, lam.x), pen.x);
                    B.row(jy, cross3(r_self, jy), hard_force(abk.z, pen.y, hr.chat.y, lam.y), pen.y);
                    B.row(jz, cross3(r_self, jz), hard_force(abk.z, pen.z, hr.chat.z, lam.z), pen.z);
                } else if (hr.ok) {
                    const float f = hard_force(abk.z, pen.x, hr.chat.x, lam.x);
                    if (abk.z != kHardRope || f > 0.0f) {
                        const f3 jl = scale(hr.dir, sign);
                        B.row(jl, cross3(r_self, jl), f, pen.x);
                    }
                }
            } else if (abk.z == kBallSocket) {
                const f3 jx = scale(f3{1.0f, 0.0f, 0.0f}, sign);
                const f3 jy = scale(f3{0.0f, 1.0f, 0.0f}, sign);
                const f3 jz = scale(f3{0.0f, 0.0f, 1.0f}, sign);
                B.row(jx, cross3(r_self, jx), delta.x * stiff, stiff);
                B.row(jy, cross3(r_self, jy), delta.y * stiff, stiff);
                B.row(jz, cross3(r_self, jz), delta.z * stiff, stiff);
            } else {
                const float len = len3(delta);
                const bool active = len > 1.0e-6f && (abk.z != kRope || len > aa4.w);
                if (active) {
                    const f3 nrm = divs(delta, len);
                    const f3 jl = scale(nrm, sign);
                    B.row(jl, cross3(r_self, jl), stiff * (len - aa4.w), stiff);
                }
            }
        }
    }

    float x[6];
    solve6(B, x);
    const f3 new_pos = add(snap_pos, f3{x[0], x[1], x[2]});
    const q4 new_rot = quat_step(snap_rot, f3{x[3], x[4], x[5]});
    v.pos[1 - parity][i] = f4(new_pos, 0.0f);
    v.rot[1 - parity][i] = f4q(new_rot);
}

template <bool FUSED>
__global__ void __launch_bounds__(kBlock) k_avbd_primal(SolveView v, AvbdParams p, uint32_t iter) {
    const uint32_t i = body_at(v, p, work_chunk(v) * kBlock + threadIdx.x);
    if (i >= p.n) return;
    primal_one<FUSED>(v, p, iter, i);
}

// ---------------------------------------------------------------------------
// Grouped primal: G lanes per body.  A body's contacts and joints ("items")
// are processed in chunks of G; lane l of the group builds the up-to-3
// constraint rows of item chunk*G + l (each row: j_lin, j_ang, force,
// stiffness, 1) and stages them in LDS.  Each lane owns a fixed subset of the
// 27 block entries and adds, for every staged row in item order, that row's
// term (a[p] * b[q]) * w, w the stiffness for the Hessian entries and 1 for
// the gradient entries (multiplying by 1 is exact) -- the same sequence of
// roundings per entry as the one-lane-per-body loop, so the result is
// bit-identical.  A row an item does not have is staged as zeros with force
// and stiffness -0, so each of its terms is exactly -0 and adding it leaves
// every sum unchanged (x + (-0) == x for every x, +0 and -0 included): the
// loop needs no per-row test.  The 27 entries are then gathered through LDS
// and one thread per body runs the 6x6 LDL^T.
//
// Work table: in a step's solve (v.work_tab set) iteration 0 stores each work
// item's body and CSR ranges; iterations >= 1 read them in one load, so the
// body's snapshot and its first CSR entries are fetched together, one
// dependent global load fewer per launch than the order lookup followed by
// the CSR offsets.  Each lane also fetches the CSR entry of its item in the
// next chunk while the current chunk is accumulated.
namespace {
// entry e: term = (R[kP[e]] * R[kQ[e]]) * R[kW[e]];
// R = (j_lin.xyz, j_ang.xyz, force, stiffness, 1)
__constant__ int8_t kP[27] = {0, 1, 2, 1, 2, 2, 3, 4, 5, 4, 5, 5, 3, 3, 3, 4, 4, 4, 5, 5, 5, 0, 1, 2, 3, 4, 5};
__constant__ int8_t kQ[27] = {0, 0, 0, 1, 1, 2, 3, 3, 3, 4, 4, 5, 0, 1, 2, 0, 1, 2, 0, 1, 2, 6, 6, 6, 6, 6, 6};
constexpr int kRowFloats = 9;
constexpr int kItemFloats = 28;  // 3 rows of 9, padded to 16 bytes

// stage row r of an item: (j_lin, j_ang, force, stiffness, 1); RF = 8 leaves
// out the trailing 1 (a reader that applies the row as a whole needs none)
template <int RF = kRowFloats>
__device__ __forceinline__ void stage_row(float* st, int r, f3 jl, f3 ja, float f, float k) {
    float* R = st + RF * r;
    R[0] = jl.x; R[1] = jl.y; R[2] = jl.z;
    R[3] = ja.x; R[4] = ja.y; R[5] = ja.z;
    R[6] = f; R[7] = k;
    if (RF > 8) R[8] = 1.0f;
}
// a row the item does not have: every term it contributes is -0
template <int RF = kRowFloats>
__device__ __forceinline__ void stage_empty_row(float* st, int r) {
    const f3 z{0.0f, 0.0f, 0.0f};
    stage_row<RF>(st, r, z, z, -0.0f, -0.0f);
}

// The body an item's rows are built for: its pose at the iteration's start
// and its displacement from the step's initial pose.
struct RowBody {
    f3 pos;
    q4 rot;
    f3 dsl, dsa;
    uint32_t i;
};

// A contact item's inputs (the CSR entry's contact and the other body),
// loaded apart from the arithmetic so a kernel can fetch them early.
struct ContactIn {
    uint32_t ci, other;
    bool is_a, wall;
    float4 n4, ra4, pen, lam;
    f3 rb;
    q4 orr, oia;  // the other body's rotation and initial rotation
    f3 orp, oil;  // its position and initial position
    float oinv;   // its inverse mass, read when the fused update needs it
};
__device__ __forceinline__ ContactIn load_contact(const SolveView& v, const float4* __restrict__ lam_in,
                                                  const float4* __restrict__ pen_in, const float4* __restrict__ rp,
                                                  const float4* __restrict__ rr, bool fuse, uint32_t ci,
                                                  uint32_t pc) {
    ContactIn c;
    c.ci = ci;
    c.wall = pc == kContainer;
    c.is_a = c.wall || (pc & kPeerIsB) == 0u;
    c.other = c.wall ? kContainer : (pc & ~kPeerIsB);
    c.n4 = v.c_n[ci];
    c.ra4 = v.c_ra[ci];
    c.pen = pen_in[ci];
    c.lam = lam_in[ci];
    c.rb = f3{0.0f, 0.0f, 0.0f};  // a wall contact's r_b is zero
    c.orr = q4{0.0f, 0.0f, 0.0f, 1.0f};
    c.oia = c.orr;
    c.orp = c.rb;
    c.oil = c.rb;
    c.oinv = 1.0f;
    if (!c.wall) {
        c.rb = xyz(v.c_rb[ci]);
        c.orr = as_q(rr[c.other]);
        c.orp = xyz(rp[c.other]);
        c.oil = xyz(v.init_lin[c.other]);
        c.oia = as_q(v.init_ang[c.other]);
        if (fuse && !c.is_a) c.oinv = v.inv_mm[c.other].w;
    }
    return c;
}

// The three rows of a contact item for body b (avbd/src/lib.rs:290-418), with
// the fused schedule's dual update k - 1 first (stored by body a, or by b when
// a is static).
template <int RF = kRowFloats>
__device__ __forceinline__ void stage_contact(const SolveView& v, const AvbdParams& p, float kk, bool fuse,
                                              float4* lam_out, float4* pen_out, const RowBody& b,
                                              const ContactIn& c, float* st) {
    const f3 zero{0.0f, 0.0f, 0.0f};
    const f3 n0 = xyz(c.n4);
    const f3 ra = xyz(c.ra4);
    const ContactFrame cf = contact_frame(v, c.ci, n0, c.ra4, c.wall);
    const f3 t1 = cf.t1, t2 = cf.t2;
    const f3 r0 = c.is_a ? n0 : neg(n0);
    const f3 r1 = c.is_a ? t1 : neg(t1);
    const f3 r2 = c.is_a ? t2 : neg(t2);
    const f3 r_world = qrot(b.rot, c.is_a ? ra : c.rb);
    f3 dol = zero, doa = zero, r_other = zero;
    if (!c.wall) {
        dol = sub(c.orp, c.oil);
        doa = ang_diff(c.orr, c.oia);
        r_other = qrot(c.orr, c.is_a ? c.rb : ra);
    }
    const float4 c04 = cf.c0;
    float4 pen = c.pen;
    float4 lam = c.lam;
    if (fuse) {
        // dual update k - 1 of the contact, in its own (a, b) order
        const DualState d = dual_update(p, kk, n0, t1, t2, c.n4.w, c.is_a ? r_world : r_other,
                                        c.is_a ? r_other : r_world, c.is_a ? b.dsl : dol, c.is_a ? b.dsa : doa,
                                        c.is_a ? dol : b.dsl, c.is_a ? doa : b.dsa, c04, pen, lam);
        pen = d.pen;
        lam = d.lam;
        if (c.is_a || c.oinv == 0.0f) {
            lam_out[c.ci] = lam;
            pen_out[c.ci] = pen;
        }
    }
    const f3 j0 = cross3(r_world, r0), j1 = cross3(r_world, r1), j2 = cross3(r_world, r2);
    float c0 = c04.x * kk, c1 = c04.y * kk, c2 = c04.z * kk;
    c0 = c0 + (dot3(r0, b.dsl) + dot3(j0, b.dsa) - dot3(r0, dol) - dot3(cross3(r_other, r0), doa));
    c1 = c1 + (dot3(r1, b.dsl) + dot3(j1, b.dsa) - dot3(r1, dol) - dot3(cross3(r_other, r1), doa));
    c2 = c2 + (dot3(r2, b.dsl) + dot3(j2, b.dsa) - dot3(r2, dol) - dot3(cross3(r_other, r2), doa));
    const Force3 F = contact_force(pen, lam, c.n4.w, c0, c1, c2);
    stage_row<RF>(st, 0, r0, j0, F.f0, pen.x);
    stage_row<RF>(st, 1, r1, j1, F.f1, pen.y);
    stage_row<RF>(st, 2, r2, j2, F.f2, pen.z);
}

// The rows of joint cidx for body b (avbd/src/lib.rs:422-539; hard and
// bending kinds, with their fused dual update, when HARD); returns how many
// rows it staged (0, 1 or 3).
template <bool HARD>
__device__ __forceinline__ int stage_joint(const SolveView& v, const AvbdParams& p, uint32_t iter, bool fuse,
                                           const float4* __restrict__ rp, const float4* __restrict__ rr,
                                           const RowBody& b, uint32_t cidx, float* st) {
    const f3 zero{0.0f, 0.0f, 0.0f};
    const uint4 abk = v.j_abk[cidx];
    const float4 aa4 = v.j_anc_a[cidx];
    const float4 ab4 = v.j_anc_b[cidx];
    const bool is_a = abk.x == b.i;
    const f3 aa = xyz(aa4), abn = xyz(ab4);
    const f3 pa = is_a ? b.pos : xyz(rp[abk.x]);
    const f3 pb = !is_a ? b.pos : xyz(rp[abk.y]);
    const q4 qa = is_a ? b.rot : as_q(rr[abk.x]);
    const q4 qb = !is_a ? b.rot : as_q(rr[abk.y]);
    const f3 delta = sub(add(pa, qrot(qa, aa)), add(pb, qrot(qb, abn)));
    const f3 r_self = qrot(b.rot, is_a ? aa : abn);
    const float sign = is_a ? 1.0f : -1.0f;
    const float stiff = ab4.w;
    if (HARD && is_bend(abk.z)) {
        // three angular rows: j_lin = 0, j_ang = +-e_r
        const f3 err = bend_error(qa, as_q(aa4), qb);
        float4 lam = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pen = make_float4(stiff, stiff, stiff, 0.0f);
        float fs[3] = {err.x * stiff, err.y * stiff, err.z * stiff};
        if (abk.z == kHardBend) {
            const f3 d0 = joint_delta(v.init_lin, v.init_ang, abk, aa4, abn);
            const HardRows hr = hard_rows(abk.z, err, d0, aa4.w, p.alpha);
            const uint32_t jin = fuse ? (iter - 1) & 1u : 0u;
            lam = v.j_lam[jin][cidx];
            pen = v.j_pen[jin][cidx];
            if (fuse) {
                hard_update(p, abk.z, hr, lam, pen);
                if (is_a || v.inv_mm[abk.x].w == 0.0f) {
                    v.j_lam[iter & 1u][cidx] = lam;
                    v.j_pen[iter & 1u][cidx] = pen;
                }
            }
            fs[0] = hard_force(abk.z, pen.x, hr.chat.x, lam.x);
            fs[1] = hard_force(abk.z, pen.y, hr.chat.y, lam.y);
            fs[2] = hard_force(abk.z, pen.z, hr.chat.z, lam.z);
        }
        const float ks[3] = {pen.x, pen.y, pen.z};
        // e_r * sign, as the one-lane loop forms it (zeros keep the sign's sign)
        const f3 js[3] = {scale(f3{1.0f, 0.0f, 0.0f}, sign), scale(f3{0.0f, 1.0f, 0.0f}, sign),
                          scale(f3{0.0f, 0.0f, 1.0f}, sign)};
#pragma unroll
        for (int r = 0; r < 3; ++r) stage_row(st, r, zero, js[r], fs[r], ks[r]);
        return 3;
    }
    if (HARD && is_hard(abk.z)) {
        const f3 d0 = joint_delta(v.init_lin, v.init_ang, abk, aa4, abn);
        const HardRows hr = hard_rows(abk.z, delta, d0, aa4.w, p.alpha);
        const uint32_t jin = fuse ? (iter - 1) & 1u : 0u;
        float4 lam = v.j_lam[jin][cidx];
        float4 pen = v.j_pen[jin][cidx];
        if (fuse) {
            // dual update k - 1 of the joint (the same on both bodies)
            hard_update(p, abk.z, hr, lam, pen);
            if (is_a || v.inv_mm[abk.x].w == 0.0f) {
                v.j_lam[iter & 1u][cidx] = lam;
                v.j_pen[iter & 1u][cidx] = pen;
            }
        }
        if (abk.z == kHardSocket) {
            const f3 js[3] = {scale(f3{1.0f, 0.0f, 0.0f}, sign), scale(f3{0.0f, 1.0f, 0.0f}, sign),
                              scale(f3{0.0f, 0.0f, 1.0f}, sign)};
            const float fs[3] = {hard_force(abk.z, pen.x, hr.chat.x, lam.x),
                                 hard_force(abk.z, pen.y, hr.chat.y, lam.y),
                                 hard_force(abk.z, pen.z, hr.chat.z, lam.z)};
            const float ks[3] = {pen.x, pen.y, pen.z};
#pragma unroll
            for (int r = 0; r < 3; ++r) stage_row(st, r, js[r], cross3(r_self, js[r]), fs[r], ks[r]);
            return 3;
        }
        if (hr.ok) {
            const float f = hard_force(abk.z, pen.x, hr.chat.x, lam.x);
            if (abk.z != kHardRope || f > 0.0f) {
                const f3 jl = scale(hr.dir, sign);
                stage_row(st, 0, jl, cross3(r_self, jl), f, pen.x);
                return 1;
            }
        }
        return 0;
    }
    if (abk.z == kBallSocket) {
        const f3 js[3] = {scale(f3{1.0f, 0.0f, 0.0f}, sign), scale(f3{0.0f, 1.0f, 0.0f}, sign),
                          scale(f3{0.0f, 0.0f, 1.0f}, sign)};
        const float fs[3] = {delta.x * stiff, delta.y * stiff, delta.z * stiff};
#pragma unroll
        for (int r = 0; r < 3; ++r) stage_row(st, r, js[r], cross3(r_self, js[r]), fs[r], stiff);
        return 3;
    }
    const float len = len3(delta);
    if (len > 1.0e-6f && (abk.z != kRope || len > aa4.w)) {
        const f3 jl = scale(divs(delta, len), sign);
        stage_row(st, 0, jl, cross3(r_self, jl), stiff * (len - aa4.w), stiff);
        return 1;
    }
    return 0;
}
}  // namespace

// HARD: the world has hard joints (their rows and fused dual updates are
// compiled only into that instantiation).
template <int G, bool FUSED, bool HARD>
__global__ void __launch_bounds__(kBlock) k_avbd_primal_grouped(SolveView v, AvbdParams p, uint32_t iter) {
    const uint32_t parity = iter & 1u;
    const bool fuse = FUSED && iter > 0;
    const float4* __restrict__ lam_in = lam_buf(v, fuse ? iter - 1 : 0);
    const float4* __restrict__ pen_in = pen_buf(v, fuse ? iter - 1 : 0);
    float4* lam_out = lam_buf(v, iter);
    float4* pen_out = pen_buf(v, iter);
    constexpr int kBodies = kBlock / G;
    constexpr int kOwned = (27 + G - 1) / G;
    __shared__ __align__(16) float stage[kBodies][G][kItemFloats];
    __shared__ float gathered[kBodies][28];  // the accumulated block
    __shared__ float4 body_f[kBodies][2];    // snapshot pos (.w: active), rot: for the solving thread
    __shared__ uint32_t body_i[kBodies];     // the body of each group
    const int slot = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    const uint32_t chunk = work_chunk(v);
    if (past_colour(v, chunk * kBodies)) return;  // uniform over the workgroup, before any barrier
    const uint32_t wk = chunk * kBodies + slot;   // this group's work item
    const float4* __restrict__ rp = v.pos[parity];
    const float4* __restrict__ rr = v.rot[parity];
    const float dt = p.dt;

    // ---- the body and its item ranges: from the work table (iterations >= 1
    // of a step's solve), or from the work order and the CSR offsets
    uint32_t i, jb = 0, ncontacts = 0, kb = 0, njoints = 0;
    if (v.work_tab && iter > 0) {
        const uint4 w4 = wk < p.n ? v.work_tab[wk] : make_uint4(p.n, 0u, 0u, 0u);
        i = w4.x;
        jb = w4.y;
        ncontacts = w4.z;
        kb = w4.w;
        if (v.j_off && i < p.n) njoints = v.j_off[i + 1] - kb;
    } else {
        i = body_at(v, p, wk);
        if (i < p.n) {
            jb = v.offsets[i];
            ncontacts = v.offsets[i + 1] - jb;
            if (v.j_off) {
                kb = v.j_off[i];
                njoints = v.j_off[i + 1] - kb;
            }
        }
        if (v.work_tab && lane == 0 && wk < p.n) v.work_tab[wk] = make_uint4(i, jb, ncontacts, kb);
    }
    // the CSR entry of the lane's item in the first chunk (fetched beside the
    // snapshot; a static body's entries are fetched and not used)
    uint32_t ci_c = 0u, pc_c = 0u;
    if ((uint32_t)lane < ncontacts) {
        ci_c = v.indices[jb + lane];
        pc_c = v.peers[jb + lane];
    }

    // ---- every lane of the group loads its body's snapshot and computes the
    // initial block entries it owns (no set-up barrier)
    bool active = i < p.n;
    float4 imm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 sp4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sr4 = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float4 il4 = sp4, ia4 = sr4, nl4 = sp4, na4 = sr4;
    if (active) {
        imm = v.inv_mm[i];
        sp4 = rp[i];
        sr4 = rr[i];
        il4 = v.init_lin[i];
        ia4 = v.init_ang[i];
        nl4 = v.iner_lin[i];
        na4 = v.iner_ang[i];
        active = imm.w != 0.0f;
    }
    if (!active) { ncontacts = 0; njoints = 0; }
    const f3 snap_pos = xyz(sp4);
    const q4 snap_rot = as_q(sr4);
    const f3 dsl = active ? sub(snap_pos, xyz(il4)) : f3{0.0f, 0.0f, 0.0f};
    const f3 dsa = active ? ang_diff(snap_rot, as_q(ia4)) : f3{0.0f, 0.0f, 0.0f};
    if (lane == 0) {
        body_f[slot][0] = f4(snap_pos, active ? 1.0f : 0.0f);
        body_f[slot][1] = sr4;
        body_i[slot] = i;
    }
    float acc[kOwned];
    {
        // initial block: diag(M / dt^2, I / dt^2) and the inertial gradient
        // (entry order: lin 0-5, ang 6-11, cross 12-20, rhs lin 21-23, rhs ang 24-26)
        float m_dt2 = 0.0f;
        f3 i_dt2{0.0f, 0.0f, 0.0f}, gl{0.0f, 0.0f, 0.0f}, ga{0.0f, 0.0f, 0.0f};
        if (active) {
            const float mass = 1.0f / imm.w;
            const f3 moment{1.0f / imm.x, 1.0f / imm.y, 1.0f / imm.z};
            const float dt2 = dt * dt;
            m_dt2 = mass / dt2;
            i_dt2 = divs(moment, dt2);
            const f3 dl = sub(snap_pos, xyz(nl4));
            const f3 ad = ang_diff(snap_rot, as_q(na4));
            gl = f3{m_dt2 * dl.x, m_dt2 * dl.y, m_dt2 * dl.z};
            ga = f3{i_dt2.x * ad.x, i_dt2.y * ad.y, i_dt2.z * ad.z};
        }
#pragma unroll
        for (int t = 0; t < kOwned; ++t) {
            const int e = lane + G * t;
            float x = 0.0f;
            x = (e == 0 || e == 3 || e == 5) ? m_dt2 : x;
            x = e == 6 ? i_dt2.x : x;
            x = e == 9 ? i_dt2.y : x;
            x = e == 11 ? i_dt2.z : x;
            x = e == 21 ? gl.x : x;
            x = e == 22 ? gl.y : x;
            x = e == 23 ? gl.z : x;
            x = e == 24 ? ga.x : x;
            x = e == 25 ? ga.y : x;
            x = e == 26 ? ga.z : x;
            acc[t] = x;
        }
    }
    const uint32_t nitems = ncontacts + njoints;
    const float kk = 1.0f - p.alpha;
    const f3 zero{0.0f, 0.0f, 0.0f};
    // offsets of each owned entry's three factors within a row, loaded once
    // (a table lookup inside the row loop would be a memory load per row);
    // lanes past entry 26 read row offset 0 and keep nothing
    int pofs[kOwned], qofs[kOwned], wofs[kOwned];
#pragma unroll
    for (int t = 0; t < kOwned; ++t) {
        const int e = lane + G * t;
        pofs[t] = e < 27 ? (int)kP[e] : 0;
        qofs[t] = e < 27 ? (int)kQ[e] : 0;
        wofs[t] = e < 21 ? 7 : 8;
    }

    for (uint32_t base = 0;; base += G) {
        const uint32_t item = base + (uint32_t)lane;
        // the CSR entry of the lane's item in the next chunk, fetched while
        // this chunk is built and accumulated
        uint32_t ci_n = 0u, pc_n = 0u;
        if (item + G < ncontacts) {
            ci_n = v.indices[jb + item + G];
            pc_n = v.peers[jb + item + G];
        }
        // ---- build the rows of item base + lane
        float* st = stage[slot][lane];
        int nrows = 0;
        if (active && item < nitems) {
            const RowBody rb{snap_pos, snap_rot, dsl, dsa, i};
            if (item < ncontacts) {
                const ContactIn cin = load_contact(v, lam_in, pen_in, rp, rr, fuse, ci_c, pc_c);
                stage_contact(v, p, kk, fuse, lam_out, pen_out, rb, cin, st);
                nrows = 3;
            } else {
                nrows = stage_joint<HARD>(v, p, iter, fuse, rp, rr, rb, v.j_idx[kb + (item - ncontacts)], st);
            }
        }
        if (nrows < 1) stage_empty_row(st, 0);
        if (nrows < 2) stage_empty_row(st, 1);
        if (nrows < 3) stage_empty_row(st, 2);
        __syncthreads();
        // ---- accumulate the owned entries over the staged rows, in item order
        if (active) {
            const int nin = (int)min((uint32_t)G, nitems > base ? nitems - base : 0u);
            // items taken four at a time with one lane per entry: all their
            // LDS reads are issued before the first add
            constexpr int kUnroll = kOwned == 1 ? 4 : 2;
#pragma unroll kUnroll
            for (int c = 0; c < nin; ++c) {
                const float* it = stage[slot][c];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float* R = it + kRowFloats * r;
#pragma unroll
                    for (int t = 0; t < kOwned; ++t) acc[t] = acc[t] + (R[pofs[t]] * R[qofs[t]]) * R[wofs[t]];
                }
            }
        }
        const bool more = active && base + G < nitems;
        if (!__syncthreads_or(more)) break;
        ci_c = ci_n;
        pc_c = pc_n;
    }
    // ---- gather the 27 entries; one thread per body solves (wave 0)
    if (active) {
#pragma unroll
        for (int t = 0; t < kOwned; ++t) {
            const int e = lane + G * t;
            if (e < 27) gathered[slot][e] = acc[t];
        }
    }
    __syncthreads();
    if (threadIdx.x >= (unsigned)kBodies) return;
    const int t = threadIdx.x;
    const float4 f0 = body_f[t][0];
    if (f0.w == 0.0f) return;
    const uint32_t bi = body_i[t];
    const float* gs = gathered[t];
    Block6 B;
#pragma unroll
    for (int k = 0; k < 6; ++k) { B.l[k] = gs[k]; B.a[k] = gs[6 + k]; }
#pragma unroll
    for (int k = 0; k < 9; ++k) B.c[k] = gs[12 + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) { B.rl[k] = gs[21 + k]; B.ra[k] = gs[24 + k]; }
    float x[6];
    solve6(B, x);
    const q4 sr = as_q(body_f[t][1]);
    const f3 new_pos = add(xyz(f0), f3{x[0], x[1], x[2]});
    const q4 new_rot = quat_step(sr, f3{x[3], x[4], x[5]});
    v.pos[1 - parity][bi] = f4(new_pos, 0.0f);
    v.rot[1 - parity][bi] = f4q(new_rot);
}

// Packed primal (the Jacobi solve of a step; no counterpart in the
// reference's thread-per-body kernel).  A workgroup of kPackThreads threads
// takes kPackBodies consecutive work items.  Their items -- each body's
// contacts, then its joints, body after body -- form one list, and in round r
// thread t builds the rows of list entry r * kPackThreads + t, so nearly every
// lane of a round builds an item however unevenly the items fall on the
// bodies (with G lanes per body, the lanes past a body's item count sit idle
// through the costliest phase).  Each body's kPackLanes lanes then add its
// entries over its own range of the list in item order, as in the grouped
// kernel (bit-identical to the one-lane loop), and one thread per body
// solves.
//
// Iteration 0 of a solve stores, per work item, the body, its range in the
// list and its CSR ranges (v.work_tab, two entries per work item), and per
// workgroup the identity of its first kPackCap list entries (v.item_tab:
// contact and peer code, or constraint).  Iterations >= 1 read both in their
// first load.
namespace {
constexpr int kPackBodies = 8;
constexpr int kPackLanes = 32;
constexpr int kPackThreads = kPackBodies * kPackLanes;
constexpr int kPackCap = kPackThreads;  // list entries per workgroup in the item table (round 0)
}  // namespace

template <bool FUSED, bool HARD>
__global__ void __launch_bounds__(kPackThreads) k_avbd_primal_packed(SolveView v, AvbdParams p, uint32_t iter) {
    constexpr int G = kPackLanes;
    constexpr int kOwned = (27 + G - 1) / G;
    const uint32_t parity = iter & 1u;
    const bool fuse = FUSED && iter > 0;
    const float4* __restrict__ lam_in = lam_buf(v, fuse ? iter - 1 : 0);
    const float4* __restrict__ pen_in = pen_buf(v, fuse ? iter - 1 : 0);
    float4* lam_out = lam_buf(v, iter);
    float4* pen_out = pen_buf(v, iter);
    __shared__ __align__(16) float stage[kPackThreads][kItemFloats];
    __shared__ float gathered[kPackBodies][28];
    __shared__ float4 body_f[kPackBodies][4];  // pos (.w: active), rot, dsl, dsa of each slot's body
    __shared__ uint4 slot_csr[kPackBodies];    // body, CSR start, contacts, joint-CSR start
    __shared__ uint32_t slot_n[kPackBodies];   // list entries of each slot (0: inactive)
    const int slot = threadIdx.x / G;
    const int lane = threadIdx.x % G;
    const uint32_t chunk = work_chunk(v);
    const uint32_t wk = chunk * kPackBodies + slot;
    const float4* __restrict__ rp = v.pos[parity];
    const float4* __restrict__ rr = v.rot[parity];
    const bool from_tab = iter > 0;
    uint4* __restrict__ wt = v.work_tab;
    uint2* __restrict__ itab = v.item_tab + (size_t)chunk * kPackCap;

    // ---- the slot's body and ranges, and (iterations >= 1) this thread's
    // list entries of rounds 0 and 1, in one round of loads
    uint32_t i = p.n, jb = 0, nc = 0, kb = 0, nj = 0, pg = 0, ng = 0, total = 0;
    uint2 pt0 = make_uint2(0u, 0u);
    if (from_tab) {
        const uint4 a = wt[2 * wk], b = wt[2 * wk + 1];
        pt0 = itab[threadIdx.x];
        i = a.x; pg = a.y; ng = a.z; total = a.w;
        jb = b.x; nc = b.y; kb = b.z; nj = b.w;
    } else {
        i = body_at(v, p, wk);
        if (i < p.n) {
            jb = v.offsets[i];
            nc = v.offsets[i + 1] - jb;
            if (v.j_off) {
                kb = v.j_off[i];
                nj = v.j_off[i + 1] - kb;
            }
        }
    }
    // ---- the body's snapshot and the initial block entries this lane owns
    bool active = i < p.n;
    float4 imm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 sp4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sr4 = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float4 il4 = sp4, ia4 = sr4, nl4 = sp4, na4 = sr4;
    if (active) {
        imm = v.inv_mm[i];
        sp4 = rp[i];
        sr4 = rr[i];
        il4 = v.init_lin[i];
        ia4 = v.init_ang[i];
        nl4 = v.iner_lin[i];
        na4 = v.iner_ang[i];
        active = imm.w != 0.0f;
    }
    if (!active) { nc = 0; nj = 0; }
    const f3 snap_pos = xyz(sp4);
    const q4 snap_rot = as_q(sr4);
    const f3 dsl = active ? sub(snap_pos, xyz(il4)) : f3{0.0f, 0.0f, 0.0f};
    const f3 dsa = active ? ang_diff(snap_rot, as_q(ia4)) : f3{0.0f, 0.0f, 0.0f};
    if (lane == 0) {
        body_f[slot][0] = f4(snap_pos, active ? 1.0f : 0.0f);
        body_f[slot][1] = sr4;
        body_f[slot][2] = f4(dsl, 0.0f);
        body_f[slot][3] = f4(dsa, 0.0f);
        slot_csr[slot] = make_uint4(i, jb, nc, kb);
        slot_n[slot] = nc + nj;
    }
    float acc[kOwned];
    {
        // initial block: diag(M / dt^2, I / dt^2) and the inertial gradient
        // (entry order: lin 0-5, ang 6-11, cross 12-20, rhs lin 21-23, rhs ang 24-26)
        float m_dt2 = 0.0f;
        f3 i_dt2{0.0f, 0.0f, 0.0f}, gl{0.0f, 0.0f, 0.0f}, ga{0.0f, 0.0f, 0.0f};
        if (active) {
            const float mass = 1.0f / imm.w;
            const f3 moment{1.0f / imm.x, 1.0f / imm.y, 1.0f / imm.z};
            const float dt2 = p.dt * p.dt;
            m_dt2 = mass / dt2;
            i_dt2 = divs(moment, dt2);
            const f3 dl = sub(snap_pos, xyz(nl4));
            const f3 ad = ang_diff(snap_rot, as_q(na4));
            gl = f3{m_dt2 * dl.x, m_dt2 * dl.y, m_dt2 * dl.z};
            ga = f3{i_dt2.x * ad.x, i_dt2.y * ad.y, i_dt2.z * ad.z};
        }
#pragma unroll
        for (int t = 0; t < kOwned; ++t) {
            const int e = lane + G * t;
            float x = 0.0f;
            x = (e == 0 || e == 3 || e == 5) ? m_dt2 : x;
            x = e == 6 ? i_dt2.x : x;
            x = e == 9 ? i_dt2.y : x;
            x = e == 11 ? i_dt2.z : x;
            x = e == 21 ? gl.x : x;
            x = e == 22 ? gl.y : x;
            x = e == 23 ? gl.z : x;
            x = e == 24 ? ga.x : x;
            x = e == 25 ? ga.y : x;
            x = e == 26 ? ga.z : x;
            acc[t] = x;
        }
    }
    int pofs[kOwned], qofs[kOwned], wofs[kOwned];
#pragma unroll
    for (int t = 0; t < kOwned; ++t) {
        const int e = lane + G * t;
        pofs[t] = e < 27 ? (int)kP[e] : 0;
        qofs[t] = e < 27 ? (int)kQ[e] : 0;
        wofs[t] = e < 21 ? 7 : 8;
    }
    __syncthreads();
    // ---- list starts of the slots (a prefix over the slots' entry counts)
    uint32_t pstart[kPackBodies];
    {
        uint32_t run = 0;
#pragma unroll
        for (int s = 0; s < kPackBodies; ++s) {
            pstart[s] = run;
            run += slot_n[s];
        }
        if (!from_tab) {
            total = run;
            pg = pstart[slot];
            ng = slot_n[slot];
            if (lane == 0) {
                wt[2 * wk] = make_uint4(i, pg, ng, total);
                wt[2 * wk + 1] = make_uint4(jb, nc, kb, nj);
            }
        }
    }
    const float kk = 1.0f - p.alpha;
    const uint32_t rounds = (total + kPackThreads - 1) / kPackThreads;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t f = r * kPackThreads + threadIdx.x;
        float* st = stage[threadIdx.x];
        int nrows = 0;
        if (f < total) {
            // ---- which item: its slot and place in the slot's range from the
            // list starts; its contact and peer code (or constraint) from the
            // item table, or from the CSR
            uint32_t s = 0;
#pragma unroll
            for (int q = 1; q < kPackBodies; ++q) s += pstart[q] <= f ? 1u : 0u;
            uint32_t ps = 0;
#pragma unroll
            for (int q = 0; q < kPackBodies; ++q) ps = (uint32_t)q == s ? pstart[q] : ps;
            const uint32_t c = f - ps;
            const uint4 sc = slot_csr[s];
            const bool joint = c >= sc.z;
            uint32_t x, y;
            if (from_tab && f < (uint32_t)kPackCap) {
                x = pt0.x;
                y = pt0.y;
            } else {
                if (!joint) {
                    x = v.indices[sc.y + c];
                    y = v.peers[sc.y + c];
                } else {
                    x = v.j_idx[sc.w + (c - sc.z)];
                    y = 0u;
                }
                if (!from_tab && f < (uint32_t)kPackCap) itab[f] = make_uint2(x, y);
            }
            RowBody rb;
            {
                const float4 b0 = body_f[s][0];
                rb.pos = xyz(b0);
                rb.rot = as_q(body_f[s][1]);
                rb.dsl = xyz(body_f[s][2]);
                rb.dsa = xyz(body_f[s][3]);
                rb.i = slot_csr[s].x;
            }
            if (!joint) {
                const ContactIn cin = load_contact(v, lam_in, pen_in, rp, rr, fuse, x, y);
                stage_contact(v, p, kk, fuse, lam_out, pen_out, rb, cin, st);
                nrows = 3;
            } else {
                nrows = stage_joint<HARD>(v, p, iter, fuse, rp, rr, rb, x, st);
            }
        }
        if (nrows < 1) stage_empty_row(st, 0);
        if (nrows < 2) stage_empty_row(st, 1);
        if (nrows < 3) stage_empty_row(st, 2);
        __syncthreads();
        // ---- accumulate the owned entries over this round's part of the
        // slot's range, in item order
        if (active) {
            const uint32_t r0 = r * kPackThreads;
            const uint32_t lo = pg > r0 ? pg : r0;
            const uint32_t end = pg + ng, r1 = r0 + kPackThreads;
            const uint32_t hi = end < r1 ? end : r1;
#pragma unroll 2
            for (uint32_t q = lo; q < hi; ++q) {
                const float* it = stage[q - r0];
#pragma unroll
                for (int rw = 0; rw < 3; ++rw) {
                    const float* R = it + kRowFloats * rw;
#pragma unroll
                    for (int t = 0; t < kOwned; ++t) acc[t] = acc[t] + (R[pofs[t]] * R[qofs[t]]) * R[wofs[t]];
                }
            }
        }
        __syncthreads();
    }
    // ---- gather the 27 entries; one thread per body solves
    if (active) {
#pragma unroll
        for (int t = 0; t < kOwned; ++t) {
            const int e = lane + G * t;
            if (e < 27) gathered[slot][e] = acc[t];
        }
    }
    __syncthreads();
    if (threadIdx.x >= (unsigned)kPackBodies) return;
    const int t = threadIdx.x;
    const float4 f0 = body_f[t][0];
    if (f0.w == 0.0f) return;
    const uint32_t bi = slot_csr[t].x;
    const float* gs = gathered[t];
    Block6 B;
#pragma unroll
    for (int k = 0; k < 6; ++k) { B.l[k] = gs[k]; B.a[k] = gs[6 + k]; }
#pragma unroll
    for (int k = 0; k < 9; ++k) B.c[k] = gs[12 + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) { B.rl[k] = gs[21 + k]; B.ra[k] = gs[24 + k]; }
    float x[6];
    solve6(B, x);
    const q4 sr = as_q(body_f[t][1]);
    const f3 new_pos = add(xyz(f0), f3{x[0], x[1], x[2]});
    const q4 new_rot = quat_step(sr, f3{x[3], x[4], x[5]});
    v.pos[1 - parity][bi] = f4(new_pos, 0.0f);
    v.rot[1 - parity][bi] = f4q(new_rot);
}

// Span primal (the Jacobi solve of a step for bodies with few contacts; no
// counterpart in the reference's thread-per-body kernel).  A workgroup takes
// kSpanBodies consecutive bodies in index order, one thread each, and with
// them one contiguous range of the CSR, [offsets[i0], offsets[i0 +
// kSpanBodies]).  In round r thread t builds the rows of range entry
// r * kSpanBodies + t (the next round's CSR entry fetched ahead) and stages
// them in LDS, without each row's trailing 1: consecutive threads read
// consecutive CSR entries, and a body's
// contacts with the container or as body a are consecutive records, so the
// gathers that one thread per body would make entry by entry become
// neighbouring loads.  Each thread then adds the staged rows of its own
// body's entries, in CSR order, into its block (Block6::row: per entry the
// operations of the one-lane loop, in its order) and, after the last round,
// solves.  Worlds without joints.
constexpr int kSpanBodies = 256;  // bodies (threads) per workgroup of the span primal

template <bool FUSED>
__global__ void __launch_bounds__(kSpanBodies) k_avbd_primal_span(SolveView v, AvbdParams p, uint32_t iter) {
    constexpr uint32_t SB = kSpanBodies;
    constexpr int kSteps = __builtin_ctz(SB);  // binary-search steps over the slot starts
    const uint32_t parity = iter & 1u;
    const bool fuse = FUSED && iter > 0;
    const float4* __restrict__ lam_in = lam_buf(v, fuse ? iter - 1 : 0);
    const float4* __restrict__ pen_in = pen_buf(v, fuse ? iter - 1 : 0);
    float4* lam_out = lam_buf(v, iter);
    float4* pen_out = pen_buf(v, iter);
    constexpr int kSpanRow = 8;                  // rows staged without their trailing 1
    __shared__ __align__(16) float stage[SB][3 * kSpanRow];
    __shared__ float4 body_f[SB][3];   // rot; dsl, dsa.x; dsa.y, dsa.z, active of each thread's body
    __shared__ uint32_t offs[SB + 1];  // the bodies' CSR starts, and the range's end
    const float4* __restrict__ rp = v.pos[parity];
    const float4* __restrict__ rr = v.rot[parity];
    const uint32_t t = threadIdx.x;
    const uint32_t i0 = work_chunk(v) * SB;
    const uint32_t i = i0 + t;
    offs[t] = v.offsets[i < p.n ? i : p.n];
    if (t == 0) offs[SB] = v.offsets[i0 + SB < p.n ? i0 + SB : p.n];
    // ---- the body's snapshot and initial block (as the one-lane loop)
    bool active = i < p.n;
    float4 imm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 sp4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sr4 = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float4 il4 = sp4, ia4 = sr4, nl4 = sp4, na4 = sr4;
    if (active) {
        imm = v.inv_mm[i];
        sp4 = rp[i];
        sr4 = rr[i];
        il4 = v.init_lin[i];
        ia4 = v.init_ang[i];
        nl4 = v.iner_lin[i];
        na4 = v.iner_ang[i];
    }
    active = active && imm.w != 0.0f;
    const f3 snap_pos = xyz(sp4);
    const q4 snap_rot = as_q(sr4);
    Block6 B;
    {
        const float mass = 1.0f / imm.w;
        const f3 moment{1.0f / imm.x, 1.0f / imm.y, 1.0f / imm.z};
        const float dt2 = p.dt * p.dt;
        const float m_dt2 = mass / dt2;
        const f3 i_dt2 = divs(moment, dt2);
#pragma unroll
        for (int k = 0; k < 6; ++k) { B.l[k] = 0.0f; B.a[k] = 0.0f; }
#pragma unroll
        for (int k = 0; k < 9; ++k) B.c[k] = 0.0f;
        B.l[0] = m_dt2; B.l[3] = m_dt2; B.l[5] = m_dt2;
        B.a[0] = i_dt2.x; B.a[3] = i_dt2.y; B.a[5] = i_dt2.z;
        const f3 rl = scale(sub(snap_pos, xyz(nl4)), 1.0f);
        B.rl[0] = m_dt2 * rl.x; B.rl[1] = m_dt2 * rl.y; B.rl[2] = m_dt2 * rl.z;
        const f3 ad = ang_diff(snap_rot, as_q(na4));
        B.ra[0] = i_dt2.x * ad.x; B.ra[1] = i_dt2.y * ad.y; B.ra[2] = i_dt2.z * ad.z;
    }
    const f3 dsl = active ? sub(snap_pos, xyz(il4)) : f3{0.0f, 0.0f, 0.0f};
    const f3 dsa = active ? ang_diff(snap_rot, as_q(ia4)) : f3{0.0f, 0.0f, 0.0f};
    body_f[t][0] = sr4;
    body_f[t][1] = f4(dsl, dsa.x);
    body_f[t][2] = make_float4(dsa.y, dsa.z, active ? 1.0f : 0.0f, 0.0f);
    __syncthreads();
    const uint32_t o0 = offs[0], o1 = offs[SB];
    const uint32_t my0 = offs[t], my1 = offs[t + 1];
    const float kk = 1.0f - p.alpha;
    // the CSR entry of this thread's first round, fetched ahead
    uint32_t nci = 0, npc = 0;
    if (o0 + t < o1) {
        nci = v.indices[o0 + t];
        npc = v.peers[o0 + t];
    }
    for (uint32_t r0 = o0; r0 < o1; r0 += SB) {
        const uint32_t f = r0 + t;
        const uint32_t ci = nci, pc = npc;
        // the next round's entry, in flight while this round is built and added
        if (f + SB < o1) {
            nci = v.indices[f + SB];
            npc = v.peers[f + SB];
        }
        float* st = stage[t];
        bool built = false;
        if (f < o1) {
            // ---- the entry's body: the last slot whose start is at or below f
            // (offs[0] <= f < offs[SB]; an empty slot never ends the search)
            uint32_t s = 0, hi = SB;
#pragma unroll
            for (int step = 0; step < kSteps; ++step) {
                const uint32_t mid = (s + hi) >> 1;
                if (offs[mid] <= f) s = mid;
                else hi = mid;
            }
            const float4 b2 = body_f[s][2];
            if (b2.z != 0.0f) {
                const float4 b1 = body_f[s][1];
                RowBody rb;  // (a contact's rows use no position)
                rb.pos = f3{0.0f, 0.0f, 0.0f};
                rb.rot = as_q(body_f[s][0]);
                rb.dsl = xyz(b1);
                rb.dsa = f3{b1.w, b2.x, b2.y};
                rb.i = i0 + s;
                const ContactIn cin = load_contact(v, lam_in, pen_in, rp, rr, fuse, ci, pc);
                stage_contact<kSpanRow>(v, p, kk, fuse, lam_out, pen_out, rb, cin, st);
                built = true;
            }
        }
        if (!built) {
            stage_empty_row<kSpanRow>(st, 0);
            stage_empty_row<kSpanRow>(st, 1);
            stage_empty_row<kSpanRow>(st, 2);
        }
        __syncthreads();
        // ---- this body's entries of the round, in CSR order
        if (active) {
            const uint32_t lo = my0 > r0 ? my0 : r0;
            const uint32_t end = r0 + SB;
            const uint32_t hi = my1 < end ? my1 : end;
            for (uint32_t q = lo; q < hi; ++q) {
                const float* it = stage[q - r0];
#pragma unroll
                for (int rw = 0; rw < 3; ++rw) {
                    const float* R = it + kSpanRow * rw;
                    B.row(f3{R[0], R[1], R[2]}, f3{R[3], R[4], R[5]}, R[6], R[7]);
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    float x[6];
    solve6(B, x);
    const f3 new_pos = add(snap_pos, f3{x[0], x[1], x[2]});
    const q4 new_rot = quat_step(snap_rot, f3{x[3], x[4], x[5]});
    v.pos[1 - parity][i] = f4(new_pos, 0.0f);
    v.rot[1 - parity][i] = f4q(new_rot);
}

// One dual update of contact ci from the Jacobi poses written by a primal of
// parity `parity` (pos[1 - parity]): multipliers read from buffer src, stored
// to buffer dst.  static_updates > 0 (fused schedule, final launch): a
// contact between two static bodies starts from buffer 0 and takes that many
// updates instead (its poses never move, no primal carried its updates).
__device__ __forceinline__ void dual_one(const SolveView& v, const AvbdParams& p, uint32_t ci, uint32_t parity,
                                         uint32_t src, uint32_t dst, uint32_t static_updates) {
    const float4* __restrict__ wp = v.pos[1 - parity];
    const float4* __restrict__ wr = v.rot[1 - parity];
    const float kk = 1.0f - p.alpha;
    const f3 zero{0.0f, 0.0f, 0.0f};
    const uint2 ab = v.c_ab[ci];
    const q4 rot_a = as_q(wr[ab.x]);
    const f3 dal = sub(xyz(wp[ab.x]), xyz(v.init_lin[ab.x]));
    const f3 daa = ang_diff(rot_a, as_q(v.init_ang[ab.x]));
    f3 dbl = zero, dba = zero, rbw = zero;
    if (ab.y != kContainer) {
        const q4 rot_b = as_q(wr[ab.y]);
        dbl = sub(xyz(wp[ab.y]), xyz(v.init_lin[ab.y]));
        dba = ang_diff(rot_b, as_q(v.init_ang[ab.y]));
        rbw = qrot(rot_b, xyz(v.c_rb[ci]));
    }
    const float4 ra4 = v.c_ra[ci];
    const f3 raw = qrot(rot_a, xyz(ra4));
    const float4 n4 = v.c_n[ci];
    const f3 r0 = xyz(n4);
    const ContactFrame cf = contact_frame(v, ci, r0, ra4, ab.y == kContainer);
    const f3 r1 = cf.t1, r2 = cf.t2;
    const float4 c04 = cf.c0;
    uint32_t reps = 1, from = src;
    if (static_updates > 0u) {
        const bool moving = v.inv_mm[ab.x].w != 0.0f || (ab.y != kContainer && v.inv_mm[ab.y].w != 0.0f);
        if (!moving) {
            reps = static_updates;
            from = 0;
        }
    }
    DualState d;
    d.lam = lam_buf(v, from)[ci];
    d.pen = pen_buf(v, from)[ci];
    for (uint32_t k = 0; k < reps; ++k)
        d = dual_update(p, kk, r0, r1, r2, n4.w, raw, rbw, dal, daa, dbl, dba, c04, d.pen, d.lam);
    lam_buf(v, dst)[ci] = d.lam;
    pen_buf(v, dst)[ci] = d.pen;
}

// One dual update of hard joint cj from the poses pos[1 - parity], from
// multiplier buffer src (static_updates > 0: a joint between two static
// bodies starts from buffer 0 and takes that many, as dual_one).
struct JointMults {
    float4 lam, pen;
};
__device__ __forceinline__ JointMults joint_dual_one(const SolveView& v, const AvbdParams& p, uint32_t cj,
                                                     uint32_t parity, uint32_t src, uint32_t static_updates) {
    const uint4 abk = v.j_abk[cj];
    const float4 aa4 = v.j_anc_a[cj];
    const f3 ab = xyz(v.j_anc_b[cj]);
    const f3 delta = joint_delta(v.pos[1 - parity], v.rot[1 - parity], abk, aa4, ab);
    const f3 d0 = joint_delta(v.init_lin, v.init_ang, abk, aa4, ab);
    const HardRows hr = hard_rows(abk.z, delta, d0, aa4.w, p.alpha);
    uint32_t reps = 1, from = src;
    if (static_updates > 0u && v.inv_mm[abk.x].w == 0.0f && v.inv_mm[abk.y].w == 0.0f) {
        reps = static_updates;
        from = 0;
    }
    JointMults m{v.j_lam[from][cj], v.j_pen[from][cj]};
    for (uint32_t k = 0; k < reps; ++k) hard_update(p, abk.z, hr, m.lam, m.pen);
    return m;
}

// standalone dual update (unfused schedule)
__global__ void __launch_bounds__(kBlock) k_avbd_dual(SolveView v, AvbdParams p, uint32_t parity, uint32_t src,
                                                      uint32_t dst) {
    const uint32_t m = v.m_dev ? *v.m_dev : v.m_host;
    for (uint32_t ci = blockIdx.x * kBlock + threadIdx.x; ci < m; ci += gridDim.x * kBlock)
        dual_one(v, p, ci, parity, src, dst, 0u);
    if (v.j_hard)
        for (uint32_t cj = blockIdx.x * kBlock + threadIdx.x; cj < v.j_count; cj += gridDim.x * kBlock) {
            if (!is_hard(v.j_abk[cj].z)) continue;
            const JointMults jm = joint_dual_one(v, p, cj, parity, src, 0u);
            v.j_lam[dst][cj] = jm.lam;
            v.j_pen[dst][cj] = jm.pen;
        }
}

__device__ __forceinline__ void recover_one(const SolveView& v, const AvbdParams& p, uint32_t i, uint32_t parity);

// The last dual update (into the contact set's own lam / pen) and recover in
// one launch: the dual reads only the final Jacobi poses and the initial
// poses, which recover does not write.
// recover alone (the replay keeps every iteration's multipliers)
__global__ void __launch_bounds__(kBlock) k_avbd_recover(SolveView v, AvbdParams p, uint32_t parity) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < p.n) recover_one(v, p, i, parity);
}

__global__ void __launch_bounds__(kBlock) k_avbd_dual_recover(SolveView v, AvbdParams p, uint32_t iters,
                                                              uint32_t fused) {
    const uint32_t m = v.m_dev ? *v.m_dev : v.m_host;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t parity = (iters - 1) & 1u;
    const uint32_t src = fused ? parity : 0u;
    if (v.stat_acc && blockIdx.x == 0 && threadIdx.x == 0) {
        v.stat_acc[0] += *v.stat_pairs;
        v.stat_acc[1] += m;
        v.stat_acc[2] += v.offsets[p.n];
    }
    for (uint32_t ci = blockIdx.x * kBlock + threadIdx.x; ci < m; ci += stride)
        dual_one(v, p, ci, parity, src, 0u, fused ? iters : 0u);
    // hard joints: the last dual update into the state carried to the next
    // step, and that step's warm start when this launch prepares its init
    if (v.j_hard)
        for (uint32_t cj = blockIdx.x * kBlock + threadIdx.x; cj < v.j_count; cj += stride) {
            if (!is_hard(v.j_abk[cj].z)) continue;
            const JointMults jm = joint_dual_one(v, p, cj, parity, src, fused ? iters : 0u);
            v.j_st_lam[cj] = jm.lam;
            v.j_st_pen[cj] = jm.pen;
            if (v.next.init_lin) {
                float4 lam, pen;
                hard_warm(p, jm.lam, jm.pen, lam, pen);
                v.j_lam[0][cj] = lam;
                v.j_pen[0][cj] = pen;
            }
        }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < p.n; i += stride) recover_one(v, p, i, parity);
}

__device__ __forceinline__ void recover_one(const SolveView& v, const AvbdParams& p, uint32_t i, uint32_t parity) {
    const float inv_mass = v.inv_mm[i].w;
    if (inv_mass == 0.0f) {
        if (v.next.init_lin)
            init_body(v.next, p, i, v.cur_pos[i], v.cur_rot[i], xyz(v.vel[i]), xyz(v.ang[i]), xyz(v.prev_vel[i]),
                      inv_mass);
        return;
    }
    const float4 pos4 = v.pos[1 - parity][i];
    const float4 rot4 = v.rot[1 - parity][i];
    const float dt = p.dt;
    const float4 vel = v.vel[i];
    v.prev_vel[i] = vel;
    const f3 nv = divs(sub(xyz(pos4), xyz(v.init_lin[i])), dt);
    v.vel[i] = f4(nv, 0.0f);
    const q4 rel = qmul(as_q(rot4), qconj(as_q(v.init_ang[i])));
    const float4 nang = make_float4(2.0f * rel.x / dt, 2.0f * rel.y / dt, 2.0f * rel.z / dt, 0.0f);
    v.ang[i] = nang;
    if (v.out_pos) {
        v.out_pos[i] = pos4;
        v.out_rot[i] = rot4;
    }
    if (v.next.init_lin) init_body(v.next, p, i, pos4, rot4, nv, xyz(nang), xyz(vel), inv_mass);
}

// The whole solve of a scene table in one launch (no counterpart in the
// reference).  Every contact of a batched world joins two bodies of one scene
// (or a body and its scene's floor), so a scene's primal, dual and recover
// read and write only that scene's bodies and contacts, and the launch-wide
// order of the Jacobi schedule (every primal before any dual of the same
// iteration) is needed only within each scene: a workgroup barrier stands in
// for the kernel boundary.  A workgroup takes a bundle (a body range made of
// whole consecutive scenes) from a queue, most bodies first so the long ones
// start early, runs its whole solve, and takes the next until the queue is
// empty; the last workgroup to finish resets the queue for the next launch.
// Per body and per contact the arithmetic is that of the separate launches,
// so the result is bit-identical.  A contact's dual update is run by the
// thread that holds its body a (each contact is in a's CSR list once, with
// the a side flagged).  The step's init runs before (the previous step's
// final launch, or its own launch).  On the 4096-scene table this launch is
// slower than the separate ones (0.63-0.68 ms against 0.44 ms of solve per
// step at I = 4): a workgroup waits at every barrier for its slowest body
// while the separate launches overlap every workgroup's latency, the largest
// scenes' chain of phases bounds the launch, and the dual walks the CSR per
// body instead of the contiguous contact arrays.
template <bool FUSED, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4)))
k_avbd_scenes(SolveView v, AvbdParams p, SceneQueue q, uint32_t iters) {
    __shared__ uint32_t s_slot;
    const uint32_t parity = (iters - 1) & 1u;
    if (v.stat_acc && blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t m = v.m_dev ? *v.m_dev : v.m_host;
        v.stat_acc[0] += *v.stat_pairs;
        v.stat_acc[1] += m;
        v.stat_acc[2] += v.offsets[p.n];
    }
    for (;;) {
        __syncthreads();  // every thread has read the previous slot
        if (threadIdx.x == 0) s_slot = atomicAdd(&q.next[0], 1u);
        __syncthreads();
        const uint32_t slot = s_slot;
        if (slot >= q.count) break;  // uniform over the workgroup
        const uint32_t sc = q.order[slot];
        const uint32_t b0 = q.body_off[sc], b1 = q.body_off[sc + 1];
        for (uint32_t k = 0; k < iters; ++k) {
            for (uint32_t i = b0 + threadIdx.x; i < b1; i += BS) primal_one<FUSED>(v, p, k, i);
            __syncthreads();
            if (!FUSED && k + 1 < iters) {
                for (uint32_t i = b0 + threadIdx.x; i < b1; i += BS)
                    for (uint32_t j = v.offsets[i]; j < v.offsets[i + 1]; ++j) {
                        const uint32_t pc = v.peers[j];
                        if (pc == kContainer || (pc & kPeerIsB) == 0u) dual_one(v, p, v.indices[j], k & 1u, 0u, 0u, 0u);
                    }
                __syncthreads();
            }
        }
        // the last dual update and recover (as k_avbd_dual_recover: the dual
        // reads only the final Jacobi poses and the initial poses)
        for (uint32_t i = b0 + threadIdx.x; i < b1; i += BS) {
            for (uint32_t j = v.offsets[i]; j < v.offsets[i + 1]; ++j) {
                const uint32_t pc = v.peers[j];
                if (pc == kContainer || (pc & kPeerIsB) == 0u)
                    dual_one(v, p, v.indices[j], parity, FUSED ? parity : 0u, 0u, FUSED ? iters : 0u);
            }
            recover_one(v, p, i, parity);
        }
    }
    if (threadIdx.x == 0 && atomicAdd(&q.next[1], 1u) + 1u == gridDim.x) {
        // every other workgroup has taken its last slot
        q.next[0] = 0u;
        q.next[1] = 0u;
    }
}

// Colours first .. ncol - 1 (ncol = *ncol_at) one after another in one
// workgroup, with a barrier between colours: the one-lane primal of every
// body of the colour (bit-identical to every grouped shape).
__global__ void __launch_bounds__(kBlock) k_avbd_colour_tail(SolveView v, AvbdParams p, uint32_t iter,
                                                           const uint32_t* starts, uint32_t first,
                                                           const uint32_t* ncol_at) {
    const uint32_t ncol = *ncol_at;
    for (uint32_t c = first; c < ncol; ++c) {
        const uint32_t b0 = starts[c], b1 = starts[c + 1];
        for (uint32_t t = b0 + threadIdx.x; t < b1; t += kBlock) primal_one<false>(v, p, iter, v.body_list[t]);
        __syncthreads();  // the next colour reads these poses
    }
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

namespace {
// Launch with optional start / stop events carried by the dispatch itself
// (hipExtLaunchKernelGGL), so profile mode times the kernel alone, with no
// marker packets between kernels.
template <typename F, typename... Args>
void launch_k(F kernel, dim3 grid, dim3 block, hipStream_t s, const LaunchEvents& ev, Args... args) {
    if (ev.start)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev.start, ev.stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}
}  // namespace

size_t primal_work_entries(uint32_t n) { return 2 * (((size_t)n + kPackBodies - 1) / kPackBodies) * kPackBodies; }
size_t primal_item_entries(uint32_t n) { return (((size_t)n + kPackBodies - 1) / kPackBodies) * kPackCap; }

void launch_avbd_init(const SolveView& v, const AvbdParams& p, hipStream_t s, const LaunchEvents& ev) {
    if (p.n == 0) return;
    launch_k(k_avbd_init, dim3(blocks_for(p.n)), dim3(kBlock), s, ev, v, p);
}
namespace {
template <bool FUSED, bool HARD>
void launch_primal_t(const SolveView& v, const AvbdParams& p, uint32_t iter, hipStream_t s, int lanes,
                     const LaunchEvents& ev, uint32_t work) {
    switch (lanes) {
        case 4:
            launch_k(k_avbd_primal_grouped<4, FUSED, HARD>, dim3((work + kBlock / 4 - 1) / (kBlock / 4)), dim3(kBlock), s, ev,
                     v, p, iter);
            break;
        case 8:
            launch_k(k_avbd_primal_grouped<8, FUSED, HARD>, dim3((work + kBlock / 8 - 1) / (kBlock / 8)), dim3(kBlock), s, ev,
                     v, p, iter);
            break;
        case 16:
            launch_k(k_avbd_primal_grouped<16, FUSED, HARD>, dim3((work + kBlock / 16 - 1) / (kBlock / 16)), dim3(kBlock), s,
                     ev, v, p, iter);
            break;
        case 32:
            launch_k(k_avbd_primal_grouped<32, FUSED, HARD>, dim3((work + kBlock / 32 - 1) / (kBlock / 32)), dim3(kBlock), s,
                     ev, v, p, iter);
            break;
        case 64:
            launch_k(k_avbd_primal_grouped<64, FUSED, HARD>, dim3((work + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), s,
                     ev, v, p, iter);
            break;
        default:
            launch_k(k_avbd_primal<FUSED>, dim3(blocks_for(work)), dim3(kBlock), s, ev, v, p, iter);
    }
}
template <bool FUSED>
void launch_primal(const SolveView& v, const AvbdParams& p, uint32_t iter, hipStream_t s, int lanes,
                   const LaunchEvents& ev, uint32_t work) {
    if (v.j_hard) launch_primal_t<FUSED, true>(v, p, iter, s, lanes, ev, work);
    else launch_primal_t<FUSED, false>(v, p, iter, s, lanes, ev, work);
}
}  // namespace

void launch_avbd_primal(const SolveView& v, const AvbdParams& p, uint32_t iter, bool fused, hipStream_t s, int lanes,
                        const LaunchEvents& ev) {
    if (p.n == 0) return;
    if (v.span_primal && v.j_count == 0) {
        const dim3 grid((p.n + kSpanBodies - 1) / kSpanBodies);
        if (fused && iter > 0) launch_k(k_avbd_primal_span<true>, grid, dim3(kSpanBodies), s, ev, v, p, iter);
        else launch_k(k_avbd_primal_span<false>, grid, dim3(kSpanBodies), s, ev, v, p, iter);
        return;
    }
    if (v.work_tab && v.item_tab) {
        const dim3 grid((p.n + kPackBodies - 1) / kPackBodies);
        const bool f = fused && iter > 0;
        if (v.j_hard) {
            if (f) launch_k(k_avbd_primal_packed<true, true>, grid, dim3(kPackThreads), s, ev, v, p, iter);
            else launch_k(k_avbd_primal_packed<false, true>, grid, dim3(kPackThreads), s, ev, v, p, iter);
        } else {
            if (f) launch_k(k_avbd_primal_packed<true, false>, grid, dim3(kPackThreads), s, ev, v, p, iter);
            else launch_k(k_avbd_primal_packed<false, false>, grid, dim3(kPackThreads), s, ev, v, p, iter);
        }
        return;
    }
    if (fused && iter > 0) launch_primal<true>(v, p, iter, s, lanes, ev, p.n);
    else launch_primal<false>(v, p, iter, s, lanes, ev, p.n);
}
void launch_avbd_primal_colour(const SolveView& v, const AvbdParams& p, uint32_t iter, uint32_t count, hipStream_t s,
                               int lanes, const LaunchEvents& ev) {
    if (count == 0) return;
    launch_primal<false>(v, p, iter, s, lanes, ev, count);
}
void launch_avbd_primal_colour_range(const SolveView& v, const AvbdParams& p, uint32_t iter, hipStream_t s,
                                     int lanes) {
    if (p.n == 0) return;
    launch_primal<false>(v, p, iter, s, lanes, LaunchEvents{}, p.n);
}
void launch_avbd_colour_tail(const SolveView& v, const AvbdParams& p, uint32_t iter, const uint32_t* starts,
                             uint32_t first, uint32_t ncol_at, hipStream_t s) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_avbd_colour_tail, dim3(1), dim3(kBlock), 0, s, v, p, iter, starts, first, starts + ncol_at);
}
static uint32_t contact_blocks(uint32_t work) {
    uint32_t blocks = blocks_for(work == 0 ? 1 : work);
    return blocks > 4096 ? 4096 : blocks;
}
void launch_avbd_dual(const SolveView& v, const AvbdParams& p, uint32_t parity, uint32_t src, uint32_t dst,
                      uint32_t m_grid, hipStream_t s, const LaunchEvents& ev) {
    launch_k(k_avbd_dual, dim3(contact_blocks(m_grid)), dim3(kBlock), s, ev, v, p, parity, src, dst);
}
void launch_avbd_recover(const SolveView& v, const AvbdParams& p, uint32_t parity, hipStream_t s) {
    if (p.n == 0) return;
    hipLaunchKernelGGL(k_avbd_recover, dim3(blocks_for(p.n)), dim3(kBlock), 0, s, v, p, parity);
}
void launch_avbd_scenes(const SolveView& v, const AvbdParams& p, uint32_t iters, bool fused, const SceneQueue& q,
                        hipStream_t s, const LaunchEvents& ev) {
    if (p.n == 0 || q.count == 0 || iters == 0) return;
    // (one-wave workgroups for small bundles measured slower still: the
    // largest scenes then take 8 bodies per lane through every phase)
    const dim3 grid(std::min<uint32_t>(q.count, kSceneWorkgroups));
    if (fused) launch_k(k_avbd_scenes<true, kBlock>, grid, dim3(kBlock), s, ev, v, p, q, iters);
    else launch_k(k_avbd_scenes<false, kBlock>, grid, dim3(kBlock), s, ev, v, p, q, iters);
}

void launch_avbd_dual_recover(const SolveView& v, const AvbdParams& p, uint32_t iters, bool fused, uint32_t m_grid,
                              hipStream_t s, const LaunchEvents& ev) {
    launch_k(k_avbd_dual_recover, dim3(contact_blocks(m_grid > p.n ? m_grid : p.n)), dim3(kBlock), s, ev, v, p, iters,
             fused ? 1u : 0u);
}

}  // namespace dio
