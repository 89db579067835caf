// adjoint_math.hpp — vector-Jacobian products of the AVBD step's primitives.
//
// Each *_vjp takes the forward inputs and the adjoint of the output and
// returns (or accumulates) the adjoints of the inputs.  Piecewise functions
// (min, clamps, the friction cone, |x|) use the piece the forward value lies
// on, with the same tie rules as the forward code, so the results are the
// derivatives of the branch the forward took.
#pragma once

#include "dio_math.hpp"

namespace dio {

DIO_HD f3 f3_zero() { return f3{0.0f, 0.0f, 0.0f}; }
DIO_HD q4 q4_zero() { return q4{0.0f, 0.0f, 0.0f, 0.0f}; }
DIO_HD q4 qadd4(q4 a, q4 b) { return q4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
DIO_HD q4 qscale4(q4 a, float s) { return q4{a.x * s, a.y * s, a.z * s, a.w * s}; }
DIO_HD float qdot4(q4 a, q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// qrot(q, v) = v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w), b = q.xyz
DIO_HD void qrot_vjp(q4 q, f3 v, f3 go, q4& gq, f3& gv) {
    const f3 b{q.x, q.y, q.z};
    const float w = q.w;
    const float bb = dot3(b, b), vb = dot3(v, b), ob = dot3(go, b), ov = dot3(go, v);
    gv = add(add(scale(go, w * w - bb), scale(b, 2.0f * ob)), scale(cross3(go, b), 2.0f * w));
    const f3 gb = add(add(scale(b, -2.0f * ov), scale(go, 2.0f * vb)), add(scale(v, 2.0f * ob), scale(cross3(v, go), 2.0f * w)));
    const float gw = 2.0f * w * ov + 2.0f * dot3(go, cross3(b, v));
    gq = q4{gb.x, gb.y, gb.z, gw};
}

// Hamilton product: (a.w b.v + b.w a.v + a.v x b.v, a.w b.w - a.v . b.v)
DIO_HD void qmul_vjp(q4 a, q4 b, q4 go, q4& ga, q4& gb) {
    const f3 av{a.x, a.y, a.z}, bv{b.x, b.y, b.z}, ov{go.x, go.y, go.z};
    const float os = go.w;
    const f3 gav = sub(add(scale(ov, b.w), cross3(bv, ov)), scale(bv, os));
    const f3 gbv = sub(add(scale(ov, a.w), cross3(ov, av)), scale(av, os));
    ga = q4{gav.x, gav.y, gav.z, dot3(ov, bv) + os * b.w};
    gb = q4{gbv.x, gbv.y, gbv.z, dot3(ov, av) + os * a.w};
}

// u = s / |s|
DIO_HD q4 qnormalize_vjp(q4 s, q4 gu) {
    const float len = sqrtf(qdot4(s, s));
    const q4 u = qscale4(s, 1.0f / len);
    const float ug = qdot4(u, gu);
    return qscale4(q4{gu.x - u.x * ug, gu.y - u.y * ug, gu.z - u.z * ug, gu.w - u.w * ug}, 1.0f / len);
}

// quat_step(q, d) = normalize(q + 0.5 (d, 0) q)
DIO_HD void quat_step_vjp(q4 q, f3 d, q4 gu, q4& gq, f3& gd) {
    const q4 p{d.x, d.y, d.z, 0.0f};
    const q4 pq = qmul(p, q);
    const q4 s{q.x + pq.x * 0.5f, q.y + pq.y * 0.5f, q.z + pq.z * 0.5f, q.w + pq.w * 0.5f};
    const q4 gs = qnormalize_vjp(s, gu);
    q4 gp, gq2;
    qmul_vjp(p, q, qscale4(gs, 0.5f), gp, gq2);
    gq = qadd4(gs, gq2);
    gd = f3{gp.x, gp.y, gp.z};
}

// ang_diff(to, from) = 2 (to conj(from)).xyz
DIO_HD void ang_diff_vjp(q4 to, q4 from, f3 go, q4& gto, q4& gfrom) {
    const q4 grel{2.0f * go.x, 2.0f * go.y, 2.0f * go.z, 0.0f};
    q4 gc;
    qmul_vjp(to, qconj(from), grel, gto, gc);
    gfrom = q4{-gc.x, -gc.y, -gc.z, gc.w};
}

// d|x|/dx with |0| taken on the +x piece
DIO_HD float abs_slope(float x) { return x < 0.0f ? -1.0f : 1.0f; }

// contact_force (avbd/src/lib.rs: the clamped normal force and the friction
// cone): adjoints of penalty, lambda, the constraint values and mu from the
// adjoint of (f0, f1, f2), accumulated.
DIO_HD void contact_force_vjp(float4 pen, float4 lam, float mu, float c0, float c1, float c2, float gf0, float gf1,
                              float gf2, float gk[3], float gl[3], float gc[3], float& gmu) {
    const float t0 = pen.x * c0 + lam.x;
    const float f0 = rmin(t0, 0.0f);
    const float f1r = pen.y * c1 + lam.y;
    const float f2r = pen.z * c2 + lam.z;
    const float bound = fabsf(f0) * mu;
    const float tang = sqrtf(f1r * f1r + f2r * f2r);
    float g1r = gf1, g2r = gf2, gbound = 0.0f;
    if (tang > bound && tang > 0.0f) {
        const float s = bound / tang;
        const float u1 = f1r / tang, u2 = f2r / tang;
        const float ug = u1 * gf1 + u2 * gf2;
        g1r = s * (gf1 - u1 * ug);
        g2r = s * (gf2 - u2 * ug);
        gbound = ug;
    }
    gmu += gbound * fabsf(f0);
    const float gf0t = gf0 + gbound * mu * abs_slope(f0);
    const float gt0 = (0.0f < t0 || t0 != t0) ? 0.0f : gf0t;  // rmin(t0, 0) picked the constant
    gk[0] += gt0 * c0; gc[0] += gt0 * pen.x; gl[0] += gt0;
    gk[1] += g1r * c1; gc[1] += g1r * pen.y; gl[1] += g1r;
    gk[2] += g2r * c2; gc[2] += g2r * pen.z; gl[2] += g2r;
}

// The dual's penalty ramp: k' = rmin(k + beta |c|, kmax) on active rows.
DIO_HD void penalty_ramp_vjp(float k, float c, bool active, float beta, float kmax, float gk_out, float& gk,
                             float& gc) {
    if (!active) {
        gk += gk_out;
        return;
    }
    const float v = k + beta * fabsf(c);
    if (kmax < v || v != v) return;  // clamped at kmax
    gk += gk_out;
    gc += gk_out * beta * abs_slope(c);
}

}  // namespace dio
