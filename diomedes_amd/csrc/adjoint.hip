// adjoint.hip — recording, frozen-contact replay and the hand-written reverse
// mode of the AVBD solve (avbd/src/lib.rs:202-699, scheduled as in
// gpu_physics.rs:541-550) for gfx950.
//
// Forward of one step (per dynamic body; contacts and CSR frozen):
//   init      x0 = x, q0 = q, x~ = x + v dt + g dt^2, q~ = quat_step(q, w dt),
//             Y0 = x + v dt + g wgt dt^2 (adaptive warm start), R0 = q~
//   primal k  H = diag(M, I) / dt^2 + sum_rows pen J J^T,
//             g = (M / dt^2)(y - x~), (I / dt^2) ang_diff(r, q~) + sum_rows J f,
//             D = -H^-1 g, Y' = y + D_lin, R' = quat_step(r, D_ang)
//   dual k    lam', pen' = dual_update(lam, pen, c(Y_{k+1}))
//   recover   v' = (Y_I - x0) / dt, w' = 2 (R_I q0^-1).xyz / dt, vp' = v
// Reverse of a primal with output adjoint Db: z = H^-1 Db (H is symmetric),
// g_bar = -z, and each row (J, pen, f) gets J_bar += g_bar f - pen (z (J.D) +
// D (J.z)), pen_bar -= (J.z)(J.D), f_bar = J.g_bar, then through the contact
// force, the constraint value and the rotations by the VJPs of
// adjoint_math.hpp.  A row's constraint value also depends on the other
// body's pose: its adjoint is stored per (contact, side) and gathered by the
// other body in a second launch, so every sum has a fixed order and no
// atomics are used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "adjoint.hpp"
#include "adjoint_math.hpp"
#include "prim.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }
inline uint32_t grid_for(uint32_t work) {
    const uint32_t b = blocks_for(work == 0 ? 1 : work);
    return b > 4096 ? 4096 : b;
}

// ------------------------------------------------------------------ record
__global__ void k_record_head(AdjRecord r, const uint32_t* m_dev, const uint32_t* off, uint32_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t* c = r.counters;
    const uint32_t s = c[kRecSteps];
    if (s >= r.max_steps) {
        c[kRecOverflow] |= 2u;
        c[kRecValid] = 0u;
        return;
    }
    uint32_t m = *m_dev;
    uint32_t k = off[n];
    const uint32_t c0 = c[kRecContacts], i0 = c[kRecIdx];
    if ((uint64_t)c0 + m > r.rec_cap || (uint64_t)i0 + k > r.idx_cap) {
        c[kRecOverflow] |= 1u;
        m = 0;
        k = 0;
    }
    r.step_coff[s] = c0;
    r.step_ioff[s] = i0;
    r.step_m[s] = m;
    r.step_coff[s + 1] = c0 + m;
    r.step_ioff[s + 1] = i0 + k;
    c[kRecContacts] = c0 + m;
    c[kRecIdx] = i0 + k;
    c[kRecSteps] = s + 1;
    c[kRecValid] = 1u;
}

__global__ void __launch_bounds__(kBlock) k_record_copy(AdjRecord r, ContactArrays src, const uint32_t* off,
                                                        const uint32_t* idx, const uint32_t* peer, BodyState st,
                                                        uint32_t n) {
    if (r.counters[kRecValid] == 0u) return;
    const uint32_t s = r.counters[kRecSteps] - 1;
    const uint32_t cb = r.step_coff[s], m = r.step_m[s];
    const uint32_t ib = r.step_ioff[s], k = r.step_ioff[s + 1] - ib;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t t0 = blockIdx.x * kBlock + threadIdx.x;
    for (uint32_t i = t0; i < m; i += stride) {
        const uint32_t d = cb + i;
        r.c.ab[d] = src.ab[i];
        r.c.n[d] = src.n[i];
        r.c.t1[d] = src.t1[i];
        r.c.t2[d] = src.t2[i];
        r.c.ra[d] = src.ra[i];
        r.c.rb[d] = src.rb[i];
        r.c.c0[d] = src.c0[i];
        r.c.pen[d] = src.pen[i];
        r.c.lam[d] = src.lam[i];
    }
    for (uint32_t j = t0; j < k; j += stride) {
        r.csr_idx[ib + j] = idx[j];
        r.csr_peer[ib + j] = peer[j];
    }
    uint32_t* ro = r.csr_off + (size_t)s * (r.n_cap + 1);
    for (uint32_t j = t0; j <= n; j += stride) ro[j] = off[j];
    if (s % r.every == 0) {
        float4* ck = r.ck + (size_t)(s / r.every) * 5 * r.n_cap;
        for (uint32_t i = t0; i < n; i += stride) {
            ck[i] = st.pos[i];
            ck[r.n_cap + i] = st.rot[i];
            ck[2 * (size_t)r.n_cap + i] = st.vel[i];
            ck[3 * (size_t)r.n_cap + i] = st.ang[i];
            ck[4 * (size_t)r.n_cap + i] = st.prev[i];
        }
    }
}

// ------------------------------------------------------------------ replay
__global__ void __launch_bounds__(kBlock) k_copy4(const float4* src, float4* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock) k_set_friction(float4* nrm, uint32_t m, float mu) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) nrm[i].w = mu;
}

void copy4(const float4* src, float4* dst, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy4, dim3(blocks_for(n)), dim3(kBlock), 0, s, src, dst, n);
}

// tape indexing
inline float4* tstate(const AdjWork& w, uint32_t j, int f) { return w.tape_state + ((size_t)j * 5 + f) * w.n_cap; }
inline float4* tiner(const AdjWork& w, uint32_t j, int f) { return w.tape_iner + ((size_t)j * 2 + f) * w.n_cap; }
inline float4* tpose(const AdjWork& w, uint32_t j, uint32_t k, int f) {
    return w.tape_pose + (((size_t)j * (w.iters + 1) + k) * 2 + f) * w.n_cap;
}
inline float4* tmult(const AdjWork& w, uint32_t j, uint32_t k, int f) {
    return w.tape_mult + (((size_t)j * w.iters + k) * 2 + f) * w.m_max;
}

// The solve view of recorded step s at segment slot j (contacts, CSR and
// tapes bound; the multiplier / pose buffers are set per launch).
SolveView step_view(const AdjRecord& r, const AdjHostIndex& h, const AdjWork& w, const AdjProblem& pb, uint32_t s,
                    uint32_t j) {
    SolveView v{};
    v.cur_pos = w.st.pos;
    v.cur_rot = w.st.rot;
    v.vel = w.st.vel;
    v.ang = w.st.ang;
    v.prev_vel = w.st.prev;
    v.inv_mm = pb.inv_mm;
    v.init_lin = tstate(w, j, 0);
    v.init_ang = tstate(w, j, 1);
    v.iner_lin = tiner(w, j, 0);
    v.iner_ang = tiner(w, j, 1);
    const uint32_t c0 = h.coff[s];
    v.c_ab = r.c.ab + c0;
    v.c_n = r.c.n + c0;
    v.c_t1 = r.c.t1 + c0;
    v.c_t2 = r.c.t2 + c0;
    v.c_ra = r.c.ra + c0;
    v.c_rb = r.c.rb + c0;
    v.c_c0 = r.c.c0 + c0;
    v.derived_frame = 1;  // recorded from the world's contact build (ra.w holds c0)
    v.m_dev = nullptr;
    v.m_host = h.m[s];
    v.offsets = r.csr_off + (size_t)s * (r.n_cap + 1);
    v.indices = r.csr_idx + h.ioff[s];
    v.peers = r.csr_peer + h.ioff[s];
    v.j_off = nullptr;
    v.out_pos = w.st.pos;
    v.out_rot = w.st.rot;
    return v;
}

// ------------------------------------------------------------------ reverse
// One recorded step's data for the reverse kernels.
struct AdjStepView {
    uint32_t n, m, n_cap, m_max, iters;
    uint32_t frozen;  // the dual updates' outputs are constants: no dual reverse, no DPOSE
    AvbdParams p;
    float mu;
    const float4* inv_mm;
    const uint2* ab;
    const float4 *cn, *ct1, *ct2, *cra, *crb, *cc0;
    const uint32_t *off, *idx, *peer;
    const float4 *x0, *q0, *vel, *ang, *prev, *xt, *qt;  // step tapes
    const float4* pose;                                 // (I + 1) x 2 x n_cap
    const float4* mult;                                 // I x 2 x m_max
    float4* sbar;
    float4* ybar;
    float4* acc;
    float4* mbar;
    float* mubar_body;
    double* mubar_contact;
    float4* slots;
    float4* dpart;
    float4* dpose;
    float4* opose;
};

__device__ __forceinline__ float4 Yk(const AdjStepView& v, uint32_t k, uint32_t i) {
    return v.pose[((size_t)k * 2) * v.n_cap + i];
}
__device__ __forceinline__ float4 Rk(const AdjStepView& v, uint32_t k, uint32_t i) {
    return v.pose[((size_t)k * 2 + 1) * v.n_cap + i];
}
__device__ __forceinline__ float4 LAMk(const AdjStepView& v, uint32_t k, uint32_t c) {
    return v.mult[((size_t)k * 2) * v.m_max + c];
}
__device__ __forceinline__ float4 PENk(const AdjStepView& v, uint32_t k, uint32_t c) {
    return v.mult[((size_t)k * 2 + 1) * v.m_max + c];
}
__device__ __forceinline__ float4& YB(const AdjStepView& v, uint32_t buf, int f, uint32_t i) {
    return v.ybar[((size_t)buf * 2 + f) * v.n_cap + i];
}
__device__ __forceinline__ float4& ACC(const AdjStepView& v, int f, uint32_t i) {
    return v.acc[(size_t)f * v.n_cap + i];
}
constexpr int kSlotPen = 0, kSlotLam = 1;
__device__ __forceinline__ float4& SLOT(const AdjStepView& v, uint32_t slot, int f) { return v.slots[(size_t)slot * 2 + f]; }
__device__ __forceinline__ float4& DPOSE(const AdjStepView& v, uint32_t slot, int f) {
    return v.dpose[(size_t)slot * 4 + f];
}
__device__ __forceinline__ float4& OPOSE(const AdjStepView& v, uint32_t slot, int f) {
    return v.opose[(size_t)slot * 4 + f];
}
__device__ __forceinline__ float4& DPART(const AdjStepView& v, int f, uint32_t c) {
    return v.dpart[(size_t)f * v.m_max + c];
}

__device__ __forceinline__ f3 add3(f3 a, float4 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ q4 add4(q4 a, float4 b) { return q4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ float4 q4f(q4 q) { return make_float4(q.x, q.y, q.z, q.w); }
__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// step begin: recover reverse per body, clear the step's per-contact sums
__global__ void __launch_bounds__(kBlock) k_adj_recover_rev(AdjStepView v) {
    const uint32_t t0 = blockIdx.x * kBlock + threadIdx.x, stride = gridDim.x * kBlock;
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t c = t0; c < v.m; c += stride) {
        DPART(v, 0, c) = z4;
        DPART(v, 1, c) = z4;
    }
    for (uint32_t i = t0; i < v.n; i += stride) {
        const float4 gx = v.sbar[i], gq = v.sbar[v.n_cap + i], gv = v.sbar[2 * (size_t)v.n_cap + i];
        const float4 gw = v.sbar[3 * (size_t)v.n_cap + i], gvp = v.sbar[4 * (size_t)v.n_cap + i];
        float4 yb = gx, rb = gq, x0b = z4, q0b = z4;
        float4 sv = z4, sw = z4, svp = z4;
        if (v.inv_mm[i].w != 0.0f) {
            const float dt = v.p.dt;
            yb = make_float4(gx.x + gv.x / dt, gx.y + gv.y / dt, gx.z + gv.z / dt, 0.0f);
            x0b = make_float4(-gv.x / dt, -gv.y / dt, -gv.z / dt, 0.0f);
            const q4 R = as_q(Rk(v, v.iters, i));
            const q4 q0 = as_q(v.q0[i]);
            const q4 grel{2.0f * gw.x / dt, 2.0f * gw.y / dt, 2.0f * gw.z / dt, 0.0f};
            q4 gR, gc;
            qmul_vjp(R, qconj(q0), grel, gR, gc);
            rb = make_float4(gq.x + gR.x, gq.y + gR.y, gq.z + gR.z, gq.w + gR.w);
            q0b = make_float4(-gc.x, -gc.y, -gc.z, gc.w);
            sv = gvp;  // vp' = v
        } else {
            sv = gv;
            sw = gw;
            svp = gvp;
        }
        YB(v, 0, 0, i) = yb;
        YB(v, 0, 1, i) = rb;
        ACC(v, 0, i) = x0b;
        ACC(v, 1, i) = q0b;
        ACC(v, 2, i) = z4;
        ACC(v, 3, i) = z4;
        v.sbar[i] = z4;
        v.sbar[v.n_cap + i] = z4;
        v.sbar[2 * (size_t)v.n_cap + i] = sv;
        v.sbar[3 * (size_t)v.n_cap + i] = sw;
        v.sbar[4 * (size_t)v.n_cap + i] = svp;
    }
}

// Reverse of dual update k (lam^(k+1), pen^(k+1) from lam^(k), pen^(k) and the
// poses Y_(k+1)), for k < I - 1.
__global__ void __launch_bounds__(kBlock) k_adj_dual_rev(AdjStepView v, uint32_t k) {
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < v.m; c += gridDim.x * kBlock) {
        const uint2 ab = v.ab[c];
        const bool dyn_a = v.inv_mm[ab.x].w != 0.0f;
        const bool has_b = ab.y != kContainer;
        const bool dyn_b = has_b && v.inv_mm[ab.y].w != 0.0f;
        // adjoint of lam^(k+1), pen^(k+1): the next dual's part and both primal sides
        float4 gl = DPART(v, 0, c), gk = DPART(v, 1, c);
        if (dyn_a) {
            const float4 a1 = SLOT(v, 2 * c, kSlotLam), a2 = SLOT(v, 2 * c, kSlotPen);
            gl = make_float4(gl.x + a1.x, gl.y + a1.y, gl.z + a1.z, 0.0f);
            gk = make_float4(gk.x + a2.x, gk.y + a2.y, gk.z + a2.z, 0.0f);
        }
        if (dyn_b) {
            const float4 b1 = SLOT(v, 2 * c + 1, kSlotLam), b2 = SLOT(v, 2 * c + 1, kSlotPen);
            gl = make_float4(gl.x + b1.x, gl.y + b1.y, gl.z + b1.z, 0.0f);
            gk = make_float4(gk.x + b2.x, gk.y + b2.y, gk.z + b2.z, 0.0f);
        }
        // forward of dual k (the dual kernel's arithmetic)
        const float kk = 1.0f - v.p.alpha;
        const f3 zero = f3_zero();
        const q4 rot_a = as_q(Rk(v, k + 1, ab.x));
        const q4 q0a = as_q(v.q0[ab.x]);
        const f3 dal = sub(xyz(Yk(v, k + 1, ab.x)), xyz(v.x0[ab.x]));
        const f3 daa = ang_diff(rot_a, q0a);
        f3 dbl = zero, dba = zero, rbw = zero;
        q4 rot_b{0.0f, 0.0f, 0.0f, 1.0f}, q0b{0.0f, 0.0f, 0.0f, 1.0f};
        if (has_b) {
            rot_b = as_q(Rk(v, k + 1, ab.y));
            q0b = as_q(v.q0[ab.y]);
            dbl = sub(xyz(Yk(v, k + 1, ab.y)), xyz(v.x0[ab.y]));
            dba = ang_diff(rot_b, q0b);
            rbw = qrot(rot_b, xyz(v.crb[c]));
        }
        const f3 ra_l = xyz(v.cra[c]);
        const f3 raw = qrot(rot_a, ra_l);
        const float4 n4 = v.cn[c];
        const f3 r[3] = {xyz(n4), xyz(v.ct1[c]), xyz(v.ct2[c])};
        const float4 c04 = v.cc0[c];
        const float c0s[3] = {c04.x, c04.y, c04.z};
        float cv[3];
        for (int q = 0; q < 3; ++q) {
            float x = c0s[q] * kk;
            cv[q] = x + (dot3(r[q], dal) + dot3(cross3(raw, r[q]), daa) - dot3(r[q], dbl) - dot3(cross3(rbw, r[q]), dba));
        }
        const float4 pen = PENk(v, k, c), lam = LAMk(v, k, c);
        const float mu = n4.w;
        // forward force and the ramp conditions
        const float t0 = pen.x * cv[0] + lam.x;
        const float f0 = rmin(t0, 0.0f);
        const float f1r = pen.y * cv[1] + lam.y, f2r = pen.z * cv[2] + lam.z;
        const float bound = fabsf(f0) * mu;
        const float tang = sqrtf(f1r * f1r + f2r * f2r);
        float gk_in[3] = {0.0f, 0.0f, 0.0f}, gl_in[3] = {0.0f, 0.0f, 0.0f}, gc[3] = {0.0f, 0.0f, 0.0f};
        float gmu = 0.0f;
        penalty_ramp_vjp(pen.x, cv[0], f0 < 0.0f, v.p.beta_lin, v.p.penalty_max, gk.x, gk_in[0], gc[0]);
        penalty_ramp_vjp(pen.y, cv[1], tang <= bound, v.p.beta_lin, v.p.penalty_max, gk.y, gk_in[1], gc[1]);
        penalty_ramp_vjp(pen.z, cv[2], tang <= bound, v.p.beta_lin, v.p.penalty_max, gk.z, gk_in[2], gc[2]);
        contact_force_vjp(pen, lam, mu, cv[0], cv[1], cv[2], gl.x, gl.y, gl.z, gk_in, gl_in, gc, gmu);
        DPART(v, 0, c) = make_float4(gl_in[0], gl_in[1], gl_in[2], 0.0f);
        DPART(v, 1, c) = make_float4(gk_in[0], gk_in[1], gk_in[2], 0.0f);
        v.mubar_contact[c] += (double)gmu;
        // constraint values -> poses of both bodies at Y_(k+1)
        f3 gdal = zero, gdaa = zero, graw = zero, gdbl = zero, gdba = zero, grbw = zero;
        for (int q = 0; q < 3; ++q) {
            gdal = add(gdal, scale(r[q], gc[q]));
            gdaa = add(gdaa, scale(cross3(raw, r[q]), gc[q]));
            graw = add(graw, cross3(r[q], scale(daa, gc[q])));
            gdbl = sub(gdbl, scale(r[q], gc[q]));
            gdba = sub(gdba, scale(cross3(rbw, r[q]), gc[q]));
            grbw = add(grbw, cross3(r[q], scale(dba, -gc[q])));
        }
        {
            q4 gto, gfrom, gq;
            f3 gv;
            ang_diff_vjp(rot_a, q0a, gdaa, gto, gfrom);
            qrot_vjp(rot_a, ra_l, graw, gq, gv);
            DPOSE(v, 2 * c, 0) = f4(gdal, 0.0f);
            DPOSE(v, 2 * c, 1) = q4f(qadd4(gto, gq));
            DPOSE(v, 2 * c, 2) = f4(neg(gdal), 0.0f);
            DPOSE(v, 2 * c, 3) = q4f(gfrom);
        }
        if (has_b) {
            q4 gto, gfrom, gq;
            f3 gv;
            ang_diff_vjp(rot_b, q0b, gdba, gto, gfrom);
            qrot_vjp(rot_b, xyz(v.crb[c]), grbw, gq, gv);
            DPOSE(v, 2 * c + 1, 0) = f4(gdbl, 0.0f);
            DPOSE(v, 2 * c + 1, 1) = q4f(qadd4(gto, gq));
            DPOSE(v, 2 * c + 1, 2) = f4(neg(gdbl), 0.0f);
            DPOSE(v, 2 * c + 1, 3) = q4f(gfrom);
        } else {
            for (int f = 0; f < 4; ++f) DPOSE(v, 2 * c + 1, f) = z4;
        }
    }
}

// The 6x6 block system (the forward's Block6 layout and add order).
struct Blk {
    float l[6], a[6], c[9], rl[3], ra[3];
};
__device__ __forceinline__ void blk_row(Blk& B, f3 jl, f3 ja, float k, float f) {
    B.l[0] = B.l[0] + (jl.x * jl.x) * k; B.l[1] = B.l[1] + (jl.y * jl.x) * k; B.l[2] = B.l[2] + (jl.z * jl.x) * k;
    B.l[3] = B.l[3] + (jl.y * jl.y) * k; B.l[4] = B.l[4] + (jl.z * jl.y) * k; B.l[5] = B.l[5] + (jl.z * jl.z) * k;
    B.a[0] = B.a[0] + (ja.x * ja.x) * k; B.a[1] = B.a[1] + (ja.y * ja.x) * k; B.a[2] = B.a[2] + (ja.z * ja.x) * k;
    B.a[3] = B.a[3] + (ja.y * ja.y) * k; B.a[4] = B.a[4] + (ja.z * ja.y) * k; B.a[5] = B.a[5] + (ja.z * ja.z) * k;
    B.c[0] = B.c[0] + (ja.x * jl.x) * k; B.c[1] = B.c[1] + (ja.x * jl.y) * k; B.c[2] = B.c[2] + (ja.x * jl.z) * k;
    B.c[3] = B.c[3] + (ja.y * jl.x) * k; B.c[4] = B.c[4] + (ja.y * jl.y) * k; B.c[5] = B.c[5] + (ja.y * jl.z) * k;
    B.c[6] = B.c[6] + (ja.z * jl.x) * k; B.c[7] = B.c[7] + (ja.z * jl.y) * k; B.c[8] = B.c[8] + (ja.z * jl.z) * k;
    (void)f;
}

// solve6 of avbd.hip restated (LDL^T of the lower triangle; H x = -rhs)
__device__ __forceinline__ void solve_blk(const Blk& B, float x[6]) {
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

// A contact entry of body i's primal at iteration k: the forward rows.
struct EntryRows {
    f3 r[3], j[3];       // lin / ang Jacobian rows (signed for this body)
    float cv[3], f[3];   // constraint values, forces
    float4 pen, lam;
    f3 rl_self, rl_other, r_world, r_other, dol, doa;
    q4 rot_other, q0_other;
    uint32_t ci, other;
    bool is_a;
};

__device__ __forceinline__ void entry_rows(const AdjStepView& v, uint32_t k, uint32_t i, uint32_t e, f3 dsl, f3 dsa,
                                           q4 snap_rot, EntryRows& E) {
    const f3 zero = f3_zero();
    const uint32_t ci = v.idx[e];
    const uint32_t pc = v.peer[e];
    E.ci = ci;
    E.is_a = pc == kContainer || (pc & kPeerIsB) == 0u;
    E.other = pc == kContainer ? kContainer : (pc & ~kPeerIsB);
    const float4 n4 = v.cn[ci];
    const f3 n0 = xyz(n4), t1 = xyz(v.ct1[ci]), t2 = xyz(v.ct2[ci]);
    const f3 ra = xyz(v.cra[ci]), rb = xyz(v.crb[ci]);
    E.r[0] = E.is_a ? n0 : neg(n0);
    E.r[1] = E.is_a ? t1 : neg(t1);
    E.r[2] = E.is_a ? t2 : neg(t2);
    E.rl_self = E.is_a ? ra : rb;
    E.rl_other = E.is_a ? rb : ra;
    E.r_world = qrot(snap_rot, E.rl_self);
    E.dol = zero;
    E.doa = zero;
    E.r_other = zero;
    E.rot_other = q4{0.0f, 0.0f, 0.0f, 1.0f};
    E.q0_other = E.rot_other;
    if (E.other != kContainer) {
        E.rot_other = as_q(Rk(v, k, E.other));
        E.q0_other = as_q(v.q0[E.other]);
        E.dol = sub(xyz(Yk(v, k, E.other)), xyz(v.x0[E.other]));
        E.doa = ang_diff(E.rot_other, E.q0_other);
        E.r_other = qrot(E.rot_other, E.rl_other);
    }
    for (int q = 0; q < 3; ++q) E.j[q] = cross3(E.r_world, E.r[q]);
    const float kk = 1.0f - v.p.alpha;
    const float4 c04 = v.cc0[ci];
    const float c0s[3] = {c04.x, c04.y, c04.z};
    for (int q = 0; q < 3; ++q) {
        float x = c0s[q] * kk;
        E.cv[q] = x + (dot3(E.r[q], dsl) + dot3(E.j[q], dsa) - dot3(E.r[q], E.dol) -
                       dot3(cross3(E.r_other, E.r[q]), E.doa));
    }
    E.pen = PENk(v, k, ci);
    E.lam = LAMk(v, k, ci);
    const float mu = n4.w;
    const float t0 = E.pen.x * E.cv[0] + E.lam.x;
    const float f0 = rmin(t0, 0.0f);
    const float f1r = E.pen.y * E.cv[1] + E.lam.y, f2r = E.pen.z * E.cv[2] + E.lam.z;
    const float bound = fabsf(f0) * mu;
    const float tang = sqrtf(f1r * f1r + f2r * f2r);
    float f1 = f1r, f2 = f2r;
    if (tang > bound && tang > 0.0f) {
        const float s = bound / tang;
        f1 = f1r * s;
        f2 = f2r * s;
    }
    E.f[0] = f0;
    E.f[1] = f1;
    E.f[2] = f2;
}

// Body i's share of the pose adjoints stored per (contact, side): from the
// dual update (DPOSE, i's own side) and from the other body's primal rows
// (OPOSE, the other side's slot, written only when that body is dynamic).
__device__ __forceinline__ void gather_pose(const AdjStepView& v, uint32_t e0, uint32_t e1, uint32_t step, f3& gy,
                                            q4& gr, f3& gx0, q4& gq0) {
    for (uint32_t e = e0; e < e1; e += step) {
        const uint32_t ci = v.idx[e];
        const uint32_t pc = v.peer[e];
        const bool is_b = pc != kContainer && (pc & kPeerIsB);
        const uint32_t sl = 2 * ci + (is_b ? 1u : 0u);
        if (!v.frozen) {
            gy = add3(gy, DPOSE(v, sl, 0));
            gr = add4(gr, DPOSE(v, sl, 1));
            gx0 = add3(gx0, DPOSE(v, sl, 2));
            gq0 = add4(gq0, DPOSE(v, sl, 3));
        }
        if (pc == kContainer || v.inv_mm[pc & ~kPeerIsB].w == 0.0f) continue;
        const uint32_t so = 2 * ci + (is_b ? 0u : 1u);
        gy = add3(gy, OPOSE(v, so, 0));
        gr = add4(gr, OPOSE(v, so, 1));
        gx0 = add3(gx0, OPOSE(v, so, 2));
        gq0 = add4(gq0, OPOSE(v, so, 3));
    }
}

// The other bodies' rows of primal 0 read Y_0: their share only.
__device__ __forceinline__ void gather_other_rows(const AdjStepView& v, uint32_t e0, uint32_t e1, f3& gy, q4& gr,
                                                  f3& gx0, q4& gq0) {
    for (uint32_t e = e0; e < e1; ++e) {
        const uint32_t ci = v.idx[e];
        const uint32_t pc = v.peer[e];
        if (pc == kContainer || v.inv_mm[pc & ~kPeerIsB].w == 0.0f) continue;
        const uint32_t so = 2 * ci + ((pc & kPeerIsB) ? 0u : 1u);
        gy = add3(gy, OPOSE(v, so, 0));
        gr = add4(gr, OPOSE(v, so, 1));
        gx0 = add3(gx0, OPOSE(v, so, 2));
        gq0 = add4(gq0, OPOSE(v, so, 3));
    }
}

// Reverse of primal k: the output adjoint (with the contributions of dual k
// and of the other bodies' rows of primal k + 1), the own-pose, initial /
// inertial pose and mass adjoints, per (contact, side) the adjoints of c,
// pen and lam, and the adjoint of the other body's pose through this body's
// rows (OPOSE, gathered by that body in the next launch).
template <int G> __device__ __forceinline__ float gsum(float x) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, G);
    return x;
}
template <int G> __device__ __forceinline__ f3 gsum3(f3 a) { return f3{gsum<G>(a.x), gsum<G>(a.y), gsum<G>(a.z)}; }
template <int G> __device__ __forceinline__ q4 gsum4(q4 a) {
    return q4{gsum<G>(a.x), gsum<G>(a.y), gsum<G>(a.z), gsum<G>(a.w)};
}

// G lanes per body (a group within one wavefront): the body's CSR entries
// are dealt to the lanes, per-lane partial sums are combined by butterfly
// shuffles (every lane ends with the same total), the two 6x6 solves run on
// every lane, and lane 0 stores the body's results.
template <int G>
__global__ void __launch_bounds__(kBlock) k_adj_primal(AdjStepView v, uint32_t k, uint32_t cur, uint32_t nxt,
                                                       uint32_t has_dual) {
    const uint32_t lane = threadIdx.x % G;
    const uint32_t i = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    if (i >= v.n) return;  // uniform over the group
    const uint32_t e0 = v.off[i], e1 = v.off[i + 1];
    f3 gout_y = f3_zero(), gx0 = f3_zero();
    q4 gout_r = q4_zero(), gq0 = q4_zero();
    // primal k + 1 and dual k read Y_(k+1): gather their contributions
    if (has_dual) {
        gather_pose(v, e0 + lane, e1, G, gout_y, gout_r, gx0, gq0);
        gout_y = gsum3<G>(gout_y);
        gout_r = gsum4<G>(gout_r);
        gx0 = gsum3<G>(gx0);
        gq0 = gsum4<G>(gq0);
    }
    gout_y = add3(gout_y, YB(v, cur, 0, i));
    gout_r = add4(gout_r, YB(v, cur, 1, i));
    const float4 imm = v.inv_mm[i];
    if (imm.w == 0.0f) {  // static: poses pass through
        if (lane == 0) {
            YB(v, nxt, 0, i) = f4(gout_y, 0.0f);
            YB(v, nxt, 1, i) = q4f(gout_r);
            float4& a0 = ACC(v, 0, i);
            a0 = make_float4(a0.x + gx0.x, a0.y + gx0.y, a0.z + gx0.z, 0.0f);
            float4& a1 = ACC(v, 1, i);
            a1 = make_float4(a1.x + gq0.x, a1.y + gq0.y, a1.z + gq0.z, a1.w + gq0.w);
        }
        return;
    }
    // ---- forward of the primal (k_avbd_primal's arithmetic; the rows are
    // summed per lane, then over the lanes, then added to the inertia terms)
    const f3 snap_pos = xyz(Yk(v, k, i));
    const q4 snap_rot = as_q(Rk(v, k, i));
    const float dt = v.p.dt;
    const float mass = 1.0f / imm.w;
    const f3 moment{1.0f / imm.x, 1.0f / imm.y, 1.0f / imm.z};
    const float dt2 = dt * dt;
    const float m_dt2 = mass / dt2;
    const f3 i_dt2 = divs(moment, dt2);
    const f3 x0i = xyz(v.x0[i]);
    const q4 q0i = as_q(v.q0[i]);
    const f3 dsl = sub(snap_pos, x0i);
    const f3 dsa = ang_diff(snap_rot, q0i);
    Blk B;
    for (int q = 0; q < 6; ++q) { B.l[q] = 0.0f; B.a[q] = 0.0f; }
    for (int q = 0; q < 9; ++q) B.c[q] = 0.0f;
    for (int q = 0; q < 3; ++q) { B.rl[q] = 0.0f; B.ra[q] = 0.0f; }
    for (uint32_t e = e0 + lane; e < e1; e += G) {
        EntryRows E;
        entry_rows(v, k, i, e, dsl, dsa, snap_rot, E);
        const float kq[3] = {E.pen.x, E.pen.y, E.pen.z};
        for (int q = 0; q < 3; ++q) blk_row(B, E.r[q], E.j[q], kq[q], E.f[q]);
        for (int q = 0; q < 3; ++q) {
            B.rl[0] = B.rl[0] + E.r[q].x * E.f[q]; B.rl[1] = B.rl[1] + E.r[q].y * E.f[q]; B.rl[2] = B.rl[2] + E.r[q].z * E.f[q];
            B.ra[0] = B.ra[0] + E.j[q].x * E.f[q]; B.ra[1] = B.ra[1] + E.j[q].y * E.f[q]; B.ra[2] = B.ra[2] + E.j[q].z * E.f[q];
        }
    }
    for (int q = 0; q < 6; ++q) { B.l[q] = gsum<G>(B.l[q]); B.a[q] = gsum<G>(B.a[q]); }
    for (int q = 0; q < 9; ++q) B.c[q] = gsum<G>(B.c[q]);
    for (int q = 0; q < 3; ++q) { B.rl[q] = gsum<G>(B.rl[q]); B.ra[q] = gsum<G>(B.ra[q]); }
    B.l[0] = m_dt2 + B.l[0]; B.l[3] = m_dt2 + B.l[3]; B.l[5] = m_dt2 + B.l[5];
    B.a[0] = i_dt2.x + B.a[0]; B.a[3] = i_dt2.y + B.a[3]; B.a[5] = i_dt2.z + B.a[5];
    const f3 xt = xyz(v.xt[i]);
    const q4 qt = as_q(v.qt[i]);
    const f3 dl = sub(snap_pos, xt);
    const f3 ad = ang_diff(snap_rot, qt);
    B.rl[0] = m_dt2 * dl.x + B.rl[0]; B.rl[1] = m_dt2 * dl.y + B.rl[1]; B.rl[2] = m_dt2 * dl.z + B.rl[2];
    B.ra[0] = i_dt2.x * ad.x + B.ra[0]; B.ra[1] = i_dt2.y * ad.y + B.ra[1]; B.ra[2] = i_dt2.z * ad.z + B.ra[2];
    float D[6];
    solve_blk(B, D);
    const f3 Dl{D[0], D[1], D[2]}, Da{D[3], D[4], D[5]};
    // ---- reverse: y' = y + D_lin, r' = quat_step(r, D_ang)
    q4 gr;
    f3 gDa;
    quat_step_vjp(snap_rot, Da, gout_r, gr, gDa);
    // z = H^-1 D_bar
    Blk Bz = B;
    Bz.rl[0] = -gout_y.x; Bz.rl[1] = -gout_y.y; Bz.rl[2] = -gout_y.z;
    Bz.ra[0] = -gDa.x; Bz.ra[1] = -gDa.y; Bz.ra[2] = -gDa.z;
    float z[6];
    solve_blk(Bz, z);
    const f3 zl{z[0], z[1], z[2]}, za{z[3], z[4], z[5]};
    const f3 ggl = neg(zl), gga = neg(za);  // gradient adjoint
    // rows (this lane's entries): per-lane partial adjoints
    f3 gdsl = f3_zero(), gdsa = f3_zero();
    q4 grr = q4_zero();
    float gmu = 0.0f;
    for (uint32_t e = e0 + lane; e < e1; e += G) {
        EntryRows E;
        entry_rows(v, k, i, e, dsl, dsa, snap_rot, E);
        const float kq[3] = {E.pen.x, E.pen.y, E.pen.z};
        float gpen[3], gf[3];
        f3 gJa[3];
        for (int q = 0; q < 3; ++q) {
            const float Jz = dot3(E.r[q], zl) + dot3(E.j[q], za);
            const float JD = dot3(E.r[q], Dl) + dot3(E.j[q], Da);
            gpen[q] = -Jz * JD;
            gf[q] = -Jz;
            gJa[q] = sub(scale(gga, E.f[q]), scale(add(scale(za, JD), scale(Da, Jz)), kq[q]));
        }
        float gk[3] = {gpen[0], gpen[1], gpen[2]}, gl[3] = {0.0f, 0.0f, 0.0f}, gc[3] = {0.0f, 0.0f, 0.0f};
        contact_force_vjp(E.pen, E.lam, v.cn[E.ci].w, E.cv[0], E.cv[1], E.cv[2], gf[0], gf[1], gf[2], gk, gl, gc, gmu);
        f3 grw = f3_zero();
        for (int q = 0; q < 3; ++q) {
            gdsl = add(gdsl, scale(E.r[q], gc[q]));
            gJa[q] = add(gJa[q], scale(dsa, gc[q]));
            gdsa = add(gdsa, scale(E.j[q], gc[q]));
            grw = add(grw, cross3(E.r[q], gJa[q]));
        }
        q4 gq;
        f3 gvv;
        qrot_vjp(snap_rot, E.rl_self, grw, gq, gvv);
        grr = qadd4(grr, gq);
        const uint32_t sl = 2 * E.ci + (E.is_a ? 0u : 1u);
        SLOT(v, sl, kSlotPen) = make_float4(gk[0], gk[1], gk[2], 0.0f);
        SLOT(v, sl, kSlotLam) = make_float4(gl[0], gl[1], gl[2], 0.0f);
        if (E.other != kContainer) {
            // c depends on the other body's pose: - r.(Y_o - x0_o) - (R_o r_o x r).ang_diff(R_o, q0_o)
            f3 gdol = f3_zero(), gdoa = f3_zero(), gro = f3_zero();
            for (int q = 0; q < 3; ++q) {
                gdol = sub(gdol, scale(E.r[q], gc[q]));
                gro = add(gro, cross3(E.r[q], scale(E.doa, -gc[q])));
                gdoa = sub(gdoa, scale(cross3(E.r_other, E.r[q]), gc[q]));
            }
            q4 gto, gfrom, gqo;
            f3 gvo;
            ang_diff_vjp(E.rot_other, E.q0_other, gdoa, gto, gfrom);
            qrot_vjp(E.rot_other, E.rl_other, gro, gqo, gvo);
            OPOSE(v, sl, 0) = f4(gdol, 0.0f);
            OPOSE(v, sl, 1) = q4f(qadd4(gto, gqo));
            OPOSE(v, sl, 2) = f4(neg(gdol), 0.0f);
            OPOSE(v, sl, 3) = q4f(gfrom);
        }
    }
    gdsl = gsum3<G>(gdsl);
    gdsa = gsum3<G>(gdsa);
    grr = gsum4<G>(grr);
    gmu = gsum<G>(gmu);
    if (lane != 0) return;
    // mass terms
    const float gm = -(z[0] * D[0] + z[1] * D[1] + z[2] * D[2]) / dt2 + dot3(ggl, dl) / dt2;
    const f3 gI{(-z[3] * D[3] + gga.x * ad.x) / dt2, (-z[4] * D[4] + gga.y * ad.y) / dt2,
                (-z[5] * D[5] + gga.z * ad.z) / dt2};
    f3 gy = add(gout_y, scale(ggl, m_dt2));
    const f3 gxt = scale(ggl, -m_dt2);
    q4 gqt;
    {
        const f3 gad{i_dt2.x * gga.x, i_dt2.y * gga.y, i_dt2.z * gga.z};
        q4 gto;
        ang_diff_vjp(snap_rot, qt, gad, gto, gqt);
        gr = qadd4(gr, gto);
    }
    gr = qadd4(gr, grr);
    gy = add(gy, gdsl);
    gx0 = sub(gx0, gdsl);
    {
        q4 gto, gfrom;
        ang_diff_vjp(snap_rot, q0i, gdsa, gto, gfrom);
        gr = qadd4(gr, gto);
        gq0 = qadd4(gq0, gfrom);
    }
    YB(v, nxt, 0, i) = f4(gy, 0.0f);
    YB(v, nxt, 1, i) = q4f(gr);
    float4& a0 = ACC(v, 0, i);
    a0 = make_float4(a0.x + gx0.x, a0.y + gx0.y, a0.z + gx0.z, 0.0f);
    float4& a1 = ACC(v, 1, i);
    a1 = make_float4(a1.x + gq0.x, a1.y + gq0.y, a1.z + gq0.z, a1.w + gq0.w);
    float4& a2 = ACC(v, 2, i);
    a2 = make_float4(a2.x + gxt.x, a2.y + gxt.y, a2.z + gxt.z, 0.0f);
    float4& a3 = ACC(v, 3, i);
    a3 = make_float4(a3.x + gqt.x, a3.y + gqt.y, a3.z + gqt.z, a3.w + gqt.w);
    // mass = 1 / inv_mass, moment = 1 / inv_moment
    float4& mb = v.mbar[i];
    mb = make_float4(mb.x - moment.x * moment.x * gI.x, mb.y - moment.y * moment.y * gI.y,
                     mb.z - moment.z * moment.z * gI.z, mb.w - mass * mass * gm);
    v.mubar_body[i] += gmu;
}

// Reverse of init: the adjoint of the step's initial state.
__global__ void __launch_bounds__(kBlock) k_adj_init_rev(AdjStepView v, uint32_t cur) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= v.n) return;
    f3 gy0 = xyz(YB(v, cur, 0, i));
    q4 gr0 = as_q(YB(v, cur, 1, i));
    f3 gx0 = xyz(ACC(v, 0, i));
    q4 gq0 = as_q(ACC(v, 1, i));
    gather_other_rows(v, v.off[i], v.off[i + 1], gy0, gr0, gx0, gq0);
    f3 gv = xyz(v.sbar[2 * (size_t)v.n_cap + i]);
    f3 gw = xyz(v.sbar[3 * (size_t)v.n_cap + i]);
    f3 gvp = xyz(v.sbar[4 * (size_t)v.n_cap + i]);
    f3 gx;
    q4 gq;
    if (v.inv_mm[i].w != 0.0f) {
        const float dt = v.p.dt;
        const f3 gxt = xyz(ACC(v, 2, i));
        const q4 gqt = as_q(ACC(v, 3, i));
        gx = add(add(gy0, gxt), gx0);
        gv = add(gv, scale(add(gy0, gxt), dt));
        // adaptive warm-start weight (k_avbd_init's arithmetic)
        const f3 g{v.p.gx, v.p.gy, v.p.gz};
        const f3 vel = xyz(v.vel[i]), prev = xyz(v.prev[i]);
        const f3 accel = divs(sub(vel, prev), dt);
        const float gmag = len3(g);
        if (gmag > 0.0f) {
            const float s = dot3(accel, divs(g, gmag)) / gmag;
            if (!(s < 0.0f) && !(s > 1.0f)) {
                const float gwt = dot3(gy0, g) * (dt * dt);
                const f3 gacc = scale(divs(g, gmag), gwt / gmag);
                gv = add(gv, divs(gacc, dt));
                gvp = sub(gvp, divs(gacc, dt));
            }
        }
        // R0 = q~ = quat_step(q, w dt), also the inertial rotation
        const q4 q = as_q(v.q0[i]);
        q4 gqs;
        f3 gd;
        quat_step_vjp(q, scale(xyz(v.ang[i]), dt), qadd4(gr0, gqt), gqs, gd);
        gw = add(gw, scale(gd, dt));
        gq = qadd4(gqs, gq0);
    } else {
        gx = add(gy0, gx0);
        gq = qadd4(gr0, gq0);
    }
    v.sbar[i] = f4(gx, 0.0f);
    v.sbar[v.n_cap + i] = q4f(gq);
    v.sbar[2 * (size_t)v.n_cap + i] = f4(gv, 0.0f);
    v.sbar[3 * (size_t)v.n_cap + i] = f4(gw, 0.0f);
    v.sbar[4 * (size_t)v.n_cap + i] = f4(gvp, 0.0f);
}

// fixed-order sum of x[0..n) added to *out (one workgroup)
template <class T>
__global__ void __launch_bounds__(1024) k_sum_into(const T* x, uint32_t n, double* out) {
    __shared__ double part[1024];
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) s += (double)x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += part[0];
}

void launch_adj_primal(int lanes, const AdjStepView& v, uint32_t k, uint32_t cur, uint32_t nxt, uint32_t has_dual,
                       hipStream_t s) {
    const uint32_t n = v.n;
    switch (lanes) {
        case 2: hipLaunchKernelGGL(k_adj_primal<2>, dim3((n + 127) / 128), dim3(kBlock), 0, s, v, k, cur, nxt, has_dual); break;
        case 4: hipLaunchKernelGGL(k_adj_primal<4>, dim3((n + 63) / 64), dim3(kBlock), 0, s, v, k, cur, nxt, has_dual); break;
        case 8: hipLaunchKernelGGL(k_adj_primal<8>, dim3((n + 31) / 32), dim3(kBlock), 0, s, v, k, cur, nxt, has_dual); break;
        case 16: hipLaunchKernelGGL(k_adj_primal<16>, dim3((n + 15) / 16), dim3(kBlock), 0, s, v, k, cur, nxt, has_dual); break;
        default: hipLaunchKernelGGL(k_adj_primal<1>, dim3(blocks_for(n)), dim3(kBlock), 0, s, v, k, cur, nxt, has_dual); break;
    }
}

}  // namespace

void launch_record_step(const AdjRecord& r, const ContactSet& cs, const uint32_t* off, const uint32_t* idx,
                        const uint32_t* peer, const BodyState& st, uint32_t n, uint32_t m_cap, hipStream_t s) {
    hipLaunchKernelGGL(k_record_head, dim3(1), dim3(64), 0, s, r, (const uint32_t*)cs.m_dev, off, n);
    const uint32_t work = std::max(std::max(m_cap, 2 * m_cap), n + 1);
    hipLaunchKernelGGL(k_record_copy, dim3(grid_for(work)), dim3(kBlock), 0, s, r, cs.c, off, idx, peer, st, n);
}

void adj_work_free(AdjWork& w) {
    if (w.base) (void)hipFree(w.base);
    w = AdjWork{};
}

void adj_work_reserve(AdjWork& w, uint32_t n_cap, uint32_t every, uint32_t iters, uint32_t m_max, uint32_t n_ck) {
    if (w.base && w.n_cap == n_cap && w.every >= every && w.iters == iters && w.m_max >= m_max && w.n_ck >= n_ck)
        return;
    adj_work_free(w);
    if (m_max == 0) m_max = 1;
    const size_t n = n_cap, m = m_max;
    size_t f4 = 0;
    auto take = [&](size_t count) {
        const size_t at = f4;
        f4 += (count + 15) & ~(size_t)15;
        return at;
    };
    const size_t o_st = take(5 * n), o_inv = take(n), o_ts = take((size_t)every * 5 * n), o_ti = take((size_t)every * 2 * n);
    const size_t o_tp = take((size_t)every * (iters + 1) * 2 * n), o_tm = take((size_t)every * iters * 2 * m);
    const size_t o_sb = take(5 * n), o_yb = take(4 * n), o_acc = take(4 * n), o_mb = take(n);
    const size_t o_sl = take(2 * m * 2), o_dp = take(2 * m), o_dpo = take(2 * m * 4), o_opo = take(2 * m * 4);
    const size_t o_mub = take(n / 4 + 1), o_muc = take(m / 2 + 1);  // float per body, double per contact
    const size_t o_ck = take((size_t)n_ck * 5 * n);
    const size_t bytes = f4 * 16 + 64;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("adjoint scratch allocation failed");
    float4* b4 = static_cast<float4*>(p);
    w.base = p;
    w.bytes = bytes;
    w.n_cap = n_cap;
    w.every = every;
    w.iters = iters;
    w.m_max = m_max;
    w.n_ck = n_ck;
    w.ck = b4 + o_ck;
    w.st = BodyState{b4 + o_st, b4 + o_st + n, b4 + o_st + 2 * n, b4 + o_st + 3 * n, b4 + o_st + 4 * n};
    w.inv_mm = b4 + o_inv;
    w.tape_state = b4 + o_ts;
    w.tape_iner = b4 + o_ti;
    w.tape_pose = b4 + o_tp;
    w.tape_mult = b4 + o_tm;
    w.sbar = b4 + o_sb;
    w.ybar = b4 + o_yb;
    w.acc = b4 + o_acc;
    w.mbar = b4 + o_mb;
    w.slots = b4 + o_sl;
    w.dpart = b4 + o_dp;
    w.dpose = b4 + o_dpo;
    w.opose = b4 + o_opo;
    w.mubar_body = reinterpret_cast<float*>(b4 + o_mub);
    w.mubar_contact = reinterpret_cast<double*>(b4 + o_muc);
    w.mubar_total = reinterpret_cast<double*>(static_cast<char*>(p) + f4 * 16);
}

void adj_restore(const float4* ck, AdjWork& w, uint32_t k, uint32_t n, hipStream_t s) {
    const size_t nc = w.n_cap;
    const float4* c = ck + (size_t)k * 5 * nc;
    copy4(c, w.st.pos, n, s);
    copy4(c + nc, w.st.rot, n, s);
    copy4(c + 2 * nc, w.st.vel, n, s);
    copy4(c + 3 * nc, w.st.ang, n, s);
    copy4(c + 4 * nc, w.st.prev, n, s);
}

void adj_save(AdjWork& w, uint32_t k, uint32_t n, hipStream_t s) {
    const size_t nc = w.n_cap;
    float4* c = w.ck + (size_t)k * 5 * nc;
    copy4(w.st.pos, c, n, s);
    copy4(w.st.rot, c + nc, n, s);
    copy4(w.st.vel, c + 2 * nc, n, s);
    copy4(w.st.ang, c + 3 * nc, n, s);
    copy4(w.st.prev, c + 4 * nc, n, s);
}

void adj_set_friction(const AdjRecord& r, const AdjHostIndex& h, float mu, hipStream_t s) {
    const uint32_t m = h.steps ? h.coff[h.steps] : 0u;
    if (m == 0) return;
    hipLaunchKernelGGL(k_set_friction, dim3(grid_for(m)), dim3(kBlock), 0, s, r.c.n, m, mu);
}

void adj_replay(const AdjRecord& r, const AdjHostIndex& h, AdjWork& w, const AdjProblem& pb, uint32_t first,
                uint32_t last, hipStream_t s) {
    const uint32_t n = pb.n, I = pb.iters;
    for (uint32_t st = first; st < last; ++st) {
        const uint32_t j = st - first;
        SolveView v = step_view(r, h, w, pb, st, j);
        copy4(w.st.pos, tstate(w, j, 0), n, s);
        copy4(w.st.rot, tstate(w, j, 1), n, s);
        copy4(w.st.vel, tstate(w, j, 2), n, s);
        copy4(w.st.ang, tstate(w, j, 3), n, s);
        copy4(w.st.prev, tstate(w, j, 4), n, s);
        v.pos[0] = tpose(w, j, 0, 0);
        v.rot[0] = tpose(w, j, 0, 1);
        v.pos[1] = tpose(w, j, 1, 0);
        v.rot[1] = tpose(w, j, 1, 1);
        launch_avbd_init(v, pb.p, s);
        for (uint32_t k = 2; k <= I; ++k) {
            copy4(tpose(w, j, 1, 0), tpose(w, j, k, 0), n, s);
            copy4(tpose(w, j, 1, 1), tpose(w, j, k, 1), n, s);
        }
        const uint32_t m = h.m[st];
        copy4(r.c.lam + h.coff[st], tmult(w, j, 0, 0), m, s);
        copy4(r.c.pen + h.coff[st], tmult(w, j, 0, 1), m, s);
        for (uint32_t k = 0; k < I; ++k) {
            const uint32_t par = k & 1u;
            v.pos[par] = tpose(w, j, k, 0);
            v.rot[par] = tpose(w, j, k, 1);
            v.pos[1 - par] = tpose(w, j, k + 1, 0);
            v.rot[1 - par] = tpose(w, j, k + 1, 1);
            v.c_lam = tmult(w, j, k, 0);
            v.c_pen = tmult(w, j, k, 1);
            v.c_lam_alt = k + 1 < I ? tmult(w, j, k + 1, 0) : nullptr;
            v.c_pen_alt = k + 1 < I ? tmult(w, j, k + 1, 1) : nullptr;
            launch_avbd_primal(v, pb.p, k, false, s, pb.lanes);
            if (k + 1 < I) launch_avbd_dual(v, pb.p, par, 0u, 1u, m, s);
        }
        const uint32_t fp = (I - 1) & 1u;
        v.pos[1 - fp] = tpose(w, j, I, 0);
        v.rot[1 - fp] = tpose(w, j, I, 1);
        v.pos[fp] = tpose(w, j, I - 1, 0);
        v.rot[fp] = tpose(w, j, I - 1, 1);
        launch_avbd_recover(v, pb.p, fp, s);
    }
}

void adj_backward(const AdjRecord& r, const AdjHostIndex& h, AdjWork& w, const AdjProblem& pb, const float4* ck,
                  hipStream_t s) {
    const uint32_t n = pb.n, I = pb.iters, E = r.every;
    launch_fill_u32(reinterpret_cast<uint32_t*>(w.mbar), 0u, 4 * n, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(w.mubar_body), 0u, n, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(w.mubar_contact), 0u, 2 * w.m_max, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(w.mubar_total), 0u, 2, s);
    if (h.steps == 0) return;
    const uint32_t segs = (h.steps + E - 1) / E;
    for (int sg = (int)segs - 1; sg >= 0; --sg) {
        const uint32_t first = (uint32_t)sg * E;
        const uint32_t last = std::min(h.steps, first + E);
        adj_restore(ck, w, (uint32_t)sg, n, s);
        adj_replay(r, h, w, pb, first, last, s);
        for (int st = (int)last - 1; st >= (int)first; --st) {
            const uint32_t j = (uint32_t)st - first;
            const uint32_t m = h.m[st];
            AdjStepView v{};
            v.n = n;
            v.m = m;
            v.n_cap = w.n_cap;
            v.m_max = w.m_max;
            v.iters = I;
            v.frozen = pb.frozen;
            v.p = pb.p;
            v.mu = pb.mu;
            v.inv_mm = pb.inv_mm;
            const uint32_t c0 = h.coff[st];
            v.ab = r.c.ab + c0;
            v.cn = r.c.n + c0;
            v.ct1 = r.c.t1 + c0;
            v.ct2 = r.c.t2 + c0;
            v.cra = r.c.ra + c0;
            v.crb = r.c.rb + c0;
            v.cc0 = r.c.c0 + c0;
            v.off = r.csr_off + (size_t)st * (r.n_cap + 1);
            v.idx = r.csr_idx + h.ioff[st];
            v.peer = r.csr_peer + h.ioff[st];
            v.x0 = tstate(w, j, 0);
            v.q0 = tstate(w, j, 1);
            v.vel = tstate(w, j, 2);
            v.ang = tstate(w, j, 3);
            v.prev = tstate(w, j, 4);
            v.xt = tiner(w, j, 0);
            v.qt = tiner(w, j, 1);
            v.pose = tpose(w, j, 0, 0);
            v.mult = tmult(w, j, 0, 0);
            v.sbar = w.sbar;
            v.ybar = w.ybar;
            v.acc = w.acc;
            v.mbar = w.mbar;
            v.mubar_body = w.mubar_body;
            v.mubar_contact = w.mubar_contact;
            v.slots = w.slots;
            v.dpart = w.dpart;
            v.dpose = w.dpose;
            v.opose = w.opose;
            hipLaunchKernelGGL(k_adj_recover_rev, dim3(grid_for(std::max(n, m))), dim3(kBlock), 0, s, v);
            uint32_t cur = 0;
            for (int k = (int)I - 1; k >= 0; --k) {
                const bool dual = k + 1 < (int)I;
                // frozen: dual k's outputs are constants, so nothing flows back through it
                if (dual && m > 0 && !pb.frozen)
                    hipLaunchKernelGGL(k_adj_dual_rev, dim3(grid_for(m)), dim3(kBlock), 0, s, v, (uint32_t)k);
                launch_adj_primal(pb.rev_lanes, v, (uint32_t)k, cur, 1u - cur, dual ? 1u : 0u, s);
                cur = 1u - cur;
            }
            hipLaunchKernelGGL(k_adj_init_rev, dim3(blocks_for(n)), dim3(kBlock), 0, s, v, cur);
        }
    }
    hipLaunchKernelGGL(k_sum_into<double>, dim3(1), dim3(1024), 0, s, (const double*)w.mubar_contact, w.m_max,
                       w.mubar_total);
    hipLaunchKernelGGL(k_sum_into<float>, dim3(1), dim3(1024), 0, s, (const float*)w.mubar_body, n, w.mubar_total);
}

}  // namespace dio
