// oracle_capi.cpp — C entry points of the CPU oracle, loaded with ctypes.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker / baseline, never by the
// product library.  `precision` selects the FP32 parity reference (32) or the
// FP64 oracle (64).  Record layouts are the reference's #[repr(C)] records as
// declared in include/diomedes_hip.h.
#include <type_traits>
#include <cstdint>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/diomedes_hip.h"
#include "avbd_oracle.hpp"

using namespace oracle;

namespace {

template <class T> Body<T> body_from(const dio_gpu_body& g, const dio_avbd_body_state* st) {
    Body<T> b;
    b.tag = g.shape.tag;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) b.corners[i][k] = T(g.shape.corners[i][k]);
    b.position = {T(g.position[0]), T(g.position[1]), T(g.position[2])};
    b.orientation = {T(g.orientation[0]), T(g.orientation[1]), T(g.orientation[2]), T(g.orientation[3])};
    if (st) {
        b.velocity = {T(st->vel[0]), T(st->vel[1]), T(st->vel[2])};
        b.angular_velocity = {T(st->ang[0]), T(st->ang[1]), T(st->ang[2])};
        b.prev_velocity = {T(st->prev_vel[0]), T(st->prev_vel[1]), T(st->prev_vel[2])};
        b.inv_moment = {T(st->inv_moment[0]), T(st->inv_moment[1]), T(st->inv_moment[2])};
        b.inv_mass = T(st->inv_mass);
        b.friction = T(st->friction);
    }
    return b;
}

template <class T> std::vector<V3<T>> hull_from(const float* v, uint32_t count) {
    std::vector<V3<T>> h(count);
    for (uint32_t i = 0; i < count; ++i) h[i] = {T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2])};
    return h;
}

template <class T> Options<T> options_from(const dio_avbd_options* o) {
    Options<T> r;
    r.dt = T(o->dt);
    r.gravity = {T(o->gravity[0]), T(o->gravity[1]), T(o->gravity[2])};
    r.iterations = o->iterations;
    r.alpha = T(o->alpha);
    r.beta_lin = T(o->beta_lin);
    r.margin = T(o->margin);
    r.penalty_min = T(o->penalty_min);
    r.penalty_max = T(o->penalty_max);
    r.gamma = T(o->gamma);
    r.friction = T(o->friction);
    return r;
}

template <class T> Constraint<T> constraint_from(const dio_avbd_constraint& c) {
    Constraint<T> r;
    r.a = c.a;
    r.b = c.b;
    r.kind = c.kind;
    r.anchor_a = {T(c.anchor_a[0]), T(c.anchor_a[1]), T(c.anchor_a[2])};
    r.length = T(c.anchor_a[3]);
    r.anchor_b = {T(c.anchor_b[0]), T(c.anchor_b[1]), T(c.anchor_b[2])};
    r.stiffness = T(c.anchor_b[3]);
    return r;
}

template <class T> Contact<T> contact_from(const dio_avbd_contact& c) {
    Contact<T> r;
    auto v = [](const float* p) { return V3<T>{T(p[0]), T(p[1]), T(p[2])}; };
    r.a = c.a;
    r.b = c.b;
    r.friction = T(c.friction);
    r.normal = v(c.normal);
    r.tangent1 = v(c.tangent1);
    r.tangent2 = v(c.tangent2);
    r.r_a = v(c.r_a);
    r.r_b = v(c.r_b);
    r.c0 = v(c.c0);
    r.penalty = v(c.penalty);
    r.lambda = v(c.lambda);
    return r;
}

template <class T> void put3(double* o, V3<T> v) { o[0] = double(v.x); o[1] = double(v.y); o[2] = double(v.z); }
template <class T> void put4(double* o, Q<T> q) {
    o[0] = double(q.x); o[1] = double(q.y); o[2] = double(q.z); o[3] = double(q.w);
}

// 25 doubles per contact: friction, normal, t1, t2, r_a, r_b, c0, penalty, lambda
template <class T> void put_contact(double* o, uint32_t* ab, const Contact<T>& c) {
    ab[0] = c.a;
    ab[1] = c.b;
    o[0] = double(c.friction);
    put3(o + 1, c.normal);
    put3(o + 4, c.tangent1);
    put3(o + 7, c.tangent2);
    put3(o + 10, c.r_a);
    put3(o + 13, c.r_b);
    put3(o + 16, c.c0);
    put3(o + 19, c.penalty);
    put3(o + 22, c.lambda);
}

struct WorldHandle {
    int precision;
    std::unique_ptr<World<float>> f;
    std::unique_ptr<World<double>> d;
};

template <class T>
void fill_world(World<T>& w, const dio_gpu_body* bodies, const dio_avbd_body_state* state, uint32_t n,
                const dio_avbd_constraint* cons, uint32_t c, const dio_container* ct, const float* hull,
                uint32_t hc) {
    w.bodies.resize(n);
    for (uint32_t i = 0; i < n; ++i) w.bodies[i] = body_from<T>(bodies[i], state ? &state[i] : nullptr);
    w.constraints.resize(c);
    for (uint32_t i = 0; i < c; ++i) w.constraints[i] = constraint_from<T>(cons[i]);
    if (ct && ct->enabled) {
        w.has_container = true;
        w.container.center = {T(ct->center[0]), T(ct->center[1]), T(ct->center[2])};
        w.container.rotation = {T(ct->rotation[0]), T(ct->rotation[1]), T(ct->rotation[2]), T(ct->rotation[3])};
        w.container.half_extent = T(ct->half_extent);
    }
    w.hull = hull_from<T>(hull, hc);
}

template <class F> auto with_world(WorldHandle* h, F&& f) {
    if (h->precision == 64) return f(*h->d);
    return f(*h->f);
}

}  // namespace

#define OAPI extern "C"

OAPI int oracle_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

OAPI void* oracle_world_create(int precision, const dio_gpu_body* bodies, const dio_avbd_body_state* state, uint32_t n,
                          const dio_avbd_constraint* cons, uint32_t c, const dio_container* ct, const float* hull,
                          uint32_t hc) {
    auto* h = new WorldHandle{precision, nullptr, nullptr};
    if (precision == 64) {
        h->d.reset(new World<double>());
        fill_world(*h->d, bodies, state, n, cons, c, ct, hull, hc);
    } else {
        h->f.reset(new World<float>());
        fill_world(*h->f, bodies, state, n, cons, c, ct, hull, hc);
    }
    return h;
}

OAPI void oracle_world_destroy(void* p) { delete static_cast<WorldHandle*>(p); }

// A copy of a world in its current state (bodies, contacts with their
// multipliers, the pending narrow output, joints, timers): stepping the copy
// continues the same trajectory.
OAPI void* oracle_world_clone(void* p) {
    auto* h = static_cast<WorldHandle*>(p);
    auto* c = new WorldHandle{h->precision, nullptr, nullptr};
    if (h->d) c->d.reset(new World<double>(*h->d));
    if (h->f) c->f.reset(new World<float>(*h->f));
    return c;
}

OAPI void oracle_world_step(void* p, const dio_avbd_options* o, int lag, uint32_t nsteps) {
    auto* h = static_cast<WorldHandle*>(p);
    with_world(h, [&](auto& w) {
        using T = typename std::remove_reference<decltype(w.bodies[0].inv_mass)>::type;
        Options<T> opt = options_from<T>(o);
        for (uint32_t s = 0; s < nsteps; ++s) w.step(opt, lag);
        return 0;
    });
}

OAPI void oracle_world_get_state(void* p, double* pos, double* rot, double* vel, double* ang, double* prev) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.bodies.size(); ++i) {
            const auto& b = w.bodies[i];
            if (pos) put3(pos + 3 * i, b.position);
            if (rot) put4(rot + 4 * i, b.orientation);
            if (vel) put3(vel + 3 * i, b.velocity);
            if (ang) put3(ang + 3 * i, b.angular_velocity);
            if (prev) put3(prev + 3 * i, b.prev_velocity);
        }
        return 0;
    });
}

OAPI uint32_t oracle_world_contact_count(void* p) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) { return (uint32_t)w.contacts.size(); });
}

OAPI void oracle_world_get_contacts(void* p, uint32_t* ab, double* data) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.contacts.size(); ++i) put_contact(data + 25 * i, ab + 2 * i, w.contacts[i]);
        return 0;
    });
}

OAPI uint32_t oracle_world_pair_count(void* p) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) { return (uint32_t)w.last_pairs.size(); });
}

OAPI void oracle_world_get_pairs(void* p, uint32_t* out) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.last_pairs.size(); ++i) { out[2 * i] = w.last_pairs[i].a; out[2 * i + 1] = w.last_pairs[i].b; }
        return 0;
    });
}

OAPI uint32_t oracle_world_csr_nnz(void* p) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) { return (uint32_t)w.indices.size(); });
}

OAPI void oracle_world_get_csr(void* p, uint32_t* offsets, uint32_t* indices) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.offsets.size(); ++i) offsets[i] = w.offsets[i];
        for (size_t i = 0; i < w.indices.size(); ++i) indices[i] = w.indices[i];
        return 0;
    });
}

// Each constraint's multipliers after the last step: lambda and penalty per
// row, (c, 3) each (zero for soft joints, which carry none).
OAPI uint32_t oracle_world_get_joints(void* p, double* lam, double* pen) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        const size_t c = w.constraints.size();
        for (size_t i = 0; i < c; ++i) {
            if (i < w.joint_state.size()) {
                put3(lam + 3 * i, w.joint_state[i].lam);
                put3(pen + 3 * i, w.joint_state[i].pen);
            } else {
                for (int k = 0; k < 3; ++k) lam[3 * i + k] = pen[3 * i + k] = 0.0;
            }
        }
        return (uint32_t)c;
    });
}

// 0: Jacobi (the reference); 1: coloured Gauss-Seidel
OAPI void oracle_world_set_mode(void* p, int mode) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        w.mode = mode;
        return 0;
    });
}

// The last coloured step's colouring: colour per body (kNotInGraph for static
// bodies); returns the number of colours, *rounds the Jones-Plassmann rounds.
OAPI uint32_t oracle_world_get_colours(void* p, uint32_t* colour, uint32_t* rounds) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.plan.colour.size(); ++i) colour[i] = w.plan.colour[i];
        if (rounds) *rounds = w.plan.rounds;
        return (uint32_t)(w.plan.start.empty() ? 0 : w.plan.start.size() - 1);
    });
}

// Colour an arbitrary graph: n vertices (dynamic[i] != 0 takes part), edges
// (a, b) pairs.  Colour per vertex out; returns the number of colours, or
// 0xFFFFFFFF when a vertex found all colours taken.
OAPI uint32_t oracle_colour_graph(uint32_t n, const uint8_t* dynamic, const uint32_t* edges, uint32_t ne,
                                  uint32_t* colour, uint32_t* rounds) {
    std::vector<bool> dyn(n);
    for (uint32_t i = 0; i < n; ++i) dyn[i] = dynamic[i] != 0;
    std::vector<std::vector<uint32_t>> adj(n);
    for (uint32_t e = 0; e < ne; ++e) {
        const uint32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a < n && b < n && a != b) {
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
    }
    const ColourPlan P = colour_graph(n, dyn, [&](uint32_t i, auto&& f) {
        for (uint32_t o : adj[i])
            if (dyn[o]) f(o);
    });
    for (uint32_t i = 0; i < n; ++i) colour[i] = P.colour[i];
    if (rounds) *rounds = P.rounds;
    return P.overflow ? 0xFFFFFFFFu : (uint32_t)(P.start.empty() ? 0 : P.start.size() - 1);
}

// broad, narrow, contacts, avbd, total, init, primal, dual, recover (seconds)
OAPI void oracle_world_times(void* p, double* out) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        const StepTimes& t = w.times;
        out[0] = t.broad; out[1] = t.narrow; out[2] = t.contacts; out[3] = t.avbd; out[4] = t.total;
        out[5] = t.solve.init; out[6] = t.solve.primal; out[7] = t.solve.dual; out[8] = t.solve.recover;
        return 0;
    });
}

OAPI void oracle_world_reset_times(void* p) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) { w.times = StepTimes(); return 0; });
}

// ------------------------------------------------------------ unit level
template <class T>
static void aabbs_t(const dio_gpu_body* bodies, uint32_t n, const float* hull, uint32_t hc, double* out) {
    auto h = hull_from<T>(hull, hc);
    for (uint32_t i = 0; i < n; ++i) {
        V3<T> mn, mx;
        aabb(body_from<T>(bodies[i], nullptr), h, mn, mx);
        put3(out + 6 * i, mn);
        put3(out + 6 * i + 3, mx);
    }
}
OAPI void oracle_aabbs(int precision, const dio_gpu_body* bodies, uint32_t n, const float* hull, uint32_t hc,
                  double* out) {
    if (precision == 64) aabbs_t<double>(bodies, n, hull, hc, out);
    else aabbs_t<float>(bodies, n, hull, hc, out);
}

template <class T>
static uint32_t broad_t(const dio_gpu_body* bodies, uint32_t n, const float* hull, uint32_t hc, uint32_t* out,
                        uint32_t cap) {
    std::vector<Body<T>> b(n);
    for (uint32_t i = 0; i < n; ++i) b[i] = body_from<T>(bodies[i], nullptr);
    auto pairs = broad_pairs(b, hull_from<T>(hull, hc));
    for (size_t i = 0; i < pairs.size() && i < cap; ++i) { out[2 * i] = pairs[i].a; out[2 * i + 1] = pairs[i].b; }
    return (uint32_t)pairs.size();
}
OAPI uint32_t oracle_broad_pairs(int precision, const dio_gpu_body* bodies, uint32_t n, const float* hull, uint32_t hc,
                            uint32_t* out, uint32_t cap) {
    return precision == 64 ? broad_t<double>(bodies, n, hull, hc, out, cap)
                           : broad_t<float>(bodies, n, hull, hc, out, cap);
}

// out: 4 slots x 10 doubles (normal 3, depth, point_a 3, point_b 3)
template <class T>
static uint32_t narrow_t(const dio_gpu_body* a, const dio_gpu_body* b, const float* hull, uint32_t hc, double* out) {
    RawContact<T> rc[4];
    uint32_t c = narrow_pair(body_from<T>(*a, nullptr), body_from<T>(*b, nullptr), hull_from<T>(hull, hc), rc);
    for (uint32_t k = 0; k < c; ++k) {
        put3(out + 10 * k, rc[k].normal);
        out[10 * k + 3] = double(rc[k].depth);
        put3(out + 10 * k + 4, rc[k].point_a);
        put3(out + 10 * k + 7, rc[k].point_b);
    }
    return c;
}
OAPI uint32_t oracle_narrow_pair(int precision, const dio_gpu_body* a, const dio_gpu_body* b, const float* hull,
                            uint32_t hc, double* out) {
    return precision == 64 ? narrow_t<double>(a, b, hull, hc, out) : narrow_t<float>(a, b, hull, hc, out);
}

// GJK only: 1 when the shapes intersect
template <class T> static int gjk_t(const dio_gpu_body* a, const dio_gpu_body* b, const float* hull, uint32_t hc) {
    Simplex<T> s;
    return gjk(body_from<T>(*a, nullptr), body_from<T>(*b, nullptr), hull_from<T>(hull, hc), s) ? 1 : 0;
}
OAPI int oracle_gjk(int precision, const dio_gpu_body* a, const dio_gpu_body* b, const float* hull, uint32_t hc) {
    return precision == 64 ? gjk_t<double>(a, b, hull, hc) : gjk_t<float>(a, b, hull, hc);
}

// lin6 / ang6 lower triangles (00,10,20,11,21,22), cross9 row-major cross[r][c]
template <class T> static void solve6_t(const double* lin, const double* ang, const double* cr, const double* b,
                                        double* x) {
    T l[6], a[6], c[3][3], bl[3], ba[3], r[6];
    for (int i = 0; i < 6; ++i) { l[i] = T(lin[i]); a[i] = T(ang[i]); }
    for (int i = 0; i < 9; ++i) c[i / 3][i % 3] = T(cr[i]);
    for (int i = 0; i < 3; ++i) { bl[i] = T(b[i]); ba[i] = T(b[3 + i]); }
    solve6(l, a, c, bl, ba, r);
    for (int i = 0; i < 6; ++i) x[i] = double(r[i]);
}
OAPI void oracle_solve6(int precision, const double* lin, const double* ang, const double* cr, const double* b, double* x) {
    if (precision == 64) solve6_t<double>(lin, ang, cr, b, x);
    else solve6_t<float>(lin, ang, cr, b, x);
}

template <class T>
static uint32_t container_t(const dio_gpu_body* body, const dio_avbd_body_state* st, const dio_container* ct,
                            const dio_avbd_options* o, uint32_t index, const float* hull, uint32_t hc, uint32_t* ab,
                            double* data) {
    Container<T> c;
    c.center = {T(ct->center[0]), T(ct->center[1]), T(ct->center[2])};
    c.rotation = {T(ct->rotation[0]), T(ct->rotation[1]), T(ct->rotation[2]), T(ct->rotation[3])};
    c.half_extent = T(ct->half_extent);
    std::vector<Contact<T>> out;
    container_contacts(c, body_from<T>(*body, st), index, options_from<T>(o), hull_from<T>(hull, hc), out);
    for (size_t i = 0; i < out.size(); ++i) put_contact(data + 25 * i, ab + 2 * i, out[i]);
    return (uint32_t)out.size();
}
// up to 24 contacts: ab[2 * 24], data[25 * 24]
OAPI uint32_t oracle_container_contacts(int precision, const dio_gpu_body* body, const dio_avbd_body_state* st,
                                   const dio_container* ct, const dio_avbd_options* o, uint32_t index,
                                   const float* hull, uint32_t hc, uint32_t* ab, double* data) {
    return precision == 64 ? container_t<double>(body, st, ct, o, index, hull, hc, ab, data)
                           : container_t<float>(body, st, ct, o, index, hull, hc, ab, data);
}

// raw_contact for one decoded narrow contact: 25-double contact out
template <class T>
static void raw_t(const dio_gpu_body* ba, const dio_gpu_body* bb, uint32_t a, uint32_t b, const double* nrm,
                  double depth, const double* pa, const double* pb, const dio_avbd_options* o, uint32_t* ab,
                  double* data) {
    auto v = [](const double* p) { return V3<T>{T(p[0]), T(p[1]), T(p[2])}; };
    Contact<T> c = raw_contact(a, b, body_from<T>(*ba, nullptr), body_from<T>(*bb, nullptr), v(nrm), T(depth), v(pa),
                               v(pb), options_from<T>(o));
    put_contact(data, ab, c);
}
OAPI void oracle_raw_contact(int precision, const dio_gpu_body* ba, const dio_gpu_body* bb, uint32_t a, uint32_t b,
                        const double* nrm, double depth, const double* pa, const double* pb,
                        const dio_avbd_options* o, uint32_t* ab, double* data) {
    if (precision == 64) raw_t<double>(ba, bb, a, b, nrm, depth, pa, pb, o, ab, data);
    else raw_t<float>(ba, bb, a, b, nrm, depth, pa, pb, o, ab, data);
}

// finalize_contacts: old contacts (with solved λ / penalty) + new contacts ->
// warm-started new contacts and the CSR.  Contacts are dio_avbd_contact records.
template <class T>
static void finalize_t(const dio_avbd_contact* old_c, uint32_t mo, const dio_avbd_contact* new_c, uint32_t mn,
                       uint32_t n, const dio_avbd_options* o, uint32_t* ab, double* data, uint32_t* offsets,
                       uint32_t* indices) {
    std::vector<Contact<T>> old(mo), cur(mn);
    for (uint32_t i = 0; i < mo; ++i) old[i] = contact_from<T>(old_c[i]);
    for (uint32_t i = 0; i < mn; ++i) cur[i] = contact_from<T>(new_c[i]);
    std::vector<uint32_t> off, idx;
    finalize_contacts(old, cur, n, options_from<T>(o), off, idx);
    for (uint32_t i = 0; i < mn; ++i) put_contact(data + 25 * i, ab + 2 * i, cur[i]);
    for (size_t i = 0; i < off.size(); ++i) offsets[i] = off[i];
    for (size_t i = 0; i < idx.size(); ++i) indices[i] = idx[i];
}
OAPI void oracle_finalize_contacts(int precision, const dio_avbd_contact* old_c, uint32_t mo,
                              const dio_avbd_contact* new_c, uint32_t mn, uint32_t n, const dio_avbd_options* o,
                              uint32_t* ab, double* data, uint32_t* offsets, uint32_t* indices) {
    if (precision == 64) finalize_t<double>(old_c, mo, new_c, mn, n, o, ab, data, offsets, indices);
    else finalize_t<float>(old_c, mo, new_c, mn, n, o, ab, data, offsets, indices);
}

// AvbdCompute::run (gpu_physics.rs:626-650): one solve over caller records.
template <class T>
static void avbd_run_t(const dio_avbd_body_state* st, const float* pos, const float* rot, uint32_t n,
                       const dio_avbd_contact* ct, uint32_t m, const dio_avbd_constraint* cons, uint32_t c,
                       const uint32_t* offsets, const uint32_t* indices, uint32_t nidx, const dio_run_options* ro,
                       double* o_pos, double* o_rot, double* o_vel, double* o_ang, double* o_prev, double* o_lam,
                       double* o_pen) {
    SolveState<T> S;
    S.vel.resize(n); S.ang.resize(n); S.prev_vel.resize(n); S.inv_moment.resize(n); S.inv_mass.resize(n);
    S.pos[0].resize(n); S.rot[0].resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        S.vel[i] = {T(st[i].vel[0]), T(st[i].vel[1]), T(st[i].vel[2])};
        S.ang[i] = {T(st[i].ang[0]), T(st[i].ang[1]), T(st[i].ang[2])};
        S.prev_vel[i] = {T(st[i].prev_vel[0]), T(st[i].prev_vel[1]), T(st[i].prev_vel[2])};
        S.inv_moment[i] = {T(st[i].inv_moment[0]), T(st[i].inv_moment[1]), T(st[i].inv_moment[2])};
        S.inv_mass[i] = T(st[i].inv_mass);
        S.pos[0][i] = {T(pos[3 * i]), T(pos[3 * i + 1]), T(pos[3 * i + 2])};
        S.rot[0][i] = {T(rot[4 * i]), T(rot[4 * i + 1]), T(rot[4 * i + 2]), T(rot[4 * i + 3])};
    }
    S.pos[1] = S.pos[0];
    S.rot[1] = S.rot[0];
    std::vector<Contact<T>> contacts(m);
    for (uint32_t i = 0; i < m; ++i) contacts[i] = contact_from<T>(ct[i]);
    std::vector<Constraint<T>> cs(c);
    for (uint32_t i = 0; i < c; ++i) cs[i] = constraint_from<T>(cons[i]);
    std::vector<uint32_t> off(offsets, offsets + n + 1), idx(indices, indices + nidx);
    RunOptions<T> pc{T(ro->dt), {T(ro->gravity[0]), T(ro->gravity[1]), T(ro->gravity[2])}, T(ro->alpha),
                     T(ro->beta_lin), T(ro->penalty_max), ro->iterations};
    std::vector<V3<T>> op;
    std::vector<Q<T>> orr;
    avbd_solve(S, contacts, cs, off, idx, pc, op, orr);
    for (uint32_t i = 0; i < n; ++i) {
        put3(o_pos + 3 * i, op[i]);
        put4(o_rot + 4 * i, orr[i]);
        put3(o_vel + 3 * i, S.vel[i]);
        put3(o_ang + 3 * i, S.ang[i]);
        put3(o_prev + 3 * i, S.prev_vel[i]);
    }
    for (uint32_t i = 0; i < m; ++i) {
        put3(o_lam + 3 * i, contacts[i].lambda);
        put3(o_pen + 3 * i, contacts[i].penalty);
    }
}
OAPI void oracle_avbd_run(int precision, const dio_avbd_body_state* st, const float* pos, const float* rot, uint32_t n,
                     const dio_avbd_contact* ct, uint32_t m, const dio_avbd_constraint* cons, uint32_t c,
                     const uint32_t* offsets, const uint32_t* indices, uint32_t nidx, const dio_run_options* ro,
                     double* o_pos, double* o_rot, double* o_vel, double* o_ang, double* o_prev, double* o_lam,
                     double* o_pen) {
    if (precision == 64)
        avbd_run_t<double>(st, pos, rot, n, ct, m, cons, c, offsets, indices, nidx, ro, o_pos, o_rot, o_vel, o_ang,
                           o_prev, o_lam, o_pen);
    else
        avbd_run_t<float>(st, pos, rot, n, ct, m, cons, c, offsets, indices, nidx, ro, o_pos, o_rot, o_vel, o_ang,
                          o_prev, o_lam, o_pen);
}

// avbd_init alone (avbd/src/lib.rs:202-254): the warm-started pose (pos0 /
// rot0) and the inertial pose of every body, for the unit test of the
// adaptive gravity warm start.  Outputs n x 3 / n x 4 doubles.
template <class T>
static void avbd_init_t(const dio_avbd_body_state* st, const float* pos, const float* rot, uint32_t n,
                        const dio_run_options* ro, double* warm_pos, double* warm_rot, double* iner_pos,
                        double* iner_rot) {
    SolveState<T> S;
    S.vel.resize(n); S.ang.resize(n); S.prev_vel.resize(n); S.inv_moment.resize(n); S.inv_mass.resize(n);
    S.pos[0].resize(n); S.rot[0].resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        S.vel[i] = {T(st[i].vel[0]), T(st[i].vel[1]), T(st[i].vel[2])};
        S.ang[i] = {T(st[i].ang[0]), T(st[i].ang[1]), T(st[i].ang[2])};
        S.prev_vel[i] = {T(st[i].prev_vel[0]), T(st[i].prev_vel[1]), T(st[i].prev_vel[2])};
        S.inv_moment[i] = {T(st[i].inv_moment[0]), T(st[i].inv_moment[1]), T(st[i].inv_moment[2])};
        S.inv_mass[i] = T(st[i].inv_mass);
        S.pos[0][i] = {T(pos[3 * i]), T(pos[3 * i + 1]), T(pos[3 * i + 2])};
        S.rot[0][i] = {T(rot[4 * i]), T(rot[4 * i + 1]), T(rot[4 * i + 2]), T(rot[4 * i + 3])};
    }
    S.pos[1] = S.pos[0];
    S.rot[1] = S.rot[0];
    RunOptions<T> pc{T(ro->dt), {T(ro->gravity[0]), T(ro->gravity[1]), T(ro->gravity[2])}, T(ro->alpha),
                     T(ro->beta_lin), T(ro->penalty_max), 0};
    avbd_init_only(S, pc);
    for (uint32_t i = 0; i < n; ++i) {
        put3(warm_pos + 3 * i, S.pos[0][i]);
        put4(warm_rot + 4 * i, S.rot[0][i]);
        put3(iner_pos + 3 * i, S.inertial_lin[i]);
        put4(iner_rot + 4 * i, S.inertial_ang[i]);
    }
}
OAPI void oracle_avbd_init(int precision, const dio_avbd_body_state* st, const float* pos, const float* rot,
                           uint32_t n, const dio_run_options* ro, double* warm_pos, double* warm_rot,
                           double* iner_pos, double* iner_rot) {
    if (precision == 64) avbd_init_t<double>(st, pos, rot, n, ro, warm_pos, warm_rot, iner_pos, iner_rot);
    else avbd_init_t<float>(st, pos, rot, n, ro, warm_pos, warm_rot, iner_pos, iner_rot);
}

// The renderer's instance matrices (view_proj * model, model) per slot, in
// float (precision 32; inputs rounded to float, outputs exact) or double.
OAPI void oracle_instance_transforms(int precision, const double* pos, const double* rot, const double* scale,
                                     uint32_t n, const uint32_t* order, uint32_t count, const double* view_proj,
                                     double* out) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        std::vector<T> p(pos, pos + 3 * (size_t)n), r(rot, rot + 4 * (size_t)n), s(scale, scale + 3 * (size_t)n);
        std::vector<T> vp(view_proj, view_proj + 16), o(32 * (size_t)count);
        instance_transforms<T>(p.data(), r.data(), s.data(), order, count, vp.data(), o.data());
        for (size_t k = 0; k < o.size(); ++k) out[k] = (double)o[k];
    };
    if (precision == 64) run(0.0);
    else run(0.0f);
}

// glam primitives for the known-answer tests: Quat * Vec3 and Quat * Quat
OAPI void oracle_qrot(int precision, const double* q, const double* v, double* out) {
    if (precision == 64) put3(out, qrot(Q<double>{q[0], q[1], q[2], q[3]}, V3<double>{v[0], v[1], v[2]}));
    else put3(out, qrot(Q<float>{float(q[0]), float(q[1]), float(q[2]), float(q[3])},
                        V3<float>{float(v[0]), float(v[1]), float(v[2])}));
}
OAPI void oracle_qmul(int precision, const double* a, const double* b, double* out) {
    if (precision == 64) put4(out, qmul(Q<double>{a[0], a[1], a[2], a[3]}, Q<double>{b[0], b[1], b[2], b[3]}));
    else put4(out, qmul(Q<float>{float(a[0]), float(a[1]), float(a[2]), float(a[3])},
                        Q<float>{float(b[0]), float(b[1]), float(b[2]), float(b[3])}));
}
// Shape::support (host, last-max) and the kernels' support (first-max); a
// hull body reads its vertices from the pool hull[0 .. hc)
template <class T>
static void support_t(int host, const dio_gpu_body* body, const double* d, const float* hull, uint32_t hc,
                      double* out) {
    Body<T> b = body_from<T>(*body, nullptr);
    V3<T> dd{T(d[0]), T(d[1]), T(d[2])};
    auto h = hull_from<T>(hull, hc);
    put3(out, host ? support_cpu(b, dd, h) : support_gpu(b, dd, h));
}
OAPI void oracle_support(int precision, int host, const dio_gpu_body* body, const double* d, const float* hull,
                         uint32_t hc, double* out) {
    if (precision == 64) support_t<double>(host, body, d, hull, hc, out);
    else support_t<float>(host, body, d, hull, hc, out);
}


// ------------------------------------------------- frozen-contact replay (AD)
namespace {
using D5 = Dual<5>;
using D5f = Dual<5, float>;

std::vector<ReplayStep> replay_steps(uint32_t n, uint32_t nsteps, const dio_avbd_contact* contacts,
                                     const uint64_t* coff, const uint32_t* csr_off, const uint32_t* csr_idx,
                                     const uint64_t* ioff) {
    std::vector<ReplayStep> steps(nsteps);
    for (uint32_t s = 0; s < nsteps; ++s) {
        ReplayStep& st = steps[s];
        for (uint64_t c = coff[s]; c < coff[s + 1]; ++c) st.contacts.push_back(contact_from<double>(contacts[c]));
        st.offsets.assign(csr_off + (size_t)s * (n + 1), csr_off + (size_t)(s + 1) * (n + 1));
        st.indices.assign(csr_idx + ioff[s], csr_idx + ioff[s + 1]);
    }
    return steps;
}

template <class T> void put_poses(const std::vector<Body<T>>& b, double* pos, double* rot) {
    for (size_t i = 0; i < b.size(); ++i) {
        put3(pos + 3 * i, b[i].position);
        put4(rot + 4 * i, b[i].orientation);
    }
}
}  // namespace

// Replay nsteps recorded steps (frozen contact sets) from the given bodies /
// state.  params (5 doubles, or null for the state's values and `friction`):
// the initial velocity (x, y, z) and the mass m of body seed_body (inverse
// mass 1/m, inverse moments scaled as 1/m) and the friction μ of every
// contact.  precision 32 or 64: plain replay.  precision 65: FP64 with
// forward-mode derivatives along those 5 parameters; dpos (n x 3 x 5) and
// drot (n x 4 x 5) receive d(final pose)/d(parameter).  precision 33: the
// same derivatives (in double) along the float replay's values and branches.
namespace {
void replay_impl(int precision, const dio_gpu_body* bodies, const dio_avbd_body_state* st, uint32_t n,
                 const std::vector<ReplayStep>& steps, const dio_avbd_options* o, const double* params,
                 double friction, uint32_t seed_body, double* pos, double* rot, double* dpos, double* drot,
                 const std::vector<double>* sched_in, std::vector<double>* sched_out, int freeze) {
    double pv[5] = {0.0, 0.0, 0.0, 0.0, friction};
    const bool has_seed = seed_body < n;
    double c_mom[3] = {0.0, 0.0, 0.0};
    if (has_seed) {
        const dio_avbd_body_state& sb = st[seed_body];
        for (int k = 0; k < 3; ++k) pv[k] = (double)sb.vel[k];
        pv[3] = sb.inv_mass != 0.0f ? 1.0 / (double)sb.inv_mass : 0.0;
        for (int k = 0; k < 3; ++k) c_mom[k] = sb.inv_mass != 0.0f ? (double)sb.inv_moment[k] / (double)sb.inv_mass : 0.0;
    }
    if (params)
        for (int k = 0; k < 5; ++k) pv[k] = params[k];
    auto setup = [&](auto& b, auto mk) {
        using T = std::decay_t<decltype(mk(0, 0.0))>;
        for (uint32_t i = 0; i < n; ++i) b[i] = body_from<T>(bodies[i], &st[i]);
        if (has_seed && params) {
            auto& sb = b[seed_body];
            sb.velocity = {mk(0, pv[0]), mk(1, pv[1]), mk(2, pv[2])};
            if (pv[3] != 0.0) {
                const T m = mk(3, pv[3]);
                sb.inv_mass = T(1.0) / m;
                sb.inv_moment = {T(c_mom[0]) / m, T(c_mom[1]) / m, T(c_mom[2]) / m};
            }
        }
        return mk(4, pv[4]);
    };
    if (precision == 32 || precision == 64) {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            std::vector<Body<T>> b(n);
            const T mu = setup(b, [](int, double x) { return T(x); });
            replay(b, steps, options_from<T>(o), mu, sched_in, sched_out, freeze);
            put_poses(b, pos, rot);
        };
        if (precision == 32) run(float{});
        else run(double{});
        return;
    }
    auto run_dual = [&](auto tag) {
        using DT = decltype(tag);
        std::vector<Body<DT>> b(n);
        const DT mu = setup(b, [&](int k, double x) {
            DT r(x);
            if (has_seed || k == 4) r.d[k] = 1.0;
            return r;
        });
        if (has_seed && !params) {
            // seed the state's own values
            Body<DT>& sb = b[seed_body];
            sb.velocity.x.d[0] = 1.0;
            sb.velocity.y.d[1] = 1.0;
            sb.velocity.z.d[2] = 1.0;
            if (pv[3] != 0.0) {
                DT m(pv[3]);
                m.d[3] = 1.0;
                sb.inv_mass = DT(1.0) / m;
                sb.inv_moment = {DT(c_mom[0]) / m, DT(c_mom[1]) / m, DT(c_mom[2]) / m};
            }
        }
        replay(b, steps, options_from<DT>(o), mu, sched_in, sched_out, freeze);
        for (uint32_t i = 0; i < n; ++i) {
            const V3<DT>& p = b[i].position;
            const Q<DT>& q = b[i].orientation;
            pos[3 * i] = p.x.v; pos[3 * i + 1] = p.y.v; pos[3 * i + 2] = p.z.v;
            rot[4 * i] = q.x.v; rot[4 * i + 1] = q.y.v; rot[4 * i + 2] = q.z.v; rot[4 * i + 3] = q.w.v;
            for (int k = 0; k < 5; ++k) {
                if (dpos) {
                    dpos[(3 * i + 0) * 5 + k] = p.x.d[k];
                    dpos[(3 * i + 1) * 5 + k] = p.y.d[k];
                    dpos[(3 * i + 2) * 5 + k] = p.z.d[k];
                }
                if (drot) {
                    drot[(4 * i + 0) * 5 + k] = q.x.d[k];
                    drot[(4 * i + 1) * 5 + k] = q.y.d[k];
                    drot[(4 * i + 2) * 5 + k] = q.z.d[k];
                    drot[(4 * i + 3) * 5 + k] = q.w.d[k];
                }
            }
        }
    };
    if (precision == 33) run_dual(D5f{});
    else run_dual(D5{});
}
}  // namespace

OAPI void oracle_replay(int precision, const dio_gpu_body* bodies, const dio_avbd_body_state* st, uint32_t n,
                        uint32_t nsteps, const dio_avbd_contact* contacts, const uint64_t* coff,
                        const uint32_t* csr_off, const uint32_t* csr_idx, const uint64_t* ioff,
                        const dio_avbd_options* o, const double* params, double friction, uint32_t seed_body,
                        double* pos, double* rot, double* dpos, double* drot) {
    const std::vector<ReplayStep> steps = replay_steps(n, nsteps, contacts, coff, csr_off, csr_idx, ioff);
    replay_impl(precision, bodies, st, n, steps, o, params, friction, seed_body, pos, rot, dpos, drot, nullptr,
                nullptr, 0);
}

// The frozen replay: first the replay in sched_precision (32 or 64) at
// sched_params (5 doubles as params, or null for the state's values and
// `friction`) writes its schedules of penalties and multipliers, then the
// replay in `precision` (32, 64, 33 or 65, as oracle_replay) at params runs
// with them held fixed: freeze bit 0 the penalties, bit 1 the multipliers
// after each dual update.  At params = sched_params and precision =
// sched_precision it gives the plain replay's values bit for bit; its dual
// numbers are the frozen derivatives, and its central differences
// (precision 64) check them.
OAPI void oracle_replay_frozen(int precision, int sched_precision, int freeze, const dio_gpu_body* bodies,
                               const dio_avbd_body_state* st, uint32_t n, uint32_t nsteps,
                               const dio_avbd_contact* contacts, const uint64_t* coff, const uint32_t* csr_off,
                               const uint32_t* csr_idx, const uint64_t* ioff, const dio_avbd_options* o,
                               const double* sched_params, const double* params, double friction,
                               uint32_t seed_body, double* pos, double* rot, double* dpos, double* drot) {
    const std::vector<ReplayStep> steps = replay_steps(n, nsteps, contacts, coff, csr_off, csr_idx, ioff);
    std::vector<double> sched, p0((size_t)n * 3), r0((size_t)n * 4);
    replay_impl(sched_precision == 64 ? 64 : 32, bodies, st, n, steps, o, sched_params, friction, seed_body,
                p0.data(), r0.data(), nullptr, nullptr, nullptr, &sched, 0);
    replay_impl(precision, bodies, st, n, steps, o, params, friction, seed_body, pos, rot, dpos, drot, &sched,
                nullptr, freeze);
}

// The world's contacts as its last solve received them (see keep_presolve).
OAPI void oracle_world_keep_presolve(void* p, int on) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        w.keep_presolve = on != 0;
        return 0;
    });
}
OAPI uint32_t oracle_world_presolve_count(void* p) {
    return with_world(static_cast<WorldHandle*>(p), [&](auto& w) { return (uint32_t)w.presolve.size(); });
}
OAPI void oracle_world_get_presolve(void* p, dio_avbd_contact* out) {
    with_world(static_cast<WorldHandle*>(p), [&](auto& w) {
        for (size_t i = 0; i < w.presolve.size(); ++i) {
            const auto& c = w.presolve[i];
            dio_avbd_contact& r = out[i];
            std::memset(&r, 0, sizeof(r));
            r.a = c.a;
            r.b = c.b;
            r.friction = float(c.friction);
            auto put = [](float* d, const auto& v) { d[0] = float(v.x); d[1] = float(v.y); d[2] = float(v.z); };
            put(r.normal, c.normal); put(r.tangent1, c.tangent1); put(r.tangent2, c.tangent2);
            put(r.r_a, c.r_a); put(r.r_b, c.r_b); put(r.c0, c.c0);
            put(r.penalty, c.penalty); put(r.lambda, c.lambda);
        }
        return 0;
    });
}
