// avbd_oracle.hpp — CPU restatement of the diomedes physics step.
//
// TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it.  The product library never links or calls it.
//
// It restates, function by function and in the same floating-point operation
// order, the reference's Rust code (paths relative to the reference root):
//   engine/gpu_physics/broadphase/src/lib.rs      (support, aabb, dense sweep)
//   engine/gpu_physics/narrowphase/src/*.rs       (SAT, sphere, GJK, EPA)
//   engine/gpu_physics/avbd/src/lib.rs            (init, primal, dual, recover)
//   engine/src/physics/{gpu,shape,constraints}.rs (raw_contact, container walls,
//                                                  warm start, CSR)
//   engine/src/render/gpu_physics.rs:439-622      (iteration schedule, parity)
//   demos/avbd_box/src/main.rs:189-272            (step protocol, contact lag 1)
// Vector / quaternion arithmetic follows glam 0.33.3's scalar formulas (the
// reference's third-party math crate, absent from the reference tree).
//
// Templated on the scalar: float is the FP32 parity reference and the CPU
// baseline; double is the FP64 oracle.  Build with -ffp-contract=off and
// without fast-math so every operation rounds once, as written.
//
// Parity status per function (what pins it; see tests/test_oracle_*.py):
//   support (sphere, box, tet)      pinned: shape.rs:53-94 tests (both tie rules)
//   hull support / hull aabb        pinned: numpy first argmax over the pool; rotated-vertex extent
//   aabb / dense sweep              pinned: brute force pair set
//   SAT axis, depth, normal         pinned: cuboid.rs:404-422 tests, FP64 brute-force 15-axis SAT
//   SAT manifold points             pinned: clipped face as a half-plane intersection, projection
//                                           onto the reference plane, centred contact closed form,
//                                           two resting-cube closed forms
//   sphere-sphere                   pinned: closed form
//   GJK / EPA depth                 pinned: FP64 Minkowski-difference hull (scipy), tets, boxes, hulls
//   EPA witness points              pinned: containment, (pa - pb).n = depth, support planes,
//                                           sphere resting on a box
//   raw_contact / tangents          pinned: orthonormal frame, lever arms, axis-normal closed forms
//   container_contacts              pinned: gpu.rs:421 test, closed form incl. predicted depth
//   warm start / CSR                pinned: dictionary restatement of the HashMap semantics
//   solve6                          pinned: numpy.linalg.solve on random SPD blocks
//   init / recover                  pinned: free fall, adaptive-gravity cases, torque-free spin
//   primal / dual                   pinned: resting and stacked force balance, sliding friction mu g,
//                                           spinning-neighbour sign, spring / socket / slack rope
//   hard joints, bending, colouring pinned: closed forms and invariants (no reference counterpart)
//   quaternion product / rotation   pinned: constraints.rs:158 test + rotation identities
//   instance transforms (model,     pinned: FP64 products, rotation part = qrot of the basis,
//     view_proj * model)                    orthonormality, closed-form poses
// tests/test_oracle_mutations.py plants 45 plausible mistakes in copies of
// this file; each fails at least one of the designed pins above.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <utility>
#include <vector>

namespace oracle {

constexpr uint32_t CONTAINER = 0xFFFFFFFFu;
constexpr uint32_t SPHERE = 0, CUBE = 1, TETRAHEDRON = 2, BOX = 3, HULL = 4;
constexpr uint32_t SPRING = 0, ROD = 1, BALL_SOCKET = 2, ROPE = 3;
// hard joints (no counterpart in the reference, whose joints are the soft
// kinds above): the same geometry, solved to the constraint with a
// multiplier and a ramped penalty per row, as contacts are
constexpr uint32_t HARD_ROD = 4, HARD_SOCKET = 5, HARD_ROPE = 6;
// bending joints (no counterpart either): keep b's orientation at a's times a
// rest rotation (anchor_a holds the rest quaternion), softly (BEND, stiffness in
// anchor_b.w) or with multipliers (HARD_BEND)
constexpr uint32_t BEND = 7, HARD_BEND = 8;
inline bool is_hard(uint32_t kind) { return (kind >= HARD_ROD && kind <= HARD_ROPE) || kind == HARD_BEND; }
inline bool is_bend(uint32_t kind) { return kind == BEND || kind == HARD_BEND; }
// three rows per joint (socket and bend kinds), else one
inline bool three_rows(uint32_t kind) { return kind == HARD_SOCKET || kind == HARD_BEND; }

// ---------------------------------------------------------------- glam math
template <class T> struct V3 { T x, y, z; };
template <class T> struct Q { T x, y, z, w; };

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// sqrt / |x| found by argument-dependent lookup, so the templates also run on
// the forward-mode dual numbers below (std::sqrt / std::fabs for float and
// double, unchanged)
template <class T> inline T osqrt(T x) {
    using std::sqrt;
    return sqrt(x);
}
template <class T> inline T ofabs(T x) {
    using std::fabs;
    return fabs(x);
}
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> mulv(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline bool operator==(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length_squared(V3<T> a) { return dot(a, a); }
template <class T> inline T length(V3<T> a) { return osqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / length(a)); }
template <class T> inline V3<T> normalize_or_zero(V3<T> a) {
    T rcp = T(1) / length(a);
    if (std::isfinite(rcp) && rcp > T(0)) return a * rcp;
    return {T(0), T(0), T(0)};
}
// Rust f32::min / f32::max: a NaN operand yields the other operand
template <class T> inline T rmin(T a, T b) { return (b < a || a != a) ? b : a; }
template <class T> inline T rmax(T a, T b) { return (b > a || a != a) ? b : a; }
// Rust f32::clamp
template <class T> inline T rclamp(T x, T lo, T hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

template <class T> inline Q<T> qmul(Q<T> a, Q<T> b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> inline Q<T> qinv(Q<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> inline Q<T> qadd(Q<T> a, Q<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline Q<T> qscale(Q<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline Q<T> qnormalize(Q<T> q) {
    T len = osqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return qscale(q, T(1) / len);
}
// glam Quat * Vec3: v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w)
template <class T> inline V3<T> qrot(Q<T> q, V3<T> v) {
    T w = q.w;
    V3<T> b{q.x, q.y, q.z};
    T b2 = dot(b, b);
    return v * (w * w - b2) + b * (dot(v, b) * T(2)) + cross(b, v) * (w * T(2));
}

// avbd/src/lib.rs:91-94
template <class T> inline Q<T> quat_step(Q<T> q, V3<T> d) {
    Q<T> p{d.x, d.y, d.z, T(0)};
    return qnormalize(qadd(q, qscale(qmul(p, q), T(0.5))));
}
// avbd/src/lib.rs:96-99
template <class T> inline V3<T> ang_diff(Q<T> to, Q<T> from) {
    Q<T> rel = qmul(to, qinv(from));
    return {T(2) * rel.x, T(2) * rel.y, T(2) * rel.z};
}
// A bending joint's error: the rotation from b's orientation to a's times the
// rest rotation, 2 vec(qa rest qb^-1) on the shorter arc.  Turning a by a
// small world rotation d changes it by +d, turning b by d changes it by -d.
template <class T> inline V3<T> bend_error(Q<T> qa, Q<T> rest, Q<T> qb) {
    Q<T> rel = qmul(qmul(qa, rest), qinv(qb));
    T s = rel.w < T(0) ? T(-1) : T(1);
    return {(T(2) * rel.x) * s, (T(2) * rel.y) * s, (T(2) * rel.z) * s};
}

// ---------------------------------------------------------------- records
template <class T> struct Body {
    uint32_t tag = CUBE;
    T corners[4][4] = {};
    V3<T> position{};
    Q<T> orientation{T(0), T(0), T(0), T(1)};
    V3<T> velocity{}, angular_velocity{};
    T inv_mass = T(0);
    V3<T> inv_moment{};
    T friction = T(0.8);
    V3<T> prev_velocity{};
    bool is_static() const { return inv_mass == T(0); }
};

template <class T> struct Contact {  // AvbdContact (physics/gpu.rs:103-115)
    uint32_t a = 0, b = 0;
    V3<T> normal{}, tangent1{}, tangent2{}, r_a{}, r_b{}, c0{}, penalty{}, lambda{};
    T friction = T(0);
};

template <class T> struct RawContact {  // GpuContact (compute.rs:139-146)
    uint32_t a = 0, b = 0;
    V3<T> normal{};
    T depth = T(0);
    V3<T> point_a{}, point_b{};
};

template <class T> struct Constraint {  // GpuAvbdConstraint (gpu_physics.rs:49-56)
    uint32_t a = 0, b = 0, kind = 0;
    V3<T> anchor_a{};
    T length = T(0);  // anchor_a.w
    V3<T> anchor_b{};
    T stiffness = T(0);  // anchor_b.w
};

template <class T> struct Container {
    V3<T> center{};
    Q<T> rotation{T(0), T(0), T(0), T(1)};
    T half_extent = T(0);
};

template <class T> struct Options {  // AvbdOptions (physics/gpu.rs:127-155)
    T dt = T(1) / T(60);
    V3<T> gravity{T(0), T(-10), T(0)};
    uint32_t iterations = 10;
    T alpha = T(0.99), beta_lin = T(10000), margin = T(0.01), penalty_min = T(1),
      penalty_max = T(1e10), gamma = T(0.999), friction = T(0.8);
};

struct Pair { uint32_t a, b; };

// ------------------------------------------------------- shape support (GPU)
// narrowphase/src/support.rs:21-61 and broadphase/src/lib.rs:79-114: tetrahedra
// keep the FIRST maximal corner.  Hulls (extension) also keep the first.
template <class T>
inline V3<T> support_gpu(const Body<T>& s, V3<T> d, const std::vector<V3<T>>& hull) {
    switch (s.tag) {
        case SPHERE: return normalize_or_zero(d) * s.corners[0][0];
        case CUBE:
        case BOX:
            return {d.x >= T(0) ? s.corners[0][0] : -s.corners[0][0],
                    d.y >= T(0) ? s.corners[0][1] : -s.corners[0][1],
                    d.z >= T(0) ? s.corners[0][2] : -s.corners[0][2]};
        case TETRAHEDRON: {
            V3<T> best{s.corners[0][0], s.corners[0][1], s.corners[0][2]};
            T best_dot = dot(best, d);
            for (int i = 1; i < 4; ++i) {
                V3<T> c{s.corners[i][0], s.corners[i][1], s.corners[i][2]};
                T dd = dot(c, d);
                if (dd > best_dot) { best_dot = dd; best = c; }
            }
            return best;
        }
        case HULL: {
            uint32_t first = (uint32_t)s.corners[0][0], count = (uint32_t)s.corners[0][1];
            V3<T> best = hull[first];
            T best_dot = dot(best, d);
            for (uint32_t i = 1; i < count; ++i) {
                T dd = dot(hull[first + i], d);
                if (dd > best_dot) { best_dot = dd; best = hull[first + i]; }
            }
            return best;
        }
        default: return {T(0), T(0), T(0)};
    }
}

// Rust f32::total_cmp key
inline int64_t total_key(float v) {
    int32_t i;
    std::memcpy(&i, &v, 4);
    i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
    return i;
}
inline int64_t total_key(double v) {
    int64_t i;
    std::memcpy(&i, &v, 8);
    i ^= (int64_t)(((uint64_t)(i >> 63)) >> 1);
    return i;
}

// physics/shape.rs:20-48 (host Shape::support, used by the container walls):
// tetrahedra keep the LAST maximal corner (Iterator::max_by with total_cmp).
template <class T>
inline V3<T> support_cpu(const Body<T>& s, V3<T> d, const std::vector<V3<T>>& hull) {
    switch (s.tag) {
        case SPHERE: return normalize_or_zero(d) * s.corners[0][0];
        case CUBE: {
            T h = s.corners[0][0];
            return {d.x >= T(0) ? h : -h, d.y >= T(0) ? h : -h, d.z >= T(0) ? h : -h};
        }
        case TETRAHEDRON: {
            int best = 0;
            int64_t best_key = total_key(dot(V3<T>{s.corners[0][0], s.corners[0][1], s.corners[0][2]}, d));
            for (int i = 1; i < 4; ++i) {
                int64_t k = total_key(dot(V3<T>{s.corners[i][0], s.corners[i][1], s.corners[i][2]}, d));
                if (k >= best_key) { best_key = k; best = i; }
            }
            return {s.corners[best][0], s.corners[best][1], s.corners[best][2]};
        }
        default: return support_gpu(s, d, hull);
    }
}

template <class T> inline V3<T> pos_of(const Body<T>& b) { return b.position; }

// ------------------------------------------------------------ broad phase
// broadphase/src/lib.rs:116-151
template <class T>
inline V3<T> world_support(const Body<T>& b, V3<T> d, const std::vector<V3<T>>& hull) {
    V3<T> local = qrot(qinv(b.orientation), d);
    return qrot(b.orientation, support_gpu(b, local, hull)) + b.position;
}

template <class T>
inline void aabb(const Body<T>& b, const std::vector<V3<T>>& hull, V3<T>& mn, V3<T>& mx) {
    mn = b.position;
    mx = b.position;
    V3<T> sx = world_support(b, V3<T>{T(1), T(0), T(0)}, hull);
    V3<T> snx = world_support(b, -V3<T>{T(1), T(0), T(0)}, hull);
    V3<T> sy = world_support(b, V3<T>{T(0), T(1), T(0)}, hull);
    V3<T> sny = world_support(b, -V3<T>{T(0), T(1), T(0)}, hull);
    V3<T> sz = world_support(b, V3<T>{T(0), T(0), T(1)}, hull);
    V3<T> snz = world_support(b, -V3<T>{T(0), T(0), T(1)}, hull);
    if (sx.x > mx.x) mx.x = sx.x;
    if (snx.x < mn.x) mn.x = snx.x;
    if (sy.y > mx.y) mx.y = sy.y;
    if (sny.y < mn.y) mn.y = sny.y;
    if (sz.z > mx.z) mx.z = sz.z;
    if (snz.z < mn.z) mn.z = snz.z;
}

template <class T>
inline bool aabb_overlap(const V3<T>& mna, const V3<T>& mxa, const V3<T>& mnb, const V3<T>& mxb) {
    // broadphase/src/lib.rs:195-201 (closed intervals)
    return mna.x <= mxb.x && mnb.x <= mxa.x && mna.y <= mxb.y && mnb.y <= mxa.y &&
           mna.z <= mxb.z && mnb.z <= mxa.z;
}

// broad_sweep_dense (broadphase/src/lib.rs:173-216): every (i, j), i < j, with
// overlapping AABBs.  The reference appends in atomic order; the restatement
// returns the same set sorted lexicographically (per-thread lists, in i order).
template <class T>
std::vector<Pair> broad_pairs(const std::vector<Body<T>>& bodies, const std::vector<V3<T>>& hull) {
    const int64_t n = (int64_t)bodies.size();
    std::vector<V3<T>> mn(n), mx(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) aabb(bodies[i], hull, mn[i], mx[i]);
    std::vector<std::vector<Pair>> rows(n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j)
            if (aabb_overlap(mn[i], mx[i], mn[j], mx[j])) rows[i].push_back(Pair{(uint32_t)i, (uint32_t)j});
    std::vector<Pair> out;
    for (auto& r : rows) out.insert(out.end(), r.begin(), r.end());
    return out;
}

// ------------------------------------------------------------- narrow: SAT
// narrowphase/src/cuboid.rs; BOX bodies use per-axis half extents.
template <class T> struct Manifold {
    V3<T> normal{};
    T depth = T(0);
    V3<T> point_a[4]{}, point_b[4]{};
    uint32_t count = 0;
};

template <class T> inline void body_axes(const Body<T>& b, V3<T> ax[3]) {
    ax[0] = qrot(b.orientation, V3<T>{T(1), T(0), T(0)});
    ax[1] = qrot(b.orientation, V3<T>{T(0), T(1), T(0)});
    ax[2] = qrot(b.orientation, V3<T>{T(0), T(0), T(1)});
}

// projected radius of a cube / box on a unit axis (cuboid.rs:37-40)
template <class T> inline T box_radius(const Body<T>& b, const V3<T> ax[3], V3<T> axis) {
    if (b.tag == CUBE)
        return b.corners[0][0] *
               (ofabs(dot(axis, ax[0])) + ofabs(dot(axis, ax[1])) + ofabs(dot(axis, ax[2])));
    return b.corners[0][0] * ofabs(dot(axis, ax[0])) + b.corners[0][1] * ofabs(dot(axis, ax[1])) +
           b.corners[0][2] * ofabs(dot(axis, ax[2]));
}
template <class T> inline T box_half(const Body<T>& b, uint32_t axis) {
    return b.tag == CUBE ? b.corners[0][0] : b.corners[0][axis];
}

template <class T> struct SatBest {
    T overlap;
    V3<T> normal;
    uint32_t kind, index;
};

// cuboid.rs:19-52
template <class T>
inline bool test_axis(V3<T> raw, uint32_t kind, uint32_t index, const Body<T>& a, const V3<T> aa[3],
                      const Body<T>& b, const V3<T> ba[3], V3<T> delta, SatBest<T>& best) {
    if (length_squared(raw) <= T(1.0e-8)) return true;
    V3<T> axis = normalize(raw);
    T ra = box_radius(a, aa, axis);
    T rb = box_radius(b, ba, axis);
    T overlap = ra + rb - ofabs(dot(delta, axis));
    if (overlap < T(0)) return false;
    if (overlap < best.overlap) {
        best.overlap = overlap;
        best.normal = dot(delta, axis) >= T(0) ? axis : -axis;
        best.kind = kind;
        best.index = index;
    }
    return true;
}

// cuboid.rs:54-167
template <class T> inline bool sat(const Body<T>& a, const Body<T>& b, SatBest<T>& best) {
    V3<T> aa[3], ba[3];
    body_axes(a, aa);
    body_axes(b, ba);
    V3<T> delta = b.position - a.position;
    best.overlap = std::numeric_limits<float>::max();
    best.normal = V3<T>{T(1), T(0), T(0)};
    best.kind = 0;
    best.index = 0;
    for (uint32_t i = 0; i < 3; ++i)
        if (!test_axis(aa[i], 0, i, a, aa, b, ba, delta, best)) return false;
    for (uint32_t i = 0; i < 3; ++i)
        if (!test_axis(ba[i], 1, i, a, aa, b, ba, delta, best)) return false;
    for (uint32_t i = 0; i < 3; ++i)
        for (uint32_t j = 0; j < 3; ++j)
            if (!test_axis(cross(aa[i], ba[j]), 2, i * 3 + j, a, aa, b, ba, delta, best)) return false;
    return true;
}

// cuboid.rs:169-196: returns center, u, v, extent_u, extent_v
template <class T>
inline void face_frame(const Body<T>& body, uint32_t axis_index, V3<T> outward, V3<T> verts[4], V3<T>& center,
                       V3<T>& u, V3<T>& v, T& eu, T& ev) {
    V3<T> ax[3];
    body_axes(body, ax);
    V3<T> axis = ax[axis_index];
    T hn = box_half(body, axis_index);
    V3<T> face_normal = axis * (dot(axis, outward) >= T(0) ? T(1) : T(-1));
    center = body.position + face_normal * hn;
    uint32_t iu, iv;
    if (axis_index == 0) { iu = 1; iv = 2; }
    else if (axis_index == 1) { iu = 0; iv = 2; }
    else { iu = 0; iv = 1; }
    u = ax[iu];
    v = ax[iv];
    eu = box_half(body, iu);
    ev = box_half(body, iv);
    verts[0] = center + u * eu + v * ev;
    verts[1] = center - u * eu + v * ev;
    verts[2] = center - u * eu - v * ev;
    verts[3] = center + u * eu - v * ev;
}

// cuboid.rs:198-214
template <class T> inline uint32_t incident_axis(const Body<T>& body, V3<T> n) {
    V3<T> ax[3];
    body_axes(body, ax);
    uint32_t best_axis = 0;
    T best_dot = ofabs(dot(ax[0], n));
    for (uint32_t i = 1; i < 3; ++i) {
        T d = ofabs(dot(ax[i], n));
        if (d > best_dot) { best_dot = d; best_axis = i; }
    }
    return best_axis;
}

// cuboid.rs:216-250
template <class T>
inline uint32_t clip_plane(const V3<T> in[8], uint32_t in_count, V3<T> n, T offset, V3<T> out[8]) {
    if (in_count == 0) return 0;
    uint32_t count = 0;
    V3<T> prev = in[in_count - 1];
    T prev_d = dot(n, prev) - offset;
    for (uint32_t i = 0; i < in_count; ++i) {
        V3<T> cur = in[i];
        T cur_d = dot(n, cur) - offset;
        bool prev_in = prev_d <= T(1.0e-5);
        bool cur_in = cur_d <= T(1.0e-5);
        if (prev_in != cur_in) {
            T den = prev_d - cur_d;
            if (ofabs(den) > T(1.0e-8) && count < 8) {
                T t = rclamp(prev_d / den, T(0), T(1));
                out[count++] = prev + (cur - prev) * t;
            }
        }
        if (cur_in && count < 8) out[count++] = cur;
        prev = cur;
        prev_d = cur_d;
    }
    return count;
}

// cuboid.rs:252-272
template <class T> inline Manifold<T> centered_contact(const Body<T>& a, const Body<T>& b, V3<T> n, T depth) {
    V3<T> aa[3], ba[3];
    body_axes(a, aa);
    body_axes(b, ba);
    T ra = box_radius(a, aa, n);
    T rb = box_radius(b, ba, n);
    Manifold<T> m;
    m.point_a[0] = a.position + n * ra;
    m.point_b[0] = b.position - n * rb;
    m.normal = n;
    m.depth = depth;
    m.count = 1;
    return m;
}

// cuboid.rs:276-377
template <class T> inline bool collide_manifold(const Body<T>& a, const Body<T>& b, Manifold<T>& out) {
    SatBest<T> best;
    if (!sat(a, b, best)) return false;
    V3<T> normal = best.normal;
    T depth = best.overlap;
    if (best.kind == 2) { out = centered_contact(a, b, normal, depth); return true; }
    bool ref_is_a = best.kind == 0;
    V3<T> ref_n = ref_is_a ? normal : -normal;
    const Body<T>& rb = ref_is_a ? a : b;
    const Body<T>& ib = ref_is_a ? b : a;
    uint32_t inc_axis = incident_axis(ib, ref_n);
    V3<T> ref_face[4], center, u, v;
    T eu, ev;
    face_frame(rb, best.index, ref_n, ref_face, center, u, v, eu, ev);
    V3<T> pa[8]{}, pb[8]{}, inc_face[4], c2, u2, v2;
    T eu2, ev2;
    face_frame(ib, inc_axis, -ref_n, inc_face, c2, u2, v2, eu2, ev2);
    for (int i = 0; i < 4; ++i) pa[i] = inc_face[i];
    uint32_t count = 4;
    count = clip_plane(pa, count, u, dot(u, center) + eu, pb);
    count = clip_plane(pb, count, -u, dot(-u, center) + eu, pa);
    count = clip_plane(pa, count, v, dot(v, center) + ev, pb);
    count = clip_plane(pb, count, -v, dot(-v, center) + ev, pa);
    if (count == 0) { out = centered_contact(a, b, normal, depth); return true; }
    T ref_plane = dot(ref_n, center);
    Manifold<T> m;
    uint32_t cc = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (cc >= 4) break;
        V3<T> ip = pa[i];
        T pd = ref_plane - dot(ref_n, ip);
        if (pd < T(-1.0e-5)) continue;
        V3<T> rp = ip + ref_n * rmax(pd, T(0));
        if (ref_is_a) { m.point_a[cc] = rp; m.point_b[cc] = ip; }
        else { m.point_a[cc] = ip; m.point_b[cc] = rp; }
        ++cc;
    }
    if (cc == 0) { out = centered_contact(a, b, normal, depth); return true; }
    m.normal = normal;
    m.depth = depth;
    m.count = cc;
    out = m;
    return true;
}

// ------------------------------------------------------- narrow: GJK / EPA
template <class T> struct Simplex {
    V3<T> points[4]{}, witnesses[4]{};
    uint32_t len = 0;
    void push_front(V3<T> p, V3<T> w) {  // gjk.rs:29-39
        for (uint32_t i = len; i > 0; --i) {
            if (i < 4) { points[i] = points[i - 1]; witnesses[i] = witnesses[i - 1]; }
        }
        points[0] = p;
        witnesses[0] = w;
        len = std::min(len + 1, 4u);
    }
    void remove(uint32_t index) {  // gjk.rs:41-49
        for (uint32_t i = index; i + 1 < len; ++i) { points[i] = points[i + 1]; witnesses[i] = witnesses[i + 1]; }
        len -= 1;
    }
    void swap(uint32_t i, uint32_t j) {
        std::swap(points[i], points[j]);
        std::swap(witnesses[i], witnesses[j]);
    }
};

// gjk.rs:67-73
template <class T>
inline void world_support_witness(const Body<T>& a, const Body<T>& b, V3<T> d, const std::vector<V3<T>>& hull,
                                  V3<T>& point, V3<T>& witness) {
    V3<T> al = qrot(qinv(a.orientation), d);
    V3<T> ap = qrot(a.orientation, support_gpu(a, al, hull)) + a.position;
    V3<T> bl = qrot(qinv(b.orientation), -d);
    V3<T> bp = qrot(b.orientation, support_gpu(b, bl, hull)) + b.position;
    point = ap - bp;
    witness = ap;
}

template <class T> inline bool same_direction(V3<T> a, V3<T> b) { return dot(a, b) > T(0); }
template <class T> inline V3<T> triple_product(V3<T> a, V3<T> b, V3<T> c) { return cross(cross(a, b), c); }
// gjk.rs:214-223
template <class T> inline V3<T> any_perpendicular(V3<T> v) {
    V3<T> axis;
    if (ofabs(v.x) <= ofabs(v.y) && ofabs(v.x) <= ofabs(v.z)) axis = {T(1), T(0), T(0)};
    else if (ofabs(v.y) <= ofabs(v.z)) axis = {T(0), T(1), T(0)};
    else axis = {T(0), T(0), T(1)};
    return normalize_or_zero(cross(axis, v));
}

// gjk.rs:112-201
template <class T> inline bool contains_origin(Simplex<T>& s, V3<T>& dir) {
    if (s.len == 4) {
        V3<T> a = s.points[0], b = s.points[1], c = s.points[2], d = s.points[3];
        V3<T> ao = -a, ab = b - a, ac = c - a, ad = d - a;
        V3<T> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
        if (same_direction(abc, ao)) s.remove(3);
        else if (same_direction(acd, ao)) s.remove(1);
        else if (same_direction(adb, ao)) s.remove(2);
        else return true;
    }
    if (s.len == 1) {
        dir = -s.points[0];
        return false;
    }
    if (s.len == 2) {
        V3<T> a = s.points[0], b = s.points[1];
        V3<T> ab = b - a, ao = -a;
        if (same_direction(ab, ao)) {
            V3<T> d = triple_product(ab, ao, ab);
            if (length_squared(d) < T(1e-12)) d = any_perpendicular(ab);
            dir = d;
        } else {
            s.remove(1);
            dir = ao;
        }
        return false;
    }
    if (s.len == 3) {
        V3<T> a = s.points[0], b = s.points[1], c = s.points[2];
        V3<T> ab = b - a, ac = c - a, ao = -a;
        V3<T> abc = cross(ab, ac);
        if (same_direction(cross(abc, ac), ao)) {
            if (same_direction(ac, ao)) { s.remove(1); dir = triple_product(ac, ao, ac); }
            else if (same_direction(ab, ao)) { s.remove(2); dir = triple_product(ab, ao, ab); }
            else { s.len = 1; dir = ao; }
        } else if (same_direction(cross(ab, abc), ao)) {
            if (same_direction(ab, ao)) { s.remove(2); dir = triple_product(ab, ao, ab); }
            else { s.len = 1; dir = ao; }
        } else {
            if (dot(abc, ao) > T(0)) dir = abc;
            else { s.swap(1, 2); dir = -abc; }
        }
        return false;
    }
    return false;
}

// gjk.rs:77-109
template <class T>
inline bool gjk(const Body<T>& a, const Body<T>& b, const std::vector<V3<T>>& hull, Simplex<T>& out) {
    V3<T> dir = a.position - b.position;
    if (dir == V3<T>{T(0), T(0), T(0)}) dir = {T(1), T(0), T(0)};
    Simplex<T> s;
    V3<T> p, w;
    world_support_witness(a, b, dir, hull, p, w);
    s.push_front(p, w);
    dir = -s.points[0];
    for (int it = 0; it < 64; ++it) {
        world_support_witness(a, b, dir, hull, p, w);
        if (dot(p, dir) < T(0)) return false;
        s.push_front(p, w);
        if (contains_origin(s, dir)) { out = s; return true; }
        if (length_squared(dir) < T(1e-12)) dir = any_perpendicular(s.points[0]);
    }
    return false;
}

template <class T> struct Face { uint32_t a, b, c; V3<T> normal; T distance; };

constexpr int EPA_MAX_VERTICES = 256, EPA_MAX_FACES = 512;

template <class T> struct Polytope {  // epa.rs:33-86
    V3<T> vertices[EPA_MAX_VERTICES]{};
    V3<T> witnesses[EPA_MAX_VERTICES]{};
    Face<T> faces[EPA_MAX_FACES]{};
    uint32_t vertex_count = 0, face_count = 0;
    void add_face(uint32_t a, uint32_t b, uint32_t c) {
        if (face_count >= (uint32_t)EPA_MAX_FACES) return;
        V3<T> va = vertices[a];
        V3<T> n = cross(vertices[b] - va, vertices[c] - va);
        if (length_squared(n) < T(1e-12)) return;
        n = normalize(n);
        Face<T> f{a, b, c, n, dot(n, va)};
        if (f.distance < T(0)) {
            f.a = c;
            f.c = a;
            f.normal = -f.normal;
            f.distance = -f.distance;
        }
        faces[face_count++] = f;
    }
};

template <class T> struct Penetration { V3<T> normal; T depth; V3<T> point_a, point_b; };

// epa.rs:193-240
template <class T> inline Penetration<T> face_contact(const Face<T>& f, const Polytope<T>& p) {
    V3<T> va = p.vertices[f.a], vb = p.vertices[f.b], vc = p.vertices[f.c];
    V3<T> closest = f.normal * f.distance;
    V3<T> v0 = vb - va, v1 = vc - va, v2 = closest - va;
    T d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    T denom = d00 * d11 - d01 * d01;
    T w1, w2, w3;
    if (ofabs(denom) > T(1e-12)) {
        w2 = (d11 * d20 - d01 * d21) / denom;
        w3 = (d00 * d21 - d01 * d20) / denom;
        w1 = T(1) - w2 - w3;
    } else {
        w1 = T(1); w2 = T(0); w3 = T(0);
    }
    w1 = rclamp(w1, T(0), T(1));
    w2 = rclamp(w2, T(0), T(1));
    w3 = rclamp(w3, T(0), T(1));
    T sum = w1 + w2 + w3;
    if (sum > T(1e-12)) { w1 = w1 / sum; w2 = w2 / sum; w3 = w3 / sum; }
    else { w1 = T(1); w2 = T(0); w3 = T(0); }
    V3<T> wa = p.witnesses[f.a], wb = p.witnesses[f.b], wc = p.witnesses[f.c];
    V3<T> pa = wa * w1 + wb * w2 + wc * w3;
    V3<T> mk = va * w1 + vb * w2 + vc * w3;
    return Penetration<T>{f.normal, f.distance, pa, pa - mk};
}

inline bool edge_in_face(uint32_t ea, uint32_t eb, uint32_t fa, uint32_t fb, uint32_t fc) {
    return (fa == ea && fb == eb) || (fb == ea && fc == eb) || (fc == ea && fa == eb) ||
           (fa == eb && fb == ea) || (fb == eb && fc == ea) || (fc == eb && fa == ea);
}

// epa.rs:89-189.  The polytope lives on the heap (the reference keeps it in
// private memory).  The horizon array holds MAX_FACES edges as in epa.rs:145;
// edges past it are dropped instead of indexing out of bounds.
template <class T>
inline bool epa(const Body<T>& a, const Body<T>& b, const std::vector<V3<T>>& hull, const Simplex<T>& s,
                Penetration<T>& out) {
    std::vector<Polytope<T>> holder(1);
    Polytope<T>& poly = holder[0];
    for (uint32_t i = 0; i < s.len; ++i) { poly.vertices[i] = s.points[i]; poly.witnesses[i] = s.witnesses[i]; }
    poly.vertex_count = s.len;
    poly.add_face(0, 1, 2);
    poly.add_face(0, 3, 1);
    poly.add_face(0, 2, 3);
    poly.add_face(1, 3, 2);
    std::vector<uint32_t> visible(EPA_MAX_FACES);
    std::vector<std::pair<uint32_t, uint32_t>> horizon(EPA_MAX_FACES);
    for (int it = 0; it < 64; ++it) {
        uint32_t mi = 0;
        T md = poly.faces[0].distance;
        for (uint32_t i = 1; i < poly.face_count; ++i)
            if (poly.faces[i].distance < md) { md = poly.faces[i].distance; mi = i; }
        Face<T> mf = poly.faces[mi];
        V3<T> sp, sw;
        world_support_witness(a, b, mf.normal, hull, sp, sw);
        T sd = dot(mf.normal, sp);
        if (sd - mf.distance < T(1e-4)) { out = face_contact(poly.faces[mi], poly); return true; }
        uint32_t ni = poly.vertex_count;
        if (ni >= (uint32_t)EPA_MAX_VERTICES) { out = face_contact(poly.faces[mi], poly); return true; }
        poly.vertices[ni] = sp;
        poly.witnesses[ni] = sw;
        poly.vertex_count += 1;
        for (uint32_t i = 0; i < poly.face_count; ++i) {
            const Face<T>& f = poly.faces[i];
            visible[i] = dot(f.normal, sp - poly.vertices[f.a]) > T(0) ? 1u : 0u;
        }
        uint32_t hc = 0;
        for (uint32_t i = 0; i < poly.face_count; ++i) {
            if (!visible[i]) continue;
            const Face<T>& f = poly.faces[i];
            uint32_t e[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
            for (int k = 0; k < 3; ++k) {
                bool shared = false;
                for (uint32_t j = 0; j < poly.face_count; ++j) {
                    if (j == i) continue;
                    const Face<T>& g = poly.faces[j];
                    if (visible[j] && edge_in_face(e[k][0], e[k][1], g.a, g.b, g.c)) { shared = true; break; }
                }
                if (!shared && hc < (uint32_t)EPA_MAX_FACES) horizon[hc++] = {e[k][0], e[k][1]};
            }
        }
        uint32_t w = 0;
        for (uint32_t i = 0; i < poly.face_count; ++i)
            if (!visible[i]) poly.faces[w++] = poly.faces[i];
        poly.face_count = w;
        for (uint32_t i = 0; i < hc; ++i) poly.add_face(ni, horizon[i].first, horizon[i].second);
    }
    return false;
}

// -------------------------------------------------------- narrow per pair
// narrowphase/src/lib.rs:89-158: up to four contacts, written to slots k < count
template <class T>
inline uint32_t narrow_pair(const Body<T>& a, const Body<T>& b, const std::vector<V3<T>>& hull,
                            RawContact<T> out[4]) {
    auto boxlike = [](uint32_t t) { return t == CUBE || t == BOX; };
    bool both_cubes = a.tag == CUBE && b.tag == CUBE;
    bool box_pair = boxlike(a.tag) && boxlike(b.tag);  // BOX extension routes to SAT too
    if (both_cubes || box_pair) {
        Manifold<T> m;
        if (!collide_manifold(a, b, m)) return 0;
        for (uint32_t k = 0; k < m.count; ++k)
            out[k] = RawContact<T>{0, 0, m.normal, m.depth, m.point_a[k], m.point_b[k]};
        return m.count;
    }
    if (a.tag == SPHERE && b.tag == SPHERE) {
        V3<T> delta = b.position - a.position;
        T distance = length(delta);
        T radius = a.corners[0][0] + b.corners[0][0];
        if (distance >= radius || distance <= T(std::numeric_limits<float>::epsilon())) return 0;
        V3<T> n = delta / distance;
        out[0] = RawContact<T>{0, 0, n, radius - distance, a.position + n * a.corners[0][0],
                               b.position - n * b.corners[0][0]};
        return 1;
    }
    Simplex<T> s;
    if (!gjk(a, b, hull, s)) return 0;
    Penetration<T> pen;
    if (!epa(a, b, hull, s, pen)) return 0;
    out[0] = RawContact<T>{0, 0, pen.normal, pen.depth, pen.point_a, pen.point_b};
    return 1;
}

// narrow phase over a pair list; decoded contacts in (pair, slot) order
// (compute.rs:738-750 keeps valid slots in slot order)
template <class T>
std::vector<RawContact<T>> narrow(const std::vector<Body<T>>& bodies, const std::vector<V3<T>>& hull,
                                  const std::vector<Pair>& pairs) {
    const int64_t p = (int64_t)pairs.size();
    std::vector<RawContact<T>> slots(p * 4);
    std::vector<uint32_t> counts(p);
#pragma omp parallel for schedule(dynamic, 32)
    for (int64_t i = 0; i < p; ++i) {
        RawContact<T> out[4];
        uint32_t c = narrow_pair(bodies[pairs[i].a], bodies[pairs[i].b], hull, out);
        counts[i] = c;
        for (uint32_t k = 0; k < c; ++k) {
            out[k].a = pairs[i].a;
            out[k].b = pairs[i].b;
            slots[i * 4 + k] = out[k];
        }
    }
    std::vector<RawContact<T>> res;
    for (int64_t i = 0; i < p; ++i)
        for (uint32_t k = 0; k < counts[i]; ++k) res.push_back(slots[i * 4 + k]);
    return res;
}

// --------------------------------------------------- host contact building
// physics/gpu.rs:406-414
template <class T> inline void tangents(V3<T> n, V3<T>& t1, V3<T>& t2) {
    V3<T> c = ofabs(n.x) > T(0.9) ? cross(n, V3<T>{T(0), T(1), T(0)}) : cross(n, V3<T>{T(1), T(0), T(0)});
    t1 = normalize(c);
    t2 = cross(n, t1);
}

// physics/gpu.rs:239-264
template <class T>
inline Contact<T> raw_contact(uint32_t a, uint32_t b, const Body<T>& ba, const Body<T>& bb, V3<T> normal, T depth,
                              V3<T> pa, V3<T> pb, const Options<T>& o) {
    Contact<T> c;
    V3<T> t1, t2;
    tangents(normal, t1, t2);
    c.a = a;
    c.b = b;
    c.normal = -normal;
    c.tangent1 = t1;
    c.tangent2 = t2;
    c.r_a = qrot(qinv(ba.orientation), pa - ba.position);
    c.r_b = qrot(qinv(bb.orientation), pb - bb.position);
    c.c0 = V3<T>{o.margin - depth, T(0), T(0)};
    c.penalty = V3<T>{o.penalty_min, o.penalty_min, o.penalty_min};
    c.lambda = V3<T>{T(0), T(0), T(0)};
    c.friction = o.friction;
    return c;
}

// physics/gpu.rs:314-350 (BOX bodies use their per-axis half extents)
template <class T> inline void cube_support_face(const Body<T>& body, V3<T> outward, V3<T> pts[4]) {
    V3<T> ax[3];
    body_axes(body, ax);
    uint32_t ai = 0;
    T best = ofabs(dot(ax[0], outward));
    for (uint32_t i = 1; i < 3; ++i) {
        T al = ofabs(dot(ax[i], outward));
        if (al > best) { ai = i; best = al; }
    }
    V3<T> fa = ax[ai];
    V3<T> fn = fa * (dot(fa, outward) >= T(0) ? T(1) : T(-1));
    uint32_t iu, iv;
    if (ai == 0) { iu = 1; iv = 2; }
    else if (ai == 1) { iu = 0; iv = 2; }
    else { iu = 0; iv = 1; }
    T hn = box_half(body, ai), hu = box_half(body, iu), hv = box_half(body, iv);
    V3<T> center = body.position + fn * hn;
    V3<T> u = ax[iu], v = ax[iv];
    pts[0] = center + u * hu + v * hv;
    pts[1] = center - u * hu + v * hv;
    pts[2] = center - u * hu - v * hv;
    pts[3] = center + u * hu - v * hv;
}

// physics/gpu.rs:352-404
template <class T>
inline void container_contacts(const Container<T>& ct, const Body<T>& body, uint32_t index, const Options<T>& o,
                               const std::vector<V3<T>>& hull, std::vector<Contact<T>>& out) {
    const V3<T> axes[3] = {{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}};
    for (int w = 0; w < 6; ++w) {
        T sign = (w % 2 == 0) ? T(1) : T(-1);
        V3<T> outward = qrot(ct.rotation, axes[w / 2] * sign);
        T plane = dot(ct.center, outward) + ct.half_extent;
        V3<T> pts[4];
        uint32_t count;
        if (body.tag == CUBE || body.tag == BOX) {
            cube_support_face(body, outward, pts);
            count = 4;
        } else {
            V3<T> local = qrot(qinv(body.orientation), outward);
            pts[0] = qrot(body.orientation, support_cpu(body, local, hull)) + body.position;
            count = 1;
        }
        V3<T> normal = -outward;
        V3<T> t1, t2;
        tangents(normal, t1, t2);
        for (uint32_t k = 0; k < count; ++k) {
            V3<T> p = pts[k];
            T cur = dot(p, outward) - plane;
            V3<T> pred = p + body.velocity * o.dt;
            T pd = dot(pred, outward) - plane;
            T depth = rmax(cur, pd);
            if (depth > T(0)) {
                Contact<T> c;
                c.a = index;
                c.b = CONTAINER;
                c.normal = normal;
                c.tangent1 = t1;
                c.tangent2 = t2;
                c.r_a = qrot(qinv(body.orientation), p - body.position);
                c.r_b = V3<T>{T(0), T(0), T(0)};
                c.c0 = V3<T>{o.margin - depth, T(0), T(0)};
                c.penalty = V3<T>{o.penalty_min, o.penalty_min, o.penalty_min};
                c.lambda = V3<T>{T(0), T(0), T(0)};
                c.friction = o.friction;
                out.push_back(c);
            }
        }
    }
}

// physics/gpu.rs:266-311: warm start (HashMap: last old contact of a key is
// kept, the first new contact of a key takes it), scaling, CSR in contact order.
template <class T>
inline void finalize_contacts(const std::vector<Contact<T>>& old_contacts, std::vector<Contact<T>>& contacts,
                              uint32_t n, const Options<T>& o, std::vector<uint32_t>& offsets,
                              std::vector<uint32_t>& indices) {
    // A sorted key table stands in for the HashMap; semantics are identical.
    std::vector<std::pair<uint64_t, uint32_t>> keyed;  // key, old index (ascending)
    keyed.reserve(old_contacts.size());
    for (uint32_t i = 0; i < old_contacts.size(); ++i)
        keyed.push_back({((uint64_t)old_contacts[i].a << 32) | old_contacts[i].b, i});
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<uint64_t, uint32_t>& x, const std::pair<uint64_t, uint32_t>& y) {
                         return x.first < y.first;
                     });
    std::vector<char> taken(keyed.size(), 0);
    for (auto& c : contacts) {
        uint64_t key = ((uint64_t)c.a << 32) | c.b;
        auto it = std::upper_bound(keyed.begin(), keyed.end(), key,
                                   [](uint64_t k, const std::pair<uint64_t, uint32_t>& e) { return k < e.first; });
        if (it != keyed.begin()) {
            auto last = it - 1;  // the last old contact with this key
            size_t pos = (size_t)(last - keyed.begin());
            if (last->first == key && !taken[pos]) {
                // every entry of this key is consumed by the first taker
                size_t q = pos;
                while (true) {
                    taken[q] = 1;
                    if (q == 0 || keyed[q - 1].first != key) break;
                    --q;
                }
                c.lambda = old_contacts[last->second].lambda;
                c.penalty = old_contacts[last->second].penalty;
            }
        }
        T ag = o.alpha * o.gamma;
        c.lambda = c.lambda * ag;
        V3<T> p = c.penalty * o.gamma;
        c.penalty = V3<T>{rmin(rmax(p.x, o.penalty_min), o.penalty_max), rmin(rmax(p.y, o.penalty_min), o.penalty_max),
                          rmin(rmax(p.z, o.penalty_min), o.penalty_max)};
    }
    offsets.assign(n + 1, 0);
    for (auto& c : contacts) {
        offsets[c.a + 1] += 1;
        if (c.b != CONTAINER) offsets[c.b + 1] += 1;
    }
    for (uint32_t i = 0; i < n; ++i) offsets[i + 1] += offsets[i];
    indices.assign(offsets[n], 0);
    std::vector<uint32_t> cursor(offsets.begin(), offsets.end());
    for (uint32_t i = 0; i < contacts.size(); ++i) {
        indices[cursor[contacts[i].a]++] = i;
        if (contacts[i].b != CONTAINER) indices[cursor[contacts[i].b]++] = i;
    }
}

// ------------------------------------------------------------------- solve6
// avbd/src/lib.rs:107-180.  lin/ang are lower triangles (00,10,20,11,21,22),
// cross is the full ang x lin block, cross[r][c] = j_ang[r] * j_lin[c] * k.
template <class T>
inline void solve6(const T lin[6], const T ang[6], const T cr[3][3], const T bl[3], const T ba[3], T x[6]) {
    T a11 = lin[0], a21 = lin[1], a31 = lin[2], a22 = lin[3], a32 = lin[4], a33 = lin[5];
    T a41 = cr[0][0], a51 = cr[1][0], a61 = cr[2][0];
    T a42 = cr[0][1], a52 = cr[1][1], a62 = cr[2][1];
    T a43 = cr[0][2], a53 = cr[1][2], a63 = cr[2][2];
    T a44 = ang[0], a54 = ang[1], a64 = ang[2], a55 = ang[3], a65 = ang[4], a66 = ang[5];
    T l21 = a21 / a11, l31 = a31 / a11, l41 = a41 / a11, l51 = a51 / a11, l61 = a61 / a11;
    T d1 = a11;
    T d2 = a22 - l21 * l21 * d1;
    T l32 = (a32 - l21 * l31 * d1) / d2;
    T l42 = (a42 - l21 * l41 * d1) / d2;
    T l52 = (a52 - l21 * l51 * d1) / d2;
    T l62 = (a62 - l21 * l61 * d1) / d2;
    T d3 = a33 - (l31 * l31 * d1 + l32 * l32 * d2);
    T l43 = (a43 - l31 * l41 * d1 - l32 * l42 * d2) / d3;
    T l53 = (a53 - l31 * l51 * d1 - l32 * l52 * d2) / d3;
    T l63 = (a63 - l31 * l61 * d1 - l32 * l62 * d2) / d3;
    T d4 = a44 - (l41 * l41 * d1 + l42 * l42 * d2 + l43 * l43 * d3);
    T l54 = (a54 - l41 * l51 * d1 - l42 * l52 * d2 - l43 * l53 * d3) / d4;
    T l64 = (a64 - l41 * l61 * d1 - l42 * l62 * d2 - l43 * l63 * d3) / d4;
    T d5 = a55 - (l51 * l51 * d1 + l52 * l52 * d2 + l53 * l53 * d3 + l54 * l54 * d4);
    T l65 = (a65 - l51 * l61 * d1 - l52 * l62 * d2 - l53 * l63 * d3 - l54 * l64 * d4) / d5;
    T d6 = a66 - (l61 * l61 * d1 + l62 * l62 * d2 + l63 * l63 * d3 + l64 * l64 * d4 + l65 * l65 * d5);
    T y1 = bl[0];
    T y2 = bl[1] - l21 * y1;
    T y3 = bl[2] - l31 * y1 - l32 * y2;
    T y4 = ba[0] - l41 * y1 - l42 * y2 - l43 * y3;
    T y5 = ba[1] - l51 * y1 - l52 * y2 - l53 * y3 - l54 * y4;
    T y6 = ba[2] - l61 * y1 - l62 * y2 - l63 * y3 - l64 * y4 - l65 * y5;
    T z1 = y1 / d1, z2 = y2 / d2, z3 = y3 / d3, z4 = y4 / d4, z5 = y5 / d5, z6 = y6 / d6;
    T x6 = z6;
    T x5 = z5 - l65 * x6;
    T x4 = z4 - l54 * x5 - l64 * x6;
    T x3 = z3 - l43 * x4 - l53 * x5 - l63 * x6;
    T x2 = z2 - l32 * x3 - l42 * x4 - l52 * x5 - l62 * x6;
    T x1 = z1 - l21 * x2 - l31 * x3 - l41 * x4 - l51 * x5 - l61 * x6;
    x[0] = x1; x[1] = x2; x[2] = x3; x[3] = x4; x[4] = x5; x[5] = x6;
}

// 6x6 block accumulator (Mat3 lin / ang / cross of avbd/src/lib.rs:298-300)
template <class T> struct Block {
    T lin[6] = {}, ang[6] = {}, cr[3][3] = {}, rl[3] = {}, ra[3] = {};
    // add_constraint_row / the contact stamping (avbd/src/lib.rs:182-198, 402-416):
    // lhs += outer(j, j) * k per entry, entry (r, c) = (a[r] * b[c]) * k
    void add_lin(V3<T> j, T k) {
        lin[0] = lin[0] + (j.x * j.x) * k; lin[1] = lin[1] + (j.y * j.x) * k; lin[2] = lin[2] + (j.z * j.x) * k;
        lin[3] = lin[3] + (j.y * j.y) * k; lin[4] = lin[4] + (j.z * j.y) * k; lin[5] = lin[5] + (j.z * j.z) * k;
    }
    void add_ang(V3<T> j, T k) {
        ang[0] = ang[0] + (j.x * j.x) * k; ang[1] = ang[1] + (j.y * j.x) * k; ang[2] = ang[2] + (j.z * j.x) * k;
        ang[3] = ang[3] + (j.y * j.y) * k; ang[4] = ang[4] + (j.z * j.y) * k; ang[5] = ang[5] + (j.z * j.z) * k;
    }
    void add_cross(V3<T> ja, V3<T> jl, T k) {
        const T a[3] = {ja.x, ja.y, ja.z}, l[3] = {jl.x, jl.y, jl.z};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) cr[r][c] = cr[r][c] + (a[r] * l[c]) * k;
    }
    void add_rhs_lin(V3<T> j, T f) { rl[0] = rl[0] + j.x * f; rl[1] = rl[1] + j.y * f; rl[2] = rl[2] + j.z * f; }
    void add_rhs_ang(V3<T> j, T f) { ra[0] = ra[0] + j.x * f; ra[1] = ra[1] + j.y * f; ra[2] = ra[2] + j.z * f; }
};

// ------------------------------------------------------------- AVBD solve
// The kernels of avbd/src/lib.rs run in the schedule of gpu_physics.rs:541-550.
template <class T> struct RunOptions {  // AvbdRunOptions (gpu_physics.rs:76-83)
    T dt;
    V3<T> gravity;
    T alpha, beta_lin, penalty_max;
    uint32_t iterations;
    // the warm start of hard joints' multipliers (AvbdOptions gamma and
    // penalty_min, as the contact warm start uses them: physics/gpu.rs:266-283)
    T gamma = T(0.999), penalty_min = T(1);
    // A fixed penalty schedule (no counterpart in the reference; the map the
    // frozen-penalty gradient differentiates): when pen_in is set, contact c's
    // penalty after the dual update of iteration k is pen_in[(k m + c) 3 + r]
    // (m contacts) instead of the ramped value, a constant of every dual
    // number.  pen_out, when set, receives the penalties after each dual
    // update in the same layout.
    const double* pen_in = nullptr;
    double* pen_out = nullptr;
    const double* lam_in = nullptr;
    double* lam_out = nullptr;
};

// A hard joint's multipliers: lambda and penalty per row (x, y, z for a
// socket; x alone for a rod or rope).
template <class T> struct JointMult {
    V3<T> lam{T(0), T(0), T(0)};
    V3<T> pen{T(0), T(0), T(0)};
};

// The stabilised value of a hard joint's rows, C(x) - alpha C(x0), from the
// anchor separation delta = p_a - p_b at the current poses and delta0 at the
// step's initial poses (a rope, an inequality C <= 0, subtracts only a
// violated start value: C(x) - alpha max(C(x0), 0)); dir is delta / |delta|
// for a rod or rope, whose single row is defined only when |delta| > 1e-6
// (ok).
template <class T> struct HardRows {
    V3<T> chat{T(0), T(0), T(0)};
    V3<T> dir{T(0), T(0), T(0)};
    bool ok = false;
};
template <class T> HardRows<T> hard_rows(uint32_t kind, V3<T> delta, V3<T> delta0, T rest, T alpha) {
    HardRows<T> r;
    if (three_rows(kind)) {
        r.chat = V3<T>{delta.x - alpha * delta0.x, delta.y - alpha * delta0.y, delta.z - alpha * delta0.z};
        r.ok = true;
    } else {
        T len = length(delta);
        T len0 = length(delta0);
        r.ok = len > T(1.0e-6);
        T c0 = len0 - rest;
        if (kind == HARD_ROPE) c0 = rmax(c0, T(0));
        r.chat.x = (len - rest) - alpha * c0;
        if (r.ok) r.dir = delta / len;
    }
    return r;
}
// The row force of a hard joint: penalty * C^ + lambda, kept >= 0 for a rope
// (it only pulls).
template <class T> T hard_force(uint32_t kind, T pen, T chat, T lam) {
    T f = pen * chat + lam;
    return kind == HARD_ROPE ? rmax(f, T(0)) : f;
}
// One dual update of a hard joint's rows (the augmented-Lagrangian step the
// contacts take, avbd/src/lib.rs:565-663): lambda becomes the row force, and
// the penalty of an active row ramps by beta |C^| up to penalty_max.
template <class T> void hard_update(uint32_t kind, const HardRows<T>& r, JointMult<T>& m, T beta, T pmax) {
    if (!r.ok) return;
    const int rows = three_rows(kind) ? 3 : 1;
    T* lam = &m.lam.x;
    T* pen = &m.pen.x;
    const T* ch = &r.chat.x;
    for (int k = 0; k < rows; ++k) {
        T f = hard_force(kind, pen[k], ch[k], lam[k]);
        lam[k] = f;
        if (kind != HARD_ROPE || f > T(0)) pen[k] = rmin(pen[k] + beta * ofabs(ch[k]), pmax);
    }
}
// The warm start of a hard joint at the start of a step (the contacts' rule,
// physics/gpu.rs:277-283): lambda * alpha gamma, penalty * gamma clamped to
// [penalty_min, penalty_max].
template <class T> JointMult<T> hard_warm(const JointMult<T>& s, const RunOptions<T>& pc) {
    JointMult<T> m;
    T ag = pc.alpha * pc.gamma;
    m.lam = s.lam * ag;
    V3<T> p = s.pen * pc.gamma;
    m.pen = V3<T>{rmin(rmax(p.x, pc.penalty_min), pc.penalty_max), rmin(rmax(p.y, pc.penalty_min), pc.penalty_max),
                  rmin(rmax(p.z, pc.penalty_min), pc.penalty_max)};
    return m;
}

template <class T> struct SolveState {
    // body state (GpuAvbdBodyState)
    std::vector<V3<T>> vel, ang, prev_vel, inv_moment;
    std::vector<T> inv_mass;
    std::vector<V3<T>> pos[2];
    std::vector<Q<T>> rot[2];
    std::vector<V3<T>> initial_lin, inertial_lin;
    std::vector<Q<T>> initial_ang, inertial_ang;
};

// Per-stage wall times of the last solve (CPU baseline reporting).
struct SolveTimes { double init = 0, primal = 0, dual = 0, recover = 0; };

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// avbd_init (avbd/src/lib.rs:202-254): initial and inertial poses, then the
// adaptive gravity warm start into parity buffer 0
template <class T> void avbd_init_only(SolveState<T>& S, const RunOptions<T>& pc) {
    const int64_t n = (int64_t)S.inv_mass.size();
    S.initial_lin.assign(n, V3<T>{});
    S.inertial_lin.assign(n, V3<T>{});
    S.initial_ang.assign(n, Q<T>{});
    S.inertial_ang.assign(n, Q<T>{});
    const T dt = pc.dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        V3<T> pos = S.pos[0][i];
        Q<T> rot = S.rot[0][i];
        S.initial_lin[i] = pos;
        S.initial_ang[i] = rot;
        V3<T> ip = pos + S.vel[i] * dt;
        if (S.inv_mass[i] > T(0)) ip = ip + pc.gravity * (dt * dt);
        S.inertial_lin[i] = ip;
        S.inertial_ang[i] = quat_step(rot, S.ang[i] * dt);
        V3<T> wp = pos;
        Q<T> wr = rot;
        if (S.inv_mass[i] > T(0)) {
            V3<T> accel = (S.vel[i] - S.prev_vel[i]) / dt;
            V3<T> g = pc.gravity;
            T gm = length(g);
            T weight = gm > T(0) ? rclamp(dot(accel, g / gm) / gm, T(0), T(1)) : T(0);
            wp = pos + S.vel[i] * dt + g * (weight * dt * dt);
            wr = quat_step(rot, S.ang[i] * dt);
        }
        S.pos[0][i] = wp;
        S.rot[0][i] = wr;
    }
}

// ------------------------------------------------------ graph colouring
// No counterpart in the reference (its solve is Jacobi).  Vertices are the
// dynamic bodies; two are adjacent when they share a contact or a joint.
// Jones-Plassmann rounds with hashed priorities: in round r every uncoloured
// body whose priority beats that of each uncoloured neighbour takes the
// smallest colour no neighbour coloured in an earlier round holds.  The
// decisions of a round read only the state before it, so the colouring does
// not depend on the order of evaluation.  Colours are < kMaxColours (255).
constexpr uint32_t kUncoloured = 0xFFFFFFFFu, kNotInGraph = 0xFFFFFFFEu, kMaxColours = 255;

// 32-bit integer hash (C. Wellons' "lowbias32" finaliser constants)
inline uint32_t colour_priority(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
inline bool colour_beats(uint32_t j, uint32_t i) {
    const uint32_t pj = colour_priority(j), pi = colour_priority(i);
    return pj > pi || (pj == pi && j > i);
}

struct ColourPlan {
    std::vector<uint32_t> colour;  // per body (kNotInGraph for static bodies)
    std::vector<uint32_t> order;   // dynamic bodies by colour, ascending index within a colour
    std::vector<uint32_t> start;   // colours + 1 offsets into order
    uint32_t rounds = 0;
    bool overflow = false;         // a body found all kMaxColours colours taken
};

// neighbours(i, f): calls f(j) for every dynamic neighbour j of dynamic body i
template <class Neighbours>
ColourPlan colour_graph(uint32_t n, const std::vector<bool>& dynamic, Neighbours&& neighbours) {
    ColourPlan P;
    P.colour.assign(n, kNotInGraph);
    std::vector<uint32_t> round_of(n, kNotInGraph);
    uint32_t left = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (dynamic[i]) { round_of[i] = kUncoloured; ++left; }
    std::vector<std::pair<uint32_t, uint32_t>> taken;
    for (uint32_t r = 0; left > 0; ++r) {
        taken.clear();
        for (uint32_t i = 0; i < n; ++i) {
            if (round_of[i] != kUncoloured) continue;
            uint32_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0x80000000u};  // colour 255 is never free
            bool best = true;
            neighbours(i, [&](uint32_t j) {
                if (!best) return;
                if (round_of[j] < r) mask[P.colour[j] >> 5] |= 1u << (P.colour[j] & 31);
                else if (colour_beats(j, i)) best = false;
            });
            if (!best) continue;
            uint32_t c = kMaxColours - 1;
            bool found = false;
            for (uint32_t w = 0; w < 8 && !found; ++w)
                if (mask[w] != 0xFFFFFFFFu) {
                    uint32_t b = 0;
                    while (mask[w] & (1u << b)) ++b;
                    c = 32 * w + b;
                    found = true;
                }
            if (!found) P.overflow = true;
            taken.push_back({i, c});
        }
        for (const auto& t : taken) {
            P.colour[t.first] = t.second;
            round_of[t.first] = r;
        }
        left -= (uint32_t)taken.size();
        P.rounds = r + 1;
    }
    uint32_t ncol = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (dynamic[i]) ncol = std::max(ncol, P.colour[i] + 1);
    P.start.assign(ncol + 1, 0);
    for (uint32_t i = 0; i < n; ++i)
        if (dynamic[i]) P.start[P.colour[i] + 1] += 1;
    for (uint32_t c = 0; c < ncol; ++c) P.start[c + 1] += P.start[c];
    P.order.assign(P.start[ncol], 0);
    std::vector<uint32_t> fill(P.start.begin(), P.start.end() - 1);
    for (uint32_t i = 0; i < n; ++i)
        if (dynamic[i]) P.order[fill[P.colour[i]]++] = i;
    return P;
}

// The solve's graph: contacts through the CSR, joints through their bodies.
template <class T>
ColourPlan colour_bodies(const std::vector<T>& inv_mass, const std::vector<Contact<T>>& contacts,
                         const std::vector<Constraint<T>>& cons, const std::vector<uint32_t>& offsets,
                         const std::vector<uint32_t>& indices) {
    const uint32_t n = (uint32_t)inv_mass.size();
    std::vector<bool> dyn(n);
    for (uint32_t i = 0; i < n; ++i) dyn[i] = inv_mass[i] != T(0);
    std::vector<std::vector<uint32_t>> jn(n);
    for (const auto& c : cons)
        if (c.a < n && c.b < n && c.a != c.b) {
            jn[c.a].push_back(c.b);
            jn[c.b].push_back(c.a);
        }
    return colour_graph(n, dyn, [&](uint32_t i, auto&& f) {
        for (uint32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
            const Contact<T>& ct = contacts[indices[e]];
            const uint32_t o = ct.a == i ? ct.b : ct.a;
            if (o != CONTAINER && o != i && dyn[o]) f(o);
        }
        for (uint32_t o : jn[i])
            if (dyn[o]) f(o);
    });
}

// gpu_physics.rs:439-622 + avbd/src/lib.rs:202-699.  Inputs: per-body state,
// poses, contacts (λ / penalty updated in place), joints, CSR.  Outputs: the
// pose of the final parity buffer and the recovered velocities.
//
// With a colour plan the primal runs as coloured Gauss-Seidel instead (no
// counterpart in the reference): one pose buffer, updated in place, colour
// after colour; a body reads the poses its neighbours of earlier colours
// wrote in the same iteration.  Bodies of one colour share no contact or
// joint, so within a colour the order does not matter.  Each iteration's
// dual update follows its last colour, and recover reads the same buffer.
template <class T>
void avbd_solve(SolveState<T>& S, std::vector<Contact<T>>& contacts, const std::vector<Constraint<T>>& cons,
                const std::vector<uint32_t>& offsets, const std::vector<uint32_t>& indices, const RunOptions<T>& pc,
                std::vector<V3<T>>& out_pos, std::vector<Q<T>>& out_rot, SolveTimes* times = nullptr,
                const ColourPlan* plan = nullptr, std::vector<JointMult<T>>* jstate = nullptr) {
    const int64_t n = (int64_t)S.inv_mass.size();
    const int64_t m = (int64_t)contacts.size();
    out_pos.assign(n, V3<T>{});
    out_rot.assign(n, Q<T>{});
    if (n == 0) return;
    // hard joints: multipliers warm-started from the state the last step
    // left, updated after every primal, and stored back at the end
    std::vector<uint32_t> hard;
    for (uint32_t c = 0; c < cons.size(); ++c)
        if (is_hard(cons[c].kind)) hard.push_back(c);
    if (!hard.empty() && (!jstate || jstate->size() != cons.size()))
        throw std::invalid_argument("hard joints need their multiplier state");
    std::vector<JointMult<T>> jm(cons.size());
    for (uint32_t c : hard) jm[c] = hard_warm((*jstate)[c], pc);

    const T dt = pc.dt;
    double t0 = now_s();
    avbd_init_only(S, pc);
    // the anchor separation of each hard joint at the step's initial poses
    std::vector<V3<T>> delta0(cons.size(), V3<T>{T(0), T(0), T(0)});
    for (uint32_t ci : hard) {
        const Constraint<T>& c = cons[ci];
        if (is_bend(c.kind))
            delta0[ci] = bend_error(S.initial_ang[c.a], Q<T>{c.anchor_a.x, c.anchor_a.y, c.anchor_a.z, c.length},
                                    S.initial_ang[c.b]);
        else
            delta0[ci] = (S.initial_lin[c.a] + qrot(S.initial_ang[c.a], c.anchor_a)) -
                         (S.initial_lin[c.b] + qrot(S.initial_ang[c.b], c.anchor_b));
    }
    double t1 = now_s();
    if (times) times->init += t1 - t0;
    // per-body joint lists in ascending constraint order (equivalent to the
    // reference's full scan of avbd/src/lib.rs:422-539, which visits each
    // constraint once per body in index order)
    std::vector<std::vector<uint32_t>> joints(n);
    for (uint32_t c = 0; c < cons.size(); ++c) {
        if (cons[c].a < (uint32_t)n) joints[cons[c].a].push_back(c);
        if (cons[c].b != cons[c].a && cons[c].b < (uint32_t)n) joints[cons[c].b].push_back(c);
    }
    const V3<T> zero{T(0), T(0), T(0)};
    for (uint32_t k = 0; k < pc.iterations; ++k) {
        const uint32_t parity = plan ? 0u : k % 2;
        std::vector<V3<T>>& rp = S.pos[parity];
        std::vector<Q<T>>& rr = S.rot[parity];
        std::vector<V3<T>>& wpv = S.pos[plan ? 0u : 1 - parity];
        std::vector<Q<T>>& wrv = S.rot[plan ? 0u : 1 - parity];
        double ta = now_s();
        // avbd_primal (avbd/src/lib.rs:258-551)
        auto primal_body = [&](int64_t i) {
            if (S.inv_mass[i] == T(0)) return;
            V3<T> snap_pos = rp[i];
            Q<T> snap_rot = rr[i];
            T mass = T(1) / S.inv_mass[i];
            V3<T> moment{T(1) / S.inv_moment[i].x, T(1) / S.inv_moment[i].y, T(1) / S.inv_moment[i].z};
            V3<T> m_dt2{mass / (dt * dt), mass / (dt * dt), mass / (dt * dt)};
            V3<T> i_dt2 = moment / (dt * dt);
            Block<T> B;
            B.lin[0] = m_dt2.x; B.lin[3] = m_dt2.y; B.lin[5] = m_dt2.z;
            B.ang[0] = i_dt2.x; B.ang[3] = i_dt2.y; B.ang[5] = i_dt2.z;
            V3<T> rl = mulv(m_dt2, snap_pos - S.inertial_lin[i]);
            V3<T> ra = mulv(i_dt2, ang_diff(snap_rot, S.inertial_ang[i]));
            B.rl[0] = rl.x; B.rl[1] = rl.y; B.rl[2] = rl.z;
            B.ra[0] = ra.x; B.ra[1] = ra.y; B.ra[2] = ra.z;
            Q<T> initial_ang = S.initial_ang[i];
            V3<T> initial_lin = S.initial_lin[i];
            for (uint32_t j = offsets[i]; j < offsets[i + 1]; ++j) {
                const Contact<T>& ct = contacts[indices[j]];
                bool is_a = ct.a == (uint32_t)i;
                uint32_t other = is_a ? ct.b : ct.a;
                V3<T> n0 = ct.normal, t1 = ct.tangent1, t2 = ct.tangent2;
                V3<T> r0 = is_a ? n0 : -n0, r1 = is_a ? t1 : -t1, r2 = is_a ? t2 : -t2;
                V3<T> r_world = is_a ? qrot(snap_rot, ct.r_a) : qrot(snap_rot, ct.r_b);
                V3<T> dsl = snap_pos - initial_lin;
                V3<T> dsa = ang_diff(snap_rot, initial_ang);
                V3<T> dol = zero, doa = zero, r_other = zero;
                if (other != CONTAINER) {
                    V3<T> op = rp[other];
                    Q<T> orr = rr[other];
                    dol = op - S.initial_lin[other];
                    doa = ang_diff(orr, S.initial_ang[other]);
                    r_other = qrot(orr, is_a ? ct.r_b : ct.r_a);
                }
                V3<T> j0 = cross(r_world, r0), j1 = cross(r_world, r1), j2 = cross(r_world, r2);
                T kk = T(1) - pc.alpha;
                T c0 = ct.c0.x * kk, c1 = ct.c0.y * kk, c2 = ct.c0.z * kk;
                c0 = c0 + (dot(r0, dsl) + dot(j0, dsa) - dot(r0, dol) - dot(cross(r_other, r0), doa));
                c1 = c1 + (dot(r1, dsl) + dot(j1, dsa) - dot(r1, dol) - dot(cross(r_other, r1), doa));
                c2 = c2 + (dot(r2, dsl) + dot(j2, dsa) - dot(r2, dol) - dot(cross(r_other, r2), doa));
                T f0 = rmin(ct.penalty.x * c0 + ct.lambda.x, T(0));
                T f1r = ct.penalty.y * c1 + ct.lambda.y;
                T f2r = ct.penalty.z * c2 + ct.lambda.z;
                T bounds = ofabs(f0) * ct.friction;
                T tm = osqrt(f1r * f1r + f2r * f2r);
                T f1 = f1r, f2 = f2r;
                if (tm > bounds && tm > T(0)) {
                    T s = bounds / tm;
                    f1 = f1r * s;
                    f2 = f2r * s;
                }
                T k0 = ct.penalty.x, k1 = ct.penalty.y, k2 = ct.penalty.z;
                B.add_lin(r0, k0); B.add_lin(r1, k1); B.add_lin(r2, k2);
                B.add_ang(j0, k0); B.add_ang(j1, k1); B.add_ang(j2, k2);
                B.add_cross(j0, r0, k0); B.add_cross(j1, r1, k1); B.add_cross(j2, r2, k2);
                B.add_rhs_lin(r0, f0); B.add_rhs_lin(r1, f1); B.add_rhs_lin(r2, f2);
                B.add_rhs_ang(j0, f0); B.add_rhs_ang(j1, f1); B.add_rhs_ang(j2, f2);
            }
            for (uint32_t ci : joints[i]) {
                const Constraint<T>& c = cons[ci];
                bool is_a = c.a == (uint32_t)i;
                V3<T> aa = c.anchor_a, ab = c.anchor_b;
                V3<T> pa = is_a ? snap_pos : rp[c.a];
                V3<T> pb = !is_a ? snap_pos : rp[c.b];
                Q<T> qa = is_a ? snap_rot : rr[c.a];
                Q<T> qb = !is_a ? snap_rot : rr[c.b];
                V3<T> p_a = pa + qrot(qa, aa);
                V3<T> p_b = pb + qrot(qb, ab);
                V3<T> self_anchor = is_a ? aa : ab;
                T sign = is_a ? T(1) : T(-1);
                V3<T> delta = p_a - p_b;
                V3<T> r_self = qrot(snap_rot, self_anchor);
                T stiff = c.stiffness;
                if (is_bend(c.kind)) {
                    // three angular rows: j_lin = 0, j_ang = +-e_r
                    const V3<T> err = bend_error(qa, Q<T>{c.anchor_a.x, c.anchor_a.y, c.anchor_a.z, c.length}, qb);
                    const V3<T> zl{T(0), T(0), T(0)};
                    const V3<T> js[3] = {V3<T>{T(1), T(0), T(0)} * sign, V3<T>{T(0), T(1), T(0)} * sign,
                                         V3<T>{T(0), T(0), T(1)} * sign};
                    const T* e = &err.x;
                    T fs[3], ks[3];
                    if (c.kind == HARD_BEND) {
                        const HardRows<T> hr = hard_rows(c.kind, err, delta0[ci], c.length, pc.alpha);
                        const JointMult<T>& mu = jm[ci];
                        for (int r = 0; r < 3; ++r) {
                            fs[r] = hard_force(c.kind, (&mu.pen.x)[r], (&hr.chat.x)[r], (&mu.lam.x)[r]);
                            ks[r] = (&mu.pen.x)[r];
                        }
                    } else {
                        for (int r = 0; r < 3; ++r) {
                            fs[r] = e[r] * stiff;
                            ks[r] = stiff;
                        }
                    }
                    for (int r = 0; r < 3; ++r) {
                        B.add_lin(zl, ks[r]);
                        B.add_ang(js[r], ks[r]);
                        B.add_cross(js[r], zl, ks[r]);
                        B.add_rhs_lin(zl, fs[r]);
                        B.add_rhs_ang(js[r], fs[r]);
                    }
                } else if (is_hard(c.kind)) {
                    const HardRows<T> hr = hard_rows(c.kind, delta, delta0[ci], c.length, pc.alpha);
                    const JointMult<T>& mu = jm[ci];
                    if (c.kind == HARD_SOCKET) {
                        const V3<T> js[3] = {V3<T>{T(1), T(0), T(0)} * sign, V3<T>{T(0), T(1), T(0)} * sign,
                                             V3<T>{T(0), T(0), T(1)} * sign};
                        const T* ch = &hr.chat.x;
                        const T* pn = &mu.pen.x;
                        const T* lm = &mu.lam.x;
                        for (int r = 0; r < 3; ++r) {
                            V3<T> ja = cross(r_self, js[r]);
                            T f = hard_force(c.kind, pn[r], ch[r], lm[r]);
                            B.add_lin(js[r], pn[r]);
                            B.add_ang(ja, pn[r]);
                            B.add_cross(ja, js[r], pn[r]);
                            B.add_rhs_lin(js[r], f);
                            B.add_rhs_ang(ja, f);
                        }
                    } else if (hr.ok) {
                        // a rope whose row force is zero (slack) adds nothing,
                        // as the soft rope is inactive within its length
                        T f = hard_force(c.kind, mu.pen.x, hr.chat.x, mu.lam.x);
                        if (c.kind != HARD_ROPE || f > T(0)) {
                            V3<T> jl = hr.dir * sign;
                            V3<T> ja = cross(r_self, jl);
                            B.add_lin(jl, mu.pen.x);
                            B.add_ang(ja, mu.pen.x);
                            B.add_cross(ja, jl, mu.pen.x);
                            B.add_rhs_lin(jl, f);
                            B.add_rhs_ang(ja, f);
                        }
                    }
                } else if (c.kind == BALL_SOCKET) {
                    V3<T> jx = V3<T>{T(1), T(0), T(0)} * sign;
                    V3<T> jy = V3<T>{T(0), T(1), T(0)} * sign;
                    V3<T> jz = V3<T>{T(0), T(0), T(1)} * sign;
                    const V3<T> js[3] = {jx, jy, jz};
                    const T fs[3] = {delta.x * stiff, delta.y * stiff, delta.z * stiff};
                    for (int r = 0; r < 3; ++r) {
                        V3<T> ja = cross(r_self, js[r]);
                        B.add_lin(js[r], stiff);
                        B.add_ang(ja, stiff);
                        B.add_cross(ja, js[r], stiff);
                        B.add_rhs_lin(js[r], fs[r]);
                        B.add_rhs_ang(ja, fs[r]);
                    }
                } else {
                    T len = length(delta);
                    bool active = len > T(1.0e-6) && (c.kind != ROPE || len > c.length);
                    if (active) {
                        V3<T> nrm = delta / len;
                        V3<T> jl = nrm * sign;
                        V3<T> ja = cross(r_self, jl);
                        T force = stiff * (len - c.length);
                        B.add_lin(jl, stiff);
                        B.add_ang(ja, stiff);
                        B.add_cross(ja, jl, stiff);
                        B.add_rhs_lin(jl, force);
                        B.add_rhs_ang(ja, force);
                    }
                }
            }
            T bl[3] = {-B.rl[0], -B.rl[1], -B.rl[2]}, ba[3] = {-B.ra[0], -B.ra[1], -B.ra[2]};
            T x[6];
            solve6(B.lin, B.ang, B.cr, bl, ba, x);
            wpv[i] = snap_pos + V3<T>{x[0], x[1], x[2]};
            wrv[i] = quat_step(snap_rot, V3<T>{x[3], x[4], x[5]});
        };
        if (plan) {
            for (size_t c = 0; c + 1 < plan->start.size(); ++c) {
                const int64_t b0 = plan->start[c], b1 = plan->start[c + 1];
#pragma omp parallel for schedule(dynamic, 64)
                for (int64_t t = b0; t < b1; ++t) primal_body((int64_t)plan->order[t]);
            }
        } else {
#pragma omp parallel for schedule(dynamic, 64)
            for (int64_t i = 0; i < n; ++i) primal_body(i);
        }
        double tb = now_s();
        // avbd_dual (avbd/src/lib.rs:565-663): reads the just-written buffer
#pragma omp parallel for schedule(static)
        for (int64_t ci = 0; ci < m; ++ci) {
            Contact<T>& ct = contacts[ci];
            V3<T> pos_a = wpv[ct.a];
            Q<T> rot_a = wrv[ct.a];
            V3<T> dal = pos_a - S.initial_lin[ct.a];
            V3<T> daa = ang_diff(rot_a, S.initial_ang[ct.a]);
            V3<T> dbl = zero, dba = zero, rbw = zero;
            if (ct.b != CONTAINER) {
                V3<T> pos_b = wpv[ct.b];
                Q<T> rot_b = wrv[ct.b];
                dbl = pos_b - S.initial_lin[ct.b];
                dba = ang_diff(rot_b, S.initial_ang[ct.b]);
                rbw = qrot(rot_b, ct.r_b);
            }
            V3<T> raw = qrot(rot_a, ct.r_a);
            V3<T> r0 = ct.normal, r1 = ct.tangent1, r2 = ct.tangent2;
            T kk = T(1) - pc.alpha;
            T c0 = ct.c0.x * kk, c1 = ct.c0.y * kk, c2 = ct.c0.z * kk;
            c0 = c0 + (dot(r0, dal) + dot(cross(raw, r0), daa) - dot(r0, dbl) - dot(cross(rbw, r0), dba));
            c1 = c1 + (dot(r1, dal) + dot(cross(raw, r1), daa) - dot(r1, dbl) - dot(cross(rbw, r1), dba));
            c2 = c2 + (dot(r2, dal) + dot(cross(raw, r2), daa) - dot(r2, dbl) - dot(cross(rbw, r2), dba));
            T f0 = rmin(ct.penalty.x * c0 + ct.lambda.x, T(0));
            T f1r = ct.penalty.y * c1 + ct.lambda.y;
            T f2r = ct.penalty.z * c2 + ct.lambda.z;
            T bounds = ofabs(f0) * ct.friction;
            T tm = osqrt(f1r * f1r + f2r * f2r);
            T f1 = f1r, f2 = f2r;
            if (tm > bounds && tm > T(0)) {
                T s = bounds / tm;
                f1 = f1r * s;
                f2 = f2r * s;
            }
            ct.lambda = V3<T>{f0, f1, f2};
            if (pc.lam_in) {
                const double* fixed = pc.lam_in + ((size_t)k * (size_t)m + (size_t)ci) * 3;
                ct.lambda = V3<T>{T(fixed[0]), T(fixed[1]), T(fixed[2])};
            }
            if (pc.lam_out) {
                double* rec = pc.lam_out + ((size_t)k * (size_t)m + (size_t)ci) * 3;
                rec[0] = (double)ct.lambda.x; rec[1] = (double)ct.lambda.y; rec[2] = (double)ct.lambda.z;
            }
            if (pc.pen_in) {
                const double* fixed = pc.pen_in + ((size_t)k * (size_t)m + (size_t)ci) * 3;
                ct.penalty = V3<T>{T(fixed[0]), T(fixed[1]), T(fixed[2])};
            } else {
                if (f0 < T(0)) ct.penalty.x = rmin(ct.penalty.x + pc.beta_lin * ofabs(c0), pc.penalty_max);
                if (tm <= bounds) {
                    ct.penalty.y = rmin(ct.penalty.y + pc.beta_lin * ofabs(c1), pc.penalty_max);
                    ct.penalty.z = rmin(ct.penalty.z + pc.beta_lin * ofabs(c2), pc.penalty_max);
                }
            }
            if (pc.pen_out) {
                double* rec = pc.pen_out + ((size_t)k * (size_t)m + (size_t)ci) * 3;
                rec[0] = (double)ct.penalty.x; rec[1] = (double)ct.penalty.y; rec[2] = (double)ct.penalty.z;
            }
        }
        // the hard joints' dual update, from the same poses
        for (uint32_t ci : hard) {
            const Constraint<T>& c = cons[ci];
            V3<T> delta = is_bend(c.kind)
                              ? bend_error(wrv[c.a], Q<T>{c.anchor_a.x, c.anchor_a.y, c.anchor_a.z, c.length}, wrv[c.b])
                              : (wpv[c.a] + qrot(wrv[c.a], c.anchor_a)) - (wpv[c.b] + qrot(wrv[c.b], c.anchor_b));
            hard_update(c.kind, hard_rows(c.kind, delta, delta0[ci], c.length, pc.alpha), jm[ci], pc.beta_lin,
                        pc.penalty_max);
        }
        double tc = now_s();
        if (times) { times->primal += tb - ta; times->dual += tc - tb; }
    }
    for (uint32_t c : hard) (*jstate)[c] = jm[c];
    double td = now_s();
    // avbd_recover (avbd/src/lib.rs:667-699) with parity (I-1)%2: reads the
    // buffer the last primal wrote
    const uint32_t fp = (pc.iterations - 1) % 2;
    std::vector<V3<T>>& fpos = S.pos[plan ? 0u : 1 - fp];
    std::vector<Q<T>>& frot = S.rot[plan ? 0u : 1 - fp];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        if (S.inv_mass[i] != T(0)) {
            V3<T> pos = fpos[i];
            Q<T> rot = frot[i];
            S.prev_vel[i] = S.vel[i];
            S.vel[i] = (pos - S.initial_lin[i]) / dt;
            Q<T> rel = qmul(rot, qinv(S.initial_ang[i]));
            S.ang[i] = V3<T>{T(2) * rel.x / dt, T(2) * rel.y / dt, T(2) * rel.z / dt};
        }
        out_pos[i] = fpos[i];
        out_rot[i] = frot[i];
    }
    if (times) times->recover += now_s() - td;
}

// ------------------------------------------------------------ the world
struct StepTimes { double broad = 0, narrow = 0, contacts = 0, avbd = 0, total = 0; SolveTimes solve; };

template <class T> struct World {
    std::vector<Body<T>> bodies;
    bool has_container = false;
    Container<T> container;
    std::vector<Constraint<T>> constraints;
    std::vector<V3<T>> hull;
    std::vector<Contact<T>> contacts;  // AvbdSolver::contacts with solved λ / penalty
    std::vector<uint32_t> offsets, indices;
    std::vector<RawContact<T>> pending;  // narrow output awaiting the next step
    std::vector<Pair> last_pairs;
    bool broad_primed = false;
    StepTimes times;
    // with keep_presolve, the contacts as the last solve received them (warm
    // λ / penalty, before the dual updates): the frozen contact set of a replay
    bool keep_presolve = false;
    std::vector<Contact<T>> presolve;
    // 0: Jacobi (the reference); 1: coloured Gauss-Seidel (colour_bodies)
    int mode = 0;
    ColourPlan plan;
    // hard joints' multipliers as the last step left them (zero at the start)
    std::vector<JointMult<T>> joint_state;

    // prepare_contacts (physics/gpu.rs:188-197)
    void prepare_contacts(std::vector<Contact<T>> body_contacts, const Options<T>& o) {
        if (has_container)
            for (uint32_t i = 0; i < bodies.size(); ++i)
                if (!bodies[i].is_static()) container_contacts(container, bodies[i], i, o, hull, body_contacts);
        std::vector<Contact<T>> old = std::move(contacts);
        finalize_contacts(old, body_contacts, (uint32_t)bodies.size(), o, offsets, indices);
        contacts = std::move(body_contacts);
    }

    std::vector<Contact<T>> raw_contacts(const std::vector<RawContact<T>>& raw, const Options<T>& o) const {
        std::vector<Contact<T>> out;
        out.reserve(raw.size());
        for (const auto& c : raw)
            out.push_back(raw_contact(c.a, c.b, bodies[c.a], bodies[c.b], c.normal, c.depth, c.point_a, c.point_b, o));
        return out;
    }

    // One demo frame (demos/avbd_box/src/main.rs:189-272).  With lag 1 the
    // body contacts were found on the previous state; lag 0 finds them on the
    // current one.  (The reference always runs lag 1; lag 0 and switching
    // between the two are extensions.)
    void step(const Options<T>& o, int lag = 1) {
        double t0 = now_s();
        double tb = 0, tn = 0, tc = 0;
        if (lag == 1) {
            double a = now_s();
            std::vector<Contact<T>> bc = raw_contacts(broad_primed ? pending : std::vector<RawContact<T>>{}, o);
            prepare_contacts(std::move(bc), o);
            double b = now_s();
            last_pairs = broad_pairs(bodies, hull);
            double c = now_s();
            pending = narrow(bodies, hull, last_pairs);
            double d = now_s();
            broad_primed = true;
            tc += b - a; tb += c - b; tn += d - c;
        } else {
            double a = now_s();
            last_pairs = broad_pairs(bodies, hull);
            double b = now_s();
            std::vector<RawContact<T>> raw = narrow(bodies, hull, last_pairs);
            double c = now_s();
            prepare_contacts(raw_contacts(raw, o), o);
            // a lag-1 step after this one builds its contacts from this
            // narrow output, as if this step had run with lag 1
            pending = std::move(raw);
            broad_primed = true;
            double d = now_s();
            tb += b - a; tn += c - b; tc += d - c;
        }
        double e = now_s();
        solve(o);
        double f = now_s();
        times.broad += tb;
        times.narrow += tn;
        times.contacts += tc;
        times.avbd += f - e;
        times.total += f - t0;
    }

    // state_from_bodies + submit + read + sync_bodies_from_gpu
    void solve(const Options<T>& o) {
        const size_t n = bodies.size();
        if (n == 0) return;
        SolveState<T> S;
        S.vel.resize(n); S.ang.resize(n); S.prev_vel.resize(n); S.inv_moment.resize(n); S.inv_mass.resize(n);
        S.pos[0].resize(n); S.rot[0].resize(n);
        for (size_t i = 0; i < n; ++i) {
            const Body<T>& b = bodies[i];
            S.vel[i] = b.velocity; S.ang[i] = b.angular_velocity; S.prev_vel[i] = b.prev_velocity;
            S.inv_moment[i] = b.inv_moment; S.inv_mass[i] = b.inv_mass;
            S.pos[0][i] = b.position; S.rot[0][i] = b.orientation;
        }
        S.pos[1] = S.pos[0];
        S.rot[1] = S.rot[0];
        RunOptions<T> pc{o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, o.iterations, o.gamma, o.penalty_min};
        std::vector<V3<T>> op;
        std::vector<Q<T>> orr;
        if (keep_presolve) presolve = contacts;
        if (mode == 1) plan = colour_bodies(S.inv_mass, contacts, constraints, offsets, indices);
        joint_state.resize(constraints.size());
        avbd_solve(S, contacts, constraints, offsets, indices, pc, op, orr, &times.solve, mode == 1 ? &plan : nullptr,
                   &joint_state);
        for (size_t i = 0; i < n; ++i) {
            Body<T>& b = bodies[i];
            if (b.is_static()) continue;
            b.position = op[i]; b.orientation = orr[i];
            b.velocity = S.vel[i]; b.angular_velocity = S.ang[i]; b.prev_velocity = S.prev_vel[i];
        }
    }
};

// ------------------------------------------------- frozen-contact replay (AD)
// The differentiable step of the adjoint: each step's contact set (geometry
// and warm λ / penalty, as the forward's solve received it) and CSR are
// replayed instead of being found again; only the AVBD solve runs, with the
// contacts' friction replaced by the parameter μ.  Instantiated with
// Dual<5>, it gives forward-mode derivatives (the FP64 oracle gradient);
// with Dual<5, float> the values are the float replay's, bit for bit, and
// the derivatives (in double) are those of the piece the float run lies on.

// Forward-mode dual number: value (type V) and N directional derivatives
// (double).  Branches (min / max / clamps, the friction cone) follow the
// value, so the result is the derivative of the piece the value lies on.
template <int N, class V = double> struct Dual {
    V v = V(0);
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(V(x)) {}  // NOLINT: constants convert implicitly
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    friend Dual operator+(const Dual& a, const Dual& b) {
        Dual r(a.v + b.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
        return r;
    }
    friend Dual operator-(const Dual& a, const Dual& b) {
        Dual r(a.v - b.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
        return r;
    }
    friend Dual operator-(const Dual& a) {
        Dual r(-a.v);
        for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
        return r;
    }
    friend Dual operator*(const Dual& a, const Dual& b) {
        Dual r(a.v * b.v);
        for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
        return r;
    }
    friend Dual operator/(const Dual& a, const Dual& b) {
        Dual r(a.v / b.v);
        for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
        return r;
    }
    friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
    friend Dual sqrt(const Dual& a) {
        Dual r;
        r.v = std::sqrt(a.v);
        for (int k = 0; k < N; ++k) r.d[k] = r.v > V(0) ? a.d[k] / (2.0 * (double)r.v) : 0.0;
        return r;
    }
    friend Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
};

// One AVBD solve of `bodies` on a given contact set and CSR (World::solve
// without joints); contacts' λ / penalty are updated in place.  pen_in /
// lam_in: fixed penalty / multiplier schedules to read, pen_out / lam_out:
// the schedules to write (RunOptions), I m 3 values each.
template <class T>
void solve_bodies(std::vector<Body<T>>& bodies, std::vector<Contact<T>>& contacts,
                  const std::vector<uint32_t>& offsets, const std::vector<uint32_t>& indices, const Options<T>& o,
                  const double* pen_in = nullptr, double* pen_out = nullptr, const double* lam_in = nullptr,
                  double* lam_out = nullptr) {
    const size_t n = bodies.size();
    if (n == 0) return;
    SolveState<T> S;
    S.vel.resize(n); S.ang.resize(n); S.prev_vel.resize(n); S.inv_moment.resize(n); S.inv_mass.resize(n);
    S.pos[0].resize(n); S.rot[0].resize(n);
    for (size_t i = 0; i < n; ++i) {
        const Body<T>& b = bodies[i];
        S.vel[i] = b.velocity; S.ang[i] = b.angular_velocity; S.prev_vel[i] = b.prev_velocity;
        S.inv_moment[i] = b.inv_moment; S.inv_mass[i] = b.inv_mass;
        S.pos[0][i] = b.position; S.rot[0][i] = b.orientation;
    }
    S.pos[1] = S.pos[0];
    S.rot[1] = S.rot[0];
    RunOptions<T> pc{o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, o.iterations};
    pc.pen_in = pen_in;
    pc.pen_out = pen_out;
    pc.lam_in = lam_in;
    pc.lam_out = lam_out;
    std::vector<V3<T>> op;
    std::vector<Q<T>> orr;
    const std::vector<Constraint<T>> none;
    avbd_solve(S, contacts, none, offsets, indices, pc, op, orr);
    for (size_t i = 0; i < n; ++i) {
        Body<T>& b = bodies[i];
        if (b.is_static()) continue;
        b.position = op[i]; b.orientation = orr[i];
        b.velocity = S.vel[i]; b.angular_velocity = S.ang[i]; b.prev_velocity = S.prev_vel[i];
    }
}

// A recorded step of the forward: the contact set its solve received and the CSR.
struct ReplayStep {
    std::vector<Contact<double>> contacts;
    std::vector<uint32_t> offsets, indices;
};

template <class T> Contact<T> contact_cast(const Contact<double>& c, T friction) {
    auto v = [](V3<double> a) { return V3<T>{T(a.x), T(a.y), T(a.z)}; };
    Contact<T> r;
    r.a = c.a;
    r.b = c.b;
    r.friction = friction;
    r.normal = v(c.normal); r.tangent1 = v(c.tangent1); r.tangent2 = v(c.tangent2);
    r.r_a = v(c.r_a); r.r_b = v(c.r_b); r.c0 = v(c.c0);
    r.penalty = v(c.penalty); r.lambda = v(c.lambda);
    return r;
}

// Replay the recorded steps from `bodies` (their state at the first step)
// with friction μ; the contact values are exact in float and double.
// sched_in / sched_out: the schedules of every step to hold fixed, or to
// write (step after step, the penalties then the multipliers, each I m 3
// values in solve_bodies' layout; sched_out is resized to fit).  freeze bit
// 0 holds the penalties of sched_in fixed, bit 1 the multipliers.  Holding a
// run's schedules fixed gives a map with the same values at that run's
// parameters, whose derivatives treat them as constants.
template <class T>
void replay(std::vector<Body<T>>& bodies, const std::vector<ReplayStep>& steps, const Options<T>& o, T friction,
            const std::vector<double>* sched_in = nullptr, std::vector<double>* sched_out = nullptr,
            int freeze = 0) {
    size_t need = 0;
    for (const ReplayStep& st : steps) need += (size_t)o.iterations * st.contacts.size() * 6;
    if (sched_in && sched_in->size() != need) throw std::invalid_argument("schedule of the wrong size");
    if (sched_out) sched_out->assign(need, 0.0);
    size_t at = 0;
    for (const ReplayStep& st : steps) {
        std::vector<Contact<T>> cs;
        cs.reserve(st.contacts.size());
        for (const auto& c : st.contacts) cs.push_back(contact_cast<T>(c, friction));
        const size_t blk = (size_t)o.iterations * st.contacts.size() * 3;
        const double* in = sched_in ? sched_in->data() + at : nullptr;
        double* out = sched_out ? sched_out->data() + at : nullptr;
        solve_bodies(bodies, cs, st.offsets, st.indices, o, (freeze & 1) && in ? in : nullptr, out,
                     (freeze & 2) && in ? in + blk : nullptr, out ? out + blk : nullptr);
        at += 2 * blk;
    }
}

// ------------------------------------------------------- instance transforms
// The renderer's per-instance matrices (engine/src/render/renderer.rs:250-255
// and engine/src/render/context.rs:188-193): model =
// Mat4::from_scale_rotation_translation(scale, rotation, translation), then
// (projection * view) * model.  Both are glam 0.33.3 (the reference's
// third-party math crate, absent from the reference tree), restated in its
// operation order: the rotation's axes from the doubled components, each
// axis scaled, the translation column with w = 1; a product's column j is
// ((a0 * b_j.x + a1 * b_j.y) + a2 * b_j.z) + a3 * b_j.w component by
// component.  Matrices are column-major, 16 scalars.
template <class T> inline void model_matrix(const T pos[3], const T q[4], const T s[3], T m[16]) {
    const T x = q[0], y = q[1], z = q[2], w = q[3];
    const T x2 = x + x, y2 = y + y, z2 = z + z;
    const T xx = x * x2, xy = x * y2, xz = x * z2;
    const T yy = y * y2, yz = y * z2, zz = z * z2;
    const T wx = w * x2, wy = w * y2, wz = w * z2;
    const T ax[4] = {T(1) - (yy + zz), xy + wz, xz - wy, T(0)};
    const T ay[4] = {xy - wz, T(1) - (xx + zz), yz + wx, T(0)};
    const T az[4] = {xz + wy, yz - wx, T(1) - (xx + yy), T(0)};
    for (int r = 0; r < 4; ++r) {
        m[r] = ax[r] * s[0];
        m[4 + r] = ay[r] * s[1];
        m[8 + r] = az[r] * s[2];
    }
    m[12] = pos[0];
    m[13] = pos[1];
    m[14] = pos[2];
    m[15] = T(1);
}
template <class T> inline void mat4_mul(const T a[16], const T b[16], T out[16]) {
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            out[4 * j + r] = ((a[r] * b[4 * j] + a[4 + r] * b[4 * j + 1]) + a[8 + r] * b[4 * j + 2]) +
                             a[12 + r] * b[4 * j + 3];
}
// Instance slot k < count draws body order[k] (order null: body k); out holds
// (view_proj * model, model) per slot, 32 scalars.
template <class T>
void instance_transforms(const T* pos, const T* rot, const T* scale, const uint32_t* order, uint32_t count,
                         const T view_proj[16], T* out) {
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t b = order ? order[k] : k;
        T m[16];
        model_matrix(pos + 3 * (size_t)b, rot + 4 * (size_t)b, scale + 3 * (size_t)b, m);
        mat4_mul(view_proj, m, out + 32 * (size_t)k);
        std::memcpy(out + 32 * (size_t)k + 16, m, sizeof(m));
    }
}

}  // namespace oracle
