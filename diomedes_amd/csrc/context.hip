// context.hip — host side of the C ABI (include/diomedes_hip.h).
//
// A dio_ctx owns one HIP stream and fixed-capacity device buffers (the
// reference's NarrowPhaseCompute / AvbdCompute objects own fixed Vulkan
// buffers the same way, render/compute.rs:176-397, render/gpu_physics.rs:
// 224-433).  Three groups of buffers:
//   - the granular broad / narrow pipeline (dio_broad_*, dio_narrow_*),
//   - the granular AVBD solve (dio_avbd_*),
//   - the fused world: SoA bodies, joints, two pending narrow outputs and two
//     contact sets used in alternation, the CSR and the solver's pose
//     buffers.  dio_step runs the demo step protocol on the device; with
//     graphs enabled each step is two graph launches (contact build + broad +
//     narrow, then the AVBD iterations) with no host synchronisation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/diomedes_hip.h"
#include "abi_checks.hpp"
#include "adjoint.hpp"
#include "broad.hpp"
#include "tree_refit.hpp"
#include "colour.hpp"
#include "contacts.hpp"
#include "dio_device.hpp"
#include "instances.hpp"
#include "narrow.hpp"
#include "prim.hpp"

using namespace dio;

namespace {

// Traversal lanes per query: the tuned value, or by world size when it is 0.
// Up to 64k bodies the step is short of work and the chain of dependent tree
// steps sets the time, so 16 lanes pop two nodes a step; above, the GPU is
// full and 8 lanes do less work per query.
constexpr uint32_t kWideTraversalMaxBodies = 65536;
inline uint32_t traversal_lanes(uint32_t tuned, uint32_t n) {
    return tuned ? tuned : (n <= kWideTraversalMaxBodies ? 16u : 8u);
}

struct HipError {
    hipError_t code;
    const char* what;
};

#define DIO_HIP_CHECK(expr)                                  \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) throw HipError{e_, #expr};     \
    } while (0)

struct StatusError {
    dio_status status;
    std::string msg;
    uint64_t required;
};

// Device allocations owned by a context; freed together.
struct DevArena {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t count) {
        void* p = nullptr;
        DIO_HIP_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct PinnedArena {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t count) {
        void* p = nullptr;
        DIO_HIP_CHECK(hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    void release() {
        for (void* p : ptrs) (void)hipHostFree(p);
        ptrs.clear();
    }
};

BroadBuffers alloc_broad(DevArena& A, uint32_t n_cap, uint32_t p_max) {
    BroadBuffers b{};
    const size_t nodes = BroadBuffers::tree_nodes(n_cap);
    const size_t nc = n_cap;
    b.aabb_min = A.alloc<float4>(n_cap);
    b.aabb_max = A.alloc<float4>(n_cap);
    b.bounds = A.alloc<float4>(2);
    b.keys0 = A.alloc<uint32_t>(n_cap);
    b.keys1 = A.alloc<uint32_t>(n_cap);
    b.vals1 = A.alloc<uint32_t>(n_cap);
    b.perm = A.alloc<uint32_t>(n_cap);
    b.sort_scratch = A.alloc<uint32_t>(BroadBuffers::sort_scratch_words(n_cap));
    b.node_min = A.alloc<float4>(nodes);
    b.node_max = A.alloc<float4>(nodes);
    b.counts = A.alloc<uint32_t>(nc);
    b.offsets = A.alloc<uint32_t>(nc + 1);
    b.scan_partials = A.alloc<uint32_t>(scan_partials_size((uint32_t)nc + 1));
    b.pair_cache = A.alloc<uint32_t>((size_t)kPairCache * n_cap);
    b.pairs = A.alloc<uint2>(p_max);
    b.p_dev = A.alloc<uint32_t>(1);
    b.overflow = A.alloc<uint32_t>(1);
    b.err = A.alloc<uint32_t>(1);
    return b;
}

// narrow slots for a pair list of capacity p_max (pairs / p_dev set by caller)
NarrowOut alloc_narrow(DevArena& A, uint32_t p_max) {
    NarrowOut o{};
    o.n = A.alloc<float4>((size_t)p_max * 4);
    o.pa = A.alloc<float4>((size_t)p_max * 4);
    o.pb = A.alloc<float4>((size_t)p_max * 4);
    o.cnt = A.alloc<uint32_t>(p_max);
    o.off = A.alloc<uint32_t>((size_t)p_max + 1);
    o.scan_partials = A.alloc<uint32_t>(scan_partials_size(p_max + 1));
    o.gjk_list = A.alloc<uint32_t>(p_max);
    o.gjk_count = A.alloc<uint32_t>(1);
    o.epa_retry_list = A.alloc<uint32_t>(p_max);
    o.epa_retry_count = A.alloc<uint32_t>(1);
    o.p_max = p_max;
    return o;
}

ContactArrays alloc_contacts(DevArena& A, uint32_t m_cap) {
    ContactArrays c{};
    c.ab = A.alloc<uint2>(m_cap);
    c.n = A.alloc<float4>(m_cap);
    c.t1 = A.alloc<float4>(m_cap);
    c.t2 = A.alloc<float4>(m_cap);
    c.ra = A.alloc<float4>(m_cap);
    c.rb = A.alloc<float4>(m_cap);
    c.c0 = A.alloc<float4>(m_cap);
    c.pen = A.alloc<float4>(m_cap);
    c.lam = A.alloc<float4>(m_cap);
    return c;
}

// joint CSR in ascending constraint order (a self-joint is listed once)
void joint_csr(const dio_avbd_constraint* cons, uint32_t c, uint32_t n, std::vector<uint32_t>& off,
               std::vector<uint32_t>& idx) {
    off.assign(n + 1, 0);
    for (uint32_t k = 0; k < c; ++k) {
        off[cons[k].a + 1] += 1;
        if (cons[k].b != cons[k].a) off[cons[k].b + 1] += 1;
    }
    for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
    idx.assign(off[n], 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t k = 0; k < c; ++k) {
        idx[cur[cons[k].a]++] = k;
        if (cons[k].b != cons[k].a) idx[cur[cons[k].b]++] = k;
    }
}

struct JointDev {
    uint4* abk = nullptr;
    float4* anc_a = nullptr;
    float4* anc_b = nullptr;
    uint32_t* off = nullptr;
    uint32_t* idx = nullptr;
    uint32_t c = 0;
    // hard joints (world only): multipliers of a solve (two buffers) and the
    // state carried between steps; the hard joints, and those between two
    // dynamic bodies
    float4* lam[2] = {nullptr, nullptr};
    float4* pen[2] = {nullptr, nullptr};
    float4* st_lam = nullptr;
    float4* st_pen = nullptr;
    uint32_t hard = 0, hard_dynamic = 0;
    uint32_t extended = 0;  // joints of kinds past the reference's four (hard and bending joints)
};

}  // namespace

// Device memory imported from another API: the HIP import of a duplicate of
// the caller's file descriptor (the duplicate's identity kept to close it
// once the import is gone) and the mapped range.
struct dio_external_memory {
    hipExternalMemory_t ext = nullptr;
    void* ptr = nullptr;
    int fd = -1;
    dev_t st_dev = 0;
    ino_t st_ino = 0;
};

namespace {
void release_external(dio_external_memory* m) {
    if (m->ext) (void)hipDestroyExternalMemory(m->ext);
    // the HIP runtime may or may not have closed the descriptor it was given:
    // close it only if it still names the imported object
    struct stat st;
    if (m->fd >= 0 && fcntl(m->fd, F_GETFD) != -1 && fstat(m->fd, &st) == 0 && st.st_dev == m->st_dev &&
        st.st_ino == m->st_ino)
        (void)close(m->fd);
    delete m;
}
}  // namespace

struct dio_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the second stream of a step (the next step's broad and narrow phases
    // beside the contact build and CSR) and its fork / join events
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    uint64_t required = 0;
    dio_limits lim{};
    uint32_t n_cap = 0, p_max = 0, m_cap = 0, c_cap = 0, h_cap = 0;
    DevArena dev;
    PinnedArena pin;
    // reusable scan scratch (sized for the largest scanned array)
    uint32_t* scan_partials = nullptr;

    // ---- granular broad / narrow (NarrowPhaseCompute)
    bool g_ready = false;
    dio_gpu_body* g_rec = nullptr;          // device staging
    dio_gpu_body* g_rec_host = nullptr;     // pinned
    float4 *g_pos = nullptr, *g_rot = nullptr, *g_corners = nullptr, *g_hull = nullptr;
    uint32_t* g_tag = nullptr;
    BroadBuffers gb{};
    NarrowOut gn{};
    float4* g_simplex = nullptr;
    uint2* g_pairs_up = nullptr;            // caller-given pairs (dio_narrow_submit_pairs)
    uint2* g_pairs_host = nullptr;
    uint32_t* g_np_dev = nullptr;           // pair count for the narrow dispatch
    dio_gpu_contact_out* g_out = nullptr;   // packed records
    uint32_t g_n = 0, g_last_pairs = 0;
    bool g_broad_submitted = false, g_narrow_submitted = false;

    // ---- granular AVBD (AvbdCompute)
    bool a_ready = false;
    dio_avbd_body_state* a_state_host = nullptr;
    float* a_pose_host = nullptr;  // n x 8 (pos xyz0, rot xyzw)
    dio_avbd_contact* a_con_host = nullptr;
    dio_avbd_constraint* a_cons_host = nullptr;
    uint32_t* a_csr_host = nullptr;  // offsets (n_cap + 1) then indices (2 m_cap)
    dio_avbd_body_state* a_state = nullptr;
    float* a_pose = nullptr;
    dio_avbd_contact* a_con_rec = nullptr;
    float4 *a_cur_pos = nullptr, *a_cur_rot = nullptr;
    float4 *a_vel = nullptr, *a_ang = nullptr, *a_prev = nullptr, *a_inv_mm = nullptr;
    float4 *a_pos[2] = {nullptr, nullptr}, *a_rot[2] = {nullptr, nullptr};
    float4 *a_init_lin = nullptr, *a_init_ang = nullptr, *a_iner_lin = nullptr, *a_iner_ang = nullptr;
    ContactArrays a_c{};
    float4 *a_lam_alt = nullptr, *a_pen_alt = nullptr;  // scratch multipliers of a solve
    uint32_t *a_off = nullptr, *a_idx = nullptr, *a_peer = nullptr;
    JointDev a_j;
    dio_avbd_constraint* a_cons_stage = nullptr;
    dio_tuning tuning{32, 0, 16, 1, 8, 0, 1, 0, 0, 1};
    uint32_t g_lanes = 4;     // traversal lanes of the last granular broad phase
    uint64_t w_sorted_at = 0; // step of the last Morton re-sort of the world
    bool w_sort_pending = true;
    uint32_t a_n = 0, a_m = 0, a_iters = 0;
    bool a_submitted = false;

    // ---- fused world
    bool w_ready = false;
    uint32_t w_n = 0;
    std::vector<float> w_friction;
    float4 *w_pos = nullptr, *w_rot = nullptr, *w_vel = nullptr, *w_ang = nullptr, *w_prev = nullptr,
           *w_inv_mm = nullptr, *w_corners = nullptr, *w_hull = nullptr;
    uint32_t* w_tag = nullptr;
    uint32_t w_hull_count = 0;
    // the hull pool size the world's hull bodies need (checked at each step,
    // since the pool may be uploaded before or after the bodies)
    uint64_t w_hull_need = 0;
    uint32_t last_lag = 1;
    uint32_t w_n_dyn = 0;
    // false when every body is a box (or cube), or every body a sphere: no
    // pair can reach GJK / EPA
    bool w_gjk_possible = true;
    dio_gpu_body* w_stage_b = nullptr;
    dio_avbd_body_state* w_stage_s = nullptr;
    dio_avbd_constraint* w_stage_c = nullptr;
    JointDev w_j;
    ContainerParams w_ct{};
    // batched world (dio_batch_upload): scene of each body, body offsets per
    // scene, per-scene containers; w_scene_bits = 0 for a single scene
    uint32_t* w_scene_of = nullptr;
    uint32_t* w_scene_off = nullptr;
    ContainerParams* w_scene_ct = nullptr;
    uint32_t w_num_scenes = 1, w_scene_bits = 0;
    // the queue of the per-scene solve: runs of whole consecutive scenes
    // ("bundles", each at least bundle_bodies bodies, or one scene), their
    // body offsets, the order they are taken in (most bodies first) and the
    // queue's two counters (zero between launches)
    uint32_t* w_scene_order = nullptr;
    uint32_t* w_bundle_off = nullptr;
    uint32_t* w_scene_q = nullptr;
    uint32_t w_bundles = 0, w_bundle_bodies = 0;
    std::vector<uint32_t> w_scene_off_host;
    BroadBuffers wb{};
    struct Pending {
        uint2* pairs;
        uint32_t* p_dev;
        NarrowOut out;
        float4* simplex;
    } pend[2]{};
    ContactSet cs[2]{};
    CsrBuffers csr{};
    // two sets of solver buffers, used by alternate steps: the final launch
    // of a Jacobi step runs the next step's init into the other set.
    // init_set: the set holding a prepared init (-1: none), for init_dt /
    // init_g (the options the init depends on)
    InitTarget ss[2]{};
    int init_set = -1;
    float init_dt = 0.0f, init_g[3] = {0.0f, 0.0f, 0.0f};
    // and, with hard joints, the options of their warm start
    float init_alpha = 0.0f, init_gamma = 0.0f, init_pmin = 0.0f, init_pmax = 0.0f;
    uint32_t* w_overflow = nullptr;  // [0] candidate pairs past max_pairs, [1] contacts past max_contacts
    unsigned long long* w_acc = nullptr;  // sums of pairs, contacts, CSR nnz over a dio_step call
    float4 *w_lam_alt = nullptr, *w_pen_alt = nullptr;  // scratch multipliers of a solve
    uint4* w_work = nullptr;  // the Jacobi primal's work table (body and CSR ranges per work item)
    uint2* w_items = nullptr;  // the packed primal's item table
    // instanced-renderer hand-off: per-body scale (xyz) and the slot order
    float4* inst_scale = nullptr;
    uint32_t* inst_order = nullptr;  // grown by dio_world_set_instances, freed with the context
    uint32_t inst_order_cap = 0;
    uint32_t inst_count = 0;
    bool inst_ordered = false;
    // device memory imported from other APIs (dio_import_external_memory)
    std::vector<dio_external_memory*> ext_mems;
    std::vector<hipEvent_t> prof_events;
    uint64_t w_step = 0;
    dio_step_stats stats{};
    // graph cache [variant][parity]: 0 = contacts + broad + narrow, 1 = AVBD,
    // 2 = contacts + broad (with a Morton re-sort) + narrow
    hipGraphExec_t graphs[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    dio_avbd_options g_opts{};
    dio_step_opts g_sopts{};
    bool graphs_valid = false;
    std::vector<hipEvent_t> events;

    void fail(const std::string& m) { err = m; }
    // the coloured mode's captured AVBD segment (per parity), built with caps
    // on the colouring rounds and on the colours launched one by one, and the
    // device-side colour ranges it reads
    hipGraphExec_t col_graphs[2] = {nullptr, nullptr};
    bool col_valid = false;
    uint32_t col_rcap = 0, col_ccap = 0;
    dio_avbd_options col_opts{};
    dio_step_opts col_sopts{};
    uint32_t* col_starts = nullptr;
    void drop_graphs() {
        for (auto& seg : graphs)
            for (auto& g : seg)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
        graphs_valid = false;
        for (auto& g : col_graphs)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        col_valid = false;
    }
    // adjoint: the recording of dio_step calls and the replay / backward scratch
    AdjRecord adj_rec{};
    AdjWork adj_work{};
    void* adj_rec_base = nullptr;
    bool adj_on = false;
    uint32_t adj_mode = DIO_ADJOINT_EXACT;  // dio_adjoint_set_mode
    // coloured Gauss-Seidel mode: colouring buffers, a pinned read-back area
    // and the last step's colouring summary
    ColourBuffers wc{};
    uint32_t* wc_host = nullptr;  // kColourRoundSlots + kMaxColours + 2 words
    uint32_t last_colours = 0, last_rounds = 0;
    std::vector<uint8_t> adj_sweep_key;  // what w.ck was swept from (empty: nothing)
    void free_adjoint() {
        if (adj_rec_base) (void)hipFree(adj_rec_base);
        adj_rec_base = nullptr;
        adj_rec = AdjRecord{};
        adj_work_free(adj_work);
        adj_on = false;
        adj_sweep_key.clear();
    }
    ~dio_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream2) (void)hipStreamSynchronize(stream2);
        for (dio_external_memory* m : ext_mems) release_external(m);
        ext_mems.clear();
        if (inst_order) (void)hipFree(inst_order);
        free_adjoint();
        drop_graphs();
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto e : prof_events) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        dev.release();
        pin.release();
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class F> dio_status guarded(dio_ctx* ctx, F&& f) {
    if (!ctx) return DIO_INVALID;
    try {
        return f();
    } catch (const HipError& e) {
        ctx->err = std::string(e.what) + ": " + hipGetErrorString(e.code);
        (void)hipGetLastError();  // a failed call leaves no error for the next one to report
        return DIO_HIP;
    } catch (const StatusError& e) {
        ctx->err = e.msg;
        if (e.status == DIO_CAPACITY) ctx->required = e.required;
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return DIO_HIP;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DIO_INVALID;
    }
}

[[noreturn]] void invalid(const std::string& m) { throw StatusError{DIO_INVALID, m, 0}; }
[[noreturn]] void capacity(const std::string& m, uint64_t need) { throw StatusError{DIO_CAPACITY, m, need}; }

// The hull pool size the bodies need (hull_pool_need); a malformed hull
// range is DIO_INVALID.  The caller compares the size with the pool uploaded
// when it launches.
uint64_t hull_extent(const dio_gpu_body* bodies, uint32_t n) {
    const int64_t need = hull_pool_need(bodies, n);
    if (need < 0)
        invalid("body " + std::to_string(-1 - need) + ": a hull's first vertex and vertex count must be whole "
                "numbers, the count at least 1");
    return (uint64_t)need;
}

void ensure_granular_broad(dio_ctx* c) {
    if (c->g_ready) return;
    DevArena& A = c->dev;
    c->g_rec = A.alloc<dio_gpu_body>(c->n_cap);
    c->g_rec_host = c->pin.alloc<dio_gpu_body>(c->n_cap);
    c->g_pos = A.alloc<float4>(c->n_cap);
    c->g_rot = A.alloc<float4>(c->n_cap);
    c->g_tag = A.alloc<uint32_t>(c->n_cap);
    c->g_corners = A.alloc<float4>((size_t)c->n_cap * 4);
    c->g_hull = A.alloc<float4>(1);
    c->gb = alloc_broad(A, c->n_cap, c->p_max);
    c->gn = alloc_narrow(A, c->p_max);
    c->g_simplex = A.alloc<float4>((size_t)c->p_max * 8);
    c->g_pairs_up = A.alloc<uint2>(c->p_max);
    c->g_pairs_host = c->pin.alloc<uint2>(c->p_max);
    c->g_np_dev = A.alloc<uint32_t>(1);
    c->g_out = A.alloc<dio_gpu_contact_out>((size_t)c->p_max * 4);
    c->g_ready = true;
}

// The granular broad and narrow phases read the pool dio_world_upload_hulls
// uploaded (none before that call).
void check_granular_hulls(dio_ctx* c, const dio_gpu_body* bodies, uint32_t n) {
    const uint64_t have = (c->w_hull && c->g_hull == c->w_hull) ? c->w_hull_count : 0;
    const uint64_t need = hull_extent(bodies, n);
    if (need > have)
        invalid("the hull bodies name " + std::to_string(need) + " hull vertices; the uploaded pool holds " +
                std::to_string(have));
}

void ensure_granular_avbd(dio_ctx* c) {
    if (c->a_ready) return;
    DevArena& A = c->dev;
    PinnedArena& P = c->pin;
    const uint32_t n = c->n_cap, m = c->m_cap;
    c->a_state_host = P.alloc<dio_avbd_body_state>(n);
    c->a_pose_host = P.alloc<float>((size_t)n * 8);
    c->a_con_host = P.alloc<dio_avbd_contact>(m);
    c->a_cons_host = P.alloc<dio_avbd_constraint>(c->c_cap);
    c->a_csr_host = P.alloc<uint32_t>((size_t)n + 1 + 2 * (size_t)m);
    c->a_state = A.alloc<dio_avbd_body_state>(n);
    c->a_pose = A.alloc<float>((size_t)n * 8);
    c->a_con_rec = A.alloc<dio_avbd_contact>(m);
    c->a_cur_pos = A.alloc<float4>(n);
    c->a_cur_rot = A.alloc<float4>(n);
    c->a_vel = A.alloc<float4>(n);
    c->a_ang = A.alloc<float4>(n);
    c->a_prev = A.alloc<float4>(n);
    c->a_inv_mm = A.alloc<float4>(n);
    for (int k = 0; k < 2; ++k) {
        c->a_pos[k] = A.alloc<float4>(n);
        c->a_rot[k] = A.alloc<float4>(n);
    }
    c->a_init_lin = A.alloc<float4>(n);
    c->a_init_ang = A.alloc<float4>(n);
    c->a_iner_lin = A.alloc<float4>(n);
    c->a_iner_ang = A.alloc<float4>(n);
    c->a_c = alloc_contacts(A, m);
    c->a_lam_alt = A.alloc<float4>(m);
    c->a_pen_alt = A.alloc<float4>(m);
    c->a_off = A.alloc<uint32_t>((size_t)n + 1);
    c->a_idx = A.alloc<uint32_t>(2 * (size_t)m);
    c->a_peer = A.alloc<uint32_t>(2 * (size_t)m);
    c->a_j.abk = A.alloc<uint4>(c->c_cap);
    c->a_j.anc_a = A.alloc<float4>(c->c_cap);
    c->a_j.anc_b = A.alloc<float4>(c->c_cap);
    c->a_j.off = A.alloc<uint32_t>((size_t)n + 1);
    c->a_j.idx = A.alloc<uint32_t>(2 * (size_t)c->c_cap);
    c->a_cons_stage = A.alloc<dio_avbd_constraint>(c->c_cap);
    c->a_ready = true;
}

// unpack GpuAvbdBodyState + poses (granular upload)
__global__ void k_unpack_state(const dio_avbd_body_state* st, const float* pose, uint32_t n, float4* vel, float4* ang,
                               float4* prev, float4* inv_mm, float4* pos, float4* rot) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const dio_avbd_body_state s = st[i];
    vel[i] = make_float4(s.vel[0], s.vel[1], s.vel[2], 0.0f);
    ang[i] = make_float4(s.ang[0], s.ang[1], s.ang[2], 0.0f);
    prev[i] = make_float4(s.prev_vel[0], s.prev_vel[1], s.prev_vel[2], 0.0f);
    inv_mm[i] = make_float4(s.inv_moment[0], s.inv_moment[1], s.inv_moment[2], s.inv_mass);
    if (pose) {
        pos[i] = make_float4(pose[8 * i], pose[8 * i + 1], pose[8 * i + 2], 0.0f);
        rot[i] = make_float4(pose[8 * i + 4], pose[8 * i + 5], pose[8 * i + 6], pose[8 * i + 7]);
    }
}

__global__ void k_unpack_joints(const dio_avbd_constraint* rec, uint32_t c, uint4* abk, float4* a, float4* b) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c) return;
    const dio_avbd_constraint r = rec[i];
    abk[i] = make_uint4(r.a, r.b, r.kind, 0u);
    a[i] = make_float4(r.anchor_a[0], r.anchor_a[1], r.anchor_a[2], r.anchor_a[3]);
    b[i] = make_float4(r.anchor_b[0], r.anchor_b[1], r.anchor_b[2], r.anchor_b[3]);
}

inline uint32_t blocks256(uint32_t n) { return (n + 255) / 256; }

// upload constraints (records staged in pinned host memory) and their CSR
void upload_joints(dio_ctx* c, JointDev& J, const dio_avbd_constraint* cons, uint32_t nc, uint32_t n,
                   dio_avbd_constraint* stage_dev) {
    for (uint32_t k = 0; k < nc; ++k)
        if (cons[k].a >= n || cons[k].b >= n) invalid("constraint body index out of range");
    std::vector<uint32_t> off, idx;
    joint_csr(cons, nc, n, off, idx);
    J.c = nc;
    DIO_HIP_CHECK(hipMemcpyAsync(J.off, off.data(), off.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!idx.empty())
        DIO_HIP_CHECK(hipMemcpyAsync(J.idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (nc > 0) {
        DIO_HIP_CHECK(hipMemcpyAsync(stage_dev, cons, nc * sizeof(dio_avbd_constraint), hipMemcpyHostToDevice,
                                     c->stream));
        hipLaunchKernelGGL(k_unpack_joints, dim3(blocks256(nc)), dim3(256), 0, c->stream, stage_dev, nc, J.abk,
                           J.anc_a, J.anc_b);
    }
    // off / idx are host vectors: make the copies complete before they go
    DIO_HIP_CHECK(hipStreamSynchronize(c->stream));
}

AvbdParams params_of(float dt, const float g[3], float alpha, float beta, float pmax, uint32_t n,
                     float gamma = 0.999f, float pmin = 1.0f) {
    AvbdParams p;
    p.gamma = gamma;
    p.penalty_min = pmin;
    p.dt = dt;
    p.alpha = alpha;
    p.beta_lin = beta;
    p.penalty_max = pmax;
    p.gx = g[0];
    p.gy = g[1];
    p.gz = g[2];
    p.n = n;
    return p;
}

// Event pairs around each AVBD kernel (profile mode): kind 0 init, 1 primal,
// 2 dual, 3 recover.
struct KernelTimer {
    std::vector<hipEvent_t>* pool = nullptr;
    size_t next = 0;
    std::vector<std::pair<int, size_t>> marks;  // kind, index of the start event
    hipEvent_t take() {
        if (next >= pool->size()) {
            hipEvent_t e;
            DIO_HIP_CHECK(hipEventCreate(&e));
            pool->push_back(e);
        }
        return (*pool)[next++];
    }
    // the launch records its own start / stop events (no marker packets)
    template <class F> void time(int kind, F&& launch) {
        const size_t i = next;
        LaunchEvents ev;
        ev.start = take();
        ev.stop = take();
        launch(ev);
        marks.push_back({kind, i});
    }
};

// One AVBD solve (gpu_physics.rs:541-550): init; I x (primal, dual), the
// dual fused into the next primal when `fused`; the last dual with recover.
void schedule_avbd(const SolveView& v, const AvbdParams& p, uint32_t iters, uint32_t m_grid, hipStream_t s,
                   KernelTimer* kt, const dio_tuning& t, bool run_init = true) {
    const int lanes = (int)t.primal_lanes;
    const bool fused = t.fuse_dual != 0u;
    auto run = [&](int kind, auto&& launch) {
        if (kt) kt->time(kind, launch);
        else launch(LaunchEvents{});
    };
    if (run_init) run(0, [&](const LaunchEvents& ev) { launch_avbd_init(v, p, s, ev); });
    for (uint32_t k = 0; k < iters; ++k) {
        run(1, [&](const LaunchEvents& ev) { launch_avbd_primal(v, p, k, fused, s, lanes, ev); });
        if (!fused && k + 1 < iters)
            run(2, [&](const LaunchEvents& ev) { launch_avbd_dual(v, p, k & 1u, 0u, 0u, m_grid, s, ev); });
    }
    run(3, [&](const LaunchEvents& ev) { launch_avbd_dual_recover(v, p, iters, fused, m_grid, s, ev); });
}

// The coloured Gauss-Seidel solve (no counterpart in the reference): init;
// per iteration one primal launch per colour over that colour's body list,
// poses updated in place (v.pos[0] == v.pos[1]), then the dual update; the
// last dual update with recover.  start[c] .. start[c + 1] index the list.
// Lanes per body of the coloured primal.
int colour_lanes_of(const dio_tuning& t) { return (int)(t.colour_lanes ? t.colour_lanes : t.primal_lanes); }

void schedule_avbd_coloured(SolveView v, const AvbdParams& p, uint32_t iters, uint32_t m_grid, hipStream_t s,
                            KernelTimer* kt, const dio_tuning& t, const uint32_t* list,
                            const std::vector<uint32_t>& start) {
    const int lanes = colour_lanes_of(t);
    auto run = [&](int kind, auto&& launch) {
        if (kt) kt->time(kind, launch);
        else launch(LaunchEvents{});
    };
    v.pos[1] = v.pos[0];
    v.rot[1] = v.rot[0];
    run(0, [&](const LaunchEvents& ev) { launch_avbd_init(v, p, s, ev); });
    for (uint32_t k = 0; k < iters; ++k) {
        for (size_t c = 0; c + 1 < start.size(); ++c) {
            const uint32_t count = start[c + 1] - start[c];
            if (count == 0) continue;
            SolveView vc = v;
            vc.body_list = list + start[c];
            vc.list_count = count;
            run(1, [&](const LaunchEvents& ev) { launch_avbd_primal_colour(vc, p, k, count, s, lanes, ev); });
        }
        if (k + 1 < iters)
            run(2, [&](const LaunchEvents& ev) { launch_avbd_dual(v, p, k & 1u, 0u, 0u, m_grid, s, ev); });
    }
    run(3, [&](const LaunchEvents& ev) { launch_avbd_dual_recover(v, p, iters, false, m_grid, s, ev); });
}

__global__ void k_stats_zero(unsigned long long* acc) {
    if (threadIdx.x < 4) acc[threadIdx.x] = 0ull;
}

}  // namespace

// =========================================================================
extern "C" {

uint32_t dio_abi_version(void) { return DIO_ABI_VERSION; }

void dio_default_options(dio_avbd_options* o) {
    if (!o) return;
    o->dt = 1.0f / 60.0f;
    o->gravity[0] = 0.0f;
    o->gravity[1] = -10.0f;
    o->gravity[2] = 0.0f;
    o->iterations = 10;
    o->alpha = 0.99f;
    o->beta_lin = 10000.0f;
    o->margin = 0.01f;
    o->penalty_min = 1.0f;
    o->penalty_max = 1e10f;
    o->gamma = 0.999f;
    o->friction = 0.8f;
}

dio_status dio_create(int device, const dio_limits* limits, dio_ctx** out) {
    if (!out || !limits || limits->max_bodies == 0) return DIO_INVALID;
    *out = nullptr;
    dio_ctx* c = nullptr;
    try {
        c = new dio_ctx();
    } catch (...) {
        return DIO_HIP;
    }
    dio_status st = guarded(c, [&]() -> dio_status {
        c->device = device;
        DIO_HIP_CHECK(hipSetDevice(device));
        DIO_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        DIO_HIP_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        DIO_HIP_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        DIO_HIP_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        c->lim = *limits;
        if (limits->max_bodies > kMaxBroadBodies) capacity("max_bodies exceeds the broad phase limit", kMaxBroadBodies);
        c->n_cap = limits->max_bodies;
        c->p_max = limits->max_pairs ? limits->max_pairs : std::max<uint32_t>(1024u, 16u * c->n_cap);
        c->m_cap = limits->max_contacts ? limits->max_contacts : 4u * c->p_max + 24u * c->n_cap;
        c->c_cap = limits->max_constraints ? limits->max_constraints : std::max<uint32_t>(16u, c->n_cap);
        c->h_cap = limits->max_hull_vertices ? limits->max_hull_vertices : 4096u;
        const uint32_t big = std::max({c->n_cap + 1, c->p_max + 1, c->m_cap + 1});
        c->scan_partials = c->dev.alloc<uint32_t>(scan_partials_size(big));
        return DIO_OK;
    });
    if (st != DIO_OK) {
        delete c;
        return st;
    }
    *out = c;
    return DIO_OK;
}

void dio_destroy(dio_ctx* ctx) { delete ctx; }

const char* dio_last_error(const dio_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

uint64_t dio_required_capacity(const dio_ctx* ctx) { return ctx ? ctx->required : 0; }

dio_status dio_synchronize(dio_ctx* ctx) {
    return guarded(ctx, [&]() {
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return DIO_OK;
    });
}

// ------------------------------------------------------------ broad phase
dio_status dio_broad_submit(dio_ctx* ctx, const dio_gpu_body* bodies, uint32_t n) {
    return guarded(ctx, [&]() {
        if (n > 0 && !bodies) invalid("bodies is null");
        if (n > ctx->n_cap) capacity("too many bodies for the context", n);
        check_granular_hulls(ctx, bodies, n);
        ensure_granular_broad(ctx);
        DIO_HIP_CHECK(hipSetDevice(ctx->device));
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // staging buffer reuse
        ctx->g_n = n;
        if (n > 0) {
            std::memcpy(ctx->g_rec_host, bodies, (size_t)n * sizeof(dio_gpu_body));
            DIO_HIP_CHECK(hipMemcpyAsync(ctx->g_rec, ctx->g_rec_host, (size_t)n * sizeof(dio_gpu_body),
                                         hipMemcpyHostToDevice, ctx->stream));
            launch_unpack_gpu_bodies(ctx->g_rec, n, ctx->g_pos, ctx->g_rot, ctx->g_tag, ctx->g_corners, ctx->stream);
        }
        launch_fill_u32(ctx->gb.overflow, 0u, 1, ctx->stream);
        launch_fill_u32(ctx->gb.err, 0u, 1, ctx->stream);
        BodyView bv{ctx->g_pos, ctx->g_rot, ctx->g_tag, ctx->g_corners, ctx->g_hull};
        ctx->g_lanes = traversal_lanes(ctx->tuning.broad_lanes, n);
        run_broad(bv, n, ctx->gb, ctx->p_max, ctx->stream, true, (int)ctx->g_lanes);
        DIO_HIP_CHECK(hipGetLastError());
        ctx->g_broad_submitted = true;
        return DIO_OK;
    });
}

dio_status dio_broad_count(dio_ctx* ctx, uint32_t* count) {
    return guarded(ctx, [&]() {
        if (!count) invalid("count is null");
        if (!ctx->g_broad_submitted) throw StatusError{DIO_NOT_READY, "no broad phase submitted", 0};
        uint32_t v[3] = {0, 0, 0};
        DIO_HIP_CHECK(hipMemcpyAsync(&v[0], ctx->gb.offsets + ctx->g_n, 4, hipMemcpyDeviceToHost, ctx->stream));
        DIO_HIP_CHECK(hipMemcpyAsync(&v[1], ctx->gb.overflow, 4, hipMemcpyDeviceToHost, ctx->stream));
        DIO_HIP_CHECK(hipMemcpyAsync(&v[2], ctx->gb.err, 4, hipMemcpyDeviceToHost, ctx->stream));
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (v[2]) throw StatusError{DIO_HIP, "LBVH build / traversal error (malformed tree or stack overflow)", 0};
        *count = v[0];
        if (v[0] > ctx->p_max) capacity("candidate pairs exceed max_pairs", v[0]);
        return DIO_OK;
    });
}

dio_status dio_broad_read_pairs(dio_ctx* ctx, dio_gpu_pair* pairs, uint32_t cap, uint32_t* n) {
    return guarded(ctx, [&]() {
        if (!n) invalid("n is null");
        uint32_t count = 0;
        dio_status st = dio_broad_count(ctx, &count);
        if (st != DIO_OK) return st;
        *n = count;
        if (count > cap) capacity("output array too small", count);
        if (count > 0) {
            if (!pairs) invalid("pairs is null");
            DIO_HIP_CHECK(hipMemcpy(pairs, ctx->gb.pairs, (size_t)count * 8, hipMemcpyDeviceToHost));
        }
        return DIO_OK;
    });
}

dio_status dio_broad_read_aabbs(dio_ctx* ctx, float* aabbs, uint32_t cap_bodies) {
    return guarded(ctx, [&]() {
        if (!ctx->g_broad_submitted) throw StatusError{DIO_NOT_READY, "no broad phase submitted", 0};
        if (cap_bodies < ctx->g_n) capacity("output array too small", ctx->g_n);
        if (ctx->g_n == 0) return DIO_OK;
        if (!aabbs) invalid("aabbs is null");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        std::vector<float4> mn(ctx->g_n), mx(ctx->g_n);
        DIO_HIP_CHECK(hipMemcpy(mn.data(), ctx->gb.aabb_min, ctx->g_n * 16, hipMemcpyDeviceToHost));
        DIO_HIP_CHECK(hipMemcpy(mx.data(), ctx->gb.aabb_max, ctx->g_n * 16, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < ctx->g_n; ++i) {
            std::memcpy(aabbs + 8 * i, &mn[i], 16);
            std::memcpy(aabbs + 8 * i + 4, &mx[i], 16);
        }
        return DIO_OK;
    });
}

// ------------------------------------------------------------ narrow phase
static void narrow_dispatch(dio_ctx* ctx, const uint2* pairs, uint32_t count) {
    NarrowOut o = ctx->gn;
    o.pairs = pairs;
    o.p_dev = ctx->g_np_dev;
    DIO_HIP_CHECK(hipMemcpyAsync(ctx->g_np_dev, &ctx->g_last_pairs, 4, hipMemcpyHostToDevice, ctx->stream));
    BodyView bv{ctx->g_pos, ctx->g_rot, ctx->g_tag, ctx->g_corners, ctx->g_hull,
                ctx->g_hull == ctx->w_hull ? ctx->w_hull_count : 0u};
    run_narrow(bv, o, ctx->g_simplex, ctx->stream);
    launch_pack_contact_out(o, count, ctx->g_out, ctx->stream);
    DIO_HIP_CHECK(hipGetLastError());
    // g_last_pairs is a host word read by the async copy above
    DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->g_narrow_submitted = true;
}

dio_status dio_narrow_submit(dio_ctx* ctx, uint32_t pair_count) {
    return guarded(ctx, [&]() {
        if (!ctx->g_broad_submitted) throw StatusError{DIO_NOT_READY, "no broad phase submitted", 0};
        if (pair_count > ctx->p_max) capacity("too many pairs for the compute buffers", pair_count);
        ctx->g_last_pairs = pair_count;
        narrow_dispatch(ctx, ctx->gb.pairs, pair_count);
        return DIO_OK;
    });
}

dio_status dio_narrow_submit_pairs(dio_ctx* ctx, const dio_gpu_body* bodies, uint32_t n, const dio_gpu_pair* pairs,
                                   uint32_t pair_count) {
    return guarded(ctx, [&]() {
        if (n > ctx->n_cap) capacity("too many bodies for the compute buffers", n);
        if (pair_count > ctx->p_max) capacity("too many pairs for the compute buffers", pair_count);
        if ((n > 0 && !bodies) || (pair_count > 0 && !pairs)) invalid("null input");
        for (uint32_t k = 0; k < pair_count; ++k)
            if (pairs[k].a >= n || pairs[k].b >= n) invalid("pair index out of range");
        check_granular_hulls(ctx, bodies, n);
        ensure_granular_broad(ctx);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->g_n = n;
        if (n > 0) {
            std::memcpy(ctx->g_rec_host, bodies, (size_t)n * sizeof(dio_gpu_body));
            DIO_HIP_CHECK(hipMemcpyAsync(ctx->g_rec, ctx->g_rec_host, (size_t)n * sizeof(dio_gpu_body),
                                         hipMemcpyHostToDevice, ctx->stream));
            launch_unpack_gpu_bodies(ctx->g_rec, n, ctx->g_pos, ctx->g_rot, ctx->g_tag, ctx->g_corners, ctx->stream);
        }
        if (pair_count > 0) {
            std::memcpy(ctx->g_pairs_host, pairs, (size_t)pair_count * 8);
            DIO_HIP_CHECK(hipMemcpyAsync(ctx->g_pairs_up, ctx->g_pairs_host, (size_t)pair_count * 8,
                                         hipMemcpyHostToDevice, ctx->stream));
        }
        ctx->g_last_pairs = pair_count;
        ctx->g_broad_submitted = false;
        narrow_dispatch(ctx, ctx->g_pairs_up, pair_count);
        return DIO_OK;
    });
}

dio_status dio_narrow_read(dio_ctx* ctx, dio_gpu_contact_out* slots, uint32_t cap_slots, uint32_t* n_slots) {
    return guarded(ctx, [&]() {
        if (!n_slots) invalid("n_slots is null");
        if (!ctx->g_narrow_submitted) {
            *n_slots = 0;
            return DIO_OK;
        }
        const uint32_t ns = ctx->g_last_pairs * 4;
        *n_slots = ns;
        if (ns > cap_slots) capacity("output array too small", ns);
        if (ns > 0) {
            if (!slots) invalid("slots is null");
            DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            DIO_HIP_CHECK(hipMemcpy(slots, ctx->g_out, (size_t)ns * sizeof(dio_gpu_contact_out),
                                    hipMemcpyDeviceToHost));
        }
        return DIO_OK;
    });
}

// ------------------------------------------------------------ AVBD solve
dio_status dio_avbd_submit(dio_ctx* ctx, const dio_avbd_body_state* state, const float* positions,
                           const float* orientations, uint32_t n, const dio_avbd_contact* contacts, uint32_t m,
                           const dio_avbd_constraint* constraints, uint32_t c, const uint32_t* offsets,
                           const uint32_t* indices, uint32_t nidx, const dio_run_options* options) {
    return guarded(ctx, [&]() {
        if (!options) invalid("options is null");
        if (n > ctx->n_cap) capacity("too many bodies for the AVBD buffers", n);
        if (m > ctx->m_cap) capacity("too many contacts for the AVBD buffers", m);
        if (nidx > 2 * ctx->m_cap) capacity("too many CSR indices", nidx);
        if (c > ctx->c_cap) capacity("too many constraints", c);
        if (options->iterations == 0) invalid("iterations must be at least 1");
        ensure_granular_avbd(ctx);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->a_n = n;
        ctx->a_m = m;
        ctx->a_iters = options->iterations;
        ctx->a_submitted = true;
        if (n == 0) return DIO_OK;  // gpu_physics.rs:462-464
        if (!state || !positions || !orientations || !offsets || (m > 0 && !contacts) || (nidx > 0 && !indices) ||
            (c > 0 && !constraints))
            invalid("null input");
        if (offsets[n] != nidx) invalid("offsets[n] must equal nidx");
        for (uint32_t k = 0; k < m; ++k)
            if (contacts[k].a >= n || (contacts[k].b != DIO_CONTAINER && contacts[k].b >= n))
                invalid("contact body index out of range");
        for (uint32_t k = 0; k < c; ++k)
            if (constraints[k].kind > kRope)
                invalid("hard and bending joints need a world (dio_world_upload): hard joints' multipliers persist "
                        "between steps");
        for (uint32_t k = 0; k < nidx; ++k)
            if (indices[k] >= m) invalid("CSR index out of range");
        for (uint32_t i = 0; i < n; ++i)
            if (offsets[i] > offsets[i + 1]) invalid("CSR offsets must be non-decreasing");
        std::memcpy(ctx->a_state_host, state, (size_t)n * sizeof(dio_avbd_body_state));
        for (uint32_t i = 0; i < n; ++i) {
            float* p = ctx->a_pose_host + 8 * (size_t)i;
            p[0] = positions[3 * i]; p[1] = positions[3 * i + 1]; p[2] = positions[3 * i + 2]; p[3] = 0.0f;
            p[4] = orientations[4 * i]; p[5] = orientations[4 * i + 1];
            p[6] = orientations[4 * i + 2]; p[7] = orientations[4 * i + 3];
        }
        if (m > 0) std::memcpy(ctx->a_con_host, contacts, (size_t)m * sizeof(dio_avbd_contact));
        std::memcpy(ctx->a_csr_host, offsets, ((size_t)n + 1) * 4);
        if (nidx > 0) std::memcpy(ctx->a_csr_host + n + 1, indices, (size_t)nidx * 4);
        hipStream_t s = ctx->stream;
        DIO_HIP_CHECK(hipMemcpyAsync(ctx->a_state, ctx->a_state_host, (size_t)n * sizeof(dio_avbd_body_state),
                                     hipMemcpyHostToDevice, s));
        DIO_HIP_CHECK(hipMemcpyAsync(ctx->a_pose, ctx->a_pose_host, (size_t)n * 32, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_unpack_state, dim3(blocks256(n)), dim3(256), 0, s, ctx->a_state, ctx->a_pose, n,
                           ctx->a_vel, ctx->a_ang, ctx->a_prev, ctx->a_inv_mm, ctx->a_cur_pos, ctx->a_cur_rot);
        if (m > 0) {
            DIO_HIP_CHECK(hipMemcpyAsync(ctx->a_con_rec, ctx->a_con_host, (size_t)m * sizeof(dio_avbd_contact),
                                         hipMemcpyHostToDevice, s));
            launch_unpack_contacts(ctx->a_con_rec, m, ctx->a_c, s);
        }
        DIO_HIP_CHECK(hipMemcpyAsync(ctx->a_off, ctx->a_csr_host, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, s));
        if (nidx > 0)
            DIO_HIP_CHECK(hipMemcpyAsync(ctx->a_idx, ctx->a_csr_host + n + 1, (size_t)nidx * 4, hipMemcpyHostToDevice,
                                         s));
        if (c > 0) std::memcpy(ctx->a_cons_host, constraints, (size_t)c * sizeof(dio_avbd_constraint));
        upload_joints(ctx, ctx->a_j, ctx->a_cons_host, c, n, ctx->a_cons_stage);
        SolveView v{};
        v.cur_pos = ctx->a_cur_pos;
        v.cur_rot = ctx->a_cur_rot;
        v.pos[0] = ctx->a_pos[0]; v.pos[1] = ctx->a_pos[1];
        v.rot[0] = ctx->a_rot[0]; v.rot[1] = ctx->a_rot[1];
        v.init_lin = ctx->a_init_lin; v.init_ang = ctx->a_init_ang;
        v.iner_lin = ctx->a_iner_lin; v.iner_ang = ctx->a_iner_ang;
        v.vel = ctx->a_vel; v.ang = ctx->a_ang; v.prev_vel = ctx->a_prev; v.inv_mm = ctx->a_inv_mm;
        v.c_ab = ctx->a_c.ab; v.c_n = ctx->a_c.n; v.c_t1 = ctx->a_c.t1; v.c_t2 = ctx->a_c.t2;
        v.c_ra = ctx->a_c.ra; v.c_rb = ctx->a_c.rb; v.c_c0 = ctx->a_c.c0; v.c_pen = ctx->a_c.pen;
        v.c_lam = ctx->a_c.lam;
        v.c_lam_alt = ctx->a_lam_alt;
        v.c_pen_alt = ctx->a_pen_alt;
        v.m_dev = nullptr;
        v.m_host = m;
        v.offsets = ctx->a_off;
        v.indices = ctx->a_idx;
        v.peers = ctx->a_peer;
        launch_csr_peers(n, ctx->a_off, ctx->a_idx, ctx->a_c.ab, ctx->a_peer, s);
        v.j_off = ctx->a_j.off; v.j_idx = ctx->a_j.idx; v.j_abk = ctx->a_j.abk;
        v.j_anc_a = ctx->a_j.anc_a; v.j_anc_b = ctx->a_j.anc_b;
        v.out_pos = nullptr;
        v.out_rot = nullptr;
        const AvbdParams p = params_of(options->dt, options->gravity, options->alpha, options->beta_lin,
                                       options->penalty_max, n);
        schedule_avbd(v, p, options->iterations, m, s, nullptr, ctx->tuning);
        DIO_HIP_CHECK(hipGetLastError());
        return DIO_OK;
    });
}

dio_status dio_avbd_read(dio_ctx* ctx, float* positions, float* orientations, float* velocities,
                         float* angular_velocities, float* prev_velocities, float* lambda, float* penalty) {
    return guarded(ctx, [&]() {
        if (!ctx->a_submitted) throw StatusError{DIO_NOT_READY, "no AVBD solve submitted", 0};
        const uint32_t n = ctx->a_n, m = ctx->a_m;
        if (n == 0) return DIO_OK;
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        const uint32_t target = 1 - (ctx->a_iters - 1) % 2;  // gpu_physics.rs:584-585
        std::vector<float4> buf(std::max(n, m));
        auto get3 = [&](const float4* src, uint32_t count, float* dst) {
            if (!dst || count == 0) return;
            DIO_HIP_CHECK(hipMemcpy(buf.data(), src, (size_t)count * 16, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < count; ++i) {
                dst[3 * i] = buf[i].x; dst[3 * i + 1] = buf[i].y; dst[3 * i + 2] = buf[i].z;
            }
        };
        get3(ctx->a_pos[target], n, positions);
        if (orientations) {
            DIO_HIP_CHECK(hipMemcpy(orientations, ctx->a_rot[target], (size_t)n * 16, hipMemcpyDeviceToHost));
        }
        get3(ctx->a_vel, n, velocities);
        get3(ctx->a_ang, n, angular_velocities);
        get3(ctx->a_prev, n, prev_velocities);
        get3(ctx->a_c.lam, m, lambda);
        get3(ctx->a_c.pen, m, penalty);
        return DIO_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------ fused world
namespace {

void ensure_world(dio_ctx* c) {
    if (c->w_ready) return;
    DevArena& A = c->dev;
    const uint32_t n = c->n_cap, m = c->m_cap, p = c->p_max;
    c->w_pos = A.alloc<float4>(n);
    c->w_rot = A.alloc<float4>(n);
    c->w_vel = A.alloc<float4>(n);
    c->w_ang = A.alloc<float4>(n);
    c->w_prev = A.alloc<float4>(n);
    c->w_inv_mm = A.alloc<float4>(n);
    c->w_tag = A.alloc<uint32_t>(n);
    c->w_corners = A.alloc<float4>((size_t)n * 4);
    c->w_hull = A.alloc<float4>(c->h_cap);
    c->w_scene_of = A.alloc<uint32_t>(n);
    c->w_scene_off = A.alloc<uint32_t>((size_t)n + 1);
    c->w_scene_ct = A.alloc<ContainerParams>(n);
    c->w_scene_order = A.alloc<uint32_t>(n);
    c->w_bundle_off = A.alloc<uint32_t>((size_t)n + 1);
    c->w_scene_q = A.alloc<uint32_t>(2);
    c->w_j.abk = A.alloc<uint4>(c->c_cap);
    c->w_j.anc_a = A.alloc<float4>(c->c_cap);
    c->w_j.anc_b = A.alloc<float4>(c->c_cap);
    c->w_j.off = A.alloc<uint32_t>((size_t)n + 1);
    c->w_j.idx = A.alloc<uint32_t>(2 * (size_t)c->c_cap);
    for (int k = 0; k < 2; ++k) {
        c->w_j.lam[k] = A.alloc<float4>(c->c_cap);
        c->w_j.pen[k] = A.alloc<float4>(c->c_cap);
    }
    c->w_j.st_lam = A.alloc<float4>(c->c_cap);
    c->w_j.st_pen = A.alloc<float4>(c->c_cap);
    c->wb = alloc_broad(A, n, p);
    for (int k = 0; k < 2; ++k) {
        auto& pd = c->pend[k];
        pd.pairs = A.alloc<uint2>(p);
        pd.p_dev = A.alloc<uint32_t>(1);
        pd.out = alloc_narrow(A, p);
        pd.out.pairs = pd.pairs;
        pd.out.p_dev = pd.p_dev;
        pd.simplex = A.alloc<float4>((size_t)p * 8);
        ContactSet& cs = c->cs[k];
        cs.c = alloc_contacts(A, m);
        cs.m_dev = A.alloc<uint32_t>(1);
        cs.m_narrow = A.alloc<uint32_t>(2);
        cs.cont_cnt = A.alloc<uint32_t>(n);
        cs.cont_off = A.alloc<uint32_t>((size_t)n + 1);
        cs.m_cap = m;
    }
    c->csr.cnt = A.alloc<uint32_t>(n);
    c->csr.cur = A.alloc<uint32_t>(n);
    c->csr.off = A.alloc<uint32_t>((size_t)n + 1);
    c->csr.idx = A.alloc<uint32_t>(2 * (size_t)m);
    c->csr.peer = A.alloc<uint32_t>(2 * (size_t)m);
    c->csr.tmp = A.alloc<uint32_t>(2 * (size_t)m);
    c->csr.scan_partials = A.alloc<uint32_t>(scan_partials_size(n + 1));
    c->csr.idx_cap = 2 * m;
    for (int set = 0; set < 2; ++set) {
        InitTarget& t = c->ss[set];
        for (int k = 0; k < 2; ++k) {
            t.pos[k] = A.alloc<float4>(n);
            t.rot[k] = A.alloc<float4>(n);
        }
        t.init_lin = A.alloc<float4>(n);
        t.init_ang = A.alloc<float4>(n);
        t.iner_lin = A.alloc<float4>(n);
        t.iner_ang = A.alloc<float4>(n);
    }
    c->w_overflow = A.alloc<uint32_t>(2);
    c->w_acc = A.alloc<unsigned long long>(4);
    c->w_lam_alt = A.alloc<float4>(m);
    c->w_pen_alt = A.alloc<float4>(m);
    c->w_work = A.alloc<uint4>(primal_work_entries(n));
    c->w_items = A.alloc<uint2>(primal_item_entries(n));
    c->inst_scale = A.alloc<float4>(n);
    c->w_stage_b = A.alloc<dio_gpu_body>(n);
    c->w_stage_s = A.alloc<dio_avbd_body_state>(n);
    c->w_stage_c = A.alloc<dio_avbd_constraint>(c->c_cap);
    colour_buffers_bind(c->wc, A.alloc<uint32_t>(colour_buffer_words(n)), n);
    c->wc_host = c->pin.alloc<uint32_t>(kColourRoundSlots + kMaxColours + 2);
    c->col_starts = A.alloc<uint32_t>(kColourStartsWords);
    c->w_ready = true;
}

NarrowSource source_of(const dio_ctx::Pending& p) {
    return NarrowSource{p.pairs, p.out.off, p.p_dev, p.out.n, p.out.pa, p.out.pb};
}

// Copy what the contact build reads of pending buffer `from` (the pair count,
// pairs, per-pair contact offsets and the four slots of each pair) into `to`.
void copy_pending(dio_ctx* c, int from, int to) {
    hipStream_t st = c->stream;
    const dio_ctx::Pending& a = c->pend[from];
    const dio_ctx::Pending& b = c->pend[to];
    uint32_t np = 0;
    DIO_HIP_CHECK(hipStreamSynchronize(st));
    DIO_HIP_CHECK(hipMemcpy(&np, a.p_dev, 4, hipMemcpyDeviceToHost));
    np = std::min(np, c->p_max);
    DIO_HIP_CHECK(hipMemcpyAsync(b.p_dev, a.p_dev, 4, hipMemcpyDeviceToDevice, st));
    DIO_HIP_CHECK(hipMemcpyAsync(b.out.off, a.out.off, ((size_t)np + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (np > 0) {
        DIO_HIP_CHECK(hipMemcpyAsync(b.pairs, a.pairs, (size_t)np * 8, hipMemcpyDeviceToDevice, st));
        DIO_HIP_CHECK(hipMemcpyAsync(b.out.n, a.out.n, (size_t)np * 64, hipMemcpyDeviceToDevice, st));
        DIO_HIP_CHECK(hipMemcpyAsync(b.out.pa, a.out.pa, (size_t)np * 64, hipMemcpyDeviceToDevice, st));
        DIO_HIP_CHECK(hipMemcpyAsync(b.out.pb, a.out.pb, (size_t)np * 64, hipMemcpyDeviceToDevice, st));
    }
    DIO_HIP_CHECK(hipStreamSynchronize(st));
}

// The world's solve over buffer set s; prepare_next: the final launch also
// runs the next step's init into set 1 - s.
SolveView world_solve_view(dio_ctx* c, int s, bool prepare_next = false);

// Enqueue segment `seg` of one step with parity s (0: contacts + broad +
// narrow, 1: AVBD).
void enqueue_segment(dio_ctx* c, int seg, int s, const dio_avbd_options& o, const dio_step_opts& so,
                     KernelTimer* kt = nullptr, bool resort = true) {
    hipStream_t st = c->stream;
    const uint32_t n = c->w_n;
    WorldView wv{c->w_pos, c->w_rot, c->w_vel, c->w_inv_mm, c->w_tag, c->w_corners, c->w_hull, n};
    SceneTable scenes;
    if (c->w_scene_bits) {
        wv.scene_of = c->w_scene_of;
        wv.scene_ct = c->w_scene_ct;
        scenes.of = c->w_scene_of;
        scenes.off = c->w_scene_off;
        scenes.bits = c->w_scene_bits;
    }
    BodyView bv{c->w_pos, c->w_rot, c->w_tag, c->w_corners, c->w_hull, c->w_hull_count};
    BuildOptions bo{o.dt, o.alpha, o.margin, o.penalty_min, o.penalty_max, o.gamma, o.friction};
    const int lag = so.contact_lag ? 1 : 0;
    if (seg == 0) {
        // lag 1: contacts come from pend[s] (previous step's narrow phase), the
        // broad / narrow output of this step goes to pend[1 - s].  lag 0: this
        // step's narrow output goes to pend[s] and is used at once.
        const int src = s;
        const int dst_pend = lag ? 1 - s : s;
        const bool fork = lag && c->tuning.overlap_broad;
        // On a refit step with lag 1 on one stream, the broad phase's first
        // refit launch rides on the contact build's first launch: it reads only
        // the poses and the sorted order, and writes only boxes.
        const bool carry_refit = lag && !fork && !resort && n > 0;
        auto broad_narrow = [&](hipStream_t bs) {
            BroadBuffers b = c->wb;
            b.pairs = c->pend[dst_pend].pairs;
            b.p_dev = c->pend[dst_pend].p_dev;
            b.overflow = c->w_overflow;
            b.clear_words[0] = c->pend[dst_pend].out.gjk_count;
            b.clear_words[1] = c->pend[dst_pend].out.epa_retry_count;
            run_broad(bv, n, b, c->p_max, bs, resort, (int)traversal_lanes(c->tuning.broad_lanes, n), scenes,
                      carry_refit);
            run_narrow(bv, c->pend[dst_pend].out, c->pend[dst_pend].simplex, bs, true, c->w_gjk_possible);
        };
        if (!lag) broad_narrow(st);
        const TreeLower refit = tree_lower_args(bv, n, c->wb);
        run_contact_pairs(wv, c->w_ct, bo, source_of(c->pend[src]), c->cs[s], source_of(c->pend[1 - s]), c->cs[1 - s],
                          c->w_overflow + 1, st, c->csr.cnt, carry_refit ? &refit : nullptr);
        // With lag 1, the next step's broad and narrow phases read only the
        // poses, which nothing before the AVBD segment writes, and write
        // pend[1 - s], which the contact build has finished reading: with
        // overlap_broad they run on the second stream beside the container
        // contacts and CSR (both chains are short launches on few
        // workgroups), joined before the segment ends.
        if (fork) {
            DIO_HIP_CHECK(hipEventRecord(c->ev_fork, st));
            DIO_HIP_CHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
            broad_narrow(c->stream2);
        }
        run_contact_walls(wv, c->w_ct, bo, c->cs[s], c->cs[1 - s], c->scan_partials, c->w_overflow + 1, st,
                          c->csr.cnt);
        run_csr(c->cs[s], n, c->csr, c->w_overflow + 1, st);
        if (c->adj_on)
            launch_record_step(c->adj_rec, c->cs[s], c->csr.off, c->csr.idx, c->csr.peer,
                               BodyState{c->w_pos, c->w_rot, c->w_vel, c->w_ang, c->w_prev}, n, c->m_cap, st);
        if (fork) {
            DIO_HIP_CHECK(hipEventRecord(c->ev_join, c->stream2));
            DIO_HIP_CHECK(hipStreamWaitEvent(st, c->ev_join, 0));
        } else if (lag) {
            broad_narrow(st);
        }
    } else {
        // the init of this step ran already (the previous step's final launch,
        // or dio_step before this segment)
        SolveView v = world_solve_view(c, s, true);
        const int bn_pend = so.contact_lag ? 1 - s : s;
        v.stat_pairs = c->pend[bn_pend].p_dev;
        v.stat_acc = c->w_acc;
        v.work_tab = c->w_work;  // iteration 0 of the solve below fills it
        if (c->tuning.primal_packed == 2 && c->w_j.c == 0) v.span_primal = true;  // worlds without joints
        else if (c->tuning.primal_packed) v.item_tab = c->w_items;
        const AvbdParams p = params_of(o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, n, o.gamma, o.penalty_min);
        if (c->tuning.scene_solve && c->w_scene_bits && c->w_j.c == 0) {
            if (c->w_bundle_bodies != c->tuning.scene_solve)
                throw StatusError{DIO_HIP, "internal: scene bundles not built for this tuning", 0};
            const SceneQueue q{c->w_scene_order, c->w_bundle_off, c->w_bundles, c->w_scene_q};
            const bool fused = c->tuning.fuse_dual != 0u;
            auto launch = [&](const LaunchEvents& ev) { launch_avbd_scenes(v, p, o.iterations, fused, q, st, ev); };
            if (kt) kt->time(1, launch);
            else launch(LaunchEvents{});
        } else {
            schedule_avbd(v, p, o.iterations, c->m_cap, st, kt, c->tuning, false);
        }
    }
}

// The AVBD segment of a coloured step: colour the step's body graph, read
// back the colour counts (the one host synchronisation of the mode), then
// launch the solve colour by colour.
// The step's body graph for the colouring.
ColourGraph colour_graph_of(const dio_ctx* c) {
    ColourGraph g;
    g.n = c->w_n;
    g.inv_mm = c->w_inv_mm;
    g.off = c->csr.off;
    g.peer = c->csr.peer;
    if (c->w_j.c > 0) {
        g.j_off = c->w_j.off;
        g.j_idx = c->w_j.idx;
        g.j_abk = c->w_j.abk;
    }
    if (c->w_scene_bits) {
        g.scene_of = c->w_scene_of;
        g.scene_off = c->w_scene_off;
    }
    return g;
}

void run_coloured_segment(dio_ctx* c, int s, const dio_avbd_options& o, const dio_step_opts& so, KernelTimer* kt) {
    hipStream_t st = c->stream;
    const uint32_t n = c->w_n;
    const ColourGraph g = colour_graph_of(c);
    launch_colour_init(g, c->wc, st);
    // the first batch covers the previous coloured step's rounds and two more
    // (the graph changes little from step to step); later batches 16 each
    uint32_t r = 0, batch = c->last_rounds ? std::min<uint32_t>(24u, std::max<uint32_t>(4u, c->last_rounds + 2)) : 24u;
    uint32_t* h = c->wc_host;
    for (;;) {
        if (r + batch > kColourRoundSlots - 1) capacity("the colouring needs more rounds than it can track", r);
        launch_colour_rounds(g, c->wc, r, batch, st);
        DIO_HIP_CHECK(hipMemcpyAsync(h, c->wc.pending + r, (size_t)batch * 4, hipMemcpyDeviceToHost, st));
        DIO_HIP_CHECK(hipStreamSynchronize(st));
        uint32_t k = 0;
        while (k < batch && h[k] != 0u) ++k;
        r += k;
        if (k < batch) break;
        batch = 16;
    }
    c->last_rounds = n > 0 ? r + 1 : 0;
    const int lb = launch_colour_lists(g, c->wc, st);
    DIO_HIP_CHECK(hipMemcpyAsync(h, c->wc.counts, (kMaxColours + 1) * 4, hipMemcpyDeviceToHost, st));
    DIO_HIP_CHECK(hipMemcpyAsync(h + kMaxColours + 1, c->wc.pending + kColourRoundSlots - 1, 4,
                                 hipMemcpyDeviceToHost, st));
    DIO_HIP_CHECK(hipStreamSynchronize(st));
    if (h[kMaxColours + 1]) capacity("a body needs more than 254 colours", kMaxColours);
    uint32_t ncol = 0;
    for (uint32_t k = 0; k < kMaxColours; ++k)
        if (h[k]) ncol = k + 1;
    std::vector<uint32_t> start(ncol + 1, 0);
    for (uint32_t k = 0; k < ncol; ++k) start[k + 1] = start[k] + h[k];
    c->last_colours = ncol;
    SolveView v = world_solve_view(c, s);
    const int bn_pend = so.contact_lag ? 1 - s : s;
    v.stat_pairs = c->pend[bn_pend].p_dev;
    v.stat_acc = c->w_acc;
    const AvbdParams p = params_of(o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, n, o.gamma, o.penalty_min);
    schedule_avbd_coloured(v, p, o.iterations, c->m_cap, st, kt, c->tuning, c->wc.vals[lb], start);
}

// The coloured AVBD segment with no host synchronisation, for capture: the
// first col_rcap colouring rounds as separate launches (each exits at once
// when the colouring is done) and one workgroup finishing any later rounds;
// the per-colour ranges computed on the device; then per iteration one
// launch for each of colours 0 .. col_ccap - 1 (a grid for every body,
// workgroups past the colour's end exiting at once) and one workgroup
// running any colours past those, one by one.  The caps come from an
// uncaptured step, so the finishing launches do nothing unless the
// colouring needs more rounds or colours than then; the results are the
// uncaptured step's either way.
void enqueue_coloured_captured(dio_ctx* c, int s, const dio_avbd_options& o, const dio_step_opts& so) {
    hipStream_t st = c->stream;
    const uint32_t n = c->w_n;
    const ColourGraph g = colour_graph_of(c);
    launch_colour_init(g, c->wc, st);
    launch_colour_rounds(g, c->wc, 0, c->col_rcap, st);
    launch_colour_finish(g, c->wc, c->col_rcap, st);
    const int lb = launch_colour_lists(g, c->wc, st);
    launch_colour_starts(c->wc, c->col_starts, st);
    SolveView v = world_solve_view(c, s);
    const int bn_pend = so.contact_lag ? 1 - s : s;
    v.stat_pairs = c->pend[bn_pend].p_dev;
    v.stat_acc = c->w_acc;
    const AvbdParams p = params_of(o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, n, o.gamma, o.penalty_min);
    const uint32_t* list = c->wc.vals[lb];
    const int lanes = colour_lanes_of(c->tuning);
    v.pos[1] = v.pos[0];
    v.rot[1] = v.rot[0];
    launch_avbd_init(v, p, st);
    for (uint32_t k = 0; k < o.iterations; ++k) {
        for (uint32_t col = 0; col < c->col_ccap; ++col) {
            SolveView vc = v;
            vc.body_list = list;
            vc.list_range = c->col_starts + col;
            launch_avbd_primal_colour_range(vc, p, k, st, lanes);
        }
        SolveView vt = v;
        vt.body_list = list;
        launch_avbd_colour_tail(vt, p, k, c->col_starts, c->col_ccap, kMaxColours + 1, st);
        if (k + 1 < o.iterations) launch_avbd_dual(v, p, k & 1u, 0u, 0u, c->m_cap, st);
    }
    launch_avbd_dual_recover(v, p, o.iterations, false, c->m_cap, st);
}

// Capture the coloured segment of both parities with caps from the last
// uncaptured coloured step: its rounds plus 4 (at least 8) and its colours
// plus 2.
void build_colour_graphs(dio_ctx* c, const dio_avbd_options& o, const dio_step_opts& so) {
    for (auto& g : c->col_graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    c->col_valid = false;
    c->col_rcap = std::max<uint32_t>(8u, c->last_rounds + 4u);
    c->col_ccap = c->last_colours + 2u;
    DIO_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int s = 0; s < 2; ++s) {
        hipGraph_t g = nullptr;
        DIO_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_coloured_captured(c, s, o, so);
        DIO_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
        hipGraphExec_t ge = nullptr;
        hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DIO_HIP_CHECK(e);
        c->col_graphs[s] = ge;
    }
    c->col_opts = o;
    c->col_sopts = so;
    c->col_valid = true;
}

// After captured coloured steps: the last step's colour and round counts
// (read back once the stream is idle); caps that were passed invalidate the
// graphs, so the next call captures with larger ones.
void read_colour_summary(dio_ctx* c) {
    uint32_t* h = c->wc_host;
    DIO_HIP_CHECK(hipStreamSynchronize(c->stream));
    DIO_HIP_CHECK(hipMemcpy(h, c->wc.pending, (size_t)kColourRoundSlots * 4, hipMemcpyDeviceToHost));
    DIO_HIP_CHECK(hipMemcpy(h + kColourRoundSlots, c->col_starts, (size_t)kColourStartsWords * 4,
                            hipMemcpyDeviceToHost));
    if (h[kColourRoundSlots - 1])
        capacity("the colouring needs more rounds than it can track, or a body more than 254 colours", kMaxColours);
    uint32_t r = 0;
    while (r < kColourRoundSlots - 1 && h[r] != 0u) ++r;
    c->last_rounds = c->w_n > 0 ? r + 1 : 0;
    c->last_colours = h[kColourRoundSlots + kMaxColours + 1];
    if (c->last_rounds > c->col_rcap || c->last_colours > c->col_ccap) c->col_valid = false;
}

// The per-scene solve's bundles for tuning.scene_solve (a batched world
// only): consecutive scenes are joined until a bundle holds at least that
// many bodies (a larger scene is a bundle of its own), so the workgroups of
// small scenes have work for their lanes; bundles are taken most bodies
// first.  Synchronous (outside graph capture).
void build_bundles(dio_ctx* c) {
    const uint32_t want = c->tuning.scene_solve;
    if (!c->w_scene_bits || want == 0 || c->w_bundle_bodies == want) return;
    const std::vector<uint32_t>& so = c->w_scene_off_host;
    std::vector<uint32_t> off{so[0]};
    for (size_t k = 1; k < so.size(); ++k)
        if (so[k] - off.back() >= want || (k + 1 == so.size() && so[k] > off.back())) off.push_back(so[k]);
    const uint32_t count = (uint32_t)off.size() - 1;
    std::vector<uint32_t> order(count);
    for (uint32_t k = 0; k < count; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return off[x + 1] - off[x] > off[y + 1] - off[y]; });
    DIO_HIP_CHECK(hipStreamSynchronize(c->stream));
    DIO_HIP_CHECK(hipMemcpy(c->w_bundle_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    if (count > 0) DIO_HIP_CHECK(hipMemcpy(c->w_scene_order, order.data(), (size_t)count * 4, hipMemcpyHostToDevice));
    c->w_bundles = count;
    c->w_bundle_bodies = want;
}

// Whether buffer set s holds this step's init, prepared for these options.
bool init_prepared(const dio_ctx* c, int s, const dio_avbd_options& o) {
    return c->init_set == s && c->init_dt == o.dt && c->init_g[0] == o.gravity[0] && c->init_g[1] == o.gravity[1] &&
           c->init_g[2] == o.gravity[2] &&
           (c->w_j.hard == 0 || (c->init_alpha == o.alpha && c->init_gamma == o.gamma &&
                                 c->init_pmin == o.penalty_min && c->init_pmax == o.penalty_max));
}

// The standalone init of buffer set s from the world state.
void run_world_init(dio_ctx* c, int s, const dio_avbd_options& o, KernelTimer* kt) {
    const SolveView v = world_solve_view(c, s);
    const AvbdParams p = params_of(o.dt, o.gravity, o.alpha, o.beta_lin, o.penalty_max, c->w_n, o.gamma,
                                   o.penalty_min);
    if (kt) kt->time(0, [&](const LaunchEvents& ev) { launch_avbd_init(v, p, c->stream, ev); });
    else launch_avbd_init(v, p, c->stream);
}

SolveView world_solve_view(dio_ctx* c, int s, bool prepare_next) {
    {
        SolveView v{};
        v.cur_pos = c->w_pos;
        v.cur_rot = c->w_rot;
        const InitTarget& t = c->ss[s];
        v.pos[0] = t.pos[0]; v.pos[1] = t.pos[1];
        v.rot[0] = t.rot[0]; v.rot[1] = t.rot[1];
        v.init_lin = t.init_lin; v.init_ang = t.init_ang;
        v.iner_lin = t.iner_lin; v.iner_ang = t.iner_ang;
        if (prepare_next) v.next = c->ss[1 - s];
        v.vel = c->w_vel; v.ang = c->w_ang; v.prev_vel = c->w_prev; v.inv_mm = c->w_inv_mm;
        const ContactSet& cs = c->cs[s];
        v.c_ab = cs.c.ab; v.c_n = cs.c.n; v.c_t1 = cs.c.t1; v.c_t2 = cs.c.t2; v.c_ra = cs.c.ra; v.c_rb = cs.c.rb;
        v.c_c0 = cs.c.c0; v.c_pen = cs.c.pen; v.c_lam = cs.c.lam;
        v.derived_frame = 1;  // the world's contacts come from the device's contact build
        v.c_lam_alt = c->w_lam_alt; v.c_pen_alt = c->w_pen_alt;
        v.m_dev = cs.m_dev;
        v.m_host = 0;
        v.offsets = c->csr.off;
        v.indices = c->csr.idx;
        v.peers = c->csr.peer;
        v.j_off = c->w_j.off; v.j_idx = c->w_j.idx; v.j_abk = c->w_j.abk;
        v.j_anc_a = c->w_j.anc_a; v.j_anc_b = c->w_j.anc_b;
        v.j_count = c->w_j.c;
        v.j_hard = c->w_j.extended > 0 ? 1u : 0u;  // the primal with the hard and bending rows
        if (c->w_j.hard > 0) {
            v.j_lam[0] = c->w_j.lam[0]; v.j_lam[1] = c->w_j.lam[1];
            v.j_pen[0] = c->w_j.pen[0]; v.j_pen[1] = c->w_j.pen[1];
            v.j_st_lam = c->w_j.st_lam; v.j_st_pen = c->w_j.st_pen;
        }
        v.out_pos = c->w_pos;
        v.out_rot = c->w_rot;
        // with primal_order 1 the Jacobi primal takes the bodies in the broad
        // phase's Morton order (each body's result does not depend on the
        // order), so bodies in contact run close together (a coloured step
        // replaces the list)
        if (c->w_n > 0 && c->tuning.primal_order == 1) {
            v.body_list = c->wb.perm;
            v.list_count = c->w_n;
        }
        return v;
    }
}

bool same_opts(const dio_avbd_options& a, const dio_avbd_options& b, const dio_step_opts& x,
               const dio_step_opts& y) {
    return std::memcmp(&a, &b, sizeof(a)) == 0 && std::memcmp(&x, &y, sizeof(x)) == 0;
}

bool graphs_disabled() {
    const char* e = std::getenv("DIO_NO_GRAPH");
    return e && e[0] == '1';
}

void build_graphs(dio_ctx* c, const dio_avbd_options& o, const dio_step_opts& so) {
    c->drop_graphs();
    for (int var = 0; var < 3; ++var)
        for (int s = 0; s < 2; ++s) {
            const int seg = var == 1 ? 1 : 0;
            hipGraph_t g = nullptr;
            DIO_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_segment(c, seg, s, o, so, nullptr, var == 2);
            DIO_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
            hipGraphExec_t ge = nullptr;
            hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            DIO_HIP_CHECK(e);
            c->graphs[var][s] = ge;
        }
    c->g_opts = o;
    c->g_sopts = so;
    c->graphs_valid = true;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {
ContainerParams container_params(const dio_container* container) {
    ContainerParams ct{};
    if (container && container->enabled) {
        ct.cx = container->center[0]; ct.cy = container->center[1]; ct.cz = container->center[2];
        ct.qx = container->rotation[0]; ct.qy = container->rotation[1];
        ct.qz = container->rotation[2]; ct.qw = container->rotation[3];
        ct.half = container->half_extent;
        ct.enabled = 1;
    }
    return ct;
}

// The world upload; scene_offsets / containers describe a batched world
// (num_scenes > 1) or are null (one scene, one optional container).
dio_status world_upload(dio_ctx* ctx, const dio_gpu_body* bodies, const dio_avbd_body_state* state, uint32_t n,
                        const dio_avbd_constraint* constraints, uint32_t c, const uint32_t* scene_offsets,
                        uint32_t num_scenes, const dio_container* containers) {
    if (n > ctx->n_cap) capacity("too many bodies for the context", n);
    if (c > ctx->c_cap) capacity("too many constraints for the context", c);
    if ((n > 0 && (!bodies || !state)) || (c > 0 && !constraints)) invalid("null input");
    std::vector<uint32_t> scene_of;
    if (scene_offsets) {
        if (num_scenes == 0 || num_scenes > 65536u || num_scenes > ctx->n_cap) invalid("num_scenes out of range");
        if (!containers) invalid("containers is null");
        if (scene_offsets[0] != 0 || scene_offsets[num_scenes] != n) invalid("scene offsets must run from 0 to n");
        scene_of.resize(n);
        for (uint32_t k = 0; k < num_scenes; ++k) {
            if (scene_offsets[k + 1] < scene_offsets[k]) invalid("scene offsets must not decrease");
            for (uint32_t i = scene_offsets[k]; i < scene_offsets[k + 1]; ++i) scene_of[i] = k;
        }
        for (uint32_t k = 0; k < c; ++k) {
            const uint32_t a = constraints[k].a, b = constraints[k].b;
            if (a >= n || b >= n || scene_of[a] != scene_of[b]) invalid("a constraint joins two scenes");
        }
    }
    const uint64_t hull_need = hull_extent(bodies, n);
    ensure_world(ctx);
    hipStream_t s = ctx->stream;
    DIO_HIP_CHECK(hipStreamSynchronize(s));
    ctx->drop_graphs();
    ctx->free_adjoint();  // a recording belongs to the world it was made in
    ctx->init_set = -1;
    ctx->w_n = n;
    ctx->inst_count = 0;  // instance slots name the previous world's bodies
    ctx->w_hull_need = hull_need;
    ctx->w_step = 0;
    ctx->w_sort_pending = true;
    ctx->w_friction.resize(n);
    ctx->w_n_dyn = 0;
    bool all_box = true, all_sphere = true;
    for (uint32_t i = 0; i < n; ++i) {
        ctx->w_friction[i] = state[i].friction;
        if (state[i].inv_mass != 0.0f) ctx->w_n_dyn += 1;
        const uint32_t tag = bodies[i].shape.tag;
        all_box = all_box && (tag == kCube || tag == kBox);
        all_sphere = all_sphere && tag == kSphere;
    }
    ctx->w_gjk_possible = !(all_box || all_sphere);
    if (n > 0) {
        dio_gpu_body* stage_b = ctx->w_stage_b;
        dio_avbd_body_state* stage_s = ctx->w_stage_s;
        DIO_HIP_CHECK(hipMemcpy(stage_b, bodies, (size_t)n * sizeof(dio_gpu_body), hipMemcpyHostToDevice));
        DIO_HIP_CHECK(hipMemcpy(stage_s, state, (size_t)n * sizeof(dio_avbd_body_state), hipMemcpyHostToDevice));
        launch_unpack_gpu_bodies(stage_b, n, ctx->w_pos, ctx->w_rot, ctx->w_tag, ctx->w_corners, s);
        hipLaunchKernelGGL(k_unpack_state, dim3(blocks256(n)), dim3(256), 0, s, stage_s, (const float*)nullptr, n,
                           ctx->w_vel, ctx->w_ang, ctx->w_prev, ctx->w_inv_mm, (float4*)nullptr, (float4*)nullptr);
        DIO_HIP_CHECK(hipStreamSynchronize(s));
    }
    upload_joints(ctx, ctx->w_j, constraints, c, n, ctx->w_stage_c);
    ctx->w_j.hard = 0;
    ctx->w_j.hard_dynamic = 0;
    ctx->w_j.extended = 0;
    for (uint32_t k = 0; k < c; ++k) {
        if (constraints[k].kind > kRope) ctx->w_j.extended += 1;
        if (!is_hard(constraints[k].kind)) continue;
        ctx->w_j.hard += 1;
        if (state[constraints[k].a].inv_mass != 0.0f && state[constraints[k].b].inv_mass != 0.0f &&
            constraints[k].a != constraints[k].b)
            ctx->w_j.hard_dynamic += 1;
    }
    if (c > 0) {
        // hard joints start with no force and the minimum penalty (the warm
        // start clamps a zero penalty up to penalty_min)
        DIO_HIP_CHECK(hipMemsetAsync(ctx->w_j.st_lam, 0, (size_t)c * sizeof(float4), s));
        DIO_HIP_CHECK(hipMemsetAsync(ctx->w_j.st_pen, 0, (size_t)c * sizeof(float4), s));
        DIO_HIP_CHECK(hipStreamSynchronize(s));
    }
    ctx->w_num_scenes = 1;
    ctx->w_scene_bits = 0;
    if (scene_offsets && num_scenes > 1) {
        std::vector<ContainerParams> cts(num_scenes);
        for (uint32_t k = 0; k < num_scenes; ++k) cts[k] = container_params(&containers[k]);
        if (n > 0) DIO_HIP_CHECK(hipMemcpy(ctx->w_scene_of, scene_of.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        DIO_HIP_CHECK(hipMemcpy(ctx->w_scene_off, scene_offsets, ((size_t)num_scenes + 1) * 4, hipMemcpyHostToDevice));
        DIO_HIP_CHECK(hipMemcpy(ctx->w_scene_ct, cts.data(), (size_t)num_scenes * sizeof(ContainerParams),
                                hipMemcpyHostToDevice));
        uint32_t bits = 1;
        while ((1u << bits) < num_scenes) ++bits;
        ctx->w_scene_off_host.assign(scene_offsets, scene_offsets + num_scenes + 1);
        ctx->w_bundle_bodies = 0;  // built for the tuning below
        launch_fill_u32(ctx->w_scene_q, 0u, 2, s);
        ctx->w_num_scenes = num_scenes;
        ctx->w_scene_bits = bits;
        ctx->w_ct = ContainerParams{};
        build_bundles(ctx);
    } else {
        ctx->w_ct = container_params(scene_offsets ? &containers[0] : containers);
    }
    // empty lagged narrow output and empty previous contact sets
    for (int k = 0; k < 2; ++k) {
        launch_fill_u32(ctx->pend[k].p_dev, 0u, 1, s);
        launch_fill_u32(ctx->pend[k].out.off, 0u, 1, s);
        launch_fill_u32(ctx->cs[k].m_dev, 0u, 1, s);
        launch_fill_u32(ctx->cs[k].m_narrow, 0u, 2, s);
        launch_fill_u32(ctx->cs[k].cont_off, 0u, n + 1, s);
    }
    launch_fill_u32(ctx->w_overflow, 0u, 2, s);
    // the CSR counters start at zero; every step's CSR build leaves them so
    if (ctx->n_cap > 0) launch_fill2_u32(ctx->csr.cnt, ctx->csr.cur, 0u, ctx->n_cap, s);
    launch_fill_u32(ctx->wb.err, 0u, 1, s);
    DIO_HIP_CHECK(hipStreamSynchronize(s));
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    return DIO_OK;
}
}  // namespace

extern "C" {

dio_status dio_world_upload(dio_ctx* ctx, const dio_gpu_body* bodies, const dio_avbd_body_state* state, uint32_t n,
                            const dio_avbd_constraint* constraints, uint32_t c, const dio_container* container) {
    return guarded(ctx, [&]() { return world_upload(ctx, bodies, state, n, constraints, c, nullptr, 1, container); });
}

dio_status dio_batch_upload(dio_ctx* ctx, const dio_gpu_body* bodies, const dio_avbd_body_state* state, uint32_t n,
                            const dio_avbd_constraint* constraints, uint32_t c, const uint32_t* scene_offsets,
                            uint32_t num_scenes, const dio_container* containers) {
    return guarded(ctx, [&]() {
        if (!scene_offsets) invalid("scene_offsets is null");
        return world_upload(ctx, bodies, state, n, constraints, c, scene_offsets, num_scenes, containers);
    });
}

dio_status dio_world_upload_hulls(dio_ctx* ctx, const float* vertices, uint32_t count) {
    return guarded(ctx, [&]() {
        if (count > ctx->h_cap) capacity("too many hull vertices", count);
        if (count > 0 && !vertices) invalid("vertices is null");
        ensure_world(ctx);
        ensure_granular_broad(ctx);
        std::vector<float4> v(count);
        for (uint32_t i = 0; i < count; ++i)
            v[i] = make_float4(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2], 0.0f);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (count > 0) DIO_HIP_CHECK(hipMemcpy(ctx->w_hull, v.data(), (size_t)count * 16, hipMemcpyHostToDevice));
        ctx->w_hull_count = count;
        ctx->g_hull = ctx->w_hull;  // the granular pipeline shares the pool
        ctx->drop_graphs();
        return DIO_OK;
    });
}

dio_status dio_step(dio_ctx* ctx, const dio_avbd_options* options, const dio_step_opts* opts, uint32_t nsteps) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (!options) invalid("options is null");
        if (options->iterations == 0) invalid("iterations must be at least 1");
        if (ctx->w_hull_need > ctx->w_hull_count)
            invalid("the world's hull bodies name " + std::to_string(ctx->w_hull_need) +
                    " hull vertices; the uploaded pool holds " + std::to_string(ctx->w_hull_count));
        dio_step_opts so{1, 0, 0};
        if (opts) so = *opts;
        if (so.mode > 1) invalid("mode must be 0 (Jacobi) or 1 (coloured Gauss-Seidel)");
        const bool coloured = so.mode == 1;
        if (coloured && ctx->adj_on) invalid("the adjoint records Jacobi steps (mode 0) only");
        if (!coloured && nsteps > 0 && ctx->w_j.hard_dynamic > 0)
            invalid("hard joints between two dynamic bodies need the coloured mode (1): the Jacobi solve "
                    "diverges on them");
        // After a lag-0 step, a lag-1 step builds its contacts from the narrow
        // output that step found on the poses at its start, as if it had run
        // with lag 1.  That output sits in pend[1 - s], where the step's warm
        // start also looks up the old contacts' pairs; its source is pend[s],
        // so the used part of the output is copied there once.
        if (nsteps > 0 && ctx->w_step > 0 && ctx->last_lag == 0 && so.contact_lag && ctx->w_n > 0)
            copy_pending(ctx, (int)(1 - ctx->w_step % 2), (int)(ctx->w_step % 2));
        if (nsteps > 0) ctx->last_lag = so.contact_lag ? 1 : 0;
        hipStream_t s = ctx->stream;
        DIO_HIP_CHECK(hipSetDevice(ctx->device));
        const bool use_graph = !graphs_disabled() && !so.profile;
        KernelTimer kt;
        kt.pool = &ctx->prof_events;
        if (use_graph && ctx->w_n > 0 && (!ctx->graphs_valid || !same_opts(ctx->g_opts, *options, ctx->g_sopts, so)))
            build_graphs(ctx, *options, so);
        if (ctx->events.size() < 2 * (size_t)nsteps + 2) {
            size_t have = ctx->events.size();
            ctx->events.resize(2 * (size_t)nsteps + 2);
            for (size_t k = have; k < ctx->events.size(); ++k) DIO_HIP_CHECK(hipEventCreate(&ctx->events[k]));
        }
        hipLaunchKernelGGL(k_stats_zero, dim3(1), dim3(64), 0, s, ctx->w_acc);
        DIO_HIP_CHECK(hipEventRecord(ctx->events[0], s));
        const bool host_timing = std::getenv("DIO_HOST_TIMING") != nullptr;
        const auto h0 = std::chrono::steady_clock::now();
        double h_launch = 0.0;
        bool colour_graph_last = false;  // the last step ran the captured coloured segment
        uint32_t colour_graph_run = 0;   // captured coloured steps since the last look at the caps
        for (uint32_t k = 0; k < nsteps && ctx->w_n > 0; ++k) {
            const int par = (int)(ctx->w_step % 2);
            const bool resort = ctx->w_sort_pending || ctx->w_step - ctx->w_sorted_at >= ctx->tuning.resort_interval;
            if (resort) {
                ctx->w_sorted_at = ctx->w_step;
                ctx->w_sort_pending = false;
            }
            // a Jacobi step's init is the previous step's final launch when that
            // prepared this buffer set with the same dt and gravity
            const bool need_init = !coloured && !init_prepared(ctx, par, *options);
            KernelTimer* tk = (!use_graph && so.profile) ? &kt : nullptr;
            if (use_graph) {
                const auto a = std::chrono::steady_clock::now();
                DIO_HIP_CHECK(hipGraphLaunch(ctx->graphs[resort ? 2 : 0][par], s));
                DIO_HIP_CHECK(hipEventRecord(ctx->events[2 + 2 * k], s));
                if (need_init) run_world_init(ctx, par, *options, nullptr);
                if (coloured) {
                    // the captured coloured segment once an uncaptured step has
                    // given its caps; an uncaptured step otherwise, which then
                    // captures it for the following steps
                    colour_graph_last = ctx->col_valid && same_opts(ctx->col_opts, *options, ctx->col_sopts, so);
                    if (colour_graph_last) {
                        DIO_HIP_CHECK(hipGraphLaunch(ctx->col_graphs[par], s));
                        // every 16 captured steps, one look at the colouring: when it
                        // has outgrown the caps (the extra rounds and colours then run
                        // in one workgroup), the next step captures with larger ones
                        if (++colour_graph_run == 16 && k + 1 < nsteps) {
                            read_colour_summary(ctx);
                            colour_graph_run = 0;
                        }
                    } else {
                        run_coloured_segment(ctx, par, *options, so, nullptr);
                        build_colour_graphs(ctx, *options, so);
                    }
                } else {
                    DIO_HIP_CHECK(hipGraphLaunch(ctx->graphs[1][par], s));
                }
                DIO_HIP_CHECK(hipEventRecord(ctx->events[3 + 2 * k], s));
                h_launch += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
            } else {
                enqueue_segment(ctx, 0, par, *options, so, nullptr, resort);
                DIO_HIP_CHECK(hipEventRecord(ctx->events[2 + 2 * k], s));
                if (need_init) run_world_init(ctx, par, *options, tk);
                if (coloured) run_coloured_segment(ctx, par, *options, so, tk);
                else enqueue_segment(ctx, 1, par, *options, so, tk);
                DIO_HIP_CHECK(hipEventRecord(ctx->events[3 + 2 * k], s));
            }
            if (coloured) {
                ctx->init_set = -1;
            } else {
                ctx->init_set = 1 - par;
                ctx->init_dt = options->dt;
                for (int q = 0; q < 3; ++q) ctx->init_g[q] = options->gravity[q];
                ctx->init_alpha = options->alpha;
                ctx->init_gamma = options->gamma;
                ctx->init_pmin = options->penalty_min;
                ctx->init_pmax = options->penalty_max;
            }
            ctx->w_step += 1;
        }
        DIO_HIP_CHECK(hipEventRecord(ctx->events[1], s));
        DIO_HIP_CHECK(hipGetLastError());
        const auto h1 = std::chrono::steady_clock::now();
        DIO_HIP_CHECK(hipStreamSynchronize(s));
        if (host_timing) {
            const auto h2 = std::chrono::steady_clock::now();
            std::fprintf(stderr, "dio_step host timing: %u steps, submit loop %.3f ms (graph launches %.3f ms), "
                         "wait %.3f ms\n", nsteps, std::chrono::duration<double, std::milli>(h1 - h0).count(),
                         h_launch, std::chrono::duration<double, std::milli>(h2 - h1).count());
        }
        if (coloured && colour_graph_last) read_colour_summary(ctx);
        // stats of the last step
        dio_step_stats& st = ctx->stats;
        std::memset(&st, 0, sizeof(st));
        st.steps = nsteps;
        st.dynamic_bodies = ctx->w_n_dyn;
        st.colours = coloured ? ctx->last_colours : 0;
        st.colour_rounds = coloured ? ctx->last_rounds : 0;
        if (ctx->w_n > 0 && nsteps > 0) {
            const int par = (int)((ctx->w_step - 1) % 2);
            const int lag = so.contact_lag ? 1 : 0;
            const int bn_pend = lag ? 1 - par : par;
            uint32_t w[4] = {0, 0, 0, 0};
            uint32_t ov[2] = {0, 0};
            DIO_HIP_CHECK(hipMemcpy(&w[0], ctx->pend[bn_pend].p_dev, 4, hipMemcpyDeviceToHost));
            DIO_HIP_CHECK(hipMemcpy(&w[1], ctx->cs[par].m_dev, 4, hipMemcpyDeviceToHost));
            DIO_HIP_CHECK(hipMemcpy(&w[2], ctx->csr.off + ctx->w_n, 4, hipMemcpyDeviceToHost));
            DIO_HIP_CHECK(hipMemcpy(&w[3], ctx->cs[par].m_narrow, 4, hipMemcpyDeviceToHost));
            DIO_HIP_CHECK(hipMemcpy(ov, ctx->w_overflow, sizeof(ov), hipMemcpyDeviceToHost));
            uint32_t terr = 0;
            DIO_HIP_CHECK(hipMemcpy(&terr, ctx->wb.err, 4, hipMemcpyDeviceToHost));
            st.pairs = w[0];
            st.contacts = w[1];
            st.csr_nnz = w[2];
            st.narrow_contacts = w[3];
            float total = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&total, ctx->events[0], ctx->events[1]));
            st.ms_total = total;
            float avbd = 0.0f;
            for (uint32_t k = 0; k < nsteps; ++k) {
                float t = 0.0f;
                DIO_HIP_CHECK(hipEventElapsedTime(&t, ctx->events[2 + 2 * k], ctx->events[3 + 2 * k]));
                avbd += t;
            }
            st.ms_avbd = avbd;
            unsigned long long acc[4] = {0, 0, 0, 0};
            DIO_HIP_CHECK(hipMemcpy(acc, ctx->w_acc, sizeof(acc), hipMemcpyDeviceToHost));
            st.sum_pairs = acc[0];
            st.sum_contacts = acc[1];
            st.sum_csr_nnz = acc[2];
            float* ms[4] = {&st.ms_init, &st.ms_primal, &st.ms_dual, &st.ms_recover};
            uint32_t* cnt[4] = {&st.n_init, &st.n_primal, &st.n_dual, &st.n_recover};
            for (const auto& mk : kt.marks) {
                float t = 0.0f;
                DIO_HIP_CHECK(hipEventElapsedTime(&t, ctx->prof_events[mk.second], ctx->prof_events[mk.second + 1]));
                *ms[mk.first] += t;
                *cnt[mk.first] += 1;
            }
            if (terr) throw StatusError{DIO_HIP, "LBVH build / traversal error (malformed tree or stack overflow)", 0};
            if (ov[0])
                capacity("a step needed " + std::to_string(ov[0]) + " candidate pairs, more than max_pairs (" +
                             std::to_string(ctx->p_max) + "); its results are truncated",
                         ov[0]);
            if (ov[1])
                capacity("a step needed " + std::to_string(ov[1]) + " contacts, more than max_contacts (" +
                             std::to_string(ctx->m_cap) + "); its results are truncated",
                         ov[1]);
        }
        return DIO_OK;
    });
}

dio_status dio_set_tuning(dio_ctx* ctx, const dio_tuning* t) {
    return guarded(ctx, [&]() {
        if (!t) invalid("tuning is null");
        dio_tuning v = *t;
        if (v.primal_lanes == 0) v.primal_lanes = 32;
        if (v.resort_interval == 0) v.resort_interval = 16;
        if (v.adjoint_lanes == 0) v.adjoint_lanes = 8;
        if (v.primal_lanes != 1 && v.primal_lanes != 4 && v.primal_lanes != 8 && v.primal_lanes != 16 &&
            v.primal_lanes != 32 && v.primal_lanes != 64)
            invalid("primal_lanes must be 1, 4, 8, 16, 32 or 64");
        if (v.broad_lanes != 0 && v.broad_lanes != 1 && v.broad_lanes != 2 && v.broad_lanes != 4 &&
            v.broad_lanes != 8 && v.broad_lanes != 16 && v.broad_lanes != 32)
            invalid("broad_lanes must be 0 (by world size), 1, 2, 4, 8, 16 or 32");
        if (v.fuse_dual > 1) invalid("fuse_dual must be 0 or 1");
        if (v.adjoint_lanes != 1 && v.adjoint_lanes != 2 && v.adjoint_lanes != 4 && v.adjoint_lanes != 8 &&
            v.adjoint_lanes != 16)
            invalid("adjoint_lanes must be 1, 2, 4, 8 or 16");
        if (v.overlap_broad > 1) invalid("overlap_broad must be 0 or 1");
        if (v.primal_order > 1) invalid("primal_order must be 0 (index) or 1 (Morton)");
        if (v.scene_solve > 65536) invalid("scene_solve must be 0 (off) or the bodies per bundle, 1 .. 65536");
        if (v.colour_lanes != 0 && v.colour_lanes != 1 && v.colour_lanes != 4 && v.colour_lanes != 8 &&
            v.colour_lanes != 16 && v.colour_lanes != 32 && v.colour_lanes != 64)
            invalid("colour_lanes must be 0 (primal_lanes), 1, 4, 8, 16, 32 or 64");
        if (v.primal_packed > 2) invalid("primal_packed must be 0, 1 or 2");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->tuning = v;
        ctx->drop_graphs();
        build_bundles(ctx);
        return DIO_OK;
    });
}

dio_status dio_get_tuning(dio_ctx* ctx, dio_tuning* out) {
    return guarded(ctx, [&]() {
        if (!out) invalid("out is null");
        *out = ctx->tuning;
        return DIO_OK;
    });
}

dio_status dio_autotune(dio_ctx* ctx, const dio_avbd_options* options, uint32_t reps, dio_tuning* chosen,
                        float* us_per_launch) {
    return guarded(ctx, [&]() {
        if (!options) invalid("options is null");
        if (!ctx->w_ready || ctx->w_step == 0) throw StatusError{DIO_NOT_READY, "run at least one dio_step first", 0};
        if (reps == 0) reps = 20;
        hipStream_t s = ctx->stream;
        DIO_HIP_CHECK(hipStreamSynchronize(s));
        const int par = (int)((ctx->w_step - 1) % 2);
        const SolveView v = world_solve_view(ctx, par);
        const AvbdParams p = params_of(options->dt, options->gravity, options->alpha, options->beta_lin,
                                       options->penalty_max, ctx->w_n);
        const int shapes[DIO_TUNE_PRIMAL_SHAPES] = {1, 4, 8, 16, 32, 64};
        float best = 3.4e38f;
        int best_lanes = 8;
        hipEvent_t e0, e1;
        DIO_HIP_CHECK(hipEventCreate(&e0));
        DIO_HIP_CHECK(hipEventCreate(&e1));
        for (int k = 0; k < DIO_TUNE_PRIMAL_SHAPES; ++k) {
            // iteration 1 writes only the Jacobi pose buffer 0 (and, fused, the
            // scratch multipliers), which the next solve overwrites before reading
            launch_avbd_primal(v, p, 1, false, s, shapes[k]);  // warm the code path
            DIO_HIP_CHECK(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps; ++r) launch_avbd_primal(v, p, 1, false, s, shapes[k]);
            DIO_HIP_CHECK(hipEventRecord(e1, s));
            DIO_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const float us = 1000.0f * ms / (float)reps;
            if (us_per_launch) us_per_launch[k] = us;
            if (us < best) {
                best = us;
                best_lanes = shapes[k];
            }
        }
        // broad phase: traversal lanes, timed on the world's poses in the
        // granular broad buffers (one re-sort first, then refits as between
        // re-sorts of the fused step)
        ensure_granular_broad(ctx);
        const BodyView bv{ctx->w_pos, ctx->w_rot, ctx->w_tag, ctx->w_corners, ctx->w_hull};
        // a batched world's traversal keeps to each scene's own bodies, as the step's does
        SceneTable scenes;
        if (ctx->w_scene_bits) {
            scenes.of = ctx->w_scene_of;
            scenes.off = ctx->w_scene_off;
            scenes.bits = ctx->w_scene_bits;
        }
        const int blanes[DIO_TUNE_BROAD_SHAPES] = {1, 2, 4, 8, 16, 32};
        float bbest = 3.4e38f;
        int best_blanes = 8;
        run_broad(bv, ctx->w_n, ctx->gb, ctx->p_max, s, true, 4, scenes);
        for (int k = 0; k < DIO_TUNE_BROAD_SHAPES; ++k) {
            run_broad(bv, ctx->w_n, ctx->gb, ctx->p_max, s, false, blanes[k], scenes);
            DIO_HIP_CHECK(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps; ++r) run_broad(bv, ctx->w_n, ctx->gb, ctx->p_max, s, false, blanes[k], scenes);
            DIO_HIP_CHECK(hipEventRecord(e1, s));
            DIO_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const float us = 1000.0f * ms / (float)reps;
            if (us_per_launch) us_per_launch[DIO_TUNE_PRIMAL_SHAPES + k] = us;
            if (us < bbest) {
                bbest = us;
                best_blanes = blanes[k];
            }
        }
        ctx->g_broad_submitted = false;  // the granular broad buffers were used as scratch
        // dual schedule, at the chosen primal shape: a fused primal against an
        // unfused primal plus a standalone dual (into the scratch multipliers)
        float sched_us[2] = {0.0f, 0.0f};
        for (int f = 0; f < 2; ++f) {
            auto once = [&]() {
                launch_avbd_primal(v, p, 1, f == 1, s, best_lanes);
                if (f == 0) launch_avbd_dual(v, p, 1u, 0u, 1u, ctx->m_cap, s);
            };
            once();
            DIO_HIP_CHECK(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps; ++r) once();
            DIO_HIP_CHECK(hipEventRecord(e1, s));
            DIO_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            sched_us[f] = 1000.0f * ms / (float)reps;
        }
        if (us_per_launch) {
            us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES] = sched_us[0];
            us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 1] = sched_us[1];
        }
        ctx->tuning.fuse_dual = sched_us[1] < sched_us[0] ? 1u : 0u;
        // body order, at the chosen shape and schedule
        float order_us[2] = {0.0f, 0.0f};
        for (int ord = 0; ord < 2; ++ord) {
            SolveView vo = v;
            vo.body_list = ord == 1 && ctx->w_n > 0 ? ctx->wb.perm : nullptr;
            vo.list_count = ord == 1 ? ctx->w_n : 0;
            const bool fz = ctx->tuning.fuse_dual == 1;
            launch_avbd_primal(vo, p, 1, fz, s, best_lanes);
            DIO_HIP_CHECK(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps; ++r) launch_avbd_primal(vo, p, 1, fz, s, best_lanes);
            DIO_HIP_CHECK(hipEventRecord(e1, s));
            DIO_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            order_us[ord] = 1000.0f * ms / (float)reps;
        }
        if (us_per_launch) {
            us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 2] = order_us[0];
            us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 3] = order_us[1];
        }
        ctx->tuning.primal_order = order_us[1] < order_us[0] ? 1u : 0u;
        // the coloured primal: a colouring of the current contact graph (in the
        // coloured mode's buffers), then one iteration's launches, colour by
        // colour, at each shape; poses are updated in place in the Jacobi
        // buffer the primal timings above write
        if (ctx->w_n > 0) {
            const ColourGraph g = colour_graph_of(ctx);
            uint32_t* h = ctx->wc_host;
            launch_colour_init(g, ctx->wc, s);
            for (uint32_t r = 0;; r += 16) {
                if (r + 16 > kColourRoundSlots - 1) capacity("the colouring needs more rounds than it can track", r);
                launch_colour_rounds(g, ctx->wc, r, 16, s);
                DIO_HIP_CHECK(hipMemcpyAsync(h, ctx->wc.pending + r, 16 * 4, hipMemcpyDeviceToHost, s));
                DIO_HIP_CHECK(hipStreamSynchronize(s));
                if (h[15] == 0u) break;
            }
            const int lb = launch_colour_lists(g, ctx->wc, s);
            DIO_HIP_CHECK(hipMemcpyAsync(h, ctx->wc.counts, (kMaxColours + 1) * 4, hipMemcpyDeviceToHost, s));
            DIO_HIP_CHECK(hipStreamSynchronize(s));
            std::vector<uint32_t> start{0u};
            uint32_t ncol = 0;
            for (uint32_t k = 0; k < kMaxColours; ++k)
                if (h[k]) ncol = k + 1;
            for (uint32_t k = 0; k < ncol; ++k) start.push_back(start.back() + h[k]);
            SolveView vc0 = v;
            vc0.pos[1] = vc0.pos[0];
            vc0.rot[1] = vc0.rot[0];
            float cbest = 3.4e38f;
            int cbest_lanes = 0;
            for (int k = 0; k < DIO_TUNE_COLOUR_SHAPES; ++k) {
                auto once = [&]() {
                    for (uint32_t col = 0; col < ncol; ++col) {
                        const uint32_t count = start[col + 1] - start[col];
                        if (count == 0) continue;
                        SolveView vc = vc0;
                        vc.body_list = ctx->wc.vals[lb] + start[col];
                        vc.list_count = count;
                        launch_avbd_primal_colour(vc, p, 1, count, s, shapes[k]);
                    }
                };
                once();
                DIO_HIP_CHECK(hipEventRecord(e0, s));
                for (uint32_t r = 0; r < reps; ++r) once();
                DIO_HIP_CHECK(hipEventRecord(e1, s));
                DIO_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                const float us = 1000.0f * ms / (float)reps;
                if (us_per_launch)
                    us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 4 + k] = us;
                if (us < cbest) {
                    cbest = us;
                    cbest_lanes = shapes[k];
                }
            }
            ctx->tuning.colour_lanes = (uint32_t)cbest_lanes;
        }
        // the packed primal at the chosen dual schedule and body order:
        // iteration 0 fills its work and item tables (and writes pose buffer
        // 1 of the same set), then iteration 1 is timed as the shapes above;
        // then the span primal; the fastest of these and the chosen grouped
        // shape is kept
        {
            SolveView vp = v;
            vp.body_list = ctx->tuning.primal_order == 1 && ctx->w_n > 0 ? ctx->wb.perm : nullptr;
            vp.list_count = ctx->tuning.primal_order == 1 ? ctx->w_n : 0;
            vp.work_tab = ctx->w_work;
            vp.item_tab = ctx->w_items;
            const bool fz = ctx->tuning.fuse_dual == 1;
            launch_avbd_primal(vp, p, 0, false, s, 32);
            launch_avbd_primal(vp, p, 1, fz, s, 32);
            DIO_HIP_CHECK(hipEventRecord(e0, s));
            for (uint32_t r = 0; r < reps; ++r) launch_avbd_primal(vp, p, 1, fz, s, 32);
            DIO_HIP_CHECK(hipEventRecord(e1, s));
            DIO_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            DIO_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            const float packed_us = 1000.0f * ms / (float)reps;
            if (us_per_launch)
                us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 4 + DIO_TUNE_COLOUR_SHAPES] = packed_us;
            // per-iteration cost: the primal, plus a dual launch when unfused
            // (the difference of the unfused schedule and the unfused primal
            // at the chosen shape)
            const float dual_us = sched_us[0] > best ? sched_us[0] - best : 0.0f;
            float best_it = order_us[ctx->tuning.primal_order] + (fz ? 0.0f : dual_us);
            ctx->tuning.primal_packed = 0;
            if (packed_us + (fz ? 0.0f : dual_us) < best_it) {
                best_it = packed_us + (fz ? 0.0f : dual_us);
                ctx->tuning.primal_packed = 1;
            }
            // the span primal (worlds without joints) in both schedules; it
            // takes the schedule it is fastest in if it wins
            float span_it = 0.0f;
            if (ctx->w_j.c == 0) {
                SolveView vs = v;
                vs.body_list = nullptr;
                vs.list_count = 0;
                vs.span_primal = true;
                span_it = 3.4e38f;
                uint32_t span_fused = 0;
                for (uint32_t f = 0; f < 2; ++f) {
                    launch_avbd_primal(vs, p, 0, false, s, 1);
                    launch_avbd_primal(vs, p, 1, f == 1, s, 1);
                    DIO_HIP_CHECK(hipEventRecord(e0, s));
                    for (uint32_t r = 0; r < reps; ++r) launch_avbd_primal(vs, p, 1, f == 1, s, 1);
                    DIO_HIP_CHECK(hipEventRecord(e1, s));
                    DIO_HIP_CHECK(hipEventSynchronize(e1));
                    float sms = 0.0f;
                    DIO_HIP_CHECK(hipEventElapsedTime(&sms, e0, e1));
                    const float it_us = 1000.0f * sms / (float)reps + (f == 1 ? 0.0f : dual_us);
                    if (it_us < span_it) {
                        span_it = it_us;
                        span_fused = f;
                    }
                }
                if (span_it < best_it) {
                    ctx->tuning.primal_packed = 2;
                    ctx->tuning.fuse_dual = span_fused;
                }
            }
            if (us_per_launch)
                us_per_launch[DIO_TUNE_PRIMAL_SHAPES + DIO_TUNE_BROAD_SHAPES + 4 + DIO_TUNE_COLOUR_SHAPES + 1] = span_it;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->tuning.primal_lanes = (uint32_t)best_lanes;
        ctx->tuning.broad_lanes = (uint32_t)best_blanes;
        ctx->drop_graphs();
        ctx->init_set = -1;  // the timed launches wrote the solver buffers
        if (chosen) *chosen = ctx->tuning;
        return DIO_OK;
    });
}

dio_status dio_step_get_stats(dio_ctx* ctx, dio_step_stats* out) {
    return guarded(ctx, [&]() {
        if (!out) invalid("out is null");
        *out = ctx->stats;
        return DIO_OK;
    });
}

dio_status dio_world_download(dio_ctx* ctx, float* positions, float* orientations, float* velocities,
                              float* angular_velocities, float* prev_velocities) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        const uint32_t n = ctx->w_n;
        if (n == 0) return DIO_OK;
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        std::vector<float4> buf(n);
        auto get3 = [&](const float4* src, float* dst) {
            if (!dst) return;
            DIO_HIP_CHECK(hipMemcpy(buf.data(), src, (size_t)n * 16, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) {
                dst[3 * i] = buf[i].x; dst[3 * i + 1] = buf[i].y; dst[3 * i + 2] = buf[i].z;
            }
        };
        get3(ctx->w_pos, positions);
        if (orientations) DIO_HIP_CHECK(hipMemcpy(orientations, ctx->w_rot, (size_t)n * 16, hipMemcpyDeviceToHost));
        get3(ctx->w_vel, velocities);
        get3(ctx->w_ang, angular_velocities);
        get3(ctx->w_prev, prev_velocities);
        return DIO_OK;
    });
}

dio_status dio_world_contacts(dio_ctx* ctx, dio_avbd_contact* out, uint32_t cap, uint32_t* m_out) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (!m_out) invalid("m is null");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->w_step == 0) {
            *m_out = 0;
            return DIO_OK;
        }
        const ContactSet& cs = ctx->cs[(ctx->w_step - 1) % 2];
        uint32_t m = 0;
        DIO_HIP_CHECK(hipMemcpy(&m, cs.m_dev, 4, hipMemcpyDeviceToHost));
        *m_out = m;
        if (m > cap) capacity("output array too small", m);
        if (m == 0) return DIO_OK;
        if (!out) invalid("out is null");
        std::vector<uint2> ab(m);
        std::vector<float4> f[8];
        const float4* src[8] = {cs.c.n, cs.c.t1, cs.c.t2, cs.c.ra, cs.c.rb, cs.c.c0, cs.c.pen, cs.c.lam};
        DIO_HIP_CHECK(hipMemcpy(ab.data(), cs.c.ab, (size_t)m * 8, hipMemcpyDeviceToHost));
        for (int k = 0; k < 8; ++k) {
            f[k].resize(m);
            DIO_HIP_CHECK(hipMemcpy(f[k].data(), src[k], (size_t)m * 16, hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < m; ++i) {
            dio_avbd_contact& r = out[i];
            std::memset(&r, 0, sizeof(r));
            r.a = ab[i].x;
            r.b = ab[i].y;
            r.friction = f[0][i].w;
            float* dst[8] = {r.normal, r.tangent1, r.tangent2, r.r_a, r.r_b, r.c0, r.penalty, r.lambda};
            for (int k = 0; k < 8; ++k) {
                dst[k][0] = f[k][i].x;
                dst[k][1] = f[k][i].y;
                dst[k][2] = f[k][i].z;
                dst[k][3] = 0.0f;
            }
        }
        return DIO_OK;
    });
}

dio_status dio_world_colours(dio_ctx* ctx, uint32_t* colour, uint32_t cap, uint32_t* num_colours) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (num_colours) *num_colours = ctx->last_colours;
        if (!colour) return DIO_OK;
        if (cap < ctx->w_n) capacity("colour buffer too small", ctx->w_n);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->w_n > 0)
            DIO_HIP_CHECK(hipMemcpy(colour, ctx->wc.colour, (size_t)ctx->w_n * 4, hipMemcpyDeviceToHost));
        return DIO_OK;
    });
}

dio_status dio_world_joints(dio_ctx* ctx, float* lambda, float* penalty, uint32_t cap, uint32_t* c_out) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        const uint32_t c = ctx->w_j.c;
        if (c_out) *c_out = c;
        if (!lambda && !penalty) return DIO_OK;
        if (cap < c) capacity("joint buffer too small", c);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        std::vector<float4> lam(c), pen(c);
        if (c > 0) {
            DIO_HIP_CHECK(hipMemcpy(lam.data(), ctx->w_j.st_lam, (size_t)c * sizeof(float4), hipMemcpyDeviceToHost));
            DIO_HIP_CHECK(hipMemcpy(pen.data(), ctx->w_j.st_pen, (size_t)c * sizeof(float4), hipMemcpyDeviceToHost));
        }
        // the state arrays hold values for hard joints only
        std::vector<uint4> abk(c);
        if (c > 0)
            DIO_HIP_CHECK(hipMemcpy(abk.data(), ctx->w_j.abk, (size_t)c * sizeof(uint4), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < c; ++k) {
            const bool h = is_hard(abk[k].z);
            const float4 l = h ? lam[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float4 q = h ? pen[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (lambda) { lambda[3 * k] = l.x; lambda[3 * k + 1] = l.y; lambda[3 * k + 2] = l.z; }
            if (penalty) { penalty[3 * k] = q.x; penalty[3 * k + 1] = q.y; penalty[3 * k + 2] = q.z; }
        }
        return DIO_OK;
    });
}

dio_status dio_world_pairs(dio_ctx* ctx, dio_gpu_pair* out, uint32_t cap, uint32_t* n_out) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (!n_out) invalid("n is null");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->w_step == 0) {
            *n_out = 0;
            return DIO_OK;
        }
        const int par = (int)((ctx->w_step - 1) % 2);
        const dio_ctx::Pending& p = ctx->pend[ctx->last_lag ? 1 - par : par];
        uint32_t np = 0;
        DIO_HIP_CHECK(hipMemcpy(&np, p.p_dev, 4, hipMemcpyDeviceToHost));
        *n_out = np;
        if (np > cap) capacity("output array too small", np);
        if (np > 0) {
            if (!out) invalid("out is null");
            DIO_HIP_CHECK(hipMemcpy(out, p.pairs, (size_t)np * 8, hipMemcpyDeviceToHost));
        }
        return DIO_OK;
    });
}


// ------------------------------------------------ instanced-renderer hand-off
dio_status dio_world_set_instances(dio_ctx* ctx, const float* scales, const uint32_t* order, uint32_t count) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        const uint32_t n = ctx->w_n;
        if (count > 0 && !scales) invalid("scales is null");
        if (!order && count > n) invalid("without an order, count must not exceed the world's bodies");
        if (order)
            for (uint32_t k = 0; k < count; ++k)
                if (order[k] >= n) invalid("order names a body past the world's bodies");
        for (uint32_t i = 0; count > 0 && i < 3 * n; ++i)
            if (!std::isfinite(scales[i])) invalid("a scale is not finite");
        hipStream_t s = ctx->stream;
        DIO_HIP_CHECK(hipStreamSynchronize(s));
        ctx->inst_count = 0;
        if (count == 0) return DIO_OK;
        std::vector<float> s4(4 * (size_t)n, 0.0f);
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) s4[4 * (size_t)i + a] = scales[3 * (size_t)i + a];
        DIO_HIP_CHECK(hipMemcpy(ctx->inst_scale, s4.data(), s4.size() * sizeof(float), hipMemcpyHostToDevice));
        if (order) {
            if (count > ctx->inst_order_cap) {
                if (ctx->inst_order) DIO_HIP_CHECK(hipFree(ctx->inst_order));
                ctx->inst_order = nullptr;
                ctx->inst_order_cap = 0;
                DIO_HIP_CHECK(hipMalloc(&ctx->inst_order, (size_t)count * 4));
                ctx->inst_order_cap = count;
            }
            DIO_HIP_CHECK(hipMemcpy(ctx->inst_order, order, (size_t)count * 4, hipMemcpyHostToDevice));
        }
        ctx->inst_ordered = order != nullptr;
        ctx->inst_count = count;
        return DIO_OK;
    });
}

dio_status dio_world_instances(dio_ctx* ctx, const float* view_proj, void* dst, uint64_t dst_bytes) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (!view_proj) invalid("view_proj is null");
        const uint32_t count = ctx->inst_count;
        if (count == 0) return DIO_OK;
        if (!dst) invalid("dst is null");
        if ((reinterpret_cast<uintptr_t>(dst) & 15u) != 0u) invalid("dst must be 16-byte aligned");
        const uint64_t need = 128ull * count;
        if (dst_bytes < need) capacity("dst holds fewer bytes than 128 per instance slot", need);
        Mat4Cols vp;
        for (int j = 0; j < 4; ++j)
            vp.c[j] = make_float4(view_proj[4 * j], view_proj[4 * j + 1], view_proj[4 * j + 2], view_proj[4 * j + 3]);
        launch_instances(ctx->w_pos, ctx->w_rot, ctx->inst_scale, ctx->inst_ordered ? ctx->inst_order : nullptr,
                         count, vp, static_cast<float4*>(dst), ctx->stream);
        DIO_HIP_CHECK(hipGetLastError());
        return DIO_OK;
    });
}

dio_status dio_import_external_memory(dio_ctx* ctx, int fd, uint64_t object_size, uint64_t offset, uint64_t size,
                                      void** dev_ptr, dio_external_memory** out) {
    return guarded(ctx, [&]() {
        if (!dev_ptr || !out) invalid("dev_ptr or out is null");
        if (fd < 0) invalid("fd is negative");
        if (size == 0 || object_size == 0) invalid("size is zero");
        if (offset > object_size || size > object_size - offset) invalid("offset + size exceeds the object");
        auto* m = new dio_external_memory();
        m->fd = fcntl(fd, F_DUPFD_CLOEXEC, 0);
        struct stat st;
        if (m->fd < 0 || fstat(m->fd, &st) != 0) {
            if (m->fd >= 0) (void)close(m->fd);
            delete m;
            invalid("fd is not an open file descriptor");
        }
        m->st_dev = st.st_dev;
        m->st_ino = st.st_ino;
        hipExternalMemoryHandleDesc hd;
        std::memset(&hd, 0, sizeof(hd));
        hd.type = hipExternalMemoryHandleTypeOpaqueFd;
        hd.handle.fd = m->fd;
        hd.size = object_size;
        hipError_t e = hipImportExternalMemory(&m->ext, &hd);
        if (e == hipSuccess) {
            hipExternalMemoryBufferDesc bd;
            std::memset(&bd, 0, sizeof(bd));
            bd.offset = offset;
            bd.size = size;
            e = hipExternalMemoryGetMappedBuffer(&m->ptr, m->ext, &bd);
        }
        if (e != hipSuccess) {
            release_external(m);
            throw HipError{e, "importing the external memory"};
        }
        ctx->ext_mems.push_back(m);
        *dev_ptr = m->ptr;
        *out = m;
        return DIO_OK;
    });
}

dio_status dio_release_external_memory(dio_ctx* ctx, dio_external_memory* mem) {
    return guarded(ctx, [&]() {
        auto it = std::find(ctx->ext_mems.begin(), ctx->ext_mems.end(), mem);
        if (!mem || it == ctx->ext_mems.end()) invalid("not an external memory of this context");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // writes into it may still be queued
        ctx->ext_mems.erase(it);
        release_external(mem);
        return DIO_OK;
    });
}

// ------------------------------------------------------------------ adjoint
}  // extern "C"

namespace {
AdjHostIndex adj_index(dio_ctx* ctx) {
    if (!ctx->adj_rec_base) throw StatusError{DIO_NOT_READY, "no recording (dio_adjoint_begin)", 0};
    AdjHostIndex h;
    uint32_t cnt[kRecCounters] = {};
    DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    DIO_HIP_CHECK(hipMemcpy(cnt, ctx->adj_rec.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    h.steps = cnt[kRecSteps];
    h.overflow = cnt[kRecOverflow];
    h.coff.resize(h.steps + 1);
    h.ioff.resize(h.steps + 1);
    h.m.resize(h.steps);
    if (h.steps > 0) {
        DIO_HIP_CHECK(hipMemcpy(h.coff.data(), ctx->adj_rec.step_coff, (h.steps + 1) * 4, hipMemcpyDeviceToHost));
        DIO_HIP_CHECK(hipMemcpy(h.ioff.data(), ctx->adj_rec.step_ioff, (h.steps + 1) * 4, hipMemcpyDeviceToHost));
        DIO_HIP_CHECK(hipMemcpy(h.m.data(), ctx->adj_rec.step_m, h.steps * 4, hipMemcpyDeviceToHost));
    } else {
        h.coff[0] = 0;
        h.ioff[0] = 0;
    }
    return h;
}

// the problem of a replay / backward, with the masses of `state` when given
AdjProblem adj_problem(dio_ctx* ctx, const dio_avbd_options* o, const dio_avbd_body_state* state) {
    AdjProblem pb;
    pb.n = ctx->w_n;
    pb.p = params_of(o->dt, o->gravity, o->alpha, o->beta_lin, o->penalty_max, ctx->w_n);
    pb.iters = o->iterations;
    pb.lanes = ctx->tuning.primal_lanes;
    pb.rev_lanes = ctx->tuning.adjoint_lanes;
    pb.mu = o->friction;
    pb.frozen = ctx->adj_mode == DIO_ADJOINT_FROZEN_MULTIPLIERS ? 1u : 0u;
    pb.inv_mm = ctx->w_inv_mm;
    if (state) {
        std::vector<float4> imm(ctx->w_n);
        for (uint32_t i = 0; i < ctx->w_n; ++i)
            imm[i] = make_float4(state[i].inv_moment[0], state[i].inv_moment[1], state[i].inv_moment[2],
                                 state[i].inv_mass);
        if (ctx->w_n > 0)
            DIO_HIP_CHECK(hipMemcpy(ctx->adj_work.inv_mm, imm.data(), (size_t)ctx->w_n * 16, hipMemcpyHostToDevice));
        pb.inv_mm = ctx->adj_work.inv_mm;
    }
    return pb;
}

void adj_ready(dio_ctx* ctx, const dio_avbd_options* o, const AdjHostIndex& h) {
    if (!ctx->adj_rec_base) throw StatusError{DIO_NOT_READY, "no recording (dio_adjoint_begin)", 0};
    if (!o) invalid("options is null");
    if (o->iterations == 0) invalid("iterations must be positive");
    if (h.overflow & 1u) capacity("the recording ran out of contact storage", h.coff.back());
    if (h.overflow & 2u) throw StatusError{DIO_CAPACITY, "more steps than the recording holds", h.steps};
    if (h.steps == 0) throw StatusError{DIO_NOT_READY, "no recorded steps (dio_step after dio_adjoint_begin)", 0};
    uint32_t m_max = 1;
    for (uint32_t v : h.m) m_max = std::max(m_max, v);
    adj_work_reserve(ctx->adj_work, ctx->n_cap, ctx->adj_rec.every, o->iterations, m_max,
                     ctx->adj_rec.max_steps / ctx->adj_rec.every + 1);
}

// The key of a sweep: the recorded step count, the options and state0.
std::vector<uint8_t> adj_key(const AdjHostIndex& h, const dio_avbd_options* o, const dio_avbd_body_state* state0,
                             uint32_t n) {
    std::vector<uint8_t> k(sizeof(uint32_t) + sizeof(dio_avbd_options));
    std::memcpy(k.data(), &h.steps, sizeof(uint32_t));
    std::memcpy(k.data() + sizeof(uint32_t), o, sizeof(dio_avbd_options));
    if (state0) {
        const uint8_t* b = reinterpret_cast<const uint8_t*>(state0);
        k.insert(k.end(), b, b + (size_t)n * sizeof(dio_avbd_body_state));
    }
    return k;
}

// Forward sweep over every recorded step from the first checkpoint, with
// state0's velocities (and, through pb, masses) when given, saving each
// segment's starting state in w.ck.  The working state ends as the final state.
void adj_sweep(dio_ctx* ctx, const AdjHostIndex& h, const AdjProblem& pb, const dio_avbd_body_state* state0) {
    hipStream_t s = ctx->stream;
    AdjWork& w = ctx->adj_work;
    const uint32_t n = ctx->w_n;
    adj_restore(ctx->adj_rec.ck, w, 0, n, s);
    if (state0 && n > 0) {
        // the state's velocities replace the checkpoint's
        std::vector<float4> v(3 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            v[i] = make_float4(state0[i].vel[0], state0[i].vel[1], state0[i].vel[2], 0.0f);
            v[n + i] = make_float4(state0[i].ang[0], state0[i].ang[1], state0[i].ang[2], 0.0f);
            v[2 * (size_t)n + i] = make_float4(state0[i].prev_vel[0], state0[i].prev_vel[1], state0[i].prev_vel[2], 0.0f);
        }
        DIO_HIP_CHECK(hipMemcpyAsync(w.st.vel, v.data(), (size_t)n * 16, hipMemcpyHostToDevice, s));
        DIO_HIP_CHECK(hipMemcpyAsync(w.st.ang, v.data() + n, (size_t)n * 16, hipMemcpyHostToDevice, s));
        DIO_HIP_CHECK(hipMemcpyAsync(w.st.prev, v.data() + 2 * (size_t)n, (size_t)n * 16, hipMemcpyHostToDevice, s));
        DIO_HIP_CHECK(hipStreamSynchronize(s));  // v is released on return
    }
    adj_set_friction(ctx->adj_rec, h, pb.mu, s);
    const uint32_t E = ctx->adj_rec.every;
    for (uint32_t first = 0, k = 0; first < h.steps; first += E, ++k) {
        adj_save(w, k, n, s);
        adj_replay(ctx->adj_rec, h, w, pb, first, std::min(h.steps, first + E), s);
    }
    DIO_HIP_CHECK(hipGetLastError());
}
}  // namespace

extern "C" {

dio_status dio_adjoint_begin(dio_ctx* ctx, uint32_t max_steps, uint32_t checkpoint_every,
                             uint64_t max_recorded_contacts) {
    return guarded(ctx, [&]() {
        if (!ctx->w_ready) throw StatusError{DIO_NOT_READY, "no world uploaded", 0};
        if (max_steps == 0 || checkpoint_every == 0) invalid("max_steps and checkpoint_every must be positive");
        if (ctx->w_j.c > 0) invalid("the adjoint does not differentiate joints: the world has constraints");
        if (max_recorded_contacts == 0) max_recorded_contacts = (uint64_t)max_steps * ctx->m_cap;
        if (max_recorded_contacts > 0x7FFFFFFFull) invalid("max_recorded_contacts above 2^31 - 1");
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->free_adjoint();
        const uint64_t rc = max_recorded_contacts, ic = 2 * max_recorded_contacts;
        const uint32_t n = ctx->n_cap, nck = max_steps / checkpoint_every + 1;
        size_t bytes = 0;
        auto take = [&](size_t b) {
            const size_t at = bytes;
            bytes += (b + 255) & ~(size_t)255;
            return at;
        };
        const size_t o_ab = take(rc * 8), o_c = take(rc * 16 * 8);
        const size_t o_coff = take(((size_t)max_steps + 1) * 4), o_ioff = take(((size_t)max_steps + 1) * 4);
        const size_t o_m = take((size_t)max_steps * 4), o_off = take((size_t)max_steps * (n + 1) * 4);
        const size_t o_idx = take(ic * 4), o_peer = take(ic * 4), o_cnt = take(kRecCounters * 4);
        const size_t o_ck = take((size_t)nck * 5 * n * 16);
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess)
            throw StatusError{DIO_CAPACITY, "the recording does not fit in device memory", bytes};
        ctx->adj_rec_base = p;
        char* b = static_cast<char*>(p);
        AdjRecord& r = ctx->adj_rec;
        r.max_steps = max_steps;
        r.every = checkpoint_every;
        r.n_cap = n;
        r.rec_cap = (uint32_t)rc;
        r.idx_cap = (uint32_t)std::min<uint64_t>(ic, 0xFFFFFFFFull);
        r.c.ab = reinterpret_cast<uint2*>(b + o_ab);
        float4* cf = reinterpret_cast<float4*>(b + o_c);
        r.c.n = cf; r.c.t1 = cf + rc; r.c.t2 = cf + 2 * rc; r.c.ra = cf + 3 * rc;
        r.c.rb = cf + 4 * rc; r.c.c0 = cf + 5 * rc; r.c.pen = cf + 6 * rc; r.c.lam = cf + 7 * rc;
        r.step_coff = reinterpret_cast<uint32_t*>(b + o_coff);
        r.step_ioff = reinterpret_cast<uint32_t*>(b + o_ioff);
        r.step_m = reinterpret_cast<uint32_t*>(b + o_m);
        r.csr_off = reinterpret_cast<uint32_t*>(b + o_off);
        r.csr_idx = reinterpret_cast<uint32_t*>(b + o_idx);
        r.csr_peer = reinterpret_cast<uint32_t*>(b + o_peer);
        r.counters = reinterpret_cast<uint32_t*>(b + o_cnt);
        r.ck = reinterpret_cast<float4*>(b + o_ck);
        launch_fill_u32(r.counters, 0u, kRecCounters, ctx->stream);
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->adj_on = true;
        ctx->drop_graphs();
        return DIO_OK;
    });
}

dio_status dio_adjoint_end(dio_ctx* ctx) {
    return guarded(ctx, [&]() {
        DIO_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->free_adjoint();
        ctx->drop_graphs();
        return DIO_OK;
    });
}

dio_status dio_adjoint_set_mode(dio_ctx* ctx, uint32_t mode) {
    return guarded(ctx, [&]() {
        if (mode != DIO_ADJOINT_EXACT && mode != DIO_ADJOINT_FROZEN_MULTIPLIERS)
            invalid("adjoint mode must be DIO_ADJOINT_EXACT (0) or DIO_ADJOINT_FROZEN_MULTIPLIERS (1)");
        ctx->adj_mode = mode;
        return DIO_OK;
    });
}

dio_status dio_adjoint_steps(dio_ctx* ctx, uint32_t* steps) {
    return guarded(ctx, [&]() {
        if (!steps) invalid("steps is null");
        if (!ctx->adj_rec_base) throw StatusError{DIO_NOT_READY, "no recording (dio_adjoint_begin)", 0};
        *steps = adj_index(ctx).steps;
        return DIO_OK;
    });
}

dio_status dio_adjoint_replay(dio_ctx* ctx, const dio_avbd_options* options, const dio_avbd_body_state* state0,
                              float* positions, float* orientations) {
    return guarded(ctx, [&]() {
        const AdjHostIndex h = adj_index(ctx);
        adj_ready(ctx, options, h);
        hipStream_t s = ctx->stream;
        AdjWork& w = ctx->adj_work;
        const AdjProblem pb = adj_problem(ctx, options, state0);
        const uint32_t n = ctx->w_n;
        ctx->adj_sweep_key.clear();
        adj_sweep(ctx, h, pb, state0);
        DIO_HIP_CHECK(hipStreamSynchronize(s));
        ctx->adj_sweep_key = adj_key(h, options, state0, n);
        std::vector<float4> buf(n);
        if (positions && n > 0) {
            DIO_HIP_CHECK(hipMemcpy(buf.data(), w.st.pos, (size_t)n * 16, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) {
                positions[3 * i] = buf[i].x; positions[3 * i + 1] = buf[i].y; positions[3 * i + 2] = buf[i].z;
            }
        }
        if (orientations && n > 0)
            DIO_HIP_CHECK(hipMemcpy(orientations, w.st.rot, (size_t)n * 16, hipMemcpyDeviceToHost));
        return DIO_OK;
    });
}

dio_status dio_adjoint_backward(dio_ctx* ctx, const dio_avbd_options* options,
                                const dio_avbd_body_state* state0, const float* d_positions,
                                const float* d_orientations, const float* d_velocities, const float* d_angular,
                                float* g_positions, float* g_orientations, float* g_velocities, float* g_angular,
                                float* g_prev_velocities, float* g_inv_mass, float* g_inv_moment,
                                double* g_friction) {
    return guarded(ctx, [&]() {
        const AdjHostIndex h = adj_index(ctx);
        adj_ready(ctx, options, h);
        hipStream_t s = ctx->stream;
        AdjWork& w = ctx->adj_work;
        const AdjProblem pb = adj_problem(ctx, options, state0);
        const uint32_t n = ctx->w_n;
        // segment checkpoints: the recording's (the recorded run), or those
        // of a sweep from state0 (kept from the last replay when it matches)
        const float4* ck = ctx->adj_rec.ck;
        if (state0) {
            const std::vector<uint8_t> key = adj_key(h, options, state0, n);
            if (key != ctx->adj_sweep_key) {
                ctx->adj_sweep_key.clear();
                adj_sweep(ctx, h, pb, state0);
                ctx->adj_sweep_key = key;
            }
            ck = w.ck;
        }
        std::vector<float4> sb(5 * (size_t)n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        for (uint32_t i = 0; i < n; ++i) {
            if (d_positions) sb[i] = make_float4(d_positions[3 * i], d_positions[3 * i + 1], d_positions[3 * i + 2], 0.0f);
            if (d_orientations)
                sb[n + i] = make_float4(d_orientations[4 * i], d_orientations[4 * i + 1], d_orientations[4 * i + 2],
                                        d_orientations[4 * i + 3]);
            if (d_velocities)
                sb[2 * (size_t)n + i] = make_float4(d_velocities[3 * i], d_velocities[3 * i + 1], d_velocities[3 * i + 2], 0.0f);
            if (d_angular)
                sb[3 * (size_t)n + i] = make_float4(d_angular[3 * i], d_angular[3 * i + 1], d_angular[3 * i + 2], 0.0f);
        }
        for (int f = 0; f < 5; ++f)
            if (n > 0)
                DIO_HIP_CHECK(hipMemcpy(w.sbar + (size_t)f * w.n_cap, sb.data() + (size_t)f * n, (size_t)n * 16,
                                        hipMemcpyHostToDevice));
        adj_set_friction(ctx->adj_rec, h, pb.mu, s);
        adj_backward(ctx->adj_rec, h, w, pb, ck, s);
        DIO_HIP_CHECK(hipGetLastError());
        DIO_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<float4> out(5 * (size_t)n), mb(n);
        for (int f = 0; f < 5; ++f)
            if (n > 0)
                DIO_HIP_CHECK(hipMemcpy(out.data() + (size_t)f * n, w.sbar + (size_t)f * w.n_cap, (size_t)n * 16,
                                        hipMemcpyDeviceToHost));
        if (n > 0) DIO_HIP_CHECK(hipMemcpy(mb.data(), w.mbar, (size_t)n * 16, hipMemcpyDeviceToHost));
        double mu_total = 0.0;
        DIO_HIP_CHECK(hipMemcpy(&mu_total, w.mubar_total, 8, hipMemcpyDeviceToHost));
        auto put3 = [&](float* dst, int f) {
            if (!dst) return;
            for (uint32_t i = 0; i < n; ++i) {
                const float4 v = out[(size_t)f * n + i];
                dst[3 * i] = v.x; dst[3 * i + 1] = v.y; dst[3 * i + 2] = v.z;
            }
        };
        put3(g_positions, 0);
        if (g_orientations)
            for (uint32_t i = 0; i < n; ++i) {
                const float4 v = out[n + i];
                g_orientations[4 * i] = v.x; g_orientations[4 * i + 1] = v.y;
                g_orientations[4 * i + 2] = v.z; g_orientations[4 * i + 3] = v.w;
            }
        put3(g_velocities, 2);
        put3(g_angular, 3);
        put3(g_prev_velocities, 4);
        for (uint32_t i = 0; i < n; ++i) {
            if (g_inv_mass) g_inv_mass[i] = mb[i].w;
            if (g_inv_moment) { g_inv_moment[3 * i] = mb[i].x; g_inv_moment[3 * i + 1] = mb[i].y; g_inv_moment[3 * i + 2] = mb[i].z; }
        }
        if (g_friction) *g_friction = mu_total;
        return DIO_OK;
    });
}
}  // extern "C"
