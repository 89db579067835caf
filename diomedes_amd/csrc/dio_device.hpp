// dio_device.hpp — device data layout shared by the step's kernels.
//
// Bodies, contacts and joints live in HBM as structure-of-arrays of float4
// (one 16-byte load per field per lane, coalesced for sweeps and a single
// cache line per field for gathers).  Counts that the GPU produces (pairs,
// contacts) stay in device memory and kernels read them there, so a whole step
// is a fixed sequence of launches that can be captured in a HIP graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dio_math.hpp"

namespace dio {

// AvbdPush (gpu_physics.rs:61-71) less the per-launch parity
struct AvbdParams {
    float dt, alpha, beta_lin, penalty_max;
    float gx, gy, gz;
    uint32_t n;  // body count
    // the warm start of hard joints' multipliers (AvbdOptions gamma and
    // penalty_min, physics/gpu.rs:277-283)
    float gamma, penalty_min;
};

// Where an init writes (the solve's own buffers, or the other buffer set of a
// world whose final launch prepares the next step's init).
struct InitTarget {
    float4* pos[2];
    float4* rot[2];
    float4* init_lin;
    float4* init_ang;
    float4* iner_lin;
    float4* iner_ang;
};

// Everything the AVBD kernels touch (gpu_physics.rs bindings 0-10).
struct SolveView {
    // current pose S_k (read by init) and the two Jacobi parity buffers
    const float4* cur_pos;
    const float4* cur_rot;
    float4* pos[2];
    float4* rot[2];
    float4* init_lin;
    float4* init_ang;
    float4* iner_lin;
    float4* iner_ang;
    // GpuAvbdBodyState: vel, ang, prev_vel; inv_moment.xyz with inv_mass in .w
    float4* vel;
    float4* ang;
    float4* prev_vel;
    const float4* inv_mm;
    // contacts (GpuAvbdContact as SoA; friction rides in normal.w)
    const uint2* c_ab;
    const float4* c_n;
    const float4* c_t1;
    const float4* c_t2;
    const float4* c_ra;
    const float4* c_rb;
    const float4* c_c0;
    // nonzero: the contacts were built by the device's contact build, so the
    // tangent frame is a function of the stored normal (the narrow normal's
    // frame, negated by the build, or for a wall contact the stored normal's,
    // physics/gpu.rs:250,379) and c0 is (ra.w, 0, 0) (gpu.rs:259,395); the
    // solve then reads neither c_t1 / c_t2 nor c_c0.  Zero for contacts a
    // caller passes in (the granular solve), whose frame and c0 are read.
    uint32_t derived_frame;
    float4* c_pen;
    float4* c_lam;
    float4* c_pen_alt;      // scratch multipliers (odd dual updates within a solve)
    float4* c_lam_alt;
    const uint32_t* m_dev;  // contact count in device memory, or null: m_host
    uint32_t m_host;
    // per-body contact CSR (physics/gpu.rs:286-310)
    const uint32_t* offsets;
    const uint32_t* indices;
    const uint32_t* peers;   // peer code per CSR entry (kPeerIsB | other, or kContainer)
    // per-body joint CSR in ascending constraint order; joint records SoA
    const uint32_t* j_off;
    const uint32_t* j_idx;
    const uint4* j_abk;      // a, b, kind, 0
    const float4* j_anc_a;   // anchor_a.xyz, rest / max length
    const float4* j_anc_b;   // anchor_b.xyz, stiffness
    // hard joints' multipliers (.xyz: one per row), indexed by constraint:
    // j_lam / j_pen[j % 2] after j dual updates of this step (as contacts'
    // lam / pen and their scratch pair), and the state carried between steps
    // (j_st_*, read by the init's warm start and written by the final launch)
    float4* j_lam[2];
    float4* j_pen[2];
    float4* j_st_lam;
    float4* j_st_pen;
    uint32_t j_count;  // constraints
    uint32_t j_hard;   // nonzero: some constraint is a hard or bending joint (kinds past the reference's four)
    // fused path: recover writes the final pose back into the world pose
    float4* out_pos;
    float4* out_rot;
    // coloured Gauss-Seidel: the primal runs over body_list[0 .. list_count)
    // (one colour) instead of every body; pos[0] / pos[1] then alias.  When
    // list_range is set (a captured step), the colour is body_list[list_range[0]
    // .. list_range[1]), bounds read on the device, and list_count is unused.
    const uint32_t* body_list;
    uint32_t list_count;
    const uint32_t* list_range;
    // Jacobi solve of a step: the grouped primal's iteration 0 stores work
    // item k's (body, CSR start, CSR count, joint-CSR start) here and
    // iterations >= 1 read it (null: every iteration reads the order and the
    // offsets).  Holds at least n entries.
    uint4* work_tab;
    // with work_tab: the packed primal's item table (kPackCap entries per
    // workgroup), written by iteration 0 (null: the grouped primal)
    uint2* item_tab;
    // the span primal instead (k_avbd_primal_span: one thread per body, a
    // workgroup's contiguous CSR range built workgroup-wide); worlds without
    // joints
    bool span_primal;
    // when next.init_lin is set, the final launch (dual + recover) also runs
    // the next step's init from the recovered state into these buffers
    InitTarget next;
    // when stat_acc is set, the final launch adds the step's pair count
    // (*stat_pairs), contact count and CSR entries to stat_acc[0..2]
    const uint32_t* stat_pairs;
    unsigned long long* stat_acc;
};

// Optional timing events of one launch: recorded by the dispatch itself at the
// kernel's start and end (null: a plain launch).
struct LaunchEvents {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

// Launchers (avbd.hip).  All run on `stream`.  A solve with I iterations is
// init; primal(k) for k = 0 .. I-1, each followed (k < I-1) by dual(k % 2,
// 0, 0) in the unfused schedule, while in the fused schedule primal(k >= 1)
// applies dual update k - 1 itself; then dual_recover(I) (the last dual
// update, then recover).  Multipliers after j updates of a fused solve live
// in buffer j % 2 (0: the contact set's lam / pen, 1: the scratch pair).
void launch_avbd_init(const SolveView& v, const AvbdParams& p, hipStream_t stream, const LaunchEvents& ev = {});
// Entries of the Jacobi primal's work table (SolveView::work_tab) and of the
// packed primal's item table (SolveView::item_tab) for n bodies.
size_t primal_work_entries(uint32_t n);
size_t primal_item_entries(uint32_t n);
// lanes: lanes per body (1 = one lane per body; 4, 8, 16, 32 = grouped kernel)
void launch_avbd_primal(const SolveView& v, const AvbdParams& p, uint32_t iter, bool fused, hipStream_t stream,
                        int lanes = 32, const LaunchEvents& ev = {});
// The primal over one colour's bodies (v.body_list, count entries): unfused,
// poses read and written in place (v.pos[0] == v.pos[1]).
void launch_avbd_primal_colour(const SolveView& v, const AvbdParams& p, uint32_t iter, uint32_t count,
                               hipStream_t s, int lanes = 32, const LaunchEvents& ev = LaunchEvents{});
// The same over the colour v.list_range names on the device; the grid covers
// every body (the most one colour can hold) and workgroups past the colour's
// end exit at once.
void launch_avbd_primal_colour_range(const SolveView& v, const AvbdParams& p, uint32_t iter, hipStream_t s,
                                     int lanes = 32);
// Colours first .. starts[ncol_at] - 1 of the sorted list body_list (ranges
// starts[c] .. starts[c + 1]), one after another in one workgroup with the
// one-lane primal: the colours a captured step has no launch for (none
// unless the colouring needs more colours than when it was captured).
void launch_avbd_colour_tail(const SolveView& v, const AvbdParams& p, uint32_t iter, const uint32_t* starts,
                             uint32_t first, uint32_t ncol_at, hipStream_t s);
void launch_avbd_dual(const SolveView& v, const AvbdParams& p, uint32_t parity, uint32_t src, uint32_t dst,
                      uint32_t m_grid, hipStream_t stream, const LaunchEvents& ev = {});
void launch_avbd_dual_recover(const SolveView& v, const AvbdParams& p, uint32_t iters, bool fused, uint32_t m_grid,
                              hipStream_t stream, const LaunchEvents& ev = {});
// recover alone, reading the final Jacobi buffer pos[1 - parity]
void launch_avbd_recover(const SolveView& v, const AvbdParams& p, uint32_t parity, hipStream_t stream);

// The scene queue of a batched world: scene order[t] (t < count, largest
// first) holds bodies body_off[s] .. body_off[s + 1]; next[0] is the next
// queue slot and next[1] the finished workgroups, both zero between launches.
struct SceneQueue {
    const uint32_t* order;
    const uint32_t* body_off;
    uint32_t count;
    uint32_t* next;
};
// Workgroups of the scene launch (each takes scenes until the queue is empty).
constexpr uint32_t kSceneWorkgroups = 2048;
// Every primal, dual and the final dual + recover of a batched world's solve
// in one launch, a workgroup per scene (the world has no joints; the step's
// init ran before).  Bit-identical to the separate launches.
void launch_avbd_scenes(const SolveView& v, const AvbdParams& p, uint32_t iters, bool fused, const SceneQueue& q,
                        hipStream_t stream, const LaunchEvents& ev = {});

}  // namespace dio
