// adjoint.hpp — recording, frozen-contact replay and the hand-written adjoint
// of the AVBD solve.
//
// The differentiated function is the replay: each recorded step's contact set
// (geometry and warm-start multipliers, as the solve received them) and CSR
// are frozen, and only the AVBD solve runs (init, I x (primal, dual),
// recover), with every contact's friction set to the parameter mu.  At the
// recorded parameters the replay reproduces the recorded run bit for bit.
// Reverse mode runs segment by segment: the state checkpoint at the segment's
// first step is restored, the segment is replayed with its per-iteration
// poses and multipliers taped, and the steps are reversed from the tapes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "contacts.hpp"
#include "dio_device.hpp"

namespace dio {

// Everything a forward run records (device memory).
struct AdjRecord {
    uint32_t max_steps = 0, every = 0, n_cap = 0;
    uint32_t rec_cap = 0, idx_cap = 0;  // contacts / CSR entries over all steps
    ContactArrays c{};                  // rec_cap records, steps packed
    uint32_t* step_coff = nullptr;      // max_steps + 1: first contact of each step
    uint32_t* step_ioff = nullptr;      // max_steps + 1: first CSR entry of each step
    uint32_t* step_m = nullptr;         // max_steps: contacts of each step
    uint32_t* csr_off = nullptr;        // max_steps x (n_cap + 1)
    uint32_t* csr_idx = nullptr;        // idx_cap
    uint32_t* csr_peer = nullptr;       // idx_cap
    uint32_t* counters = nullptr;       // kRecContacts, kRecIdx, kRecSteps, kRecOverflow, kRecValid
    float4* ck = nullptr;               // (max_steps / every + 1) checkpoints x 5 fields x n_cap
};
constexpr int kRecContacts = 0, kRecIdx = 1, kRecSteps = 2, kRecOverflow = 3, kRecValid = 4, kRecCounters = 8;

// The body state a step starts from (and recover writes).
struct BodyState {
    float4 *pos, *rot, *vel, *ang, *prev;
};

// Record one step: the contact set `cs` and CSR (`off`, `idx`, `peer`)
// about to be solved, and a state checkpoint when the step index is a
// multiple of `every`.  The step index lives on the device, so the launches
// can be captured in a graph.
void launch_record_step(const AdjRecord& r, const ContactSet& cs, const uint32_t* off, const uint32_t* idx,
                        const uint32_t* peer, const BodyState& st, uint32_t n, uint32_t m_cap, hipStream_t s);

// Host-side view of a recording (downloaded once per replay / backward).
struct AdjHostIndex {
    uint32_t steps = 0;
    uint32_t overflow = 0;
    std::vector<uint32_t> coff, ioff, m;
};

// Scratch for replay and backward, sized for n_cap bodies, `every` steps of
// tape, I iterations and m_max contacts per step.
struct AdjWork {
    uint32_t n_cap = 0, every = 0, iters = 0, m_max = 0, n_ck = 0;
    BodyState st{};         // working state of the replay
    float4* ck = nullptr;   // n_ck segment checkpoints x 5 fields x n_cap (of the last sweep)
    float4* inv_mm = nullptr;
    // tapes: per segment step j
    float4* tape_state = nullptr;  // every x 5 x n_cap
    float4* tape_iner = nullptr;   // every x 2 x n_cap
    float4* tape_pose = nullptr;   // every x (I + 1) x 2 x n_cap
    float4* tape_mult = nullptr;   // every x I x 2 x m_max (lam, pen)
    // adjoint state
    float4* sbar = nullptr;        // 5 x n_cap
    float4* ybar = nullptr;        // 2 x 2 x n_cap (pose adjoints, ping-pong)
    float4* acc = nullptr;         // 4 x n_cap: x0bar, q0bar, xtbar, qtbar
    float4* mbar = nullptr;        // n_cap: inv_moment adjoint .xyz, inv_mass adjoint .w
    float* mubar_body = nullptr;     // n_cap (float, summed over a body's rows per launch)
    double* mubar_contact = nullptr; // m_max: the dual updates' friction adjoint, summed over all steps
    double* mubar_total = nullptr;   // 1
    float4* slots = nullptr;       // 2 m_max sides x 3: c adjoint, penalty adjoint, lambda adjoint
    float4* dpart = nullptr;       // 2 x m_max: lambda / penalty adjoint from the dual
    float4* dpose = nullptr;       // 2 m_max sides x 4: pos, rot, x0, q0 adjoint from the dual update
    float4* opose = nullptr;       // 2 m_max sides x 4: the same from the other body's primal rows
    void* base = nullptr;          // one allocation
    size_t bytes = 0;
};

// Size (or grow) the scratch.  Throws on allocation failure.
void adj_work_reserve(AdjWork& w, uint32_t n_cap, uint32_t every, uint32_t iters, uint32_t m_max, uint32_t n_ck);
void adj_work_free(AdjWork& w);

// Problem description shared by replay and backward.
struct AdjProblem {
    uint32_t n = 0;
    AvbdParams p{};
    uint32_t iters = 0;
    int lanes = 1;            // primal launch shape of the tape replay (1, 4, 8, 16, 32, 64)
    int rev_lanes = 8;        // lanes per body of the reverse primal (1, 2, 4, 8, 16)
    float mu = 0.0f;          // friction of every contact
    uint32_t frozen = 0;      // 1: the multipliers and penalties each dual update writes are constants
    const float4* inv_mm = nullptr;
};

// Replay steps [first, last) of the recording from the working state
// (w.st), writing the tapes of the segment (slot j = step - first).
void adj_replay(const AdjRecord& r, const AdjHostIndex& h, AdjWork& w, const AdjProblem& pb, uint32_t first,
                uint32_t last, hipStream_t s);

// Restore the working state from checkpoint `k` (step k * every) of `ck`
// (the recording's r.ck or the sweep's w.ck; 5 fields x n_cap each).
void adj_restore(const float4* ck, AdjWork& w, uint32_t k, uint32_t n, hipStream_t s);

// Save the working state as checkpoint `k` of w.ck.
void adj_save(AdjWork& w, uint32_t k, uint32_t n, hipStream_t s);

// Set the friction of every recorded contact to mu.
void adj_set_friction(const AdjRecord& r, const AdjHostIndex& h, float mu, hipStream_t s);

// Reverse-mode sweep over all recorded steps, restoring each segment from
// checkpoints `ck`.  On entry w.sbar holds the adjoint of the final state; on
// exit the adjoint of the initial state, with w.mbar, w.mubar_body and
// w.mubar_total the parameter adjoints.
void adj_backward(const AdjRecord& r, const AdjHostIndex& h, AdjWork& w, const AdjProblem& pb, const float4* ck,
                  hipStream_t s);

}  // namespace dio
