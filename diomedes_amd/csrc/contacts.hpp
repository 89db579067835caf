// contacts.hpp — contact build on the device (the host work of the
// reference's AvbdSolver::prepare_contacts, physics/gpu.rs:188-311).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "broad.hpp"
#include "narrow.hpp"

namespace dio {

// GpuAvbdContact as SoA (friction in n.w).
struct ContactArrays {
    uint2* ab;
    float4 *n, *t1, *t2, *ra, *rb, *c0, *pen, *lam;
};

// One step's contact list: [narrow contacts in (pair, slot) order] ++
// [container contacts by body, wall, point], as prepare_contacts orders them.
struct ContactSet {
    ContactArrays c;
    uint32_t* m_dev;     // total contact count
    uint32_t* m_narrow;  // [0] count of the narrow prefix (at most m_cap), [1] narrow contacts found
    uint32_t* cont_cnt;  // per body (n_cap)
    uint32_t* cont_off;  // n_cap + 1
    uint32_t m_cap;
};

// The narrow output a contact set was (or is) built from.
struct NarrowSource {
    const uint2* pairs;
    const uint32_t* off;  // contact offsets per pair, off[*p_dev] = total
    const uint32_t* p_dev;
    const float4* n;
    const float4* pa;
    const float4* pb;
};

struct ContainerParams {
    float cx, cy, cz;
    float qx, qy, qz, qw;
    float half;
    uint32_t enabled;
};

struct WorldView {
    const float4* pos;
    const float4* rot;
    const float4* vel;
    const float4* inv_mm;  // inv_moment.xyz, inv_mass .w
    const uint32_t* tag;
    const float4* corners;
    const float4* hull;
    uint32_t n;
    // batched world: each body's scene and each scene's container (null: the
    // single container passed to run_contact_walls)
    const uint32_t* scene_of = nullptr;
    const ContainerParams* scene_ct = nullptr;
};

// AvbdOptions fields used by the contact build
struct BuildOptions {
    float dt, alpha, margin, penalty_min, penalty_max, gamma, friction;
};

struct CsrBuffers {
    uint32_t* cnt;      // n_cap
    uint32_t* cur;      // n_cap
    uint32_t* off;      // n_cap + 1
    uint32_t* idx;      // 2 * m_cap, ascending contact index within each body
    uint32_t* peer;     // 2 * m_cap, peer code of each entry (kPeerIsB flag | other body, or kContainer)
    uint32_t* tmp;      // 2 * m_cap, entries in atomic order before ranking
    uint32_t* scan_partials;
    uint32_t idx_cap;
};

// Peer codes of a sorted CSR (the granular solve's host-built CSR).
void launch_csr_peers(uint32_t n, const uint32_t* off, const uint32_t* idx, const uint2* ab, uint32_t* peer,
                      hipStream_t s);

// Build `dst` from the narrow output `src` and the container, warm-started
// from `old` (built from `old_src`); `overflow` records a capacity excess.
// Two parts: the narrow-phase contacts (the only part that reads old_src, the
// previous step's pair list and offsets) with each body's container contact
// count, then the container contacts.  Both add each written contact to
// csr_cnt (n words, zero on entry) of its bodies.  refit (optional): the
// broad phase's first refit launch, carried by the first part's launch.
struct TreeLower;
void run_contact_pairs(const WorldView& w, const ContainerParams& ct, const BuildOptions& o, const NarrowSource& src,
                       const ContactSet& dst, const NarrowSource& old_src, const ContactSet& old, uint32_t* overflow,
                       hipStream_t s, uint32_t* csr_cnt, const TreeLower* refit = nullptr);
void run_contact_walls(const WorldView& w, const ContainerParams& ct, const BuildOptions& o, const ContactSet& dst,
                       const ContactSet& old, uint32_t* scan_partials, uint32_t* overflow, hipStream_t s,
                       uint32_t* csr_cnt);

// CSR of a contact set (physics/gpu.rs:286-310) from the per-body counts the
// contact build left in b.cnt; leaves b.cnt and b.cur zero for the next build.
void run_csr(const ContactSet& cs, uint32_t n, const CsrBuffers& b, uint32_t* overflow, hipStream_t s);

// Unpack GpuAvbdContact records into SoA (granular path).
void launch_unpack_contacts(const dio_avbd_contact* rec, uint32_t m, const ContactArrays& c, hipStream_t s);

}  // namespace dio
