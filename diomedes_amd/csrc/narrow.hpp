// narrow.hpp — narrow phase buffers and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "broad.hpp"

namespace dio {

// Narrow-phase output for a pair list: 4 slots per pair (slot pair*4 + k is
// valid for k < cnt[pair]), the layout of the reference's contact buffer
// (narrowphase/src/lib.rs:63-69) as SoA.
struct NarrowOut {
    float4* n;        // normal.xyz, depth in .w
    float4* pa;       // point on a
    float4* pb;       // point on b
    uint32_t* cnt;    // contacts per pair
    uint32_t* off;    // exclusive scan of cnt (p_max + 1): contact index of each pair
    uint32_t* scan_partials;
    uint32_t* gjk_list;   // pairs routed to GJK + EPA
    uint32_t* gjk_count;
    uint32_t* epa_retry_list;   // pairs whose polytope outgrew the first EPA pass
    uint32_t* epa_retry_count;
    const uint2* pairs;   // the pair list these slots belong to
    const uint32_t* p_dev;
    uint32_t p_max;
};

// Narrow phase over pairs [0, *p_dev) and the scan of the per-pair counts;
// the contact total is left in out.off[*p_dev].
// simplex_store holds 8 float4 per pair (the GJK simplex handed to EPA).
// counters_cleared: gjk_count and epa_retry_count were zeroed by an earlier
// launch on the stream (the broad phase's emit); otherwise they are zeroed here.
// gjk_possible false: every pair is box-box or sphere-sphere (a world of boxes
// only or of spheres only), so no pair reaches GJK and the EPA launches are left
// out.
void run_narrow(const BodyView& bv, const NarrowOut& out, float4* simplex_store, hipStream_t s,
                bool counters_cleared = false, bool gjk_possible = true);

// Pack slots into the reference's GpuContactOut records (compute.rs:123-133).
void launch_pack_contact_out(const NarrowOut& out, uint32_t pair_count, dio_gpu_contact_out* rec, hipStream_t s);

}  // namespace dio
