// prim.hpp — device-wide primitives used by the step: fill, exclusive scan,
// stable LSD radix sort of (key, value) pairs.  All take a HIP stream and
// never synchronise with the host, so they can be captured in a HIP graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dio {

// Fill n words (graph-safe replacement for hipMemsetAsync).
void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s);
// Fill n words of two arrays in one launch.
void launch_fill2_u32(uint32_t* a, uint32_t* b, uint32_t value, uint32_t n, hipStream_t s);

// Exclusive scan of `in` (n elements; n read from *n_dev when n_dev is not
// null, with n_cap the largest value it can take) into out[0..n]; out[n] is
// the total.  `partials` needs scan_partials_size(n_cap) words.  Up to 32768
// elements one workgroup scans in one launch; larger arrays take three.
size_t scan_partials_size(uint32_t n_cap);
void launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n_cap, const uint32_t* n_dev,
                           uint32_t* partials, hipStream_t s);

// Stable LSD radix sort of n (key, value) pairs over the low `bits` key bits.
// Ping-pongs between (k0, v0) and (k1, v1); returns 0 if the result is in
// (k0, v0), 1 if in (k1, v1).  `scratch` needs radix_scratch_size(n) words.
size_t radix_scratch_size(uint32_t n);
int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n, int bits,
                     uint32_t* scratch, hipStream_t s);

}  // namespace dio
