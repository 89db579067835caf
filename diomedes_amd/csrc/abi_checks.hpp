// abi_checks.hpp — host-side checks of records that cross the C ABI, shared
// by the context entry points and the allocation-free narrow phase.
#pragma once

#include <stdint.h>

#include <cmath>

#include "../../include/diomedes_hip.h"

namespace dio {

// A hull body names its vertices as a range of the hull pool: first in
// corners[0].x and count in corners[0].y, both whole numbers, count >= 1.
// The kernels read hull[first .. first + count) without a check of their own.
// Returns the pool size the bodies need (the largest first + count), or -1 - i
// when body i's range is malformed.
inline int64_t hull_pool_need(const dio_gpu_body* bodies, uint32_t n) {
    int64_t need = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (bodies[i].shape.tag != 4u) continue;  // DIO_SHAPE_HULL
        const float first = bodies[i].shape.corners[0][0], count = bodies[i].shape.corners[0][1];
        const bool whole = std::isfinite(first) && std::isfinite(count) && first >= 0.0f && count >= 1.0f &&
                           first == std::floor(first) && count == std::floor(count) && first < 16777216.0f &&
                           count < 16777216.0f;
        if (!whole) return -1 - (int64_t)i;
        const int64_t end = (int64_t)first + (int64_t)count;
        if (end > need) need = end;
    }
    return need;
}

}  // namespace dio
