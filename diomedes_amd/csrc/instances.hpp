// instances.hpp — launcher of the instance-transform kernel (instances.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dio {

// A 4 x 4 float matrix by columns (glam's Mat4 storage).
struct Mat4Cols {
    float4 c[4];
};

// For slot k < count: body b = order ? order[k] : k; writes
// (view_proj * model_b, model_b) as eight float4 columns at dst + 8 k, model_b
// = Mat4::from_scale_rotation_translation(scale[b].xyz, rot[b], pos[b].xyz).
void launch_instances(const float4* pos, const float4* rot, const float4* scale, const uint32_t* order,
                      uint32_t count, const Mat4Cols& view_proj, float4* dst, hipStream_t s);

}  // namespace dio
